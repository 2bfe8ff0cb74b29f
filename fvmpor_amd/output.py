"""Solution output in the reference's formats.

* ``write_vtu`` / ``write_pvtu_header``: one time step as VTK XML unstructured
  grid, the layout of Solution::write_timestep_VTK_XML and
  write_timestep_header_VTK_XML (solution.h:178-353): ASCII arrays, point
  data "pressure_head", cell data "dom" (the subdomain; one GPU holds domain
  0), points, connectivity, offsets and VTK cell types (5 triangle, 10
  tetrahedron), numbers printed as C++ streams print them by default (%g, 6
  significant digits).  Files are named ``<name>_dom<d>_<step>.vtu`` and
  ``<name>_head_<step>.pvtu`` as there.
* ``write_run``: all stored steps as the reference's text file
  (Solution::write_to_file, solution.h:58-88): a header "steps blocksize
  nodes", then every value on its own line with 15 digits after the point.
* ``read_vtu``: the arrays of such a file (for checking and post-processing).
"""
from __future__ import annotations

import re

import numpy as np

_VTK_TYPES = {2: 5, 3: 10}  # triangle, tetrahedron by dimension


def _fmt(values) -> str:
    return " ".join(f"{float(v):g}" for v in np.asarray(values).ravel())


def write_vtu(name: str, step: int, mesh, heads, dom: int = 0, var_name: str = "pressure_head") -> str:
    """Write one step; ``mesh`` has ``xyz`` (N, 3), ``conn`` (T, d+1) and
    ``dim``; ``heads`` (N,).  Returns the file name."""
    xyz = np.asarray(mesh.xyz, dtype=np.float64)
    conn = np.asarray(mesh.conn, dtype=np.int64)
    heads = np.asarray(heads, dtype=np.float64)
    n, t = xyz.shape[0], conn.shape[0]
    if heads.shape != (n,):
        raise ValueError(f"heads must have {n} entries")
    npe = conn.shape[1]
    path = f"{name}_dom{dom}_{step}.vtu"
    out = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid">', "  <UnstructuredGrid>",
           f'    <Piece NumberOfPoints="{n}" NumberOfCells="{t}">',
           f'      <PointData Scalars="{var_name}">',
           f'        <DataArray Name="{var_name}" type="Float32" format="ascii">',
           "          " + "".join(f"{float(v):g} " for v in heads),
           "        </DataArray>", "      </PointData>",
           '      <CellData Scalars="dom">', '        <DataArray Name="dom" type="Int32" format="ascii">',
           f"{dom} " * t, "        </DataArray>", "      </CellData>",
           "      <Points>", '        <DataArray NumberOfComponents="3" type="Float32" format="ascii">',
           "          " + "".join(f"{p[0]:g} {p[1]:g} {p[2]:g} " for p in xyz),
           "        </DataArray>", "      </Points>", "      <Cells>",
           '        <DataArray Name="connectivity" type="Int32" format="ascii">',
           "          " + "".join("".join(f"{int(v)} " for v in row) + " " for row in conn),
           "        </DataArray>",
           '        <DataArray Name="offsets" type="Int32" format="ascii">',
           "          " + "".join(f"{npe * (e + 1)} " for e in range(t)),
           "        </DataArray>",
           '        <DataArray Name="types" type="UInt8" format="ascii">',
           "          " + f"{_VTK_TYPES[mesh.dim]} " * t,
           "        </DataArray>", "      </Cells>", "    </Piece>", "  </UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    return path


def write_pvtu_header(name: str, step: int, n_dom: int = 1, var_name: str = "pressure_head") -> str:
    """The parallel header naming the pieces ``<name>_dom<d>_<step>.vtu``."""
    path = f"{name}_head_{step}.pvtu"
    out = ['<?xml version="1.0"?>', '<VTKFile type="PUnstructuredGrid">', '  <PUnstructuredGrid GhostLevel="0">',
           f'    <PPointData Scalars="{var_name}">',
           f'      <PDataArray Name="{var_name}" type="Float32" format="ascii"/>', "    </PPointData>",
           '    <PCellData Scalars="dom">', '      <PDataArray Name="dom" type="Int32" format="ascii"/>',
           "    </PCellData>", "    <PPoints>",
           '      <PDataArray NumberOfComponents="3" type="Float32" format="ascii"/>', "    </PPoints>",
           "    <PCells>", '      <PDataArray Name="connectivity" type="Int32" format="ascii"/>',
           '      <PDataArray Name="offsets" type="Int32" format="ascii"/>',
           '      <PDataArray Name="types" type="UInt8" format="ascii"/>', "    </PCells>"]
    out += [f'    <Piece Source="{name}_dom{d}_{step}.vtu"/>' for d in range(n_dom)]
    out += ["  </PUnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    return path


def write_run(path: str, solutions, blocksize: int = 1) -> None:
    """Every stored solution (a list of (N,) arrays) in the .run text format."""
    sols = [np.asarray(s, dtype=np.float64).ravel() for s in solutions]
    with open(path, "w") as fh:
        fh.write(f"{len(sols)} {blocksize}")
        if sols:
            fh.write(f" {sols[0].size}\n")
            for s in sols:
                fh.write("".join(f"{v:.15e}\n" for v in s))


def read_vtu(path: str) -> dict:
    """Arrays of a VTK XML unstructured-grid file by name ("Points" for the
    coordinates), as float64."""
    with open(path) as fh:
        text = fh.read()
    out = {}
    for m in re.finditer(r"<DataArray([^>]*)>(.*?)</DataArray>", text, re.S):
        name = re.search(r'Name="([^"]+)"', m.group(1))
        out[name.group(1) if name else "Points"] = np.array(m.group(2).split(), dtype=np.float64)
    return out
