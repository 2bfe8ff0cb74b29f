"""Write the committed golden fixtures from data files the reference ships.

Inputs (reference root, read as data only):
  meshing/meshes/cassion_1_0.pmesh   cassion mesh (1599 nodes, 3064 triangles)
  meshing/Jpattern.mtx               Jacobian sparsity of that mesh (decomp.c:92)
  examples/EEM/CASS_dom0_{0..11}.vtu pressure head of the EEM cassion run at
                                     t = k * 2.5e6 / 11 s (Float32 ASCII)
Outputs (this folder): cassion_1_0.pmesh, Jpattern.mtx (copied verbatim) and
cassion_eem_vtu.npz (points, connectivity, pressure heads, times, and the
SHA-256 of each VTU file's text, which pins the VTU writer's format).

Usage: python tests/golden/make_golden.py /path/to/reference
"""
import hashlib
import os
import re
import shutil
import sys

import numpy as np


def _arrays(text):
    out = {}
    for m in re.finditer(r'<DataArray([^>]*)>(.*?)</DataArray>', text, re.S):
        attrs, body = m.group(1), m.group(2)
        name = re.search(r'Name="([^"]+)"', attrs)
        key = name.group(1) if name else "Points"
        out[key] = np.array(body.split(), dtype=np.float64)
    return out


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    shutil.copyfile(os.path.join(ref, "meshing/meshes/cassion_1_0.pmesh"), os.path.join(here, "cassion_1_0.pmesh"))
    shutil.copyfile(os.path.join(ref, "meshing/Jpattern.mtx"), os.path.join(here, "Jpattern.mtx"))
    heads = []
    digests = []
    pts = conn = None
    for k in range(12):
        with open(os.path.join(ref, f"examples/EEM/CASS_dom0_{k}.vtu")) as fh:
            text = fh.read()
        digests.append(hashlib.sha256(text.encode()).hexdigest())
        a = _arrays(text)
        heads.append(a["pressure_head"].astype(np.float32))
        if pts is None:
            pts = a["Points"].reshape(-1, 3).astype(np.float32)
            conn = a["connectivity"].astype(np.int32).reshape(-1, 3)
    times = np.arange(12) * (2.5e6 / 11)
    np.savez_compressed(os.path.join(here, "cassion_eem_vtu.npz"), points=pts, conn=conn,
                        heads=np.stack(heads), times=times, vtu_sha256=np.array(digests))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
