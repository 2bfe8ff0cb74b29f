"""Problem definitions: soil zones, boundary conditions and initial heads.

These mirror the reference's problem headers (``cassion.h:9-117``,
``vs.h:8-87``) and the workloads of the 2D slice and 3D box runs.

Zone parameter rows (``ZONE_FIELDS``): K_xx, K_yy, K_zz, phi, alpha_vG, n_vG,
S_r, S_s.  ``m_vG = 1 - 1/n_vG`` is derived where it is used (cassion.h:50).
Boundary rows (``BC_FIELDS``): value, b, eta, qx, qy, qz, indexed by the
boundary tag; types follow definitions.h:114-146 (1 Dirichlet, 3 prescribed
flux, 4 hydrostatic Dirichlet, 6 directional flux, 7 seepage, 8 shore).
Fluxes are outward-positive: infiltration is negative (cassion.h:115).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .meshgen import Mesh, box_mesh_2d, box_mesh_3d, gaussian_field, renumber_nodes, retag_boundary, tile_order

ZONE_FIELDS = ("Kxx", "Kyy", "Kzz", "phi", "alpha", "n", "Sr", "Ss")
BC_FIELDS = ("value", "b", "eta", "qx", "qy", "qz")

# van Genuchten classes (Carsel & Parrish 1988): alpha [1/m], n, theta_r, theta_s
SOIL_CLASSES = {
    "clay": (0.8, 1.09, 0.068, 0.38),
    "loam": (3.6, 1.56, 0.078, 0.43),
    "sandy_loam": (7.5, 1.89, 0.065, 0.41),
    "sand": (14.5, 2.68, 0.045, 0.43),
}


@dataclasses.dataclass
class Problem:
    name: str
    mesh: Mesh
    zones: np.ndarray      # (nz, 8)
    ezone: np.ndarray      # (T,) int32
    ekmul: np.ndarray      # (T,) float64
    bc_type: np.ndarray    # (ntags,) int32
    bc: np.ndarray         # (ntags, 6)
    rho0: float
    psi0: np.ndarray       # (N,)
    t_end: float
    time_unit: str


def _zone_row(K, phi, alpha, n, Sr, Ss=0.0):
    return [K, K, K, phi, alpha, n, Sr, Ss]


def cassion_problem(mesh: Mesh) -> Problem:
    """The reference's cassion problem (cassion.h): four zones, SI units."""
    rho0, g, mu = 1000.0, 9.80665, 1e-3
    f = rho0 * g / mu
    zones = np.array([
        _zone_row(f * 9.33e-12, 0.368, 3.34, 1.982, 0.2771),
        _zone_row(f * 5.55e-12, 0.3510, 3.63, 1.632, 0.2806),
        _zone_row(f * 4.898e-12, 0.3250, 3.45, 1.573, 0.2643),
        _zone_row(f * 4.898e-11, 0.3250, 3.45, 1.573, 0.2643),
    ])
    bc_type = np.array([0, 3, 3], dtype=np.int32)
    bc = np.zeros((3, 6))
    bc[2, 0] = -2.3148e-7
    return Problem("cassion", mesh, zones, mesh.zone.astype(np.int32),
                   np.ones(mesh.n_elements), bc_type, bc, rho0,
                   np.full(mesh.n_nodes, -7.34), 2.5e6, "s")


def slice2d_problem(nx=101, ny=41, seed=1) -> Problem:
    """2D vertical slice 10 m x 5 m, homogeneous loam, hydrostatic psi = 1 - y,
    0.05 m/day infiltration on the top strip x in [4, 6] m, to t = 1 day."""
    mesh = box_mesh_2d(nx, ny, 10.0, 5.0, seed=seed)
    alpha, n, tr, ts = SOIL_CLASSES["loam"]
    zones = np.array([_zone_row(0.25, ts, alpha, n, tr / ts, 1e-5)])
    bc_type = np.array([0, 3, 3], dtype=np.int32)
    bc = np.zeros((3, 6))
    bc[2, 0] = -0.05
    psi0 = 1.0 - mesh.xyz[:, 1]
    return Problem("slice2d", mesh, zones, np.zeros(mesh.n_elements, dtype=np.int32),
                   np.ones(mesh.n_elements), bc_type, bc, 1000.0, psi0, 1.0, "day")


def box3d_problem(nx=160, ny=160, nz=40, seed_mesh=2, seed_field=3, seed_layers=4,
                  sigma_lnK=1.0, mean_lnK=float(np.log(0.5)), rain=0.02, mesh=None, tiled=True,
                  order=None) -> Problem:
    """3D box 100 x 100 x 20 m, log-normal Ks (exponential covariance, 10 m / 10 m /
    2 m correlation lengths), four van Genuchten classes in random horizontal
    layers, water table at z = 5 m, rain on the whole top face, to t = 1 day.
    Node order of a generated mesh: ``order`` = "tile" (``tile_order``: compact
    ~1024-node tiles, the default), "rcm" (reverse Cuthill-McKee, as the
    reference renumbers at mesh load) or "lattice" (the generator's own);
    ``tiled=False`` is "lattice"."""
    lx, ly, lz = 100.0, 100.0, 20.0
    if order is None:
        order = "tile" if tiled else "lattice"
    if mesh is None:
        mesh = box_mesh_3d(nx, ny, nz, lx, ly, lz, seed=seed_mesh)
        if order == "tile":  # node ids grouped into compact tiles (locality of the residual / block precon)
            mesh = renumber_nodes(mesh, tile_order(mesh.xyz, 3))
        elif order == "rcm":
            from .lib import rcm_order
            mesh = renumber_nodes(mesh, rcm_order(mesh))
        elif order != "lattice":
            raise ValueError(f"unknown node order {order!r}")
    p = mesh.xyz[mesh.conn]
    cen = p.mean(axis=1)
    gshape = (nx, ny, nz)
    spacing = (lx / (nx - 1), ly / (ny - 1), lz / (nz - 1))
    lnK = gaussian_field(gshape, spacing, (10.0, 10.0, 2.0), sigma_lnK, seed_field) + mean_lnK
    idx = [np.clip(np.rint(cen[:, a] / spacing[a]).astype(np.int64), 0, gshape[a] - 1) for a in range(3)]
    ekmul = np.exp(lnK[idx[0], idx[1], idx[2]])
    rng = np.random.default_rng(seed_layers)
    names = list(SOIL_CLASSES)
    n_layers = 6
    cuts = np.sort(rng.uniform(0.0, lz, n_layers - 1))
    layer_class = rng.integers(0, len(names), n_layers)
    layer_class[:4] = rng.permutation(4)  # every class appears
    layer = np.searchsorted(cuts, cen[:, 2])
    ezone = layer_class[layer].astype(np.int32)
    zones = np.array([_zone_row(1.0, SOIL_CLASSES[k][3], SOIL_CLASSES[k][0], SOIL_CLASSES[k][1],
                                SOIL_CLASSES[k][2] / SOIL_CLASSES[k][3], 1e-5) for k in names])
    bc_type = np.array([0, 3, 3], dtype=np.int32)
    bc = np.zeros((3, 6))
    bc[2, 0] = -rain
    psi0 = 5.0 - mesh.xyz[:, 2]
    return Problem("box3d", mesh, zones, ezone, ekmul, bc_type, bc, 1000.0, psi0, 1.0, "day")


def _dirichlet_heads(mesh: Mesh, bc_type, bc, psi):
    """psi with every Dirichlet node set to its boundary value (the last
    type-1/4 tag of the node decides, as in fvmpor.h:404-451)."""
    psi = psi.copy()
    el = mesh.xyz[:, 1] if mesh.dim == 2 else mesh.xyz[:, 2]
    for i in range(mesh.n_nodes):
        for tag in mesh.node_tags[mesh.node_tag_ptr[i]:mesh.node_tag_ptr[i + 1]]:
            if bc_type[tag] == 1:
                psi[i] = bc[tag, 0]
            elif bc_type[tag] == 4:
                psi[i] = (1.0 + bc[tag, 2]) * (bc[tag, 1] - el[i])
    return psi


def slice2d_bc_problem(nx=41, ny=17, seed=1) -> Problem:
    """The 2D loam slice with every boundary type of definitions.h:114-146 on
    its own part of the boundary, consistent with the hydrostatic start
    psi0 = 1 - y where they fix a head: bottom Dirichlet psi = 1 (1), left
    side hydrostatic with the water table at 1 m (4), top left directional
    flux (6), top strip x in [4, 6] m prescribed flux (3), top right seepage
    (7), right side shore (8)."""
    lx, ly = 10.0, 5.0
    mesh = box_mesh_2d(nx, ny, lx, ly, seed=seed)

    def tagger(c):
        x, y = c[:, 0], c[:, 1]
        t = np.full(len(c), 1)
        top = np.isclose(y, ly)
        t[top & (x < 4.0)] = 2
        t[top & (x >= 4.0) & (x <= 6.0)] = 3
        t[top & (x > 6.0)] = 4
        t[np.isclose(x, 0.0)] = 5
        t[np.isclose(x, lx)] = 6
        return t

    mesh = retag_boundary(mesh, tagger)
    alpha, n, tr, ts = SOIL_CLASSES["loam"]
    zones = np.array([_zone_row(0.25, ts, alpha, n, tr / ts, 1e-5)])
    bc_type = np.array([0, 1, 6, 3, 7, 4, 8], dtype=np.int32)
    bc = np.zeros((7, 6))
    bc[1, 0] = 1.0                     # bottom: fixed head, psi0 there
    bc[2, 3:6] = (0.0, -0.01, 0.0)     # top left: directional flux, downwards (into the slice)
    bc[3, 0] = -0.05                   # top strip: infiltration
    bc[4, 0] = 0.005                   # top right: seepage outflow
    bc[5, 1:3] = (1.0, 0.0)            # left side: hydrostatic, water table at 1 m
    psi0 = _dirichlet_heads(mesh, bc_type, bc, 1.0 - mesh.xyz[:, 1])
    return Problem("slice2d_bc", mesh, zones, np.zeros(mesh.n_elements, dtype=np.int32),
                   np.ones(mesh.n_elements), bc_type, bc, 1000.0, psi0, 1.0, "day")


def box3d_bc_problem(nx=12, ny=10, nz=6, seed=2) -> Problem:
    """The heterogeneous 3D box with one boundary type per face: x = 0
    Dirichlet (1), x = 100 m hydrostatic (4), y = 0 directional flux (6),
    y = 100 m seepage (7), bottom shore (8), top rain (3)."""
    base = box3d_problem(nx=nx, ny=ny, nz=nz, seed_mesh=seed)
    lx, ly, lz = 100.0, 100.0, 20.0

    def tagger(c):
        t = np.full(len(c), 3)
        t[np.isclose(c[:, 2], 0.0)] = 6
        t[np.isclose(c[:, 1], ly)] = 5
        t[np.isclose(c[:, 1], 0.0)] = 4
        t[np.isclose(c[:, 0], lx)] = 2
        t[np.isclose(c[:, 0], 0.0)] = 1
        t[np.isclose(c[:, 2], lz)] = 7
        return t

    mesh = retag_boundary(base.mesh, tagger)
    bc_type = np.array([0, 1, 4, 3, 6, 7, 8, 3], dtype=np.int32)
    bc = np.zeros((8, 6))
    bc[1, 0] = -2.0                    # fixed head on x = 0
    bc[2, 1:3] = (6.0, 0.0)            # hydrostatic x = lx, water table at 6 m
    bc[3, 0] = 0.0                     # tag 3 is not used by the tagger
    bc[4, 3:6] = (0.0, 0.05, 0.01)     # directional flux through y = 0
    bc[5, 0] = 0.005                   # seepage outflow through y = ly
    bc[7, 0] = -0.02                   # rain on the top
    psi0 = _dirichlet_heads(mesh, bc_type, bc, 5.0 - mesh.xyz[:, 2])
    return Problem("box3d_bc", mesh, base.zones, base.ezone, base.ekmul, bc_type, bc, 1000.0, psi0, 1.0, "day")
