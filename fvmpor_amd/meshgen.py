"""Mesh ingest and synthetic mesh generation for the CVFE Richards path.

Meshes are plain arrays:

* ``xyz``        (N, 3) float64 node coordinates (z = 0 in 2D),
* ``conn``       (T, d+1) int32 simplex connectivity (triangles counter-clockwise,
                 as the reference's 2D shape functions assume, shape.cpp:41-42),
* ``face_tags``  (T, d+1) int32 boundary tag of each element face, 0 = interior.
                 Face order follows mesh.cpp:409-500: triangle faces (0,1),(1,2),(2,0);
                 tetrahedron faces (0,1,2),(0,1,3),(0,2,3),(1,2,3),
* ``node_tags``  CSR (ptr, tags) of boundary tags per node (the ``nbc bc...`` list
                 of a ``.pmesh`` node line, mesh.cpp:226-240),
* ``zone``       (T,) int32 physical zone = file tag - 100 (mesh.cpp:288).

``read_pmesh`` reads the reference's single-domain ``.pmesh`` format
(mesh.cpp:50-56, 211-367).  ``box_mesh_2d`` / ``box_mesh_3d`` build the jittered
lattice Delaunay meshes of the 2D slice and the 3D box workloads.
"""
from __future__ import annotations

import dataclasses

import numpy as np

TRI_FACES = np.array([[0, 1], [1, 2], [2, 0]], dtype=np.int64)
TET_FACES = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], dtype=np.int64)


@dataclasses.dataclass
class Mesh:
    dim: int
    xyz: np.ndarray
    conn: np.ndarray
    face_tags: np.ndarray
    node_tag_ptr: np.ndarray
    node_tags: np.ndarray
    zone: np.ndarray

    @property
    def n_nodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.conn.shape[0])

    def element_volumes(self) -> np.ndarray:
        p = self.xyz[self.conn]
        if self.dim == 2:
            a = p[:, 1, :2] - p[:, 0, :2]
            b = p[:, 2, :2] - p[:, 0, :2]
            return 0.5 * (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        a = p[:, 1] - p[:, 0]
        b = p[:, 2] - p[:, 0]
        c = p[:, 3] - p[:, 0]
        return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0


def _node_tags_from_faces(n_nodes: int, conn: np.ndarray, face_tags: np.ndarray, dim: int):
    faces = TRI_FACES if dim == 2 else TET_FACES
    sets = [set() for _ in range(n_nodes)]
    t_idx, f_idx = np.nonzero(face_tags)
    for t, f in zip(t_idx.tolist(), f_idx.tolist()):
        tag = int(face_tags[t, f])
        for q in faces[f]:
            sets[int(conn[t, q])].add(tag)
    ptr = np.zeros(n_nodes + 1, dtype=np.int32)
    for i, s in enumerate(sets):
        ptr[i + 1] = ptr[i] + len(s)
    tags = np.zeros(int(ptr[-1]), dtype=np.int32)
    for i, s in enumerate(sets):
        tags[ptr[i]:ptr[i + 1]] = sorted(s)
    return ptr, tags


def boundary_faces(conn: np.ndarray, dim: int):
    """(element, local face) pairs of faces that belong to one element only."""
    faces = TRI_FACES if dim == 2 else TET_FACES
    T = conn.shape[0]
    nf = faces.shape[0]
    keys = np.sort(conn[:, faces], axis=2).reshape(T * nf, -1)
    order = np.lexsort(keys.T[::-1])
    sk = keys[order]
    same_next = np.zeros(len(order), dtype=bool)
    same_next[:-1] = np.all(sk[1:] == sk[:-1], axis=1)
    same_prev = np.zeros(len(order), dtype=bool)
    same_prev[1:] = same_next[:-1]
    single = order[~(same_next | same_prev)]
    return single // nf, single % nf


def read_pmesh(path: str) -> Mesh:
    """Read a single-domain ``.pmesh`` (mesh.cpp:211-367; writer split.c:475-508)."""
    with open(path) as fh:
        tok = fh.read().split()
    pos = 0

    def nxt():
        nonlocal pos
        v = tok[pos]
        pos += 1
        return v

    n_dom = int(nxt())
    nxt()  # dom_id
    for _ in range(n_dom + 1):
        nxt()  # vtxdist
    n_gbl, n_int, n_bnd, n_ext = (int(nxt()) for _ in range(4))
    n_el_int, n_el_bnd = int(nxt()), int(nxt())
    if n_ext:
        raise ValueError("multi-domain .pmesh files are not supported on one GPU")
    n_nodes = n_int + n_bnd
    del n_gbl
    xyz = np.zeros((n_nodes, 3))
    ptr = np.zeros(n_nodes + 1, dtype=np.int32)
    tags = []
    for i in range(n_nodes):
        xyz[i] = [float(nxt()), float(nxt()), float(nxt())]
        nbc = int(nxt())
        tags.extend(int(nxt()) for _ in range(nbc))
        ptr[i + 1] = len(tags)
    n_el = n_el_int + n_el_bnd
    conn, ftags, zone = [], [], []
    dim = 3
    for _ in range(n_el):
        etype = int(nxt())
        ptag = int(nxt()) - 100
        if etype == 2:
            nn, nf, dim = 3, 3, 2
        elif etype == 4:
            nn, nf = 4, 4
        else:
            raise ValueError(f"element type {etype} is not a simplex")
        conn.append([int(nxt()) for _ in range(nn)])
        ftags.append([int(nxt()) for _ in range(nf)])
        zone.append(ptag)
    return Mesh(dim, xyz, np.asarray(conn, dtype=np.int32), np.asarray(ftags, dtype=np.int32),
                ptr, np.asarray(tags, dtype=np.int32), np.asarray(zone, dtype=np.int32))


def _jittered_lattice(extent, counts, rng, jitter=0.3):
    axes = [np.linspace(0.0, L, n) for L, n in zip(extent, counts)]
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    h = np.array([L / (n - 1) for L, n in zip(extent, counts)])
    d = rng.uniform(-jitter, jitter, size=pts.shape) * h
    # boundary nodes move tangentially only: freeze the coordinate normal to each face
    for ax, L in enumerate(extent):
        on = np.isclose(pts[:, ax], 0.0) | np.isclose(pts[:, ax], L)
        d[on, ax] = 0.0
    return pts + d


def box_mesh_2d(nx: int = 101, ny: int = 41, lx: float = 10.0, ly: float = 5.0, seed: int = 1,
                strip=(4.0, 6.0), jitter: float = 0.3) -> Mesh:
    """Delaunay triangulation of a jittered nx x ny lattice on [0,lx] x [0,ly].

    Boundary tags: 2 on the top edge for x in ``strip`` (the infiltration strip),
    1 elsewhere.  Node order is x-major, the element order sorts by lowest node.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    p2 = _jittered_lattice((lx, ly), (nx, ny), rng, jitter)
    tri = Delaunay(p2).simplices.astype(np.int64)
    a = p2[tri[:, 1]] - p2[tri[:, 0]]
    b = p2[tri[:, 2]] - p2[tri[:, 0]]
    area2 = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    keep = np.abs(area2) > 1e-12 * (lx * ly / (nx * ny))
    tri = tri[keep]
    area2 = area2[keep]
    neg = area2 < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    tri = tri[np.argsort(tri.min(axis=1), kind="stable")]
    xyz = np.zeros((p2.shape[0], 3))
    xyz[:, :2] = p2
    ftags = np.zeros(tri.shape, dtype=np.int32)
    be, bf = boundary_faces(tri, 2)
    fn = tri[be][np.arange(len(be))[:, None], TRI_FACES[bf]]
    mid = 0.5 * (xyz[fn[:, 0]] + xyz[fn[:, 1]])
    top = np.isclose(xyz[fn[:, 0], 1], ly) & np.isclose(xyz[fn[:, 1], 1], ly)
    in_strip = top & (mid[:, 0] >= strip[0]) & (mid[:, 0] <= strip[1])
    ftags[be, bf] = np.where(in_strip, 2, 1)
    conn = tri.astype(np.int32)
    ptr, tags = _node_tags_from_faces(xyz.shape[0], conn, ftags, 2)
    return Mesh(2, xyz, conn, ftags, ptr, tags, np.zeros(conn.shape[0], dtype=np.int32))


def box_mesh_3d(nx: int = 160, ny: int = 160, nz: int = 40, lx: float = 100.0, ly: float = 100.0,
                lz: float = 20.0, seed: int = 2, jitter: float = 0.3) -> Mesh:
    """Delaunay tetrahedralisation of a jittered nx x ny x nz lattice on a box.

    Flat tetrahedra that Delaunay leaves on the planar box faces are dropped.
    Boundary tags: 2 on the top face (z = lz), 1 elsewhere.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = _jittered_lattice((lx, ly, lz), (nx, ny, nz), rng, jitter)
    # Exactly coplanar boundary points make Qhull crawl; triangulate a copy whose
    # boundary points are moved by 1e-7 cells off their plane, keep exact
    # coordinates, and drop the tetrahedra that become flat again.
    h = np.array([lx / (nx - 1), ly / (ny - 1), lz / (nz - 1)])
    tmp = pts.copy()
    for ax, L in enumerate((lx, ly, lz)):
        on = np.isclose(pts[:, ax], 0.0) | np.isclose(pts[:, ax], L)
        tmp[on, ax] += rng.uniform(-1e-7, 1e-7, on.sum()) * h[ax]
    tet = Delaunay(tmp).simplices.astype(np.int64)
    p = pts[tet]
    vol6 = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    cell = (lx / (nx - 1)) * (ly / (ny - 1)) * (lz / (nz - 1))
    keep = np.abs(vol6) > 1e-9 * cell
    tet = tet[keep]
    vol6 = vol6[keep]
    neg = vol6 < 0
    tet[neg] = tet[neg][:, [0, 2, 1, 3]]
    tet = tet[np.argsort(tet.min(axis=1), kind="stable")]
    ftags = np.zeros(tet.shape, dtype=np.int32)
    be, bf = boundary_faces(tet, 3)
    fn = tet[be][np.arange(len(be))[:, None], TET_FACES[bf]]
    top = np.all(np.isclose(pts[fn, 2], lz), axis=1)
    ftags[be, bf] = np.where(top, 2, 1)
    conn = tet.astype(np.int32)
    ptr, tags = _node_tags_from_faces(pts.shape[0], conn, ftags, 3)
    return Mesh(3, pts, conn, ftags, ptr, tags, np.zeros(conn.shape[0], dtype=np.int32))


def retag_boundary(mesh: Mesh, tagger) -> Mesh:
    """Mesh whose boundary faces carry ``tagger(centroids)`` (an int array, one
    tag > 0 per face, from the (n, 3) face centroids); node tags follow."""
    faces = TRI_FACES if mesh.dim == 2 else TET_FACES
    be, bf = np.nonzero(mesh.face_tags)
    fn = mesh.conn[be][np.arange(len(be))[:, None], faces[bf]]
    cen = mesh.xyz[fn].mean(axis=1)
    tags = np.asarray(tagger(cen), dtype=np.int32)
    if tags.shape != (len(be),) or np.any(tags <= 0):
        raise ValueError("tagger must give one positive tag per boundary face")
    ftags = np.zeros_like(mesh.face_tags)
    ftags[be, bf] = tags
    ptr, ntags = _node_tags_from_faces(mesh.n_nodes, mesh.conn, ftags, mesh.dim)
    return Mesh(mesh.dim, mesh.xyz, mesh.conn, ftags, ptr, ntags, mesh.zone)


def gaussian_field(shape, spacing, corr, sigma, seed):
    """Stationary Gaussian field with exponential covariance
    sigma^2 exp(-sqrt(sum((dx_k / corr_k)^2))) on a regular grid, by FFT of a
    periodic embedding twice the grid size (negative spectral weights clipped)."""
    rng = np.random.default_rng(seed)
    ext = [2 * s for s in shape]
    axes = []
    for n, h in zip(ext, spacing):
        k = np.arange(n)
        axes.append(np.minimum(k, n - k) * h)
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum((g / c) ** 2 for g, c in zip(grids, corr)))
    cov = sigma ** 2 * np.exp(-r)
    lam = np.fft.fftn(cov).real
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal(ext) + 1j * rng.standard_normal(ext)
    field = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(np.prod(ext))
    sl = tuple(slice(0, s) for s in shape)
    return field.real[sl]


def renumber_nodes(mesh: Mesh, perm: np.ndarray) -> Mesh:
    """Mesh with node i of the result = node perm[i] of the input."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    counts = np.diff(mesh.node_tag_ptr)[perm]
    ptr = np.zeros(len(perm) + 1, dtype=np.int32)
    ptr[1:] = np.cumsum(counts)
    tags = np.zeros(int(ptr[-1]), dtype=np.int32)
    for i, o in enumerate(perm.tolist()):
        if counts[i]:
            tags[ptr[i]:ptr[i + 1]] = mesh.node_tags[mesh.node_tag_ptr[o]:mesh.node_tag_ptr[o + 1]]
    conn = inv[mesh.conn].astype(np.int32)
    return Mesh(mesh.dim, mesh.xyz[perm].copy(), conn, mesh.face_tags.copy(), ptr, tags, mesh.zone.copy())


def tile_order(xyz: np.ndarray, dim: int, tile_nodes: int = 1024) -> np.ndarray:
    """Locality ordering: nodes grouped by cells of a coarse grid holding about
    ``tile_nodes`` nodes each (cells in lattice order, nodes keep their relative
    order inside a cell).  Consecutive blocks of ``tile_nodes`` node ids are then
    spatially compact, which is what the residual's node tiles and the
    block-Jacobi preconditioner blocks want."""
    lo = xyz[:, :dim].min(axis=0)
    ext = np.maximum(xyz[:, :dim].max(axis=0) - lo, 1e-300)
    h = (np.prod(ext) / len(xyz)) ** (1.0 / dim)
    side = h * tile_nodes ** (1.0 / dim)
    cells = np.maximum(1, np.rint(ext / side)).astype(np.int64)
    idx = np.minimum((xyz[:, :dim] - lo) / ext * cells, cells - 1).astype(np.int64)
    key = idx[:, 0]
    for d in range(1, dim):
        key = key * cells[d] + idx[:, d]
    return np.argsort(key, kind="stable")
