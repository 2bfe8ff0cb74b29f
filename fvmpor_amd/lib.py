"""ctypes binding of libfvp.so (the C ABI declared in include/fvp.h).

Device vectors are passed as raw HIP device pointers; callers holding torch
tensors pass ``tensor.data_ptr()``.  The library must be built (``make -C
fvmpor_amd/csrc``); importing this module never falls back to another
implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FVP_LIBRARY selects another build of the same library (e.g. a variant being timed)
LIB_PATH = os.environ.get("FVP_LIBRARY") or os.path.join(_HERE, "libfvp.so")
_LIB = None

ZONE_FIELDS = 8
BC_FIELDS = 6

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes) for every symbol in include/fvp.h
SIGNATURES = {
    "fvp_last_error": (ctypes.c_char_p, []),
    "fvp_rcm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p]),
    "fvp_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_sizes": (ctypes.c_int, [ctypes.c_void_p, _c_int_p]),
    "fvp_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_set_bc": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p]),
    "fvp_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p]),
    "fvp_begin_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_get_upwind": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_set_upwind": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p]),
    "fvp_capacity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_jv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_csr_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_csr_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_csr_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_csr_set_values": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_csr_get_values": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p]),
    "fvp_csr_gemv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_double,
                                    ctypes.c_void_p]),
    "fvp_precon_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_precon_block_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_set_block_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_set_overlap": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_set_jacobian_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_precon_overlap": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_set_residual_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fvp_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, _c_int_p,
                                    _c_int_p, _c_int_p, _c_dbl_p]),
    "fvp_integrator_set_segments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p]),
    "fvp_precon_apply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_precon_values": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p]),
    "fvp_gmres_csr": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "fvp_gmres_csr_deflated": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.c_void_p]),
    "fvp_gmres_csr_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "fvp_precon_setup": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int]),
    "fvp_precon_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_jacobian": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_colours": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p]),
    "fvp_newton_gmres": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_integrator_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_integrator_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_integrator_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "fvp_integrator_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]),
    "fvp_integrator_state": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_integrator_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fvp_profile_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, ctypes.POINTER(ctypes.c_long)]),
    "fvp_mpnn_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, ctypes.c_int,
                                       ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]),
    "fvp_mpnn_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_mpnn_num_params": (ctypes.c_long, [ctypes.c_void_p]),
    "fvp_mpnn_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_mpnn_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "fvp_mpnn_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "fvp_mpnn_get_grads": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "fvp_mpnn_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fvp_mpnn_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _c_dbl_p]),
    "fvp_mpnn_adam": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    "fvp_mpnn_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fvp_integrator_set_learned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_double, ctypes.c_double]),
    "fvp_integrator_set_recording": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
}

PRECON_NONE, PRECON_SCHUR, PRECON_JACOBI, PRECON_ILU0, PRECON_BJILU0, PRECON_BJILU0_F32 = 0, 1, 2, 5, 7, 8
JACOBIAN_ELEMENT, JACOBIAN_COLOURED = 0, 1
RESIDUAL_OWNER, RESIDUAL_WINDOW = 0, 1
BLOCK_ROWS = 512
ERR_PIVOT, ERR_NOCONV = 1, 2


class GmresParams(ctypes.Structure):
    _fields_ = [("dim_krylov", ctypes.c_int), ("max_restarts", ctypes.c_int), ("reorthogonalise", ctypes.c_int),
                ("tol", ctypes.c_double)]


class GmresStats(ctypes.Structure):
    _fields_ = [("restarts", ctypes.c_int), ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double), ("error", ctypes.c_double)]

    def as_dict(self):
        return dict(restarts=self.restarts, iterations=self.iterations, converged=bool(self.converged),
                    residual=self.residual, error=self.error)


class FvpError(RuntimeError):
    pass


class _Problem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("n_nodes", ctypes.c_int), ("n_elements", ctypes.c_int),
        ("xyz", _c_dbl_p), ("conn", _c_int_p), ("face_tags", _c_int_p),
        ("node_tag_ptr", _c_int_p), ("node_tags", _c_int_p), ("elem_zone", _c_int_p),
        ("elem_kmul", _c_dbl_p), ("n_zones", ctypes.c_int), ("zones", _c_dbl_p),
        ("n_tags", ctypes.c_int), ("bc_type", _c_int_p), ("bc", _c_dbl_p), ("rho0", ctypes.c_double),
    ]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FvpError(f"{LIB_PATH} is not built (run make -C fvmpor_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise FvpError(f"{what} failed: {lib().fvp_last_error().decode()}")
    return rc


def _ptr(x):
    """Device pointer of a torch tensor (or an int address)."""
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    return ctypes.c_void_p(x.data_ptr())


def rcm_order(mesh) -> np.ndarray:
    """Reverse Cuthill-McKee order of the mesh's nodes (fvp_rcm): node i of the
    renumbered mesh is node perm[i] of this one (meshgen.renumber_nodes)."""
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    perm = np.empty(mesh.n_nodes, dtype=np.int32)
    _check(lib().fvp_rcm(mesh.n_nodes, conn.shape[0], conn.shape[1], conn.ctypes.data_as(_c_int_p),
                         perm.ctypes.data_as(_c_int_p)), "fvp_rcm")
    return perm


class Context:
    """One device-resident CVFE problem (fvp_create / fvp_destroy)."""

    def __init__(self, problem, device: int = 0):
        m = problem.mesh
        self._keep = dict(
            xyz=np.ascontiguousarray(m.xyz, dtype=np.float64),
            conn=np.ascontiguousarray(m.conn, dtype=np.int32),
            ft=np.ascontiguousarray(m.face_tags, dtype=np.int32),
            ntp=np.ascontiguousarray(m.node_tag_ptr, dtype=np.int32),
            nt=np.ascontiguousarray(m.node_tags if len(m.node_tags) else np.zeros(1), dtype=np.int32),
            ez=np.ascontiguousarray(problem.ezone, dtype=np.int32),
            km=np.ascontiguousarray(problem.ekmul, dtype=np.float64),
            zones=np.ascontiguousarray(problem.zones, dtype=np.float64),
            bct=np.ascontiguousarray(problem.bc_type, dtype=np.int32),
            bc=np.ascontiguousarray(problem.bc, dtype=np.float64),
        )
        k = self._keep
        P = _Problem()
        P.dim = m.dim
        P.n_nodes = k["xyz"].shape[0]
        P.n_elements = k["conn"].shape[0]
        P.xyz = k["xyz"].ctypes.data_as(_c_dbl_p)
        P.conn = k["conn"].ctypes.data_as(_c_int_p)
        P.face_tags = k["ft"].ctypes.data_as(_c_int_p)
        P.node_tag_ptr = k["ntp"].ctypes.data_as(_c_int_p)
        P.node_tags = k["nt"].ctypes.data_as(_c_int_p)
        P.elem_zone = k["ez"].ctypes.data_as(_c_int_p)
        P.elem_kmul = k["km"].ctypes.data_as(_c_dbl_p)
        P.n_zones = k["zones"].shape[0]
        P.zones = k["zones"].ctypes.data_as(_c_dbl_p)
        P.n_tags = k["bct"].shape[0]
        P.bc_type = k["bct"].ctypes.data_as(_c_int_p)
        P.bc = k["bc"].ctypes.data_as(_c_dbl_p)
        P.rho0 = float(problem.rho0)
        h = ctypes.c_void_p()
        _check(lib().fvp_create(ctypes.byref(P), int(device), ctypes.byref(h)), "fvp_create")
        self.h = h
        s = (ctypes.c_int * 6)()
        _check(lib().fvp_sizes(self.h, s), "fvp_sizes")
        self.N, self.T, self.NE, self.ndir, self.dim, self.nz = list(s)

    def close(self):
        if getattr(self, "h", None):
            lib().fvp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream_ptr: int):
        _check(lib().fvp_set_stream(self.h, ctypes.c_void_p(stream_ptr)), "fvp_set_stream")

    def synchronize(self):
        _check(lib().fvp_synchronize(self.h), "fvp_synchronize")

    def residual(self, t, psi, psip, F):
        _check(lib().fvp_residual(self.h, float(t), _ptr(psi), _ptr(psip), _ptr(F)), "fvp_residual")

    def begin_step(self, t, psi, psip):
        _check(lib().fvp_begin_step(self.h, float(t), _ptr(psi), _ptr(psip)), "fvp_begin_step")

    def get_upwind(self):
        wf = np.zeros(self.NE)
        fl = np.zeros(self.NE)
        _check(lib().fvp_get_upwind(self.h, wf.ctypes.data_as(ctypes.c_void_p),
                                    fl.ctypes.data_as(ctypes.c_void_p)), "fvp_get_upwind")
        return wf, fl

    def set_upwind(self, wf):
        wf = np.ascontiguousarray(wf, dtype=np.float64)
        _check(lib().fvp_set_upwind(self.h, wf.ctypes.data_as(_c_dbl_p)), "fvp_set_upwind")

    def set_bc(self, bc):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        _check(lib().fvp_set_bc(self.h, bc.ctypes.data_as(_c_dbl_p)), "fvp_set_bc")

    def capacity(self, out):
        _check(lib().fvp_capacity(self.h, _ptr(out)), "fvp_capacity")

    def jv(self, t, psi, psip, F0, cj, v, Jv):
        _check(lib().fvp_jv(self.h, float(t), _ptr(psi), _ptr(psip), _ptr(F0), float(cj), _ptr(v), _ptr(Jv)),
               "fvp_jv")


class CsrMatrix:
    """Device CSR matrix with its preconditioner (the decomp Tmtx_CRS + Tprecon pair)."""

    def __init__(self, ia, ja, device: int = 0, handle=None):
        self.owned = handle is None
        if handle is None:
            self._ia = np.ascontiguousarray(ia, dtype=np.int32)
            self._ja = np.ascontiguousarray(ja, dtype=np.int32)
            h = ctypes.c_void_p()
            _check(lib().fvp_csr_create(int(device), len(self._ia) - 1, self._ia.ctypes.data_as(_c_int_p),
                                        self._ja.ctypes.data_as(_c_int_p), ctypes.byref(h)), "fvp_csr_create")
            self.h = h
            self.n = len(self._ia) - 1
            self.nnz = int(self._ia[-1])
        else:
            self.h = handle
            self.n = self.nnz = None

    def close(self):
        if self.owned and getattr(self, "h", None):
            lib().fvp_csr_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream_ptr: int):
        _check(lib().fvp_csr_set_stream(self.h, ctypes.c_void_p(stream_ptr)), "fvp_csr_set_stream")

    def set_values(self, vals):
        _check(lib().fvp_csr_set_values(self.h, _ptr(vals)), "fvp_csr_set_values")

    def get_values(self, count):
        out = np.zeros(count)
        _check(lib().fvp_csr_get_values(self.h, out.ctypes.data_as(_c_dbl_p)), "fvp_csr_get_values")
        return out

    def gemv(self, alpha, x, beta, y):
        _check(lib().fvp_csr_gemv(self.h, float(alpha), _ptr(x), float(beta), _ptr(y)), "fvp_csr_gemv")

    def precon_build(self, ptype) -> int:
        return _check(lib().fvp_precon_build(self.h, int(ptype)), "fvp_precon_build")

    def precon_overlap(self, layers):
        _check(lib().fvp_precon_overlap(self.h, int(layers)), "fvp_precon_overlap")

    def precon_block_rows(self, rows):
        _check(lib().fvp_precon_block_rows(self.h, int(rows)), "fvp_precon_block_rows")

    def precon_apply(self, x, y):
        _check(lib().fvp_precon_apply(self.h, _ptr(x), _ptr(y)), "fvp_precon_apply")

    def precon_values(self, count):
        out = np.zeros(count)
        _check(lib().fvp_precon_values(self.h, out.ctypes.data_as(_c_dbl_p)), "fvp_precon_values")
        return out

    def gmres(self, b, x, m=30, max_restarts=20, tol=1e-10, reorth=1):
        p = GmresParams(m, max_restarts, reorth, tol)
        st = GmresStats()
        rc = _check(lib().fvp_gmres_csr(self.h, _ptr(b), _ptr(x), ctypes.byref(p), ctypes.byref(st)),
                    "fvp_gmres_csr")
        return rc, st.as_dict()


def _csr_gmres_noalloc(self, b, x, m=30, max_restarts=20, tol=1e-10, reorth=1):
    """fvp_gmres_csr_noalloc: GMRES in the static per-GPU workspace."""
    p = GmresParams(m, max_restarts, reorth, tol)
    st = GmresStats()
    rc = _check(lib().fvp_gmres_csr_noalloc(self.h, _ptr(b), _ptr(x), ctypes.byref(p), ctypes.byref(st)),
                "fvp_gmres_csr_noalloc")
    return rc, st.as_dict()


CsrMatrix.gmres_noalloc = _csr_gmres_noalloc


def _csr_gmres_deflated(self, b, x, k=4, m=30, max_restarts=20, tol=1e-10, reorth=1):
    """fvp_gmres_csr_deflated: GMRES keeping k harmonic Ritz vectors at each restart."""
    p = GmresParams(m, max_restarts, reorth, tol)
    st = GmresStats()
    rc = _check(lib().fvp_gmres_csr_deflated(self.h, _ptr(b), _ptr(x), ctypes.byref(p), int(k), ctypes.byref(st)),
                "fvp_gmres_csr_deflated")
    return rc, st.as_dict()


CsrMatrix.gmres_deflated = _csr_gmres_deflated


def _ctx_precon_setup(self, t, cj, h, psi, psip, F0, ewt, ptype=PRECON_ILU0) -> int:
    return _check(lib().fvp_precon_setup(self.h, float(t), float(cj), float(h), _ptr(psi), _ptr(psip), _ptr(F0),
                                         _ptr(ewt), int(ptype)), "fvp_precon_setup")


def _ctx_precon_solve(self, r, z):
    _check(lib().fvp_precon_solve(self.h, _ptr(r), _ptr(z)), "fvp_precon_solve")


def _ctx_jacobian(self) -> CsrMatrix:
    h = ctypes.c_void_p()
    _check(lib().fvp_jacobian(self.h, ctypes.byref(h)), "fvp_jacobian")
    return CsrMatrix(None, None, handle=h)


def _ctx_set_overlap(self, layers):
    _check(lib().fvp_set_overlap(self.h, int(layers)), "fvp_set_overlap")


def _ctx_set_jacobian_mode(self, mode):
    """JACOBIAN_ELEMENT (default) or JACOBIAN_COLOURED: how precon_setup forms the FD quotients."""
    _check(lib().fvp_set_jacobian_mode(self.h, int(mode)), "fvp_set_jacobian_mode")


def _ctx_set_block_rows(self, rows):
    _check(lib().fvp_set_block_rows(self.h, int(rows)), "fvp_set_block_rows")


def _ctx_set_residual_launch(self, variant=0, tile_nodes=512, xcd_map=1):
    _check(lib().fvp_set_residual_launch(self.h, int(variant), int(tile_nodes), int(xcd_map)),
           "fvp_set_residual_launch")


def _ctx_autotune(self, t, psi, psip):
    v, tn, xm, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().fvp_autotune(self.h, float(t), _ptr(psi), _ptr(psip), ctypes.byref(v), ctypes.byref(tn),
                              ctypes.byref(xm), ctypes.byref(us)), "fvp_autotune")
    return {"variant": v.value, "tile_nodes": tn.value, "xcd_map": xm.value, "residual_us": us.value}


def _ctx_colours(self):
    nc = ctypes.c_int()
    col = np.zeros(self.N, dtype=np.int32)
    _check(lib().fvp_colours(self.h, ctypes.byref(nc), col.ctypes.data_as(_c_int_p)), "fvp_colours")
    return nc.value, col


def _ctx_newton_gmres(self, t, psi, psip, F0, cj, b, x, m=30, max_restarts=10, tol=1e-8, reorth=1):
    p = GmresParams(m, max_restarts, reorth, tol)
    st = GmresStats()
    rc = _check(lib().fvp_newton_gmres(self.h, float(t), _ptr(psi), _ptr(psip), _ptr(F0), float(cj), _ptr(b),
                                       _ptr(x), ctypes.byref(p), ctypes.byref(st)), "fvp_newton_gmres")
    return rc, st.as_dict()


def _ctx_profile(self, enable=True):
    _check(lib().fvp_profile(self.h, int(bool(enable))), "fvp_profile")


def _ctx_profile_get(self, kind):
    ms = ctypes.c_double()
    n = ctypes.c_long()
    _check(lib().fvp_profile_get(self.h, int(kind), ctypes.byref(ms), ctypes.byref(n)), "fvp_profile_get")
    return ms.value, n.value


Context.profile = _ctx_profile
Context.profile_get = _ctx_profile_get
Context.precon_setup = _ctx_precon_setup
Context.precon_solve = _ctx_precon_solve
Context.jacobian = _ctx_jacobian
Context.colours = _ctx_colours
Context.set_block_rows = _ctx_set_block_rows
Context.set_overlap = _ctx_set_overlap
Context.set_jacobian_mode = _ctx_set_jacobian_mode
Context.set_residual_launch = _ctx_set_residual_launch
Context.autotune = _ctx_autotune
Context.newton_gmres = _ctx_newton_gmres


class IntegratorParams(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("max_order", ctypes.c_int),
                ("h0", ctypes.c_double), ("h_max", ctypes.c_double), ("precon_type", ctypes.c_int),
                ("dim_krylov", ctypes.c_int), ("max_restarts", ctypes.c_int), ("lin_factor", ctypes.c_double),
                ("newton_tol", ctypes.c_double), ("max_newton", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("block_rows", ctypes.c_int), ("overlap", ctypes.c_int), ("reorthogonalise", ctypes.c_int)]


class IntegratorStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_long), ("error_fails", ctypes.c_long), ("newton_fails", ctypes.c_long),
                ("residuals", ctypes.c_long), ("newton_iters", ctypes.c_long), ("gmres_iters", ctypes.c_long),
                ("precon_setups", ctypes.c_long), ("t", ctypes.c_double), ("h_last", ctypes.c_double),
                ("order_last", ctypes.c_int), ("learned_tried", ctypes.c_long), ("learned_used", ctypes.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Integrator:
    """BDF / Newton-GMRES time integrator on a Context (fvp_integrator_*)."""

    def __init__(self, ctx: Context, psi0, t0=0.0, rtol=1e-8, atol=1e-8, max_order=3, h0=0.0, h_max=0.0,
                 precon=PRECON_ILU0, dim_krylov=20, max_restarts=5, lin_factor=0.05, newton_tol=0.33,
                 max_newton=4, max_steps=1000000, block_rows=0, overlap=-1, reorthogonalise=0):
        self.ctx = ctx
        self.params = IntegratorParams(rtol, atol, max_order, h0, h_max, precon, dim_krylov, max_restarts,
                                       lin_factor, newton_tol, max_newton, max_steps, block_rows, overlap,
                                       reorthogonalise)
        h = ctypes.c_void_p()
        _check(lib().fvp_integrator_create(ctx.h, ctypes.byref(self.params), float(t0), _ptr(psi0),
                                           ctypes.byref(h)), "fvp_integrator_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fvp_integrator_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_segments(self, offsets):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        _check(lib().fvp_integrator_set_segments(self.h, len(off) - 1, off.ctypes.data_as(_c_int_p)),
               "fvp_integrator_set_segments")

    def set_learned(self, model, static_features=None, trust=2.0, band=0.5):
        """Newton initial iterate from a Mpnn (None switches it off); static
        feature columns 5..15 from a (N, 16) float32 device tensor."""
        self._learned = (model, static_features)
        _check(lib().fvp_integrator_set_learned(self.h, model.h if model is not None else None,
                                                _ptr(static_features) if static_features is not None else None,
                                                float(trust), float(band)),
               "fvp_integrator_set_learned")

    def set_recording(self, features=None, targets=None, static_features=None):
        """Write each accepted step's (features, target) pair into device
        tensors (N, 16) and (N,) float32; None switches recording off."""
        self._recording = (features, targets, static_features)
        _check(lib().fvp_integrator_set_recording(
            self.h, _ptr(static_features) if static_features is not None else None,
            _ptr(features) if features is not None else None, _ptr(targets) if targets is not None else None),
            "fvp_integrator_set_recording")

    def step(self, t_stop) -> int:
        return _check(lib().fvp_integrator_step(self.h, float(t_stop)), "fvp_integrator_step")

    def advance(self, t_out, psi_out=None) -> int:
        return _check(lib().fvp_integrator_advance(self.h, float(t_out), _ptr(psi_out) if psi_out is not None
                                                   else ctypes.c_void_p(0)), "fvp_integrator_advance")

    def state(self):
        t = ctypes.c_double()
        y = ctypes.c_void_p()
        yp = ctypes.c_void_p()
        _check(lib().fvp_integrator_state(self.h, ctypes.byref(t), ctypes.byref(y), ctypes.byref(yp)),
               "fvp_integrator_state")
        return t.value, y.value, yp.value

    def stats(self):
        st = IntegratorStats()
        _check(lib().fvp_integrator_get_stats(self.h, ctypes.byref(st)), "fvp_integrator_get_stats")
        return st.as_dict()


MPNN_FEATURES = 16
MPNN_HIDDEN = 64


class Mpnn:
    """Message-passing network on a node graph (fvp_mpnn_*): BF16 matrix-core
    forward, backward and Adam on the device.  ``ia``/``ja``: CSR adjacency
    without the diagonal (symmetric)."""

    def __init__(self, ia, ja, layers=3, seed=0, device=0):
        self._ia = np.ascontiguousarray(ia, dtype=np.int32)
        self._ja = np.ascontiguousarray(ja, dtype=np.int32)
        self.N = len(self._ia) - 1
        self.layers = int(layers)
        h = ctypes.c_void_p()
        _check(lib().fvp_mpnn_create(int(device), self.N, self._ia.ctypes.data_as(_c_int_p),
                                     self._ja.ctypes.data_as(_c_int_p), self.layers, int(seed), ctypes.byref(h)),
               "fvp_mpnn_create")
        self.h = h
        self.n_params = int(lib().fvp_mpnn_num_params(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().fvp_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream_ptr: int):
        _check(lib().fvp_mpnn_set_stream(self.h, ctypes.c_void_p(stream_ptr)), "fvp_mpnn_set_stream")

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        _check(lib().fvp_mpnn_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "fvp_mpnn_get_params")
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters, got {flat.shape}")
        _check(lib().fvp_mpnn_set_params(self.h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "fvp_mpnn_set_params")

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        _check(lib().fvp_mpnn_get_grads(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "fvp_mpnn_get_grads")
        return out

    def forward(self, x, y):
        """x: (N, 16) float32 device tensor, y: (N,) float32 device tensor."""
        _check(lib().fvp_mpnn_forward(self.h, _ptr(x), _ptr(y)), "fvp_mpnn_forward")

    def loss_grad(self, x, t) -> float:
        loss = ctypes.c_double()
        _check(lib().fvp_mpnn_loss_grad(self.h, _ptr(x), _ptr(t), ctypes.byref(loss)), "fvp_mpnn_loss_grad")
        return loss.value

    def adam(self, lr: float):
        _check(lib().fvp_mpnn_adam(self.h, float(lr)), "fvp_mpnn_adam")

    def synchronize(self):
        _check(lib().fvp_mpnn_synchronize(self.h), "fvp_mpnn_synchronize")


def mpnn_pack(p: dict, layers: int) -> np.ndarray:
    """Flat parameter array (fvp_mpnn_set_params order) from named arrays
    We (16 x 64), be, W1_l, W2_l, b_l, Wd (64 x 1), bd."""
    parts = [p["We"], p["be"]]
    for l in range(layers):
        parts += [p[f"W1_{l}"], p[f"W2_{l}"], p[f"b_{l}"]]
    parts += [p["Wd"], p["bd"]]
    return np.concatenate([np.asarray(a, dtype=np.float32).ravel() for a in parts])


def mpnn_unpack(flat, layers: int) -> dict:
    """Inverse of mpnn_pack."""
    F, H = MPNN_FEATURES, MPNN_HIDDEN
    shapes = [("We", (F, H)), ("be", (H,))]
    for l in range(layers):
        shapes += [(f"W1_{l}", (H, H)), (f"W2_{l}", (H, H)), (f"b_{l}", (H,))]
    shapes += [("Wd", (H, 1)), ("bd", (1,))]
    out, o = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    if o != len(flat):
        raise ValueError("parameter count mismatch")
    return out
