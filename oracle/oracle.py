"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU oracle (liboracle.so).

The oracle restates the reference's algorithms on the CPU (see orc.h and the
header of each .c file for the reference file:line each function follows).
Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_mesh_create.restype = _vp
        L.orc_mesh_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip,
                                      ctypes.POINTER(ctypes.c_int)]
        L.orc_mesh_free.argtypes = [_vp]
        L.orc_mesh_counts.argtypes = [_vp, _ip]
        L.orc_mesh_volumes.argtypes = [_vp, _dp, ctypes.c_void_p]
        L.orc_mesh_edges.argtypes = [_vp, _ip, _ip]
        L.orc_mesh_cvfaces.argtypes = [_vp, _ip, _dp]
        L.orc_mesh_gradients.argtypes = [_vp, _dp]
        L.orc_problem_create.restype = _vp
        L.orc_problem_create.argtypes = [_vp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _dp,
                                         _ip, _ip, ctypes.c_double]
        L.orc_problem_free.argtypes = [_vp]
        L.orc_residual.argtypes = [_vp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_begin_step.argtypes = [_vp, ctypes.c_double, _dp, _dp]
        L.orc_set_weights.argtypes = [_vp, _dp]
        L.orc_get_weights.argtypes = [_vp, _dp, _dp]
        L.orc_get_node_fields.argtypes = [_vp, _dp, _dp, _dp]
        L.orc_get_face_fields.argtypes = [_vp, _dp, _dp]
        L.orc_flux_scale.restype = ctypes.c_double
        L.orc_flux_scale.argtypes = [_vp]
        L.orc_num_dirichlet.restype = ctypes.c_int
        L.orc_num_dirichlet.argtypes = [_vp]
        L.orc_vg.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_csr_gemv.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_ilu0_factor.restype = ctypes.c_int
        L.orc_ilu0_factor.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip]
        L.orc_ilu0_apply.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _dp, _dp]
        L.orc_jacobi_factor.restype = ctypes.c_int
        L.orc_jacobi_factor.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp]
        L.orc_gmres_csr.restype = ctypes.c_int
        L.orc_gmres_csr.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, _dp]
        L.orc_arnoldi_csr.restype = ctypes.c_int
        L.orc_arnoldi_csr.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_jac_pattern.restype = ctypes.c_int
        L.orc_jac_pattern.argtypes = [_vp, _ip, ctypes.c_void_p]
        L.orc_colouring.restype = ctypes.c_int
        L.orc_colouring.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.orc_fd_jacobian.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                      _dp, _dp, _ip, _ip, _ip, ctypes.c_int, _dp]
        L.orc_jv.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_newton_gmres.restype = ctypes.c_int
        L.orc_newton_gmres.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_double, _ip, _ip,
                                       _dp, ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, _dp]
        L.orc_integrate_replay.restype = ctypes.c_int
        L.orc_integrate_replay.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _dp]
        L.orc_integrate_replay_opts.restype = ctypes.c_int
        L.orc_integrate_replay_opts.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                                ctypes.c_double, _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int,
                                                _dp, _dp]
        L.orc_rcm.restype = ctypes.c_int
        L.orc_rcm.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_int, _ip]
        L.orc_initial_derivative.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        _LIB = L
    return _LIB


def vg(h: float, zone_row) -> tuple:
    out = np.zeros(3)
    lib().orc_vg(float(h), np.ascontiguousarray(zone_row, dtype=np.float64), out)
    return tuple(out)


class OracleMesh:
    """CVFE dual of a simplex mesh built as the reference's mesh.cpp builds it."""

    def __init__(self, mesh):
        L = lib()
        self.dim = mesh.dim
        self._xyz = np.ascontiguousarray(mesh.xyz, dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        self._ft = np.ascontiguousarray(mesh.face_tags, dtype=np.int32)
        st = ctypes.c_int(0)
        self.h = L.orc_mesh_create(mesh.dim, self._xyz.shape[0], self._xyz, self._conn.shape[0],
                                   self._conn, self._ft, ctypes.byref(st))
        if not self.h:
            raise ValueError(f"oracle mesh construction failed ({st.value})")
        c = np.zeros(6, dtype=np.int32)
        L.orc_mesh_counts(self.h, c)
        self.N, self.T, self.NE, self.NFI, self.NFB, self.nnz_cv = (int(v) for v in c)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def volumes(self):
        V = np.zeros(self.N)
        scv = np.zeros(self.T * (self.dim + 1))
        lib().orc_mesh_volumes(self.h, V, scv.ctypes.data)
        return V, scv.reshape(self.T, self.dim + 1)

    def edges(self):
        f = np.zeros(self.NE, dtype=np.int32)
        b = np.zeros(self.NE, dtype=np.int32)
        lib().orc_mesh_edges(self.h, f, b)
        return f, b

    def cvfaces(self):
        nf = self.NFI + self.NFB
        ints = np.zeros(5 * nf, dtype=np.int32)
        dbl = np.zeros(4 * nf)
        lib().orc_mesh_cvfaces(self.h, ints, dbl)
        return ints.reshape(nf, 5), dbl.reshape(nf, 4)

    def gradients(self):
        g = np.zeros(self.T * (self.dim + 1) * 3)
        lib().orc_mesh_gradients(self.h, g)
        return g.reshape(self.T, self.dim + 1, 3)


class OracleProblem:
    """The reference's VarSatPhysics residual on an OracleMesh."""

    def __init__(self, problem, omesh: OracleMesh | None = None):
        L = lib()
        self.omesh = omesh if omesh is not None else OracleMesh(problem.mesh)
        m = problem.mesh
        self._keep = [
            np.ascontiguousarray(problem.zones, dtype=np.float64),
            np.ascontiguousarray(problem.ezone, dtype=np.int32),
            np.ascontiguousarray(problem.ekmul, dtype=np.float64),
            np.ascontiguousarray(problem.bc_type, dtype=np.int32),
            np.ascontiguousarray(problem.bc, dtype=np.float64),
            np.ascontiguousarray(m.node_tag_ptr, dtype=np.int32),
            np.ascontiguousarray(m.node_tags if len(m.node_tags) else np.zeros(1), dtype=np.int32),
        ]
        z, ez, ek, bt, bc, ntp, nt = self._keep
        self.h = L.orc_problem_create(self.omesh.h, z.shape[0], z, ez, ek, bt.shape[0], bt, bc,
                                      ntp, nt, float(problem.rho0))
        self.N = self.omesh.N
        self.NE = self.omesh.NE

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_free(self.h)
            self.h = None

    def residual(self, t, psi, psip):
        F = np.zeros(self.N)
        lib().orc_residual(self.h, float(t), np.ascontiguousarray(psi, dtype=np.float64),
                           np.ascontiguousarray(psip, dtype=np.float64), F)
        return F

    def begin_step(self, t, psi, psip):
        lib().orc_begin_step(self.h, float(t), np.ascontiguousarray(psi, dtype=np.float64),
                             np.ascontiguousarray(psip, dtype=np.float64))

    def weights(self):
        wf = np.zeros(self.NE)
        fl = np.zeros(self.NE)
        lib().orc_get_weights(self.h, wf, fl)
        return wf, fl

    def set_weights(self, wf):
        lib().orc_set_weights(self.h, np.ascontiguousarray(wf, dtype=np.float64))

    def node_fields(self):
        a, s, d = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        lib().orc_get_node_fields(self.h, a, s, d)
        return a, s, d

    def face_fields(self):
        nf = self.omesh.NFI + self.omesh.NFB
        q, M = np.zeros(nf), np.zeros(nf)
        lib().orc_get_face_fields(self.h, q, M)
        return q, M

    def flux_scale(self):
        return lib().orc_flux_scale(self.h)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def csr_gemv(ia, ja, a, x):
    n = len(ia) - 1
    y = np.zeros(n)
    lib().orc_csr_gemv(n, _i32(ia), _i32(ja), _f64(a), _f64(x), y)
    return y


def ilu0(ia, ja, a):
    """ILU(0) values (L unit lower, U with inverted diagonal) and diagonal positions."""
    n = len(ia) - 1
    lu = _f64(a).copy()
    dp = np.zeros(n, dtype=np.int32)
    ok = lib().orc_ilu0_factor(n, _i32(ia), _i32(ja), lu, dp)
    return (lu, dp) if ok else (None, None)


def ilu0_apply(ia, ja, lu, dp, x):
    n = len(ia) - 1
    y = np.zeros(n)
    lib().orc_ilu0_apply(n, _i32(ia), _i32(ja), _f64(lu), _i32(dp), _f64(x), y)
    return y


def jacobi(ia, ja, a):
    n = len(ia) - 1
    J = np.zeros(n)
    return J if lib().orc_jacobi_factor(n, _i32(ia), _i32(ja), _f64(a), J) else None


def gmres_csr(ia, ja, a, b, x0=None, precon=0, m=30, max_restarts=20, tol=1e-10, reorth=0):
    n = len(ia) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    st = np.zeros(5)
    ok = lib().orc_gmres_csr(n, _i32(ia), _i32(ja), _f64(a), int(precon), _f64(b), x,
                             np.array([m, max_restarts, reorth], dtype=np.int32), float(tol), st)
    if not ok:
        raise ValueError("preconditioner build failed")
    return x, dict(restarts=int(st[0]), iterations=int(st[1]), converged=bool(st[2]), residual=st[3], error=st[4])


def arnoldi_csr(ia, ja, a, b, m, reorth=0):
    n = len(ia) - 1
    Q = np.zeros(n * (m + 1))
    H = np.zeros((m + 1) * m)
    j = lib().orc_arnoldi_csr(n, _i32(ia), _i32(ja), _f64(a), _f64(b), m, reorth, Q, H)
    return j, Q.reshape(m + 1, n).T, H.reshape(m, m + 1).T


def jac_pattern(omesh):
    ia = np.zeros(omesh.N + 1, dtype=np.int32)
    nnz = lib().orc_jac_pattern(omesh.h, ia, None)
    ja = np.zeros(nnz, dtype=np.int32)
    lib().orc_jac_pattern(omesh.h, ia, ja.ctypes.data)
    return ia, ja


def colouring(ia, ja):
    n = len(ia) - 1
    col = np.zeros(n, dtype=np.int32)
    nc = lib().orc_colouring(n, _i32(ia), _i32(ja), col)
    return nc, col


def fd_jacobian(oprob, t, c, h, y, yp, F0, ewt, ia, ja, colour, ncol):
    vals = np.zeros(len(ja))
    lib().orc_fd_jacobian(oprob.h, oprob.N, float(t), float(c), float(h), _f64(y), _f64(yp), _f64(F0), _f64(ewt),
                          _i32(ia), _i32(ja), _i32(colour), int(ncol), vals)
    return vals


def jv(oprob, t, y, yp, F0, c, v):
    out = np.zeros(oprob.N)
    lib().orc_jv(oprob.h, oprob.N, float(t), _f64(y), _f64(yp), _f64(F0), float(c), _f64(v), out)
    return out


def newton_gmres(oprob, t, y, yp, F0, c, ia, ja, jvals, b, precon=5, m=30, max_restarts=10, tol=1e-8, reorth=1):
    x = np.zeros(oprob.N)
    st = np.zeros(5)
    ok = lib().orc_newton_gmres(oprob.h, oprob.N, float(t), _f64(y), _f64(yp), _f64(F0), float(c), _i32(ia),
                                _i32(ja), _f64(jvals), int(precon), _f64(b), x,
                                np.array([m, max_restarts, reorth], dtype=np.int32), float(tol), st)
    if not ok:
        raise ValueError("preconditioner build failed")
    return x, dict(restarts=int(st[0]), iterations=int(st[1]), converged=bool(st[2]), residual=st[3], error=st[4])


def integrate_replay(oprob, psi0, hs, ks, wfs, t0=0.0, rtol=1e-8, atol=1e-8, newton_tol=1e-6):
    """Replay a BDF step sequence (sizes hs, orders ks, per-step upwind front weights wfs)."""
    out = np.zeros(oprob.N)
    wfs = _f64(np.asarray(wfs).reshape(-1))
    n = lib().orc_integrate_replay(oprob.h, float(rtol), float(atol), float(newton_tol), float(t0), _f64(psi0),
                                   len(hs), _f64(hs), _i32(ks), wfs, oprob.NE, out)
    if n < 0:
        raise RuntimeError("oracle replay: Newton failed")
    return out, n


class ReplayOpts(ctypes.Structure):
    _fields_ = [("newton_tol", ctypes.c_double), ("lin_rel", ctypes.c_double), ("lin_abs", ctypes.c_double),
                ("jac_policy", ctypes.c_int), ("ida_test", ctypes.c_int), ("max_newton", ctypes.c_int),
                ("dim_krylov", ctypes.c_int), ("max_restarts", ctypes.c_int), ("own_upwind", ctypes.c_int),
                ("verbose", ctypes.c_int)]


def integrate_replay_opts(oprob, psi0, hs, ks, wfs=None, t0=0.0, rtol=1e-8, atol=1e-8, newton_tol=1e-7,
                          lin_rel=1e-12, lin_abs=0.0, jac_policy=0, ida_test=False, max_newton=30, dim_krylov=30,
                          max_restarts=20, verbose=False):
    """integrate_replay with the Newton / GMRES / Jacobian-refresh options of
    orc_integrate_replay_opts; wfs = None freezes the oracle's own upwind
    weights at each step start.  Returns (psi, stats) with stats Newton
    iterations, GMRES iterations, Jacobian set-ups, residual evaluations."""
    out = np.zeros(oprob.N)
    st = np.zeros(4)
    own = wfs is None
    wfs = _f64(np.zeros(1) if own else np.asarray(wfs).reshape(-1))
    o = ReplayOpts(newton_tol, lin_rel, lin_abs, jac_policy, int(ida_test), max_newton, dim_krylov, max_restarts,
                   int(own), int(verbose))
    n = lib().orc_integrate_replay_opts(oprob.h, float(rtol), float(atol), ctypes.byref(o), float(t0), _f64(psi0),
                                        len(hs), _f64(hs), _i32(ks), wfs, oprob.NE, out, st)
    if n < 0:
        raise RuntimeError("oracle replay: Newton failed")
    return out, dict(newton_iters=int(st[0]), gmres_iters=int(st[1]), jac_setups=int(st[2]), residuals=int(st[3]))


def rcm(ia, ja, seed=-1):
    """Reverse Cuthill-McKee order of a CSR graph (no self loops): perm[i] = old
    id of new node i; seed < 0 starts each component at a pseudo-peripheral node."""
    n = len(ia) - 1
    perm = np.zeros(n, dtype=np.int32)
    lib().orc_rcm(n, _i32(ia), _i32(ja), int(seed), perm)
    return perm


def initial_derivative(oprob, t0, psi0):
    yp = np.zeros(oprob.N)
    lib().orc_initial_derivative(oprob.h, oprob.N, float(t0), _f64(psi0), yp)
    return yp
