"""Device Schur-complement preconditioner (FVP_PRECON_SCHUR) against the
oracle's restatement of schur.c's one-level form: the apply on mesh matrices,
the exact inverse when the interior blocks are diagonal, GMRES iteration
counts, the interface-size limit, and a BDF run against the ILU(0) run."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import device
from oracle import oracle as O
from oracle.gmres_fh import arnoldi
from oracle.schur_ref import Schur, split

pytestmark = pytest.mark.gpu


def _matrix(prob, seed, diag=1.3):
    ia, ja = O.jac_pattern(O.OracleMesh(prob.mesh))
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, len(ja))
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    a[rows == ja] = diag * np.diff(ia)[rows[rows == ja]]
    return np.asarray(ia, np.int32), np.asarray(ja, np.int32), a


@pytest.fixture(scope="module")
def meshes():
    from fvmpor_amd.problems import box3d_problem, slice2d_problem
    return {"slice2d": slice2d_problem(), "box3d": box3d_problem(nx=14, ny=12, nz=8)}


def _dev(torch, a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device=device())


def _build(torch, ia, ja, a, bs):
    from fvmpor_amd.lib import CsrMatrix, PRECON_SCHUR
    A = CsrMatrix(ia, ja)
    A.precon_block_rows(bs)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_SCHUR) == 0
    return A


def _apply(torch, A, x):
    y = torch.empty(len(x), dtype=torch.float64, device=device())
    A.precon_apply(_dev(torch, x), y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("which,bs", [("slice2d", 256), ("slice2d", 1024), ("box3d", 64), ("box3d", 128)])
def test_schur_apply_matches_oracle(meshes, which, bs):
    import torch
    ia, ja, a = _matrix(meshes[which], seed=bs)
    n = len(ia) - 1
    P = Schur(ia, ja, a, bs)
    A = _build(torch, ia, ja, a, bs)
    rng = np.random.default_rng(7)
    for _ in range(2):  # the apply leaves the factors as they were
        x = rng.normal(size=n)
        ref = P.apply(x)
        y = _apply(torch, A, x)
        rel = np.abs(y - ref).max() / np.abs(ref).max()
        print(f"{which} bs={bs}: boundary rows {len(P.Gall)} of {n}, max per subdomain "
              f"{max(len(g) for g in P.G)}, rel {rel:.2e}")
        assert rel < 1e-10


def test_schur_exact_when_interior_blocks_are_diagonal(meshes):
    """B diagonal makes M_B exact, so the preconditioner is the exact inverse of
    A without the couplings between different subdomains' boundary rows."""
    import torch
    ia, ja, a = _matrix(meshes["slice2d"], seed=11)
    n, bs = len(ia) - 1, 512
    I, _ = split(ia, ja, bs)
    inter = np.zeros(n, bool)
    inter[np.concatenate(I)] = True
    rows = np.repeat(np.arange(n), np.diff(ia))
    a[inter[rows] & inter[ja] & (rows != ja)] = 0.0
    a[~inter[rows] & ~inter[ja] & (rows // bs != ja // bs)] = 0.0
    A = _build(torch, ia, ja, a, bs)
    x = np.random.default_rng(3).normal(size=n)
    y = _apply(torch, A, sp.csr_matrix((a, ja, ia), shape=(n, n)) @ x)
    assert np.abs(y - x).max() < 1e-10 * np.abs(x).max()


def test_schur_gmres_matches_oracle_arnoldi(meshes):
    """decomp's known answer b = A 1 => x = 1 with the Schur preconditioner: the
    device GMRES stops at the Arnoldi step where the oracle's Arnoldi process on
    A M (M the oracle's Schur apply) first has a least-squares residual below
    tol ||b||."""
    import torch
    ia, ja, a = _matrix(meshes["slice2d"], seed=5, diag=0.9)
    n, bs, m, tol = len(ia) - 1, 512, 30, 1e-10
    A = _build(torch, ia, ja, a, bs)
    Am = sp.csr_matrix((a, ja, ia), shape=(n, n))
    b = Am @ np.ones(n)
    x = torch.zeros(n, dtype=torch.float64, device=device())
    rc, st = A.gmres(_dev(torch, b), x, m=m, max_restarts=20, tol=tol, reorth=1)
    P = Schur(ia, ja, a, bs)
    Q, H = np.zeros((n, m + 1)), np.zeros((m + 1, m))
    beta = np.linalg.norm(b)
    Q[:, 0] = b / beta
    j = arnoldi(lambda v: Am @ P.apply(v), Q, H, 0, m, 1)
    res = []
    for k in range(1, j + 1):
        rhs = np.zeros(k + 1)
        rhs[0] = beta
        y = np.linalg.lstsq(H[:k + 1, :k], rhs, rcond=None)[0]
        res.append(np.linalg.norm(rhs - H[:k + 1, :k] @ y) / beta)
    first = 1 + next(k for k, r in enumerate(res) if r < tol)
    print(f"schur gmres: device {st}, oracle Arnoldi reaches tol at step {first}")
    assert rc == 0 and st["converged"] and st["restarts"] == 0
    assert abs(st["iterations"] - first) <= 1
    assert np.abs(x.cpu().numpy() - 1).max() < 1e-7


def test_schur_rejects_large_interfaces(meshes):
    import torch
    from fvmpor_amd.lib import CsrMatrix, FvpError, PRECON_SCHUR
    ia, ja, a = _matrix(meshes["box3d"], seed=1)
    A = CsrMatrix(ia, ja)
    A.precon_block_rows(512)  # subdomain 1 has 196 boundary rows
    A.set_values(_dev(torch, a))
    with pytest.raises(FvpError, match="196 boundary rows"):
        A.precon_build(PRECON_SCHUR)
    A.precon_block_rows(1024)  # at most 97: builds
    assert A.precon_build(PRECON_SCHUR) == 0


def test_schur_run_agrees_with_ilu0_run():
    """The slice to 0.02 day with the Schur preconditioner (1024-row
    subdomains) against the same run with ILU(0): the preconditioner changes the
    Krylov path only, so the final heads agree to the solver tolerances."""
    import torch
    from fvmpor_amd.lib import Context, Integrator, PRECON_ILU0, PRECON_SCHUR
    from fvmpor_amd.problems import slice2d_problem
    prob = slice2d_problem()
    out = {}
    for name, ptype in (("ilu0", PRECON_ILU0), ("schur", PRECON_SCHUR)):
        ctx = Context(prob)
        psi0 = torch.tensor(prob.psi0, device=device())
        integ = Integrator(ctx, psi0, rtol=1e-6, atol=1e-6, newton_tol=1e-6, precon=ptype, block_rows=1024)
        y = torch.empty_like(psi0)
        integ.advance(0.02, y)
        torch.cuda.synchronize()
        out[name] = (y.cpu().numpy(), integ.stats())
    rel = np.abs(out["ilu0"][0] - out["schur"][0]).max() / np.abs(out["ilu0"][0]).max()
    print(f"schur vs ilu0 run: rel {rel:.2e}; gmres iters {out['schur'][1]['gmres_iters']} vs "
          f"{out['ilu0'][1]['gmres_iters']}, steps {out['schur'][1]['steps']} vs {out['ilu0'][1]['steps']}")
    assert rel < 1e-5
