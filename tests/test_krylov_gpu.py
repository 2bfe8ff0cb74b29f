"""GPU parity of the sparse linear algebra and Newton-Krylov pieces against
the CPU oracle: CSR gemv, ILU(0) / Jacobi factor and apply, GMRES (decomp
known answer b = A 1 => x = 1), coloured FD Jacobian and the matrix-free
Newton solve."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import device
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _matrix(mesh, seed=0, shift=1.5):
    om = O.OracleMesh(mesh)
    ia, ja = O.jac_pattern(om)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, len(ja))
    for i in range(len(ia) - 1):
        for k in range(ia[i], ia[i + 1]):
            if ja[k] == i:
                a[k] = shift * (ia[i + 1] - ia[i])
    return ia, ja, a


@pytest.fixture(scope="module")
def mats(cassion_mesh, box3d_small):
    return {"cassion": _matrix(cassion_mesh), "box3d": _matrix(box3d_small.mesh, 1, 1.2)}


def _dev(torch, a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device=device())


@pytest.mark.parametrize("which", ["cassion", "box3d"])
def test_csr_gemv(mats, which):
    import torch
    from fvmpor_amd.lib import CsrMatrix
    ia, ja, a = mats[which]
    A = CsrMatrix(ia, ja)
    A.set_values(_dev(torch, a))
    x = np.random.default_rng(2).normal(size=len(ia) - 1)
    y0 = np.random.default_rng(3).normal(size=len(ia) - 1)
    y = _dev(torch, y0)
    A.gemv(2.0, _dev(torch, x), 0.5, y)
    torch.cuda.synchronize()
    ref = 2.0 * O.csr_gemv(ia, ja, a, x) + 0.5 * y0
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("which", ["cassion", "box3d"])
def test_ilu0_factor_and_apply(mats, which):
    import torch
    from fvmpor_amd.lib import CsrMatrix, PRECON_ILU0
    ia, ja, a = mats[which]
    A = CsrMatrix(ia, ja)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_ILU0) == 0
    lu_g = A.precon_values(len(ja))
    lu_c, dp = O.ilu0(ia, ja, a)
    np.testing.assert_allclose(lu_g, lu_c, rtol=1e-12, atol=1e-14)
    x = np.random.default_rng(4).normal(size=len(ia) - 1)
    y = torch.empty(len(ia) - 1, dtype=torch.float64, device=device())
    A.precon_apply(_dev(torch, x), y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), O.ilu0_apply(ia, ja, lu_c, dp, x), rtol=1e-11, atol=1e-13)


def test_jacobi_and_zero_pivot(mats):
    import torch
    from fvmpor_amd.lib import CsrMatrix, PRECON_JACOBI, PRECON_ILU0, ERR_PIVOT
    ia, ja, a = mats["cassion"]
    A = CsrMatrix(ia, ja)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_JACOBI) == 0
    np.testing.assert_allclose(A.precon_values(len(ia) - 1), O.jacobi(ia, ja, a), rtol=1e-15)
    b = a.copy()
    dpos = [k for k in range(ia[5], ia[6]) if ja[k] == 5][0]
    b[dpos] = 0.0
    A.set_values(_dev(torch, b))
    assert A.precon_build(PRECON_JACOBI) == ERR_PIVOT
    # ILU(0) meeting an exactly zero pivot on the first row
    c = a.copy()
    c[[k for k in range(ia[0], ia[1]) if ja[k] == 0][0]] = 0.0
    A.set_values(_dev(torch, c))
    assert A.precon_build(PRECON_ILU0) == ERR_PIVOT


@pytest.mark.parametrize("which", ["cassion", "box3d"])
@pytest.mark.parametrize("ptype", [0, 2, 5])
@pytest.mark.parametrize("reorth", [0, 1, 2])
def test_gmres_known_answer(mats, which, ptype, reorth):
    import torch
    from fvmpor_amd.lib import CsrMatrix
    ia, ja, a = mats[which]
    n = len(ia) - 1
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    xo, sto = O.gmres_csr(ia, ja, a, b, precon=ptype, m=10, max_restarts=100, tol=1e-12)
    assert sto["converged"], sto
    A = CsrMatrix(ia, ja)
    A.set_values(_dev(torch, a))
    assert A.precon_build(ptype) == 0
    x = torch.zeros(n, dtype=torch.float64, device=device())
    torch.cuda.synchronize()
    rc, st = A.gmres(_dev(torch, b), x, m=30, max_restarts=40, tol=1e-12, reorth=reorth)
    assert rc == 0 and st["converged"], st
    assert np.abs(x.cpu().numpy() - 1.0).max() < 1e-9
    np.testing.assert_allclose(x.cpu().numpy(), xo, atol=1e-9)
    # the host inputs are untouched by the device solve
    assert np.array_equal(b, O.csr_gemv(ia, ja, a, np.ones(n)))


@pytest.mark.parametrize("which", ["slice2d", "box3d_small"])
def test_fd_jacobian_and_newton_solve(which, request):
    import torch
    from fvmpor_amd.lib import Context, PRECON_ILU0
    prob = request.getfixturevalue(which)
    ctx = Context(prob)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(7)
    y = prob.psi0 - 0.3 + rng.uniform(-0.2, 0.2, op.N)
    yp = rng.normal(scale=0.05, size=op.N)
    ty, typ = _dev(torch, y), _dev(torch, yp)
    ctx.begin_step(0.0, ty, typ)
    wf, _ = ctx.get_upwind()
    op.set_weights(wf)
    F0 = torch.empty_like(ty)
    ctx.residual(0.0, ty, typ, F0)
    torch.cuda.synchronize()
    F0c = op.residual(0.0, y, yp)
    ewt = 1.0 / (1e-6 * np.abs(y) + 1e-6)
    c, h = 12.0, 0.05
    assert ctx.precon_setup(0.0, c, h, ty, typ, F0, _dev(torch, ewt), PRECON_ILU0) == 0
    # colouring is the reference's greedy distance-2 colouring, node for node
    ia, ja = O.jac_pattern(op.omesh)
    nc_o, col_o = O.colouring(ia, ja)
    nc_g, col_g = ctx.colours()
    assert nc_g == nc_o and np.array_equal(col_g, col_o)
    Jo = O.fd_jacobian(op, 0.0, c, h, y, yp, F0c, ewt, ia, ja, col_o, nc_o)
    J = ctx.jacobian()
    Jg = J.get_values(len(ja))
    # FD columns carry O(sqrt(eps)) noise: compare against each row's largest entry
    rowmax = np.repeat(np.maximum.reduceat(np.abs(Jo), ia[:-1]), np.diff(ia))
    # noise: rounding in F (~1e-16 of the largest face flux) over shifts ~1e-8 |psi|
    assert (np.abs(Jg - Jo) / rowmax).max() < 1e-4
    # Newton linear solve J x = -F0 against the oracle's
    b = -F0c
    x = torch.zeros_like(ty)
    # the FD J v carries O(sqrt(eps)) noise, so a matrix-free residual below ~1e-8 is not reachable
    rc, st = ctx.newton_gmres(0.0, ty, typ, F0, c, _dev(torch, b), x, m=30, max_restarts=10, tol=1e-7)
    assert rc == 0 and st["converged"], st
    xo, sto = O.newton_gmres(op, 0.0, y, yp, F0c, c, ia, ja, Jo, b, m=10, max_restarts=50, tol=1e-7)
    assert sto["converged"], sto
    xg = x.cpu().numpy()
    assert np.abs(xg - xo).max() <= 1e-5 * np.abs(xo).max()
    print(f"{which}: colours={nc_g} gmres its gpu={st['iterations']} oracle={sto['iterations']}")


@pytest.mark.parametrize("which", ["slice2d", "box3d_small", "cassion", "slice2d_dirichlet"])
def test_element_local_jacobian_matches_coloured(which, request):
    """The element-local quotients (the default) against the coloured ones
    (preconditioner_dss.cpp's form) on the same state: the same Jacobian up to
    the rounding of the residual sums, Dirichlet rows included."""
    import torch
    import dataclasses
    from fvmpor_amd.lib import Context, JACOBIAN_COLOURED, JACOBIAN_ELEMENT, PRECON_ILU0
    if which == "slice2d_dirichlet":  # the slice with boundary tag 2 held at a prescribed head
        base = request.getfixturevalue("slice2d")
        bc = base.bc.copy()
        bc[2, 0] = -1.0
        prob = dataclasses.replace(base, bc_type=np.array([0, 3, 1], dtype=np.int32), bc=bc)
    else:
        prob = request.getfixturevalue(which)
    ctx = Context(prob)
    if which == "slice2d_dirichlet":
        assert ctx.ndir > 0
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(11)
    y = prob.psi0 - 0.3 + rng.uniform(-0.2, 0.2, op.N)
    yp = rng.normal(scale=0.05, size=op.N)
    ty, typ = _dev(torch, y), _dev(torch, yp)
    ctx.begin_step(0.0, ty, typ)
    F0 = torch.empty_like(ty)
    ctx.residual(0.0, ty, typ, F0)
    ewt = _dev(torch, 1.0 / (1e-6 * np.abs(y) + 1e-6))
    ia, ja = O.jac_pattern(op.omesh)
    J = ctx.jacobian()
    vals = {}
    for mode in (JACOBIAN_COLOURED, JACOBIAN_ELEMENT, JACOBIAN_COLOURED, JACOBIAN_ELEMENT):
        ctx.set_jacobian_mode(mode)
        assert ctx.precon_setup(0.0, 12.0, 0.05, ty, typ, F0, ewt, PRECON_ILU0) == 0
        vals.setdefault(mode, []).append(J.get_values(len(ja)))
    col, el = vals[JACOBIAN_COLOURED][0], vals[JACOBIAN_ELEMENT][0]
    rowmax = np.repeat(np.maximum.reduceat(np.abs(col), ia[:-1]), np.diff(ia))
    err = (np.abs(el - col) / rowmax).max()
    # the residual's sums are not bitwise repeatable (atomics in the element
    # kernel): the coloured form against itself shows part of the FD noise
    noise = (np.abs(vals[JACOBIAN_COLOURED][1] - col) / rowmax).max()
    print(f"{which}: element-local vs coloured, max |diff| / row max = {err:.2e} (coloured repeat {noise:.2e})")
    # the quotients' rounding noise (heads differ by sqrt(eps)-sized shifts on top
    # of elevations) is the bar of the oracle FD test; a missing term (capacity,
    # moved k_r) would show at O(1)
    assert err < 1e-4 and noise < 1e-4
    # the element-local form sums each entry in a fixed order: repeatable bit for bit
    assert np.array_equal(vals[JACOBIAN_ELEMENT][1], el)
    if which == "slice2d_dirichlet":
        # Dirichlet rows: the same quotient on the diagonal, exact zeros elsewhere
        # (the rows whose coloured off-diagonal entries are all exactly zero)
        zero_rows = np.array([np.all(col[ia[i]:ia[i + 1]][ja[ia[i]:ia[i + 1]] != i] == 0) for i in range(op.N)])
        assert zero_rows.sum() > 0
        for i in np.nonzero(zero_rows)[0]:
            sl = slice(ia[i], ia[i + 1])
            assert np.array_equal(el[sl], col[sl]), i
    with pytest.raises(Exception):
        ctx.set_jacobian_mode(7)


@pytest.mark.parametrize("which", ["slice2d", "box3d_small", "slice2d_dirichlet"])
def test_fp32_jacobian_of_the_fp32_preconditioner(which, request):
    """The FP32 preconditioner's Jacobian: the element-local quotients in FP32
    arithmetic, each formed without cancellation from the linearity of the
    face flux in the heads.  Against the coloured FD form (the reference's) on
    the same state and increments it stays within the FD noise bar of the
    FP64 forms; against the FP64 element-local form (whose own FD noise is the
    larger part) likewise; Dirichlet rows are equal."""
    import torch
    import dataclasses
    from fvmpor_amd.lib import Context, JACOBIAN_COLOURED, PRECON_BJILU0, PRECON_BJILU0_F32
    if which == "slice2d_dirichlet":
        base = request.getfixturevalue("slice2d")
        bc = base.bc.copy()
        bc[2, 0] = -1.0
        prob = dataclasses.replace(base, bc_type=np.array([0, 3, 1], dtype=np.int32), bc=bc)
    else:
        prob = request.getfixturevalue(which)
    ctx = Context(prob)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(11)
    y = prob.psi0 - 0.3 + rng.uniform(-0.2, 0.2, op.N)
    yp = rng.normal(scale=0.05, size=op.N)
    ty, typ = _dev(torch, y), _dev(torch, yp)
    ctx.begin_step(0.0, ty, typ)
    F0 = torch.empty_like(ty)
    ctx.residual(0.0, ty, typ, F0)
    ewt = _dev(torch, 1.0 / (1e-6 * np.abs(y) + 1e-6))
    ia, ja = O.jac_pattern(op.omesh)
    J = ctx.jacobian()
    assert ctx.precon_setup(0.0, 12.0, 0.05, ty, typ, F0, ewt, PRECON_BJILU0_F32) == 0
    j32 = J.get_values(len(ja))
    assert ctx.precon_setup(0.0, 12.0, 0.05, ty, typ, F0, ewt, PRECON_BJILU0) == 0
    j64 = J.get_values(len(ja))
    ctx.set_jacobian_mode(JACOBIAN_COLOURED)
    assert ctx.precon_setup(0.0, 12.0, 0.05, ty, typ, F0, ewt, PRECON_BJILU0) == 0
    col = J.get_values(len(ja))
    rowmax = np.repeat(np.maximum.reduceat(np.abs(col), ia[:-1]), np.diff(ia))
    e_col = (np.abs(j32 - col) / rowmax).max()
    e_64 = (np.abs(j32 - j64) / rowmax).max()
    print(f"{which}: FP32 Jacobian vs coloured {e_col:.2e}, vs FP64 element-local {e_64:.2e}")
    assert e_col < 1e-4 and e_64 < 1e-4
    if which == "slice2d_dirichlet":
        rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
        drow = np.array([np.all(col[ia[i]:ia[i + 1]][ja[ia[i]:ia[i + 1]] != i] == 0) for i in range(len(ia) - 1)])
        assert drow.sum() > 0
        sel = drow[rows]
        assert np.array_equal(j32[sel], j64[sel])


@pytest.fixture(scope="module")
def box_mid():
    from fvmpor_amd.problems import box3d_problem
    return box3d_problem(nx=22, ny=20, nz=10)


def _block_diagonal(ia, ja, a, bs):
    """Entries of A inside the diagonal blocks of bs consecutive rows (others dropped)."""
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    keep = (rows // bs) == (ja // bs)
    kia = np.zeros(n + 1, dtype=np.int32)
    np.add.at(kia, rows[keep] + 1, 1)
    return np.cumsum(kia).astype(np.int32), ja[keep].astype(np.int32), a[keep], keep


@pytest.mark.parametrize("rows", [0, 256, 1024])
def test_block_jacobi_ilu0_is_blockwise_ilu0(box_mid, rows):
    import torch
    from fvmpor_amd.lib import CsrMatrix, PRECON_BJILU0, BLOCK_ROWS
    ia, ja, a = _matrix(box_mid.mesh, seed=3, shift=1.2)
    n = len(ia) - 1
    bs = rows or BLOCK_ROWS
    assert n > 2 * bs
    A = CsrMatrix(ia, ja)
    if rows:
        A.precon_block_rows(rows)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_BJILU0) == 0
    bia, bja, ba, keep = _block_diagonal(ia, ja, a, bs)
    lu_o, dp = O.ilu0(bia, bja, ba)
    lu_g = A.precon_values(len(ja))
    np.testing.assert_allclose(lu_g[keep], lu_o, rtol=1e-12, atol=1e-14)
    assert np.all(lu_g[~keep] == 0)
    x = np.random.default_rng(5).normal(size=n)
    y = torch.empty(n, dtype=torch.float64, device=device())
    A.precon_apply(_dev(torch, x), y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), O.ilu0_apply(bia, bja, lu_o, dp, x), rtol=1e-11, atol=1e-13)
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    xs = torch.zeros(n, dtype=torch.float64, device=device())
    rc, st = A.gmres(_dev(torch, b), xs, m=30, max_restarts=40, tol=1e-12, reorth=1)
    assert rc == 0 and st["converged"]
    assert np.abs(xs.cpu().numpy() - 1.0).max() < 1e-9


def test_block_jacobi_ilu0_fp32(box_mid):
    """FP32 factor and solves: within single-precision rounding of the blockwise
    FP64 ILU(0) of the oracle, and GMRES (FP64 Krylov, right preconditioning)
    still reaches the FP64 solution."""
    import torch
    from fvmpor_amd.lib import CsrMatrix, PRECON_BJILU0_F32, BLOCK_ROWS
    ia, ja, a = _matrix(box_mid.mesh, seed=3, shift=1.2)
    n = len(ia) - 1
    A = CsrMatrix(ia, ja)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_BJILU0_F32) == 0
    bia, bja, ba, keep = _block_diagonal(ia, ja, a, BLOCK_ROWS)
    lu_o, dp = O.ilu0(bia, bja, ba)
    lu_g = A.precon_values(len(ja))
    err = np.abs(lu_g[keep] - lu_o).max() / np.abs(lu_o).max()
    assert 0 < err < 1e-5
    assert np.all(lu_g[~keep] == 0)
    x = np.random.default_rng(5).normal(size=n)
    y = torch.empty(n, dtype=torch.float64, device=device())
    A.precon_apply(_dev(torch, x), y)
    torch.cuda.synchronize()
    ref = O.ilu0_apply(bia, bja, lu_o, dp, x)
    assert np.abs(y.cpu().numpy() - ref).max() < 1e-5 * np.abs(ref).max()
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    xs = torch.zeros(n, dtype=torch.float64, device=device())
    rc, st = A.gmres(_dev(torch, b), xs, m=30, max_restarts=40, tol=1e-12, reorth=1)
    assert rc == 0 and st["converged"]
    assert np.abs(xs.cpu().numpy() - 1.0).max() < 1e-9


@pytest.mark.parametrize("layers", [0, 1])
def test_block_factor_staging_modes_agree(box_mid, monkeypatch, layers):
    """The block factorisation's two forms (values staged in LDS; in place in
    memory) run the same arithmetic: their solves agree bit for bit, in FP64
    and FP32, and the FP64 one is the oracle's blockwise ILU(0)."""
    import torch
    from fvmpor_amd.lib import CsrMatrix, PRECON_BJILU0, PRECON_BJILU0_F32, BLOCK_ROWS
    ia, ja, a = _matrix(box_mid.mesh, seed=4, shift=1.2)
    n = len(ia) - 1
    x = np.random.default_rng(9).normal(size=n)
    xd = _dev(torch, x)
    for ptype in (PRECON_BJILU0, PRECON_BJILU0_F32):
        ys = []
        for cap in (1, 0):
            monkeypatch.setenv("FVP_BJ_FACTOR_MODE", str(cap))
            A = CsrMatrix(ia, ja)
            A.precon_overlap(layers)
            A.set_values(_dev(torch, a))
            assert A.precon_build(ptype) == 0
            y = torch.empty(n, dtype=torch.float64, device=device())
            A.precon_apply(xd, y)
            torch.cuda.synchronize()
            ys.append(y.cpu().numpy())
        np.testing.assert_array_equal(ys[0], ys[1])
        if ptype == PRECON_BJILU0 and layers == 0:
            bia, bja, ba, _ = _block_diagonal(ia, ja, a, BLOCK_ROWS)
            lu_o, dp = O.ilu0(bia, bja, ba)
            np.testing.assert_allclose(ys[0], O.ilu0_apply(bia, bja, lu_o, dp, x), rtol=1e-11, atol=1e-13)


def test_block_rows_rejects_invalid_sizes(box_mid):
    from fvmpor_amd.lib import CsrMatrix, FvpError
    ia, ja, a = _matrix(box_mid.mesh, seed=3, shift=1.2)
    A = CsrMatrix(ia, ja)
    for bad in (100, 32, 8192):
        with pytest.raises(FvpError):
            A.precon_block_rows(bad)


def test_gmres_full_krylov_space_is_an_exact_solve():
    """m >= n: the Krylov space fills up (h_{n,n-1} -> 0, the lucky breakdown)
    and GMRES returns the exact solution of a small nonsymmetric system."""
    import torch
    from fvmpor_amd.lib import CsrMatrix
    rng = np.random.default_rng(11)
    n = 40
    D = rng.normal(size=(n, n)) * (rng.random((n, n)) < 0.3) + np.diag(rng.uniform(2, 4, n))
    S = sp.csr_matrix(D)
    S.sort_indices()
    A = CsrMatrix(S.indptr.astype(np.int32), S.indices.astype(np.int32))
    A.set_values(_dev(torch, S.data))
    assert A.precon_build(0) == 0
    b = rng.normal(size=n)
    x = torch.zeros(n, dtype=torch.float64, device=device())
    rc, st = A.gmres(_dev(torch, b), x, m=50, max_restarts=0, tol=1e-13, reorth=1)
    assert rc == 0 and st["converged"] and st["iterations"] <= n, st
    np.testing.assert_allclose(x.cpu().numpy(), np.linalg.solve(D, b), rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("layers", [1, 2])
def test_overlapping_blocks_are_restricted_additive_schwarz(box_mid, layers):
    """Each block's owned rows plus `layers` layers of graph neighbours, ILU(0)
    on the extended block in global row order (oracle), solution kept on the
    owned rows only; GMRES still reaches the exact solution."""
    import torch
    from fvmpor_amd.lib import CsrMatrix, FvpError, PRECON_BJILU0, BLOCK_ROWS
    ia, ja, a = _matrix(box_mid.mesh, seed=3, shift=1.2)
    n = len(ia) - 1
    A = CsrMatrix(ia, ja)
    A.precon_overlap(layers)
    A.set_values(_dev(torch, a))
    assert A.precon_build(PRECON_BJILU0) == 0
    with pytest.raises(FvpError):
        A.precon_values(len(ja))
    S = sp.csr_matrix((a, ja, ia), shape=(n, n))
    x = np.random.default_rng(7).normal(size=n)
    ref = np.empty(n)
    for lo in range(0, n, BLOCK_ROWS):
        own = np.arange(lo, min(n, lo + BLOCK_ROWS))
        ext = own
        for _ in range(layers):
            ext = np.union1d(ext, S[ext].indices)
        B = S[ext][:, ext].tocsr()
        B.sort_indices()
        bia, bja = B.indptr.astype(np.int32), B.indices.astype(np.int32)
        lu, dp = O.ilu0(bia, bja, B.data)
        ref[own] = O.ilu0_apply(bia, bja, lu, dp, x[ext])[np.searchsorted(ext, own)]
    y = torch.empty(n, dtype=torch.float64, device=device())
    A.precon_apply(_dev(torch, x), y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-11, atol=1e-13)
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    xs = torch.zeros(n, dtype=torch.float64, device=device())
    rc, st = A.gmres(_dev(torch, b), xs, m=30, max_restarts=40, tol=1e-12, reorth=1)
    assert rc == 0 and st["converged"]
    assert np.abs(xs.cpu().numpy() - 1.0).max() < 1e-9


def _wide_pattern_matrix(mesh, rings, seed):
    """Diagonally dominant values on the pattern of (adjacency + I)^rings: rows
    with many more in-block entries than the usual stencil (the solve's
    overflow lists carry everything past its first 12 entries per row part)."""
    from fvmpor_amd.learn import node_graph
    ia0, ja0 = node_graph(mesh)
    n = len(ia0) - 1
    G = sp.csr_matrix((np.ones(len(ja0)), ja0, ia0), shape=(n, n)) + sp.identity(n, format="csr")
    P = G.copy()
    for _ in range(rings - 1):
        P = (P @ G).tocsr()
    P.sort_indices()
    ia, ja = P.indptr.astype(np.int32), P.indices.astype(np.int32)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 0, len(ja))
    rows = np.repeat(np.arange(n), np.diff(ia))
    diag = rows == ja
    a[diag] = 1.2 * np.bincount(rows, weights=np.abs(a) * ~diag, minlength=n)[rows[diag]]
    return ia, ja, a


@pytest.mark.parametrize("ptype", ["BJILU0", "BJILU0_F32"])
def test_block_solve_long_rows_and_wide_levels(ptype):
    """Two structures the block solve treats specially, against the oracle's
    blockwise ILU(0): rows whose L or U part has more than 12 block entries
    (a 3-ring stencil on the 2D slice), and levels wider than a wavefront (a
    matrix whose blocks are diagonal, with couplings only between blocks, so
    every block is one level of 1024 rows)."""
    import torch
    from fvmpor_amd import lib as L
    from fvmpor_amd.problems import slice2d_problem
    pt = getattr(L, "PRECON_" + ptype)
    tol = 1e-11 if ptype == "BJILU0" else 2e-5
    mesh = slice2d_problem().mesh
    ia, ja, a = _wide_pattern_matrix(mesh, 3, seed=4)
    n = len(ia) - 1
    bs = 1024
    # couplings across blocks only: row i also couples to i +- bs
    r = np.arange(n)
    cols = [np.stack([r, r]).T]
    for off in (-bs, bs):
        ok = (r + off >= 0) & (r + off < n)
        cols.append(np.stack([r[ok], r[ok] + off]).T)
    pairs = np.concatenate(cols)
    W = sp.csr_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    W.sort_indices()
    wv = np.where(W.indices == np.repeat(r, np.diff(W.indptr)), 3.0, -0.5)
    cases = [("long rows", ia, ja, a), ("wide levels", W.indptr.astype(np.int32), W.indices.astype(np.int32), wv)]
    for name, cia, cja, ca in cases:
        rows = np.repeat(np.arange(n), np.diff(cia))
        inb = (rows // bs) == (cja // bs)
        lo_cnt = np.bincount(rows[inb & (cja < rows)], minlength=n)
        A = L.CsrMatrix(cia, cja)
        A.precon_block_rows(bs)
        A.set_values(_dev(torch, ca))
        assert A.precon_build(pt) == 0
        bia, bja, ba, keep = _block_diagonal(cia, cja, ca, bs)
        lu_o, dp = O.ilu0(bia, bja, ba)
        x = np.random.default_rng(9).normal(size=n)
        y = torch.empty(n, dtype=torch.float64, device=device())
        A.precon_apply(_dev(torch, x), y)
        torch.cuda.synchronize()
        ref = O.ilu0_apply(bia, bja, lu_o, dp, x)
        err = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"{ptype} {name}: max L entries per row {lo_cnt.max()}, rel err {err:.2e}")
        assert err <= tol
        if name == "long rows":
            assert lo_cnt.max() > 12
        A.close()
