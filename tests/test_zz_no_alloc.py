"""The allocation-free GMRES entry point (fvp_gmres_csr_noalloc, workspace in
a static per-GPU device array) against fvp_gmres_csr on the same systems:
two calls in a row, the second on a smaller system, so the second call runs
on the first call's leftover workspace; plus the size limit."""
import numpy as np
import pytest

from conftest import device

pytestmark = pytest.mark.gpu


def _system(dim, size, seed):
    from fvmpor_amd.learn import node_graph
    from fvmpor_amd.problems import box3d_problem, slice2d_problem
    prob = slice2d_problem(nx=size[0], ny=size[1]) if dim == 2 else box3d_problem(*size)
    ia0, ja0 = node_graph(prob.mesh)
    n = len(ia0) - 1
    # pattern with the diagonal, diagonally dominant values
    rows = np.repeat(np.arange(n), np.diff(ia0))
    keys = np.unique(np.concatenate([rows * n + ja0, np.arange(n) * (n + 1)]))
    r, c = keys // n, keys % n
    ia = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(r, minlength=n), out=ia[1:])
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 0, len(keys))
    diag = r == c
    a[diag] = 1.3 * np.bincount(r, weights=np.abs(a) * ~diag, minlength=n)
    return ia, c.astype(np.int32), a, rng.normal(size=n)


@pytest.mark.parametrize("precon", ["BJILU0", "ILU0", "NONE"])
def test_noalloc_gmres_matches_gmres_csr_twice(precon):
    import torch
    from fvmpor_amd import lib as L
    dev = device()
    ptype = getattr(L, "PRECON_" + precon)
    for dim, size, m in ((3, (40, 40, 12), 30), (2, (31, 13), 20)):
        ia, ja, a, b = _system(dim, size, seed=dim)
        A = L.CsrMatrix(ia, ja)
        A.set_values(torch.tensor(a, device=dev))
        assert A.precon_build(ptype) == 0
        tb = torch.tensor(b, device=dev)
        x1 = torch.zeros_like(tb)
        x2 = torch.zeros_like(tb)
        rc1, s1 = A.gmres(tb, x1, m=m, max_restarts=20, tol=1e-10)
        rc2, s2 = A.gmres_noalloc(tb, x2, m=m, max_restarts=20, tol=1e-10)
        torch.cuda.synchronize()
        g1, g2 = x1.cpu().numpy(), x2.cpu().numpy()
        rel = np.abs(g1 - g2).max() / np.abs(g1).max()
        print(f"{precon} n={len(b)} m={m}: iterations {s1['iterations']} / {s2['iterations']}, "
              f"bitwise {np.array_equal(g1, g2)}, rel diff {rel:.1e}")
        assert rc1 == rc2 == 0 and s1["converged"] and s2["converged"]
        assert s1["iterations"] == s2["iterations"]
        assert rel <= 1e-12
        A.close()


def test_noalloc_gmres_rejects_a_system_beyond_its_workspace():
    import torch
    from fvmpor_amd import lib as L
    dev = device()
    n = 200000  # 200000 x (30 + 4) doubles > 4 Mi doubles
    ia = np.arange(n + 1, dtype=np.int32)
    ja = np.arange(n, dtype=np.int32)
    A = L.CsrMatrix(ia, ja)
    A.set_values(torch.ones(n, dtype=torch.float64, device=dev))
    assert A.precon_build(L.PRECON_JACOBI) == 0
    b = torch.ones(n, dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    with pytest.raises(L.FvpError, match="static workspace"):
        A.gmres_noalloc(b, x, m=30)
    # a shorter restart fits: 200000 x (10 + 4) doubles
    rc, st = A.gmres_noalloc(b, x, m=10)
    torch.cuda.synchronize()
    assert rc == 0 and st["converged"]
    assert np.allclose(x.cpu().numpy(), 1.0)
