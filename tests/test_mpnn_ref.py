"""The CPU reference of the mesh network (oracle/mpnn_ref.py): its gradient
formulas against central finite differences (no rounding, float64), and the
BF16 rounding against the bit-level definition."""
import numpy as np

from oracle import mpnn_ref as R
from oracle import oracle as O


def _graph(prob):
    ia, ja = O.jac_pattern(O.OracleMesh(prob.mesh))
    # neighbours without the node itself
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    keep = ja != rows
    ja2 = ja[keep]
    ia2 = np.zeros_like(ia)
    ia2[1:] = np.cumsum(np.bincount(rows[keep], minlength=len(ia) - 1))
    return ia2, ja2


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], dtype=np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0 and r[4] == -2.5
    # exactly representable values are unchanged, halfway cases go to even
    assert R.bf16(np.float32(1.0 + 2 ** -7))[()] == 1.0 + 2 ** -7
    assert R.bf16(np.float32(1.0 + 2 ** -8))[()] == 1.0          # halfway, ties to even (1.0)
    assert R.bf16(np.float32(1.0 + 3 * 2 ** -8))[()] == 1.0 + 2 ** -6  # halfway, ties to even
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_gradients_match_finite_differences():
    from fvmpor_amd.problems import slice2d_problem
    prob = slice2d_problem(nx=11, ny=6)
    ia, ja = _graph(prob)
    N = len(ia) - 1
    rng = np.random.default_rng(0)
    X = rng.normal(size=(N, R.F_IN))
    t = rng.normal(size=N)
    p = R.init_params(layers=2, seed=1)
    for k in p:
        p[k] = p[k] + rng.normal(scale=0.05, size=p[k].shape)  # biases off zero: no ties at relu kinks
    L0, g = R.loss_and_grad_fp64(p, X, t, ia, ja, layers=2)
    eps = 1e-6
    for name in ("We", "be", "W1_0", "W2_1", "b_1", "Wd", "bd"):
        idx = tuple(rng.integers(0, s) for s in p[name].shape)
        q = {k: v.copy() for k, v in p.items()}
        q[name][idx] += eps
        Lp, _ = R.loss_and_grad_fp64(q, X, t, ia, ja, layers=2)
        q[name][idx] -= 2 * eps
        Lm, _ = R.loss_and_grad_fp64(q, X, t, ia, ja, layers=2)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - g[name][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, g[name][idx])


def test_aggregation_adjoint(slice2d):
    ia, ja = _graph(slice2d)
    N = len(ia) - 1
    rng = np.random.default_rng(3)
    H, G = rng.normal(size=(N, 4)), rng.normal(size=(N, 4))
    lhs = np.sum(R._mean_agg(ia, ja, H) * G)
    rhs = np.sum(H * R._mean_agg_T(ia, ja, G))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
