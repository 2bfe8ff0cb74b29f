"""Device network kernels (fvp_mpnn_*) against the CPU reference
oracle/mpnn_ref.py: forward outputs, the loss and every parameter gradient,
with the reference computing in float64 on the same BF16-rounded operands
(the device accumulates in FP32, so a BF16 rounding of an activation can land
one unit apart; tolerances below), then the Adam update against its
closed-form first step and training against a learnable target."""
import numpy as np
import pytest

from conftest import device
from oracle import mpnn_ref as R

pytestmark = pytest.mark.gpu


def _graph(prob):
    from fvmpor_amd.learn import node_graph
    return node_graph(prob.mesh)


def _params(layers, seed):
    rng = np.random.default_rng(seed + 100)
    p = R.init_params(layers=layers, seed=seed)
    for k in p:
        if k.startswith("b"):
            p[k] = p[k] + rng.normal(scale=0.1, size=p[k].shape)
    # the device keeps FP32 master weights
    return {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}


@pytest.fixture(scope="module")
def graphs(box3d_small):
    from fvmpor_amd.problems import slice2d_problem
    return {"slice": _graph(slice2d_problem(nx=21, ny=9)), "box": _graph(box3d_small)}


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


@pytest.mark.parametrize("name,layers", [("slice", 1), ("slice", 3), ("box", 3)])
def test_forward_loss_and_gradients_match_reference(graphs, name, layers):
    import torch
    from fvmpor_amd.lib import Mpnn, mpnn_pack, mpnn_unpack
    dev = device()
    ia, ja = graphs[name]
    N = len(ia) - 1
    rng = np.random.default_rng(7)
    X = rng.normal(size=(N, R.F_IN)).astype(np.float32)
    t = rng.normal(size=N).astype(np.float32)
    p = _params(layers, 3)
    m = Mpnn(ia, ja, layers=layers, seed=1)
    m.set_params(mpnn_pack(p, layers))
    back = mpnn_unpack(m.get_params(), layers)
    assert all(np.array_equal(back[k], p[k]) for k in p)
    xd, td = torch.tensor(X, device=dev), torch.tensor(t, device=dev)
    y = torch.empty(N, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    m.forward(xd, y)
    m.synchronize()
    yd = y.cpu().numpy().astype(np.float64)
    yr, cache = R.forward(p, X.astype(np.float64), ia, ja, layers)
    ey = np.abs(yd - yr)
    print(f"{name} L={layers} N={N}: forward max |dy| {ey.max():.2e} (|y| max {np.abs(yr).max():.2e}), "
          f"rms {np.sqrt(np.mean(ey ** 2)):.2e}")
    # BF16 activations: one-unit rounding flips (2^-8 relative) on a few values
    assert ey.max() <= 1e-2 * np.abs(yr).max()
    assert np.sqrt(np.mean(ey ** 2)) <= 2e-3 * np.sqrt(np.mean(yr ** 2))
    loss = m.loss_grad(xd, td)
    gd = mpnn_unpack(m.get_grads(), layers)
    lr_ = R.loss(yr, t.astype(np.float64))
    gr = R.backward(p, cache, yr, t.astype(np.float64), ia, ja, layers)
    _, g64 = R.loss_and_grad_fp64(p, X.astype(np.float64), t.astype(np.float64), ia, ja, layers)
    assert abs(loss - lr_) <= 1e-3 * lr_
    worst = 0.0
    for k in gr:
        e = _rel(gd[k], gr[k])
        e64 = _rel(gd[k], g64[k])
        worst = max(worst, e)
        print(f"  grad {k:6s}: rel to BF16 reference {e:.2e}, to exact float64 {e64:.2e}")
        assert e <= 2e-2, k
        assert e64 <= 1e-1, k
    m.close()


def test_adam_first_step_is_signed_learning_rate(graphs):
    import torch
    from fvmpor_amd.lib import Mpnn
    dev = device()
    ia, ja = graphs["slice"]
    N = len(ia) - 1
    rng = np.random.default_rng(2)
    m = Mpnn(ia, ja, layers=2, seed=5)
    p0 = m.get_params()
    xd = torch.tensor(rng.normal(size=(N, 16)).astype(np.float32), device=dev)
    td = torch.tensor(rng.normal(size=N).astype(np.float32), device=dev)
    torch.cuda.synchronize()
    m.loss_grad(xd, td)
    g = m.get_grads()
    lr = 1e-3
    m.adam(lr)
    p1 = m.get_params()
    # first step: m_hat = g, v_hat = g^2, so p1 = p0 - lr g / (|g| + eps)
    expect = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - expect).max() <= 2e-6 * max(1.0, np.abs(p0).max())
    m.close()


def test_training_fits_a_learnable_target(graphs):
    import torch
    from fvmpor_amd.lib import Mpnn
    dev = device()
    ia, ja = graphs["box"]
    N = len(ia) - 1
    rng = np.random.default_rng(4)
    X = rng.normal(size=(N, 16)).astype(np.float32)
    # a target the architecture can represent: a neighbour mean plus a local nonlinearity
    deg = np.diff(ia)
    rows = np.repeat(np.arange(N), deg)
    nb_mean = np.bincount(rows, weights=X[ja, 0], minlength=N) / np.maximum(deg, 1)
    t = (nb_mean + 0.5 * np.maximum(X[:, 1], 0) - 0.3 * X[:, 2]).astype(np.float32)
    m = Mpnn(ia, ja, layers=3, seed=9)
    xd, td = torch.tensor(X, device=dev), torch.tensor(t, device=dev)
    torch.cuda.synchronize()
    losses = []
    for it in range(400):
        losses.append(m.loss_grad(xd, td))
        m.adam(3e-3)
    print(f"training loss {losses[0]:.3e} -> {losses[-1]:.3e} (target variance {t.var():.3e})")
    assert losses[-1] < 0.05 * losses[0]
    m.close()
