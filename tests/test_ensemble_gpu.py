"""Batched ensembles on the device: a batch advanced as one block-diagonal
problem with per-instance norms agrees with each instance integrated alone
(both at rtol = atol = 1e-6; the batch's shared step is never longer than
what each instance's own error test allows, so the two differ by the
integration tolerance only)."""
import numpy as np
import pytest

from conftest import device
from fvmpor_amd.ensemble import Draw, make_instance, run_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_instances_run_alone(dim):
    import torch
    from fvmpor_amd.lib import Context, Integrator, PRECON_BJILU0_F32
    device()
    params = [(700, 0.6, 0.02), (1500, 1.9, 0.09), (1000, 1.2, 0.05)]
    probs = [make_instance(Draw(i, dim, n, s, r, 1000 + 2 * i + (dim == 3))) for i, (n, s, r) in enumerate(params)]
    T = 0.02
    heads, st, wall = run_batch(probs, t_end=T)
    print(f"dim {dim}: batch steps {st['steps']}, newton {st['newton_iters']}, gmres {st['gmres_iters']}, {wall:.2f}s")
    for p, h in zip(probs, heads):
        ctx = Context(p)
        psi0 = torch.tensor(p.psi0, device=device())
        integ = Integrator(ctx, psi0, rtol=1e-6, atol=1e-6, max_order=3, precon=PRECON_BJILU0_F32)
        out = torch.empty_like(psi0)
        integ.advance(T, out)
        torch.cuda.synchronize()
        alone = out.cpu().numpy()
        rel = np.abs(h - alone).max() / (np.ptp(alone) + 1e-30)
        print(f"  {p.name}: N={p.mesh.n_nodes}, alone steps {integ.stats()['steps']}, max diff / range {rel:.2e}")
        assert rel < 1e-4


def test_segments_are_validated():
    import torch
    from fvmpor_amd.lib import Context, FvpError, Integrator
    p = make_instance(Draw(0, 2, 600, 1.0, 0.05, 1000))
    ctx = Context(p)
    integ = Integrator(ctx, torch.tensor(p.psi0, device=device()))
    N = p.mesh.n_nodes
    for bad in ([0, N // 2, N // 2, N], [1, N], [0, N - 1]):
        with pytest.raises(FvpError):
            integ.set_segments(bad)
    integ.set_segments([0, N // 3, N])


def test_concurrent_instances_match_sequential_runs():
    """Instances advanced concurrently (one context and HIP stream each, four
    in flight) give the same heads as the same instances run one after the
    other."""
    from fvmpor_amd.ensemble import run_concurrent, run_one
    device()
    params = [(2, 900, 0.7, 0.03), (3, 1200, 1.1, 0.05), (2, 1500, 1.6, 0.08), (3, 800, 0.9, 0.02),
              (2, 600, 1.3, 0.06)]
    probs = [make_instance(Draw(i, d, n, s, r, 1000 + 2 * i + (d == 3))) for i, (d, n, s, r) in enumerate(params)]
    T = 0.02
    res, wall = run_concurrent(probs, t_end=T, streams=4)
    for p, r in zip(probs, res):
        alone = run_one(p, t_end=T)
        assert r["status"] == alone["status"] == "done" and r["t"] == pytest.approx(T)
        rel = np.abs(r["heads"] - alone["heads"]).max() / (np.ptp(alone["heads"]) + 1e-30)
        print(f"  {p.name}: N={p.mesh.n_nodes}, steps {r['steps']} / {alone['steps']}, max diff / range {rel:.2e}")
        assert rel < 1e-8
