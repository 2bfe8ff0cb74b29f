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
