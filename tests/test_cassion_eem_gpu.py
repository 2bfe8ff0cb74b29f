"""End-to-end check against the reference's own output: the cassion problem
(cassion.h, cassion_1_0.pmesh) integrated on the device to the twelve output
times of the reference's EEM run (examples/EEM/CASS_dom0_{0..11}.vtu, heads
stored as Float32, EEM integrator at rtol = atol = 5e-4, fvmpor_ODE.cpp:29-30).

The reference trajectory carries its own time-integration error at that
tolerance, so this is a tolerance-level pin: the device run (BDF at a tight
tolerance) must agree with it to 1e-2 of the head range everywhere and 1e-3 in
the field-wide RMS at every output time (the CPU oracle integrated with scipy
plateaus at 5e-3 / 5e-4, the reference run's own error).  Nodes are matched by coordinate
(the VTU points are in the reference's post-RCM order)."""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

from conftest import GOLDEN, device

pytestmark = pytest.mark.gpu


def test_cassion_heads_follow_reference_eem_trajectory(cassion):
    import torch
    from fvmpor_amd.lib import Context, Integrator, PRECON_ILU0
    ref = np.load(os.path.join(GOLDEN, "cassion_eem_vtu.npz"))
    xyz = cassion.mesh.xyz
    dist, idx = cKDTree(xyz[:, :2]).query(ref["points"][:, :2].astype(np.float64))
    assert dist.max() < 1e-4 and len(np.unique(idx)) == len(idx) == cassion.mesh.n_nodes
    ctx = Context(cassion)
    psi0 = torch.tensor(cassion.psi0, device=device())
    integ = Integrator(ctx, psi0, rtol=1e-6, atol=1e-6, precon=PRECON_ILU0)
    out = torch.empty_like(psi0)
    heads = ref["heads"].astype(np.float64)
    span = heads.max() - heads.min()
    worst = []
    for k, t in enumerate(ref["times"]):
        if t > 0:
            assert integ.advance(float(t), out) == 0
        else:
            out.copy_(psi0)
        torch.cuda.synchronize()
        h = out.cpu().numpy()[idx]
        d = h - heads[k]
        worst.append((np.abs(d).max() / span, np.sqrt(np.mean(d * d)) / span))
    st = integ.stats()
    for k, (mx, rms) in enumerate(worst):
        print(f"t[{k}]: max |dh| / range = {mx:.2e}, rms / range = {rms:.2e}")
    print(f"steps {st['steps']}, newton {st['newton_iters']}, gmres {st['gmres_iters']}")
    assert max(w[0] for w in worst) < 1e-2
    assert max(w[1] for w in worst) < 1e-3
