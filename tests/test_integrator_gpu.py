"""GPU time integration against the CPU oracle: the adaptive BDF run on the
device is replayed by the oracle over the same step sizes, orders and upwind
weights, and the final head fields are compared (bar: 1e-9 relative), plus
the physical mass balance of the device run."""
import numpy as np
import pytest

from conftest import device
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _problem(which):
    from fvmpor_amd.problems import box3d_problem, slice2d_problem
    if which == "slice2d":
        p = slice2d_problem(nx=41, ny=17)
        T = 0.1
    else:
        p = box3d_problem(nx=12, ny=10, nz=6, rain=0.05)
        T = 0.05
    p.zones = p.zones.copy()
    p.zones[:, 7] = 0.0
    return p, T


def _gpu_run(prob, T, **kw):
    import torch
    from fvmpor_amd.lib import Context, Integrator
    ctx = Context(prob)
    psi0 = torch.tensor(prob.psi0, device=device())
    integ = Integrator(ctx, psi0, rtol=1e-6, atol=1e-6, newton_tol=1e-6, **kw)
    hs, ks, wfs = [], [], []
    t = 0.0
    while t < T:
        rc = integ.step(T)
        assert rc == 0
        st = integ.stats()
        hs.append(st["h_last"])
        ks.append(st["order_last"])
        wfs.append(ctx.get_upwind()[0])
        t = st["t"]
    out = torch.empty_like(psi0)
    integ.advance(T, out)
    torch.cuda.synchronize()
    return out.cpu().numpy(), np.array(hs), np.array(ks, dtype=np.int32), np.array(wfs), integ.stats()


@pytest.mark.parametrize("which", ["slice2d", "box3d"])
def test_final_head_matches_oracle_replay(which):
    prob, T = _problem(which)
    psi_g, hs, ks, wfs, st = _gpu_run(prob, T)
    assert abs(hs.sum() - T) < 1e-12 * T
    op = O.OracleProblem(prob)
    psi_o, nit = O.integrate_replay(op, prob.psi0, hs, ks, wfs, rtol=1e-6, atol=1e-6, newton_tol=1e-7)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"{which}: steps={len(hs)} orders={np.bincount(ks)} rel={rel:.2e} stats={st}")
    assert rel <= 1e-9
    assert set(ks.tolist()) <= {1, 2, 3}


def test_mass_balance_of_device_run():
    prob, T = _problem("slice2d")
    psi_g, hs, ks, wfs, st = _gpu_run(prob, T)
    op = O.OracleProblem(prob)
    V, _ = op.omesh.volumes()
    Sw = lambda psi: np.array([O.vg(h, prob.zones[0])[0] for h in psi])
    gained = (V * prob.zones[0, 3] * (Sw(psi_g) - Sw(prob.psi0))).sum()
    from fvmpor_amd.meshgen import TRI_FACES
    m = prob.mesh
    t, f = np.nonzero(m.face_tags == 2)
    n = m.conn[t][np.arange(len(t))[:, None], TRI_FACES[f]]
    strip = np.linalg.norm(m.xyz[n[:, 0]] - m.xyz[n[:, 1]], axis=1).sum()
    inflow = 0.05 * strip * T
    assert abs(gained - inflow) < 1e-3 * inflow, (gained, inflow)


def test_jacobi_and_unpreconditioned_runs_agree():
    from fvmpor_amd.lib import PRECON_JACOBI
    prob, T = _problem("slice2d")
    a, *_ = _gpu_run(prob, 0.02)
    b, *_ = _gpu_run(prob, 0.02, precon=PRECON_JACOBI)
    assert np.abs(a - b).max() < 1e-5 * np.abs(a).max()


def test_block_jacobi_run_agrees():
    from fvmpor_amd.lib import PRECON_BJILU0
    from fvmpor_amd.problems import box3d_problem
    prob = box3d_problem(nx=22, ny=20, nz=10, rain=0.05)
    a, *_ = _gpu_run(prob, 0.01)
    b, *_ = _gpu_run(prob, 0.01, precon=PRECON_BJILU0)
    assert np.abs(a - b).max() < 1e-5 * np.abs(a).max()


def test_fp32_block_jacobi_run_agrees():
    """FP32 preconditioner (factor and solves) against the FP64 one on the same
    run: the residual, J v, GMRES and error control stay FP64, so the final
    heads agree to the Newton / linear tolerances."""
    from fvmpor_amd.lib import PRECON_BJILU0, PRECON_BJILU0_F32
    from fvmpor_amd.problems import box3d_problem
    prob = box3d_problem(nx=22, ny=20, nz=10, rain=0.05)
    a, _, _, _, sa = _gpu_run(prob, 0.01, precon=PRECON_BJILU0)
    b, _, _, _, sb = _gpu_run(prob, 0.01, precon=PRECON_BJILU0_F32)
    rel = np.abs(a - b).max() / np.abs(a).max()
    print(f"fp32 vs fp64 block-Jacobi: final head rel diff {rel:.2e}, gmres iters {sb['gmres_iters']} vs {sa['gmres_iters']}")
    assert rel < 1e-6
