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
    """Replay cases: the slice and the small box with specific storage off
    (S_s = 0 reproduces the reference's capacity, fvmpor.h:943-944) and on
    (S_s = 1e-5 /m, the workloads' value), and the slice with every boundary
    type (Dirichlet, hydrostatic, directional flux, seepage, shore)."""
    from fvmpor_amd.problems import box3d_problem, slice2d_bc_problem, slice2d_problem
    base = which.replace("_ss", "")
    if base == "slice2d":
        p = slice2d_problem(nx=41, ny=17)
        T = 0.1
    elif base == "slice2d_bc":
        p = slice2d_bc_problem()
        T = 0.1
    else:
        p = box3d_problem(nx=12, ny=10, nz=6, rain=0.05)
        T = 0.05
    if not which.endswith("_ss") and base != "slice2d_bc":
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


@pytest.mark.parametrize("which", ["slice2d", "box3d", "slice2d_ss", "box3d_ss", "slice2d_bc"])
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


def test_unpreconditioned_run_matches_oracle_replay():
    """FVP_PRECON_NONE: GMRES on the error-weight-scaled operator alone; the
    Newton updates are taken undamped (the J v operator is formed at the
    current c_j), and the run replays in the oracle like the preconditioned
    ones."""
    from fvmpor_amd.lib import PRECON_NONE
    prob, _ = _problem("slice2d_ss")
    T = 0.02
    psi_g, hs, ks, wfs, st = _gpu_run(prob, T, precon=PRECON_NONE, dim_krylov=40, max_restarts=20)
    assert st["precon_setups"] == 0
    op = O.OracleProblem(prob)
    psi_o, _ = O.integrate_replay(op, prob.psi0, hs, ks, wfs, rtol=1e-6, atol=1e-6, newton_tol=1e-7)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"no preconditioner: steps={len(hs)} rel={rel:.2e} stats={st}")
    assert rel <= 1e-9


def test_slice_run_to_one_day_matches_oracle_replay():
    """The 2D slice workload as specified: 101 x 41 lattice, loam with
    S_s = 1e-5, 0.05 m/day on the strip, global ILU(0), rtol = atol = 1e-8,
    to t = 1 day; the oracle replays the whole step sequence."""
    import torch
    from fvmpor_amd.lib import Context, Integrator, PRECON_ILU0
    from fvmpor_amd.problems import slice2d_problem
    prob = slice2d_problem()
    ctx = Context(prob)
    psi0 = torch.tensor(prob.psi0, device=device())
    integ = Integrator(ctx, psi0, rtol=1e-8, atol=1e-8, precon=PRECON_ILU0, newton_tol=1e-4)
    hs, ks, wfs = [], [], []
    t = 0.0
    while t < 1.0:
        assert integ.step(1.0) == 0
        st = integ.stats()
        hs.append(st["h_last"])
        ks.append(st["order_last"])
        wfs.append(ctx.get_upwind()[0])
        t = st["t"]
    out = torch.empty_like(psi0)
    integ.advance(1.0, out)
    torch.cuda.synchronize()
    psi_g = out.cpu().numpy()
    op = O.OracleProblem(prob)
    psi_o, rs = O.integrate_replay_opts(op, prob.psi0, np.array(hs), np.array(ks, dtype=np.int32), np.array(wfs),
                                        rtol=1e-8, atol=1e-8, newton_tol=1e-5, lin_rel=1e-11, jac_policy=1)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"slice to 1 day: {len(hs)} steps, GPU stats {integ.stats()}, oracle {rs}, rel={rel:.2e}")
    assert rel <= 1e-9


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


def test_ilu0_and_jacobi_runs_agree():
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


@pytest.mark.parametrize("index", [0, 1])
def test_ensemble_draw_instance_matches_oracle_replay(index):
    """Instances of the 256-instance ensemble draw at their drawn sizes (5k to
    60k nodes; index 0 is a 2D slice, 1 a 3D box) with the ensemble's FP32
    block-Jacobi preconditioner: the device run over the first 0.005 day,
    replayed by the oracle."""
    from fvmpor_amd.ensemble import draw_instances, make_instance
    from fvmpor_amd.lib import PRECON_BJILU0_F32
    d = draw_instances()[index]
    prob = make_instance(d)
    assert 5000 <= d.n_nodes <= 60000
    T = 0.005
    psi_g, hs, ks, wfs, st = _gpu_run(prob, T, precon=PRECON_BJILU0_F32)
    op = O.OracleProblem(prob)
    psi_o, rs = O.integrate_replay_opts(op, prob.psi0, hs, ks, wfs, rtol=1e-6, atol=1e-6, newton_tol=1e-7,
                                        lin_rel=1e-11, jac_policy=1)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"draw {index}: dim {d.dim}, N={prob.mesh.n_nodes}, sigma {d.sigma_lnK:.2f}, rain {d.rain:.3f}, "
          f"steps={len(hs)}, rel={rel:.2e}, oracle {rs}")
    assert rel <= 1e-9
