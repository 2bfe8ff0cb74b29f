"""Parity at the full one-GPU size: the 1M-node 3D heterogeneous box
(160 x 160 x 40 jittered lattice, 6.6 M tetrahedra, log-normal Ks, four van
Genuchten classes, S_s = 1e-5 /m, rain 0.02 m/day) against the CPU oracle.

* residual with averaged weights, begin-step upwind decisions, the residual
  with frozen weights, the capacity coefficient and J v: the same bars as the
  small cases (1e-12 of the flux scale, capacity rtol 1e-12, J v 1e-6);
* a 3-step BDF window of the plain FP64 path (block-Jacobi ILU(0), 1024-row
  blocks, one overlap layer, tight Newton) replayed by the oracle: final head
  1e-9;
* 2-step windows of the FP32-preconditioner path and the FP32-preconditioner
  + learned-start path at the integrator's default Newton tolerance, judged
  against the oracle's replay of their own step sequences at 1e-6.

The mesh, the oracle's CVFE dual and the device context are built once per
module (about two minutes on the host)."""
import numpy as np
import pytest

from conftest import device
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
STEPS = 3


@pytest.fixture(scope="module")
def big():
    import time

    import torch
    from fvmpor_amd.lib import Context
    from fvmpor_amd.problems import box3d_problem
    device()
    t0 = time.perf_counter()
    prob = box3d_problem()
    print(f"\nmesh N={prob.mesh.n_nodes} T={prob.mesh.n_elements} in {time.perf_counter() - t0:.1f}s", flush=True)
    t0 = time.perf_counter()
    op = O.OracleProblem(prob)
    print(f"oracle dual in {time.perf_counter() - t0:.1f}s", flush=True)
    t0 = time.perf_counter()
    ctx = Context(prob)
    torch.cuda.synchronize()
    print(f"device context in {time.perf_counter() - t0:.1f}s", flush=True)
    yield prob, op, ctx
    ctx.close()


def _state(prob, seed):
    rng = np.random.default_rng(seed)
    return prob.psi0 + rng.uniform(-0.4, 0.4, prob.mesh.n_nodes), rng.normal(scale=0.1, size=prob.mesh.n_nodes)


def test_residual_upwind_capacity_and_jv(big):
    import torch
    prob, op, ctx = big
    dev = device()
    psi, psip = _state(prob, 11)
    tp, tpp = torch.tensor(psi, device=dev), torch.tensor(psip, device=dev)
    F = torch.empty_like(tp)
    # averaged weights (a fresh problem): both sides use 1/2 on every edge
    ctx.set_upwind(np.full(ctx.NE, 0.5))
    op.set_weights(np.full(ctx.NE, 0.5))
    ctx.residual(0.0, tp, tpp, F)
    torch.cuda.synchronize()
    Fc = op.residual(0.0, psi, psip)
    s = op.flux_scale()
    rel0 = np.abs(F.cpu().numpy() - Fc).max() / max(np.abs(Fc).max(), s)
    # upwind decisions of begin_step
    ctx.begin_step(0.0, tp, tpp)
    wf_g, fl_g = ctx.get_upwind()
    op.begin_step(0.0, psi, psip)
    wf_c, fl_c = op.weights()
    big_flux = np.abs(fl_c) > 1e-9 * np.abs(fl_c).max()
    fl_err = np.abs(fl_g - fl_c).max() / np.abs(fl_c).max()
    n_diff = int(np.count_nonzero(wf_g[big_flux] != wf_c[big_flux]))
    # frozen weights: the residual, the capacity and J v
    op.set_weights(wf_g)
    ctx.residual(0.25, tp, tpp, F)
    ahh = torch.empty_like(tp)
    ctx.capacity(ahh)
    v = np.random.default_rng(12).normal(size=len(psi))
    tv = torch.tensor(v, device=dev)
    Jv = torch.empty_like(tp)
    cj = 40.0
    ctx.jv(0.25, tp, tpp, F, cj, tv, Jv)
    torch.cuda.synchronize()
    Fc = op.residual(0.25, psi, psip)
    rel1 = np.abs(F.cpu().numpy() - Fc).max() / max(np.abs(Fc).max(), s)
    a_c, _, _ = op.node_fields()
    a_g = ahh.cpu().numpy()
    cap_err = np.abs(a_g - a_c).max() / np.abs(a_c).max()
    cap_rel = np.abs(a_g - a_c) <= 1e-12 * np.abs(a_c) + 1e-14 * np.abs(a_c).max()
    sig = np.sqrt(2.0 ** -52) * np.linalg.norm(psi) / np.linalg.norm(v)
    Jc = (op.residual(0.25, psi + sig * v, psip + cj * (sig * v)) - Fc) / sig
    jv_err = np.abs(Jv.cpu().numpy() - Jc).max() / np.abs(Jc).max()
    print(f"1M box: residual (averaged) {rel0:.2e}, edge flux {fl_err:.2e}, upwind decisions differing "
          f"{n_diff} of {int(big_flux.sum())}, residual (upwind) {rel1:.2e}, capacity {cap_err:.2e}, J v {jv_err:.2e}")
    assert rel0 <= 1e-12
    assert fl_err <= 1e-11
    assert n_diff == 0
    assert rel1 <= 1e-12
    assert cap_rel.all()
    assert jv_err < 1e-6


def _window(prob, ctx, precon, steps, newton_tol=None, learned=None):
    """`steps` accepted steps from t = 0 with the bench's solver settings;
    returns the heads, step sizes, orders, per-step upwind weights and stats."""
    import torch
    from fvmpor_amd.lib import Integrator
    psi0 = torch.tensor(prob.psi0, device=device())
    kw = {} if newton_tol is None else {"newton_tol": newton_tol}
    integ = Integrator(ctx, psi0, rtol=1e-6, atol=1e-6, max_order=3, precon=precon, block_rows=1024, overlap=1,
                       **kw)
    if learned is not None:
        integ.set_learned(*learned)
    hs, ks, wfs = [], [], []
    for _ in range(steps):
        assert integ.step(1.0) == 0
        st = integ.stats()
        hs.append(st["h_last"])
        ks.append(st["order_last"])
        wfs.append(ctx.get_upwind()[0].astype(np.float32))
    out = torch.empty_like(psi0)
    integ.advance(st["t"], out)
    torch.cuda.synchronize()
    st = integ.stats()
    integ.close()
    return out.cpu().numpy(), np.array(hs), np.array(ks, dtype=np.int32), np.array(wfs, dtype=np.float64), st


def _replay(prob, op, hs, ks, wfs):
    """The oracle's BDF solution of the step sequence: one FD Jacobian and
    ILU(0) per step, Newton to 1e-7 in the weighted RMS norm with GMRES to
    1e-6 relative per Newton iteration (the FD J v noise sits near 1e-8)."""
    return O.integrate_replay_opts(op, prob.psi0, hs, ks, wfs, rtol=1e-6, atol=1e-6, newton_tol=1e-7,
                                   lin_rel=1e-6, jac_policy=1, dim_krylov=20, verbose=True)


def test_plain_fp64_window_replays(big):
    from fvmpor_amd.lib import PRECON_BJILU0
    prob, op, ctx = big
    psi_g, hs, ks, wfs, st = _window(prob, ctx, PRECON_BJILU0, STEPS, newton_tol=1e-6)
    psi_o, rs = _replay(prob, op, hs, ks, wfs)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"plain FP64: h={hs}, orders={ks}, rel={rel:.2e}, gpu {st}, oracle {rs}")
    assert rel <= 1e-9


def test_fp32_preconditioner_window_replays(big):
    """The FP32 preconditioner path as benchmarked (default Newton
    tolerance): head error and mass change against the oracle's replay."""
    from fvmpor_amd.diagnostics import total_mass
    from fvmpor_amd.lib import PRECON_BJILU0_F32
    prob, op, ctx = big
    psi_g, hs, ks, wfs, st = _window(prob, ctx, PRECON_BJILU0_F32, STEPS - 1)
    psi_o, rs = _replay(prob, op, hs, ks, wfs)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    m0 = total_mass(prob, prob.psi0)
    dm = (total_mass(prob, psi_g) - total_mass(prob, psi_o)) / max(abs(total_mass(prob, psi_o) - m0), 1e-300)
    print(f"FP32 preconditioner: h={hs}, rel={rel:.2e}, mass change vs oracle {dm:.2e}, gpu {st}, oracle {rs}")
    assert rel <= 1e-6
    assert abs(dm) <= 1e-4


def test_learned_start_window_replays(big):
    """FP32 preconditioner + learned Newton start (a network trained briefly
    on a fresh-seed 16k-node box by the library's kernels), judged against
    the oracle's replay of its step sequence."""
    import torch
    from fvmpor_amd.learn import node_graph, static_features, train_on
    from fvmpor_amd.lib import PRECON_BJILU0_F32, Mpnn
    from fvmpor_amd.problems import box3d_problem
    prob, op, ctx = big
    tp = box3d_problem(nx=40, ny=40, nz=10, seed_mesh=611, seed_field=612, seed_layers=613)
    params, hist, npairs, _ = train_on([tp], 0.05, epochs=10, lr=1e-3)
    ia, ja = node_graph(prob.mesh)
    model = Mpnn(ia, ja, layers=3)
    model.set_params(params)
    stat = torch.tensor(static_features(prob), device=device())
    torch.cuda.synchronize()
    psi_g, hs, ks, wfs, st = _window(prob, ctx, PRECON_BJILU0_F32, STEPS - 1, learned=(model, stat))
    psi_o, rs = _replay(prob, op, hs, ks, wfs)
    rel = np.abs(psi_g - psi_o).max() / np.abs(psi_o).max()
    print(f"learned start ({npairs} pairs, loss {hist[0]:.2e} -> {hist[-1]:.2e}): h={hs}, rel={rel:.2e}, "
          f"network start {st['learned_used']}/{st['learned_tried']}, gpu {st}, oracle {rs}")
    assert st["learned_tried"] > 0
    assert rel <= 1e-6
    model.close()
