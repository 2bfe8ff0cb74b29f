"""GPU residual parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: rel = ||F_gpu - F_cpu||_inf / max(||F_cpu||_inf, s) <= 1e-12 with
s = max_i (1/V_i) sum_f |M_f| (the flux scale; F ~ 0 near equilibrium).
Upwind weights: identical wherever the oracle's averaged edge flux is above
rounding noise; the residual comparison then shares the GPU's weights so that
ties (|edge flux| at rounding level, e.g. hydrostatic zones) cannot differ.
"""
import copy

import numpy as np
import pytest

from conftest import device
from oracle.oracle import OracleProblem

pytestmark = pytest.mark.gpu
TOL = 1e-12
# slice2d_bc / box3d_bc carry boundary types 1, 3, 4, 6, 7 and 8 (definitions.h:114-146)
CASES = ["cassion", "slice2d", "box3d_small", "slice2d_bc", "box3d_bc"]


def _state(prob, seed):
    rng = np.random.default_rng(seed)
    psi = prob.psi0 + rng.uniform(-0.4, 0.4, prob.mesh.n_nodes)
    psip = rng.normal(scale=0.1, size=prob.mesh.n_nodes)
    return psi, psip


def _gpu(prob):
    import torch
    from fvmpor_amd.lib import Context
    dev = device()
    ctx = Context(prob)
    return ctx, dev, torch


def _rel(Fg, Fc, s):
    return np.abs(Fg - Fc).max() / max(np.abs(Fc).max(), s)


@pytest.mark.parametrize("which", CASES)
def test_residual_averaged_weights(which, request):
    prob = request.getfixturevalue(which)
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 1)
    op = OracleProblem(prob)
    Fc = op.residual(0.0, psi, psip)  # fresh problems use weights 0.5
    tp = torch.tensor(psi, device=dev)
    tpp = torch.tensor(psip, device=dev)
    F = torch.empty_like(tp)
    ctx.residual(0.0, tp, tpp, F)
    torch.cuda.synchronize()
    rel = _rel(F.cpu().numpy(), Fc, op.flux_scale())
    print(f"{which} averaged rel={rel:.3e}")
    assert rel <= TOL


@pytest.mark.parametrize("which", CASES)
def test_begin_step_upwind_and_residual(which, request):
    prob = request.getfixturevalue(which)
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 2)
    op = OracleProblem(prob)
    tp = torch.tensor(psi, device=dev)
    tpp = torch.tensor(psip, device=dev)
    ctx.begin_step(0.0, tp, tpp)
    wf_g, fl_g = ctx.get_upwind()
    op.begin_step(0.0, psi, psip)
    wf_c, fl_c = op.weights()
    # edge fluxes agree to rounding, decisions agree away from ties
    big = np.abs(fl_c) > 1e-9 * np.abs(fl_c).max()
    assert np.abs(fl_g - fl_c).max() <= 1e-11 * np.abs(fl_c).max()
    assert np.array_equal(wf_g[big], wf_c[big])
    assert set(np.unique(wf_g)) <= {0.0, 1.0}
    op.set_weights(wf_g)
    Fc = op.residual(0.5, psi, psip)
    F = torch.empty_like(tp)
    ctx.residual(0.5, tp, tpp, F)
    torch.cuda.synchronize()
    rel = _rel(F.cpu().numpy(), Fc, op.flux_scale())
    print(f"{which} upwind rel={rel:.3e}")
    assert rel <= TOL
    # capacity coefficient matches the oracle's ahh
    ahh = torch.empty_like(tp)
    ctx.capacity(ahh)
    torch.cuda.synchronize()
    a_c, _, _ = op.node_fields()
    np.testing.assert_allclose(ahh.cpu().numpy(), a_c, rtol=1e-12, atol=1e-14 * np.abs(a_c).max())


def test_set_upwind_roundtrip_and_bc_update(slice2d):
    prob = copy.copy(slice2d)
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 3)
    rng = np.random.default_rng(4)
    wf = rng.integers(0, 2, ctx.NE).astype(np.float64)
    ctx.set_upwind(wf)
    got, _ = ctx.get_upwind()
    assert np.array_equal(got, wf)
    bc = prob.bc.copy()
    bc[2, 0] = -0.2
    ctx.set_bc(bc)
    prob.bc = bc
    op = OracleProblem(prob)
    op.set_weights(wf)
    Fc = op.residual(0.0, psi, psip)
    tp = torch.tensor(psi, device=dev)
    tpp = torch.tensor(psip, device=dev)
    F = torch.empty_like(tp)
    ctx.residual(0.0, tp, tpp, F)
    torch.cuda.synchronize()
    assert _rel(F.cpu().numpy(), Fc, op.flux_scale()) <= TOL


@pytest.mark.parametrize("which", ["slice2d", "box3d_small", "slice2d_bc", "box3d_bc"])
def test_jv_matches_oracle_difference_quotient(which, request):
    prob = request.getfixturevalue(which)
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 5)
    rng = np.random.default_rng(6)
    v = rng.normal(size=prob.mesh.n_nodes)
    cj = 3.7
    op = OracleProblem(prob)
    tp, tpp, tv = (torch.tensor(a, device=dev) for a in (psi, psip, v))
    ctx.begin_step(0.0, tp, tpp)
    wf, _ = ctx.get_upwind()
    op.set_weights(wf)
    F0 = torch.empty_like(tp)
    ctx.residual(0.1, tp, tpp, F0)
    Jv = torch.empty_like(tp)
    ctx.jv(0.1, tp, tpp, F0, cj, tv, Jv)
    torch.cuda.synchronize()
    sig = np.sqrt(2.0 ** -52) * np.linalg.norm(psi) / np.linalg.norm(v)
    F0c = op.residual(0.1, psi, psip)
    F1c = op.residual(0.1, psi + sig * v, psip + cj * (sig * v))
    Jc = (F1c - F0c) / sig
    err = np.abs(Jv.cpu().numpy() - Jc).max() / np.abs(Jc).max()
    print(f"{which} Jv rel={err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("which", CASES)
def test_every_launch_shape_and_the_autotuner_keep_parity(which, request):
    """Both residual variants (element-owner with atomics, LDS window), several
    tile sizes and both tile-to-XCD mappings: residual and J v against the
    oracle; then the autotuner's pick."""
    from fvmpor_amd.lib import RESIDUAL_OWNER, RESIDUAL_WINDOW
    prob = request.getfixturevalue(which)
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 3)
    op = OracleProblem(prob)
    tp, tpp = torch.tensor(psi, device=dev), torch.tensor(psip, device=dev)
    ctx.begin_step(0.0, tp, tpp)
    wf, _ = ctx.get_upwind()
    op.set_weights(wf)
    Fc = op.residual(0.2, psi, psip)
    s = op.flux_scale()
    v = np.random.default_rng(9).normal(size=len(psi))
    cj = 2.5
    sig = np.sqrt(2.0 ** -52) * np.linalg.norm(psi) / np.linalg.norm(v)
    Jc = (op.residual(0.2, psi + sig * v, psip + cj * (sig * v)) - Fc) / sig
    tv = torch.tensor(v, device=dev)
    F, Jv = torch.empty_like(tp), torch.empty_like(tp)
    shapes = [(RESIDUAL_OWNER, t, x) for t in (64, 256, 512, 1024) for x in (0, 1)] + [(RESIDUAL_WINDOW, 512, 1)]

    def check(label):
        ctx.residual(0.2, tp, tpp, F)
        ctx.jv(0.2, tp, tpp, F, cj, tv, Jv)
        torch.cuda.synchronize()
        rel = _rel(F.cpu().numpy(), Fc, s)
        jrel = np.abs(Jv.cpu().numpy() - Jc).max() / np.abs(Jc).max()
        assert rel <= TOL, (label, rel)
        assert jrel < 1e-6, (label, jrel)
        return rel

    for shape in shapes:
        ctx.set_residual_launch(*shape)
        check(shape)
    pick = ctx.autotune(0.0, tp, tpp)
    rel = check(pick)
    print(f"{which}: autotuner picked {pick}, rel={rel:.2e}")


@pytest.mark.parametrize("which", ["slice2d_bc", "box3d_bc"])
def test_set_bc_updates_every_boundary_type(which, request):
    """New values for every tag, Dirichlet (1) and hydrostatic (4) rows
    included, through fvp_set_bc: the residual follows the oracle built with
    the new rows."""
    prob = copy.copy(request.getfixturevalue(which))
    ctx, dev, torch = _gpu(prob)
    psi, psip = _state(prob, 7)
    bc = prob.bc.copy()
    bc[:, 0] = bc[:, 0] * 1.7 - 0.3
    bc[:, 1] += 0.75
    bc[:, 2] = 0.1
    bc[:, 3:6] = bc[:, 3:6] * -2.0 + 0.01
    ctx.set_bc(bc)
    prob.bc = bc
    op = OracleProblem(prob)
    Fc = op.residual(0.0, psi, psip)
    tp, tpp = torch.tensor(psi, device=dev), torch.tensor(psip, device=dev)
    F = torch.empty_like(tp)
    ctx.residual(0.0, tp, tpp, F)
    torch.cuda.synchronize()
    Fg = F.cpu().numpy()
    assert ctx.ndir > 0
    assert _rel(Fg, Fc, op.flux_scale()) <= TOL
