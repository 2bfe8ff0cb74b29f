"""Pins of the oracle's residual (oracle/orc_phys.c): van Genuchten closed
forms, hydrostatic equilibrium, global conservation, linear saturated flow."""
import copy

import numpy as np
import pytest

from oracle.oracle import OracleProblem, vg


ZONE = np.array([0.25, 0.25, 0.25, 0.43, 3.6, 1.56, 0.078 / 0.43, 1e-5])


def _closed_form(h, zp):
    # scalar twins of the reference (fvmpor.cpp:23-74), a different algebraic form
    alpha, n, Sr = zp[4], zp[5], zp[6]
    m = 1 - 1 / n
    if h >= 0:
        return 1.0, 0.0, 1.0
    Se = 1.0 / (1.0 + (alpha * abs(h)) ** n) ** m
    Sw = Sr + (1 - Sr) * Se
    dSw = alpha ** n * (1 - Sr) * (n - 1) * abs(h) ** (n - 1) / (1 + (alpha * abs(h)) ** n) ** (m + 1)
    kr = np.sqrt(Se) * (1 - (1 - Se ** (1 / m)) ** m) ** 2
    return Sw, dSw, kr


@pytest.mark.parametrize("h", [-1e-3, -0.05, -0.3, -1.0, -7.34, -40.0])
def test_vg_matches_closed_forms(h):
    Sw, dSw, kr = vg(h, ZONE)
    cSw, cdSw, ckr = _closed_form(h, ZONE)
    assert abs(Sw - cSw) < 1e-14
    assert abs(dSw - cdSw) <= 1e-12 * abs(cdSw)
    assert abs(kr - ckr) <= 1e-9 * abs(ckr) + 1e-16


def test_vg_special_points():
    assert vg(0.0, ZONE) == (1.0, 0.0, 1.0)
    assert vg(3.5, ZONE) == (1.0, 0.0, 1.0)
    alpha, n, Sr = ZONE[4], ZONE[5], ZONE[6]
    m = 1 - 1 / n
    Sw, dSw, kr = vg(-1.0 / alpha, ZONE)  # (alpha|h|)^n = 1 => Se = 2^-m
    Se = 2.0 ** -m
    assert abs(Sw - (Sr + (1 - Sr) * Se)) < 1e-15
    assert abs(kr - np.sqrt(Se) * (1 - 0.5 ** m) ** 2) < 1e-15


def test_vg_derivative_by_finite_difference():
    for h in [-0.01, -0.2, -1.3, -5.0]:
        e = 1e-6 * abs(h)
        fd = (vg(h + e, ZONE)[0] - vg(h - e, ZONE)[0]) / (2 * e)
        assert abs(fd - vg(h, ZONE)[1]) <= 1e-6 * abs(fd)


def _no_flow(problem):
    p = copy.copy(problem)
    p.bc = problem.bc.copy()
    p.bc[:, 0] = 0.0
    return p


def _gravity_scale(op, problem):
    # max over nodes of rho0 * K * (sum of |face area|) / V: the flux size of a unit head gradient
    V, _ = op.omesh.volumes()
    ints, dbl = op.omesh.cvfaces()
    s = np.zeros(op.N)
    np.add.at(s, ints[:, 3], dbl[:, 0])
    inter = ints[:, 2] >= 0
    np.add.at(s, ints[inter, 2], dbl[inter, 0])
    K = problem.zones[:, 0].max() * problem.ekmul.max()
    return problem.rho0 * K * (s / V).max()


@pytest.mark.parametrize("which", ["slice2d", "box3d_small"])
def test_hydrostatic_equilibrium_is_steady(which, request):
    prob = _no_flow(request.getfixturevalue(which))
    op = OracleProblem(prob)
    psi = prob.psi0
    op.begin_step(0.0, psi, 0 * psi)
    F = op.residual(0.0, psi, 0 * psi)
    assert np.abs(F).max() < 1e-12 * _gravity_scale(op, prob)


@pytest.mark.parametrize("which", ["cassion", "slice2d", "box3d_small"])
def test_global_mass_balance(which, request):
    """sum_i V_i F_i = -(boundary mass flux) - sum_i V_i ahh_i psi'_i (interior fluxes cancel)."""
    prob = request.getfixturevalue(which)
    op = OracleProblem(prob)
    rng = np.random.default_rng(11)
    psi = prob.psi0 + rng.uniform(-0.5, 0.5, op.N)
    psip = rng.normal(size=op.N)
    op.begin_step(0.0, psi, psip)
    F = op.residual(0.0, psi, psip)
    V, _ = op.omesh.volumes()
    ahh, _, _ = op.node_fields()
    q, M = op.face_fields()
    NFI = op.omesh.NFI
    lhs = (V * F).sum()
    rhs = -M[NFI:].sum() - (V * ahh * psip).sum()
    scale = np.abs(M).sum() + np.abs(V * ahh * psip).sum()
    assert abs(lhs - rhs) < 1e-12 * scale


@pytest.mark.parametrize("which", ["slice2d", "box3d_small"])
def test_linear_saturated_field_is_divergence_free(which, request):
    prob = _no_flow(request.getfixturevalue(which))
    prob.ekmul = np.ones_like(prob.ekmul)
    prob.ezone = np.zeros_like(prob.ezone)
    op = OracleProblem(prob)
    m = prob.mesh
    psi = 30.0 + m.xyz @ np.array([0.3, -0.2, 0.1 if m.dim == 3 else 0.0])  # saturated: k_r = 1
    op.begin_step(0.0, psi, 0 * psi)
    F = op.residual(0.0, psi, 0 * psi)
    interior = np.diff(m.node_tag_ptr) == 0
    assert np.abs(F[interior]).max() < 1e-11 * _gravity_scale(op, prob)
    assert np.abs(F[~interior]).max() > 1e-6 * _gravity_scale(op, prob)  # boundary sees the flux


def test_upwind_weights_follow_edge_flux_sign(slice2d):
    op = OracleProblem(slice2d)
    rng = np.random.default_rng(3)
    psi = slice2d.psi0 - 2.0 + rng.uniform(-0.3, 0.3, op.N)
    op.begin_step(0.0, psi, 0 * psi)
    wf, fl = op.weights()
    assert set(np.unique(wf)) <= {0.0, 1.0}
    assert np.all((fl < 0) == (wf == 1.0))
