"""Host-side mass diagnostics (fvmpor_amd.diagnostics) against closed forms."""
import numpy as np

from fvmpor_amd.diagnostics import inflow_rate, total_mass
from fvmpor_amd.problems import box3d_problem, slice2d_problem


def test_saturated_mass_is_pore_volume():
    p = box3d_problem(nx=10, ny=9, nz=6)
    p.zones = p.zones.copy()
    p.zones[:, 7] = 0.0
    vol = np.abs(p.mesh.element_volumes())
    pore = sum(vol[p.ezone == z].sum() * p.zones[z, 3] for z in range(len(p.zones)))
    assert abs(vol.sum() - 100.0 * 100.0 * 20.0) < 1e-6 * vol.sum()
    m = total_mass(p, np.full(p.mesh.n_nodes, 0.5))
    assert abs(m - p.rho0 * pore) < 1e-12 * m


def test_unsaturated_mass_between_residual_and_full():
    p = slice2d_problem(nx=21, ny=9)
    p.zones = p.zones.copy()
    p.zones[:, 7] = 0.0
    phi, Sr = p.zones[0, 3], p.zones[0, 6]
    area = np.abs(p.mesh.element_volumes()).sum()
    m = total_mass(p, np.full(p.mesh.n_nodes, -1e14))
    assert abs(m / (p.rho0 * phi * area) - Sr) < 1e-6
    assert total_mass(p, np.full(p.mesh.n_nodes, -0.1)) < p.rho0 * phi * area


def test_inflow_rate_matches_forcing():
    p = box3d_problem(nx=10, ny=9, nz=6, rain=0.02)
    assert abs(inflow_rate(p) - 0.02 * 100.0 * 100.0 * p.rho0) < 1e-9 * inflow_rate(p)
    s = slice2d_problem(nx=41, ny=17)
    strip = inflow_rate(s) / (0.05 * s.rho0)
    assert 1.8 < strip <= 2.0  # tagged faces cover the strip x in [4, 6] up to one jittered cell
