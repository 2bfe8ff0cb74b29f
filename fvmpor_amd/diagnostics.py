"""Host-side diagnostics of a head field: total water mass and the boundary
inflow rate, for mass-balance reporting.

Mass follows the reference's Physics::compute_mass (fvmpor_ODE_impl.cpp:146-156):
sum over control volumes of V_i theta_i rho_i, with theta = phi S_w on each
zone's share of the CV (the CV-averaged fields of fvmpor.h:874-928) and
rho = rho0 (beta = 0, fvmpor.h:646-679); with the specific-storage extension
(capacity phi dS_w/dpsi + S_s S_w) the stored volume gains S_s S_w psi per unit
volume, its first-order integral.  The inflow rate follows
mass_flux_per_time (fvmpor_ODE_impl.cpp:159-174): minus the prescribed flux
(type 3) times the boundary face area.  Post-processing only (numpy, float64).
"""
from __future__ import annotations

import numpy as np

from .meshgen import TET_FACES, TRI_FACES


def _saturation(h, alpha, n, Sr):
    """van Genuchten S_w(h) (the closed form of fvmpor.h:302-343)."""
    m = 1.0 - 1.0 / n
    Se = np.ones_like(h)
    dry = h < 0
    Se[dry] = (1.0 + (alpha * -h[dry]) ** n) ** (-m)
    return Sr + (1.0 - Sr) * Se


def total_mass(problem, psi) -> float:
    """sum_i V_i theta_i rho0 (+ S_s S_w psi storage) for the head field psi."""
    mesh = problem.mesh
    psi = np.asarray(psi, dtype=np.float64)
    npe = mesh.dim + 1
    scv = np.abs(mesh.element_volumes()) / npe
    total = 0.0
    for z, row in enumerate(problem.zones):
        sel = problem.ezone == z
        if not sel.any():
            continue
        phi, alpha, n, Sr, Ss = row[3], row[4], row[5], row[6], row[7]
        h = psi[mesh.conn[sel]]
        Sw = _saturation(h.ravel(), alpha, n, Sr).reshape(h.shape)
        total += float(np.sum(scv[sel, None] * (phi * Sw + Ss * Sw * h)))
    return total * problem.rho0


def inflow_rate(problem) -> float:
    """Mass per unit time entering through prescribed-flux (type 3) faces."""
    mesh = problem.mesh
    local = TRI_FACES if mesh.dim == 2 else TET_FACES
    rate = 0.0
    for tag, (btype, row) in enumerate(zip(problem.bc_type, problem.bc)):
        if tag == 0 or btype != 3 or row[0] == 0.0:
            continue
        e, f = np.nonzero(mesh.face_tags == tag)
        if len(e) == 0:
            continue
        p = mesh.xyz[mesh.conn[e[:, None], local[f]]]
        if mesh.dim == 2:
            area = np.linalg.norm(p[:, 1, :2] - p[:, 0, :2], axis=1)
        else:
            area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        rate -= row[0] * float(area.sum())
    return rate * problem.rho0
