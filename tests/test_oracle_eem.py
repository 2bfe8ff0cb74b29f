"""The oracle's physics end to end against the reference's own output: the
cassion problem (cassion.h on cassion_1_0.pmesh) integrated on the CPU with
scipy's BDF on psi' = F(t, psi, 0) / ahh(psi) -- the ODE form of the
residual, fvmpor_ODE_impl.cpp:99-115, with upwinding refreshed from the current
state -- and compared with the twelve head fields of the reference's EEM run
(examples/EEM/CASS_dom0_{0..11}.vtu, Float32, EEM at rtol = atol = 5e-4,
fvmpor_ODE.cpp:29-30).  The agreement plateaus at the reference run's own
integration error (about 5e-3 of the head range at worst, 5e-4 RMS) as the
CPU tolerance is tightened, so the bounds below sit just above it."""
import os

import numpy as np
import scipy.sparse as sp
from scipy.integrate import solve_ivp
from scipy.spatial import cKDTree

from conftest import GOLDEN
from oracle import oracle as O


def test_oracle_follows_reference_eem_trajectory(cassion):
    ref = np.load(os.path.join(GOLDEN, "cassion_eem_vtu.npz"))
    dist, idx = cKDTree(cassion.mesh.xyz[:, :2]).query(ref["points"][:, :2].astype(np.float64))
    assert dist.max() < 1e-4 and len(np.unique(idx)) == cassion.mesh.n_nodes
    op = O.OracleProblem(cassion)
    zero = np.zeros(op.N)

    def rhs(t, y):
        op.begin_step(t, y, zero)
        F = op.residual(t, y, zero)
        ahh, _, _ = op.node_fields()
        return F / ahh

    ia, ja = O.jac_pattern(op.omesh)
    pattern = sp.csr_matrix((np.ones(len(ja)), ja, ia), shape=(op.N, op.N))
    sol = solve_ivp(rhs, (0.0, float(ref["times"][-1])), cassion.psi0.copy(), method="BDF",
                    t_eval=ref["times"], rtol=1e-5, atol=1e-5, jac_sparsity=pattern)
    assert sol.status == 0
    heads = ref["heads"].astype(np.float64)
    span = heads.max() - heads.min()
    d = sol.y[idx, :].T - heads
    assert np.abs(d).max() / span < 1e-2
    assert np.sqrt((d * d).mean(axis=1)).max() / span < 1e-3
