"""Ensemble batching (fvmpor_amd.ensemble) on the CPU: the block-diagonal union
problem reproduces each instance's own residual, and the instance draws
follow the configured distributions."""
import numpy as np

from fvmpor_amd.ensemble import Draw, draw_instances, make_instance, union_problem
from oracle import oracle as O


def _small(dim, i, n, sigma, rain):
    return make_instance(Draw(i, dim, n, sigma, rain, 1000 + 2 * i + (dim == 3)))


def test_draws_follow_distributions():
    d = draw_instances(256)
    n = np.array([x.n_nodes for x in d])
    assert n.min() >= 5000 and n.max() <= 60000 and np.ptp(n) > 40000
    assert all(0.5 <= x.sigma_lnK <= 2.0 and 0.01 <= x.rain <= 0.1 for x in d)
    assert all(x.dim == (2 if x.seed % 2 == 0 else 3) for x in d)
    assert sum(x.dim == 3 for x in d) == 128


def test_union_residual_is_the_instances_residuals():
    for dim in (2, 3):
        probs = [_small(dim, i, n, s, r) for i, (n, s, r) in enumerate([(400, 0.7, 0.02), (900, 1.8, 0.09),
                                                                         (600, 1.2, 0.05)])]
        union, off = union_problem(probs)
        assert off[-1] == union.mesh.n_nodes == sum(p.mesh.n_nodes for p in probs)
        rng = np.random.default_rng(dim)
        psi = np.concatenate([p.psi0 - rng.uniform(0, 0.5, p.mesh.n_nodes) for p in probs])
        psip = rng.normal(scale=1e-2, size=len(psi))
        ou = O.OracleProblem(union)
        ou.begin_step(0.0, psi, psip)
        Fu = ou.residual(0.0, psi, psip)
        for i, p in enumerate(probs):
            sl = slice(off[i], off[i + 1])
            op = O.OracleProblem(p)
            op.begin_step(0.0, psi[sl], psip[sl])
            Fi = op.residual(0.0, psi[sl], psip[sl])
            scale = max(np.abs(Fi).max(), op.flux_scale())
            assert np.abs(Fu[sl] - Fi).max() <= 1e-13 * scale
