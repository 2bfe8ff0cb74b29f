"""Host side of the learned initial iterate (fvmpor_amd/learn.py): the node
graph against the oracle's Jacobian pattern without its diagonal, and the
static feature columns on known problems."""
import numpy as np

from oracle import oracle as O


def test_node_graph_is_the_jacobian_pattern_without_diagonal(slice2d, box3d_small):
    from fvmpor_amd.learn import node_graph
    for prob in (slice2d, box3d_small):
        ia, ja = node_graph(prob.mesh)
        ia2, ja2 = O.jac_pattern(O.OracleMesh(prob.mesh))
        rows = np.repeat(np.arange(len(ia2) - 1), np.diff(ia2))
        keep = ja2 != rows
        assert np.array_equal(ja, ja2[keep])
        assert np.array_equal(ia[1:], np.cumsum(np.bincount(rows[keep], minlength=len(ia2) - 1)))
        # symmetric
        r = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
        assert set(zip(r.tolist(), ja.tolist())) == set(zip(ja.tolist(), r.tolist()))


def test_static_features(slice2d, box3d_small):
    from fvmpor_amd.learn import static_features
    x = static_features(slice2d)
    assert x.shape == (slice2d.mesh.n_nodes, 16) and x.dtype == np.float32
    assert np.all(x[:, :5] == 0)
    z = slice2d.mesh.xyz[:, 1]
    assert np.allclose(x[:, 5], (z - z.min()) / (z.max() - z.min()), atol=1e-7)
    # homogeneous loam: K = 0.25 m/day, alpha 3.6, n 1.56, one zone
    assert np.allclose(x[:, 6], np.log(0.25), atol=1e-6)
    assert np.all(x[:, 7] == 1) and np.all(x[:, 8:11] == 0)
    assert np.allclose(x[:, 14], np.log(3.6), atol=1e-6) and np.allclose(x[:, 15], 1.56, atol=1e-6)
    # the infiltration strip x in [4, 6] on the top surface carries the flux flag and 10 x the rate
    top = x[:, 11] == 1
    assert top.sum() > 0 and np.allclose(x[top, 12], 0.5)
    xy = slice2d.mesh.xyz[top]
    assert np.all((xy[:, 0] >= 4 - 1e-9) & (xy[:, 0] <= 6 + 1e-9)) and np.allclose(xy[:, 1], 5.0)
    xb = static_features(box3d_small)
    assert np.allclose(xb[:, 7:11].sum(axis=1), 1.0, atol=1e-6)
