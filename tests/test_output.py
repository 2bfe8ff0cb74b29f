"""Output writers against the reference's own output files.

The EEM cassion run's VTU files (examples/EEM/CASS_dom0_{0..11}.vtu) are pinned
by their SHA-256 (tests/golden/make_golden.py): writing the same points,
connectivity and heads with write_vtu reproduces each file byte for byte,
which fixes the layout, the whitespace and the number format of
Solution::write_timestep_VTK_XML (solution.h:178-301)."""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def eem():
    z = np.load(os.path.join(GOLDEN, "cassion_eem_vtu.npz"))

    class Mesh:
        dim = 2
        xyz = z["points"].astype(np.float64)
        conn = z["conn"].astype(np.int32)
    return Mesh, z


@pytest.mark.parametrize("step", range(12))
def test_vtu_writer_reproduces_the_reference_files(eem, step, tmp_path):
    from fvmpor_amd.output import read_vtu, write_vtu
    mesh, z = eem
    path = write_vtu(str(tmp_path / "CASS"), step, mesh, z["heads"][step].astype(np.float64))
    assert os.path.basename(path) == f"CASS_dom0_{step}.vtu"
    with open(path) as fh:
        assert hashlib.sha256(fh.read().encode()).hexdigest() == z["vtu_sha256"][step]
    a = read_vtu(path)
    np.testing.assert_array_equal(a["pressure_head"].astype(np.float32), z["heads"][step])
    np.testing.assert_array_equal(a["connectivity"].reshape(-1, 3), z["conn"])
    assert np.array_equal(a["offsets"], 3 * np.arange(1, len(z["conn"]) + 1))
    assert set(a["types"].tolist()) == {5.0} and set(a["dom"].tolist()) == {0.0}


def test_tetrahedra_header_and_run_file(tmp_path):
    from fvmpor_amd.output import read_vtu, write_pvtu_header, write_run, write_vtu
    from fvmpor_amd.problems import box3d_problem
    p = box3d_problem(nx=6, ny=5, nz=4)
    path = write_vtu(str(tmp_path / "box"), 3, p.mesh, p.psi0)
    a = read_vtu(path)
    assert set(a["types"].tolist()) == {10.0}
    np.testing.assert_allclose(a["Points"].reshape(-1, 3), p.mesh.xyz, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(a["pressure_head"], p.psi0, rtol=1e-5, atol=1e-5)
    head = open(write_pvtu_header(str(tmp_path / "box"), 3, n_dom=2)).read()
    assert f'Source="{tmp_path / "box"}_dom1_3.vtu"' in head and 'GhostLevel="0"' in head
    sols = [p.psi0, p.psi0 * 0.5]
    write_run(str(tmp_path / "box.run"), sols)
    lines = open(tmp_path / "box.run").read().split("\n")
    assert lines[0] == f"2 1 {p.mesh.n_nodes}"
    vals = np.array([float(v) for v in lines[1:] if v])
    np.testing.assert_allclose(vals, np.concatenate(sols), rtol=1e-15, atol=0)  # 16 significant digits
