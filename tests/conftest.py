import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfvp.so")


@pytest.fixture(scope="session")
def cassion_mesh():
    from fvmpor_amd.meshgen import read_pmesh
    return read_pmesh(os.path.join(GOLDEN, "cassion_1_0.pmesh"))


@pytest.fixture(scope="session")
def cassion(cassion_mesh):
    from fvmpor_amd.problems import cassion_problem
    return cassion_problem(cassion_mesh)


@pytest.fixture(scope="session")
def slice2d():
    from fvmpor_amd.problems import slice2d_problem
    return slice2d_problem()


@pytest.fixture(scope="session")
def slice2d_bc():
    from fvmpor_amd.problems import slice2d_bc_problem
    return slice2d_bc_problem()


@pytest.fixture(scope="session")
def box3d_bc():
    from fvmpor_amd.problems import box3d_bc_problem
    return box3d_bc_problem()


@pytest.fixture(scope="session")
def box3d_small():
    from fvmpor_amd.problems import box3d_problem
    return box3d_problem(nx=14, ny=12, nz=8)


def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")
