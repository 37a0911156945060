import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built librsx.so")


@pytest.fixture(scope="session")
def oracle():
    import rsxlib as R
    lib = R.load_oracle()
    if lib is None:
        pytest.skip("oracle/_ref/librsx_ref.so is not built (needs the reference tree at build time)")
    return lib


@pytest.fixture(scope="session")
def cpucheck():
    import rsxlib as R
    lib = R.load_cpucheck()
    if lib is None:
        pytest.skip("tests/cpu_check/librsx_cpucheck.so is not built")
    w = lib.cc_world_new(R.MESH_DIR.encode())
    assert w, "mesh load failed"
    yield lib, w
    lib.cc_world_free(w)


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch
