import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PES_DIR = os.path.join(ROOT, "md_tian2_amd", "pes")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmdt_gpu.so")


@pytest.fixture(scope="session")
def pes_hau():
    from md_tian2_amd import pesfile
    return pesfile.load_pes(os.path.join(PES_DIR, "EMT-HAu.pes"))


@pytest.fixture(scope="session")
def pes_hpt():
    from md_tian2_amd import pesfile
    return pesfile.load_pes(os.path.join(PES_DIR, "EMT-HPt.pes"))


@pytest.fixture(scope="session")
def gpu_ctx():
    from md_tian2_amd import mdt
    if mdt.device_count() < 1:
        pytest.fail("no GPU visible to libmdt_gpu.so")
    ctx = mdt.Context(0)
    yield ctx
    ctx.close()


def rng(seed=0):
    return np.random.default_rng(seed)
