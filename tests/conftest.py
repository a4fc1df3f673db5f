import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    lib = os.path.join(ROOT, "blockvecchia_amd", "lib", "libbv.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        from blockvecchia_amd import build
        build.build_lib()
    drv = os.path.join(ROOT, "blockvecchia_amd", "bin", "test_dvecchia_batch")
    if not os.path.exists(drv):
        from blockvecchia_amd import build
        build.build_driver()
    if not os.path.exists(orc):
        from oracle import build as obuild
        obuild.build_oracle()


_ensure_built()

# Relative tolerance of the log-likelihood against the oracle (FP64 both sides).
LLH_RTOL = 1e-10


@pytest.fixture(scope="session")
def c1_small():
    """A reduced version of the n=2,000 configuration: n=600 uniform points
    (seed 0), sigma^2=1.5, range=0.1, nu=0.5, exact dense sampling, 30 blocks,
    cs=40."""
    from blockvecchia_amd import datagen
    x, y = datagen.uniform_locations(600, 0)
    obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=0)
    return x, y, obs


@pytest.fixture(scope="session")
def c1_full():
    """The n=2,000 configuration: iid uniform (seed 0), sigma^2=1.5,
    range=0.1, nu=0.5, y = L z with the dense FP64 Cholesky, 100 blocks, cs=60."""
    from blockvecchia_amd import datagen
    x, y = datagen.uniform_locations(2000, 0)
    obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=0)
    return x, y, obs


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)
