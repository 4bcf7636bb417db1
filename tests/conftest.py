import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _small_link_cfg():
    # 3000 waters at 0.0334 molecules/A^3 with a 9 A cutoff: 4 x 4 x 4 link cells
    from mdbas_amd import configs
    return dict(configs.COMMON, name="tip3p_3k_link", n_waters=3000, L=(3000 / 0.0334) ** (1 / 3),
                seed=5, grid=32, cutOff=9.0, cutOn=7.5, timeStep=0.002, alpha=0.35,
                precision="fp64")


@pytest.fixture(scope="session")
def c0_box():
    from mdbas_amd import configs
    return configs.build(configs.CONFIGS["c0"])


@pytest.fixture(scope="session")
def link_box():
    from mdbas_amd import configs
    return configs.build(_small_link_cfg())


@pytest.fixture(scope="session")
def c1_box():
    from mdbas_amd import configs
    return configs.build(configs.CONFIGS["c1"])


@pytest.fixture(scope="session")
def c2_box():
    from mdbas_amd import configs
    return configs.build(configs.CONFIGS["c2"])


@pytest.fixture(scope="session")
def small_box():
    """216 waters in an 18.6 A box, 9 A cutoff (the smallest replica)."""
    from mdbas_amd import configs
    L = (216 / 0.0334) ** (1 / 3)
    cfg = dict(configs.COMMON, name="tip3p_216", n_waters=216, L=L, seed=11, grid=20,
               cutOff=9.0, cutOn=7.5, alpha=0.35, timeStep=0.002, precision="fp64")
    return configs.build(cfg)
