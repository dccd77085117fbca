import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def models():
    from pelephysics_amd.mechgen.__main__ import load
    return {m: load(m) for m in ("LiDryer", "drm19", "dodecane_lu", "heptane_lu_88sk", "CH4_lean_qss",
                                 "LuEthylene_qss", "dodecane_lu_qss")}


@pytest.fixture(scope="session")
def oracles(models):
    import oracle
    return {k: oracle.Mech(v) for k, v in models.items()}
