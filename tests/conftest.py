import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(ref=False)
    oracle.set_pi(oracle.PI_TRUNC)
    return oracle


@pytest.fixture(scope="session")
def ref(orc):
    if not orc.Ref.available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return orc.Ref()


@pytest.fixture(autouse=True)
def _reset_pi():
    yield
    try:
        import oracle
        if oracle._lib is not None:
            oracle.set_pi(oracle.PI_TRUNC)
    except Exception:
        pass
