import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden_3c10():
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "3c10.npz"))
    return {k: d[k] for k in d.files}


@pytest.fixture(autouse=True)
def _clear_sde_error():
    yield
    try:
        import sde_amd
        if os.path.exists(sde_amd.LIBPATH):
            sde_amd.clear_error()
    except Exception:
        pass
