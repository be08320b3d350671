import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("gravity-sim_amd/python", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _gpu_visible() -> bool:
    try:
        import gsim
        import numpy as np
        p = gsim.Particles(np.zeros((1, 2), np.float32), np.zeros((1, 2), np.float32),
                           np.ones(1, np.float32))
        with gsim.Simulation(p):
            return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a box without a device must fail."""
    import gsim
    gsim.load_library()
    return gsim
