import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available():
    try:
        import algorithmengineering_altproject_amd as t3d
        t3d.device_info()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def t3d():
    import algorithmengineering_altproject_amd as mod
    mod.lib()
    return mod
