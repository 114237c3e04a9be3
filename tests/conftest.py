import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        from prothlogic_amd import core
        return core.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    from prothlogic_amd import core
    if core.device_count() <= 0:
        pytest.fail("no HIP device visible but a gpu-marked test was selected")
    return core
