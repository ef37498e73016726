import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from kazan_amd import vk  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "libkazan_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP draw path)")


def _ensure_built():
    # load what is built; build only when nothing is (a fresh checkout)
    if not os.path.exists(vk.DRIVER_PATH) or not os.path.exists(ORACLE_LIB):
        from kazan_amd import build
        build.build_all()


@pytest.fixture(scope="session")
def driver():
    _ensure_built()
    return vk.Driver()


@pytest.fixture(scope="session")
def oracle():
    _ensure_built()
    from oracle import bridge
    return bridge


@pytest.fixture(scope="session")
def ctx(driver):
    from kazan_amd.scenes.renderer import Context
    return Context(driver)
