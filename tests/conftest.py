import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (libfarm device calls)")


def _built():
    return os.path.exists(os.path.join(ROOT, "farm_amd", "libfarm.so")) and os.path.exists(
        os.path.join(ROOT, "oracle", "liborc.so"))


if not _built():  # build only when nothing is built yet
    import __graft_entry__

    __graft_entry__.build()


@pytest.fixture(scope="session")
def device():
    from farm_amd import capi

    dev = capi.Device(0)
    yield dev
    dev.close()
