import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wrt-project_amd")
for p in (os.path.dirname(os.path.abspath(__file__)), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session", autouse=True)
def _tune_cache(tmp_path_factory):
    """The library's tuning cache in a file of the session's own (tests that check the cache
    set WRT_TUNE_CACHE themselves)."""
    path = str(tmp_path_factory.mktemp("tune") / "wrt_tune.txt")
    old = os.environ.get("WRT_TUNE_CACHE")
    os.environ["WRT_TUNE_CACHE"] = path
    yield path
    if old is None:
        os.environ.pop("WRT_TUNE_CACHE", None)
    else:
        os.environ["WRT_TUNE_CACHE"] = old


@pytest.fixture(scope="session")
def ctx():
    import wrt

    c = wrt.Context(0)
    yield c
    c.close()
