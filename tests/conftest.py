import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def ref_module(name):
    """The reference's own compiled helpers (oracle/_ref), or None when they
    were not built (they need the reference sources at build time)."""
    import importlib
    path = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.isdir(path):
        return None
    if path not in sys.path:
        sys.path.insert(0, path)
    try:
        return importlib.import_module(name)
    except ImportError:
        return None


@pytest.fixture(scope="session")
def fast_geometry_ref():
    m = ref_module("_fast_geometry")
    if m is None:
        pytest.skip("oracle/_ref/_fast_geometry not built (needs the reference sources)")
    return m


@pytest.fixture(scope="session")
def fasthj_ref():
    m = ref_module("_FastHJ")
    if m is None:
        pytest.skip("oracle/_ref/_FastHJ not built (needs the reference sources)")
    return m
