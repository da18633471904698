import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libamt.so)")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """The suites load what __graft_entry__.build() produced; build only if nothing is."""
    from am_ml_suite_amd import _ffi
    import oracle_ffi
    if not os.path.exists(oracle_ffi.LIB_PATH):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if not os.path.exists(_ffi.LIB_PATH):
        from am_ml_suite_amd.build import build
        build()
    yield
