"""Shared test configuration: markers and import paths."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


class LibOptions:
    """Kernel-selection options of libdnx_fdtd for one test (set, reset to
    the value the test started with, all restored at teardown)."""

    def __init__(self):
        from diffnano_amd import _lib
        self._lib = _lib
        self._saved = {}

    def set(self, name, value):
        if name not in self._saved:
            self._saved[name] = self._lib.get_option(name)
        self._lib.set_option(name, value)

    def reset(self, name):
        if name in self._saved:
            self._lib.set_option(name, self._saved[name])

    def restore(self):
        for name, v in self._saved.items():
            self._lib.set_option(name, v)
        self._saved.clear()


@pytest.fixture
def lib_opts():
    opts = LibOptions()
    try:
        yield opts
    finally:
        opts.restore()
