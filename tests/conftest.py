"""Shared fixtures: the package (loaded by path), the oracle, small stacks."""

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "i.sar.temporal_coherence_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    """Import i.sar.temporal_coherence_amd/ as module tcoh_amd."""
    if "tcoh_amd" in sys.modules:
        return sys.modules["tcoh_amd"]
    spec = importlib.util.spec_from_file_location(
        "tcoh_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tcoh_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def tcoh():
    return load_package()


@pytest.fixture(scope="session")
def cref():
    from oracle import cref as c
    c.lib()
    return c


@pytest.fixture(scope="session")
def fp64():
    from oracle import fp64 as f
    return f


def strip_of(stack, window):
    """Single-strip buffers of a (n, rows, cols) complex stack (main.c:52-127
    through the oracle's restatement)."""
    from oracle import cref as c
    n, rows, cols = stack.shape
    ha, hr = window[0] // 2, window[1] // 2
    prow = rows + 2 * ha
    slc, amp, valid = c.prepare_strip(np.ascontiguousarray(stack.real, np.float32),
                                      np.ascontiguousarray(stack.imag, np.float32),
                                      -ha, prow, hr)
    return slc, amp, valid, prow, cols + 2 * hr
