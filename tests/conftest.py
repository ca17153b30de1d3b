"""Shared test setup: import paths, the `gpu` marker, fixture loaders.

Tests marked `gpu` need an MI355X (they run through the C ABI of
libtsdx_hip.so); everything else runs on the CPU.  The oracle (oracle/) is
only ever the checker here.
"""
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "tinysolverdx-batched-small-matrix-cholesky-lu-with-numerical_amd"
GOLDEN = ROOT / "tests" / "golden"
for p in (PKG_DIR, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libtsdx_hip.so")


def golden(name):
    """Loads tests/golden/<name>.npz (plain arrays, no pickles)."""
    return np.load(GOLDEN / f"{name}.npz", allow_pickle=False)


def golden_cases(npz, suffix):
    """Names of the cases in a fixture file that carry `suffix` (e.g. '_A')."""
    return sorted({k[: -len(suffix)] for k in npz.files if k.endswith(suffix)})


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle("port")


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import Oracle, available
    if not available("reference"):
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return Oracle("reference")


@pytest.fixture(scope="session")
def tsdx():
    import tinysolverdx
    return tinysolverdx


@pytest.fixture(scope="session")
def hip():
    """The loaded C ABI; GPU tests fail loudly if no device is visible."""
    from tinysolverdx import _lib
    L = _lib.lib()
    if not L.tsdx_hip_available():
        pytest.fail("no HIP device visible to libtsdx_hip.so")
    return _lib


def _bitview(a):
    """Integer view of a float array with every NaN mapped to one pattern (NaN
    sign and payload depend on how the hardware propagates a NaN operand, so
    only "is NaN" is compared; every other value, signed zeros included, is
    compared bit for bit)."""
    a = np.ascontiguousarray(a)
    if a.dtype.kind != "f":
        return a
    v = a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize]).copy()
    v[np.isnan(a)] = 0
    return v


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(_bitview(a), _bitview(b)) and \
        np.array_equal(np.isnan(a), np.isnan(b))


def first_bit_mismatch(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    ia, ib = _bitview(a), _bitview(b)
    if a.dtype.kind == "f":
        ia = ia + np.isnan(a)
        ib = ib + np.isnan(b)
    idx = np.argwhere(ia != ib)
    if len(idx) == 0:
        return None
    i = tuple(idx[0])
    return f"{len(idx)} mismatches, first at {i}: {a[i]!r} vs {b[i]!r}"
