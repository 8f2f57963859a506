"""The library's host elementwise exports and the oracle's zero conventions against the reference's
own dataCube class (src/data_cube.cpp, compiled from the reference sources into oracle/_ref), and
against the conventions written out where that build is absent."""
import ctypes

import numpy as np
import pytest


def arrays(seed=40):
    rng = np.random.default_rng(seed)
    x = rng.random((3, 4, 5)).astype(np.float32)
    y = rng.random((3, 4, 5)).astype(np.float32)
    x[0, 0, :2] = 0.0
    y[1, 1, :3] = 0.0
    return x, y


def test_divide_rdivide_conventions(petl_lib):
    x, y = arrays()
    a = x.copy()
    assert petl_lib.divide3D(a, y.copy(), *a.shape)
    want = np.where(y == 0, 0, x / np.where(y == 0, 1, y)).astype(np.float32)
    np.testing.assert_array_equal(a, want)
    d = y.copy()
    assert petl_lib.rdivide3D(x.copy(), d, *d.shape)       # denom = num / denom
    want = np.where(y == 0, 0, x / np.where(y == 0, 1, y)).astype(np.float32)
    np.testing.assert_array_equal(d, want)
    m = x.copy()
    assert petl_lib.multiply3D(m, y.copy(), *m.shape)
    np.testing.assert_array_equal(m, x * y)


def test_against_reference_datacube(oracle, petl_lib):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference dataCube build not present")
    x, y = arrays(41)
    for which, name in ((0, "multiply3D"), (1, "divide3D")):
        a, b = x.copy(), y.copy()
        R.ref_binary3d(a, b, *a.shape, which)
        c = x.copy()
        getattr(petl_lib, name)(c, y.copy(), *c.shape)
        np.testing.assert_array_equal(a, c)
    a, b = x.copy(), y.copy()
    R.ref_binary3d(a, b, *a.shape, 2)            # b = a / b
    d = y.copy()
    petl_lib.rdivide3D(x.copy(), d, *d.shape)
    np.testing.assert_array_equal(b, d)
    a = y.copy()
    R.ref_reciprocal3d(a, *a.shape, 7.0)
    c = y.copy()
    petl_lib.reciprocal3D(c, *c.shape, 7.0)
    np.testing.assert_array_equal(a, c)
    fp = ctypes.POINTER(ctypes.c_float)
    ref_ip = R.ref_inner3d(x, y, None, *x.shape)
    mine = petl_lib.inner_product3D(x.ctypes.data_as(fp), y.ctypes.data_as(fp), None, *x.shape)
    assert mine == ref_ip


def test_oracle_divide_conventions(oracle):
    x, y = arrays(42)
    np.testing.assert_array_equal(oracle.divide(x, y), np.where(y == 0, 0, x / np.where(y == 0, 1, y)))
    np.testing.assert_array_equal(oracle.rdivide(y, x), oracle.divide(x, y))
