"""The fast-mode walk's device arithmetic (idwarp_amd/csrc/idw_walk_math.h),
through the test library tests/native/libwalkmath.so:

* sqrt_normal equals the IEEE square root bit for bit (numpy's sqrt, which
  is correctly rounded, as is the reference's Fortran sqrt) and the device
  sqrt() on normal arguments from 2^-767 up, including perfect squares and
  arguments whose roots lie next to a rounding midpoint;
* get_error_cells equals getError (kd_tree.F90:882-906) restated here with
  the linear search over rstar and the unfused interpolation, bit for bit,
  on the radii, their neighbours and log-uniform samples."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libwalkmath.so")
RSTAR = [2.0, 3.0, 4.0, 5.0, 7.5, 10.0, 20.0, 40.0, 80.0, 160.0, 320.0, 640.0]


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libwalkmath.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    for f in (L.walkmath_sqrt, L.walkmath_error):
        f.restype = ctypes.c_int
    L.walkmath_sqrt.argtypes = [dp, dp, dp, ctypes.c_int64]
    L.walkmath_error.argtypes = [dp, dp, dp, ctypes.c_int64]
    return L


def test_sqrt_normal_is_the_ieee_square_root():
    rng = np.random.default_rng(21)
    n = 1 << 21
    wide = np.ldexp(1.0 + rng.random(n), rng.integers(-767, 1023, n))
    walk = np.exp(rng.uniform(np.log(1e-24), np.log(1e6), n))  # the walk's squared distances
    roots = rng.random(1 << 18) * 1e3 + 1e-3
    squares = roots * roots
    # x = (s + ulp(s)/2)^2 rounded: roots next to a rounding midpoint
    s = np.sqrt(walk[: 1 << 18])
    mid = (s + np.spacing(s) * 0.5) ** 2
    edge = np.array([2.0 ** -767, np.nextafter(2.0 ** -767, 1.0), 1.0, 2.0, 4.0, 0.25, np.finfo(np.float64).max,
                     np.nextafter(1.0, 0.0), np.nextafter(1.0, 2.0)])
    x = np.ascontiguousarray(np.concatenate([wide, walk, squares, mid, np.nextafter(mid, 0.0), edge]))
    assert (x >= 2.0 ** -767).all() and np.isfinite(x).all()
    fast, lib = np.empty_like(x), np.empty_like(x)
    assert _lib().walkmath_sqrt(x, fast, lib, x.size) == 0
    want = np.sqrt(x)
    bad = np.nonzero(fast.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, [(x[i].hex(), fast[i].hex(), want[i].hex()) for i in bad[:5]]
    assert np.array_equal(lib.view(np.uint64), want.view(np.uint64))


def _get_error(err, d):
    """getError (kd_tree.F90:882-906) per value: first segment with
    d < rstar(i+1), the last entry from 640 on."""
    out = np.empty_like(d)
    for k, (e, v) in enumerate(zip(err, d)):
        i = next((j for j in range(11) if v < RSTAR[j + 1]), 11)
        if i == 11:
            out[k] = e[11]
            continue
        fact = (v - RSTAR[i]) * (1.0 / (RSTAR[i + 1] - RSTAR[i]))
        out[k] = (1.0 - fact) * e[i] + fact * e[i + 1]
    return out


def test_error_lookup_is_get_error():
    rng = np.random.default_rng(22)
    d = []
    for r in RSTAR:
        d += [r, np.nextafter(r, 0.0), np.nextafter(r, np.inf)]
    d += list(np.exp(rng.uniform(np.log(2.0), np.log(5e4), 60000)))
    d += [1023.9999999999999, 1024.0, 1e300]
    d = np.ascontiguousarray([v for v in d if v >= 2.0], np.float64)
    n = d.size
    # decreasing positive error tables of a few decades, then Ai
    tab = np.sort(np.exp(rng.uniform(-30.0, 0.0, (n, 12))), axis=1)[:, ::-1]
    err = np.ascontiguousarray(np.concatenate([tab, rng.random((n, 1)) * 10.0], axis=1))
    got = np.empty(n)
    assert _lib().walkmath_error(err, d, got, n) == 0
    want = _get_error(tab, d)
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, [(d[i], got[i], want[i]) for i in bad[:5]]
