"""The C-ABI library without a GPU: it loads, exports every symbol that
include/idwarp_hip.h declares, reports a missing device as an error code
(never a CPU fallback), and its host-side routines agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from idwarp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "idwarp_hip.h")).read()
    return sorted(set(re.findall(r"\b(idw_[a-z0-9_]+)\s*\(", text)))


def test_exports_match_header():
    L = _lib.load()
    decl = _declared()
    assert decl and set(decl) == set(_lib.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def test_defaults_match_reference():
    o = _lib.IdwOptions()
    _lib.load().idw_default_options(ctypes.byref(o))
    # UnstructuredMesh.py:124-145
    assert (o.aExp, o.bExp, o.LdefFact, o.alpha, o.errTol, o.cornerAngle) == (3.0, 5.0, 1.0, 0.25, 0.0005, 30.0)
    assert (o.evalMode, o.useRotations, o.zeroCornerRotations, o.bucketSize) == (1, 1, 1, 8)


def test_no_device_is_an_error_not_a_fallback():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    L = _lib.load()
    o = _lib.IdwOptions()
    L.idw_default_options(ctypes.byref(o))
    ctx = ctypes.c_void_p(0)
    rc = L.idw_create(ctypes.byref(ctx), 0, ctypes.byref(o))
    assert rc == _lib.IDW_ENODEV and not ctx.value
    assert L.idw_last_error()
    from idwarp_amd import USMesh

    with pytest.raises(_lib.IdwError):
        USMesh(options={"evalMode": "fast"})


def test_invalid_arguments():
    L = _lib.load()
    o = _lib.IdwOptions()
    L.idw_default_options(ctypes.byref(o))
    o.bucketSize = 0
    ctx = ctypes.c_void_p(0)
    assert L.idw_create(ctypes.byref(ctx), 0, ctypes.byref(o)) == _lib.IDW_EINVAL
    assert L.idw_warp_mesh(None) == _lib.IDW_EINVAL


def test_point_reduce_matches_oracle():
    from oracle import oracle as O

    rng = np.random.default_rng(2)
    base = rng.standard_normal((500, 3))
    pts = np.concatenate([base, base[rng.integers(0, 500, 200)], base[:10] + 3e-13])
    pts = pts[rng.permutation(len(pts))]
    u1, l1, n1 = _lib.point_reduce(pts, 1e-12)
    u2, l2, n2 = O.point_reduce(pts, 1e-12)
    assert n1 == n2 == 500
    assert np.array_equal(u1, u2) and np.array_equal(l1, l2)
    for m in ("sphere", "star"):
        from idwarp_amd import synthetic

        mesh = synthetic.sphere_omesh(13, 9, 5) if m == "sphere" else synthetic.star_shells(n=9, nshell=2)
        a = _lib.point_reduce(mesh["pts"])
        b = O.point_reduce(mesh["pts"])
        assert a[2] == b[2] and np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])


def test_usmesh_option_errors():
    from idwarp_amd import Error, USMesh

    with pytest.raises(Error):
        USMesh(options=None)
    with pytest.raises(Error):
        USMesh(options={"evalMode": "bogus"})
    with pytest.raises(Error):
        USMesh(options={"notAnOption": 1})
