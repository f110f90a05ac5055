"""idw_warp_deriv_dev_noalloc, the allocation-free reverse product (needs an
MI355X): through the C ABI it gives idw_warp_deriv_dev's result (exact mode
FP64 and mixed, fast mode), on two calls in a row whose second has a smaller
mesh, without changing the device's free memory; it refuses calls without a
prior warp and surfaces beyond its scratch array."""
import ctypes

import numpy as np
import pytest

from idwarp_amd import USMesh, _lib, synthetic

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _mesh(m, **opts):
    g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", **opts))
    g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    return g


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a).ravel(), dtype=torch.float64, device="cuda:0")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("opts", [dict(evalMode="exact"), dict(evalMode="exact", precision="mixed"),
                                  dict(evalMode="fast", errTol=1e-4)], ids=["exact-fp64", "exact-mixed", "fast"])
def test_noalloc_reverse_equals_the_allocating_one(opts):
    import torch

    L = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    big, small = synthetic.ellipsoid_omesh(65, 33, 25), synthetic.sphere_omesh(13, 9, 5)
    gs = {name: _mesh(m, **opts) for name, m in (("big", big), ("small", small))}
    tol = 1e-6 if opts.get("precision") == "mixed" else 1e-12
    if opts["evalMode"] == "fast":
        tol = max(tol, 1e-10)  # the fast reverse sums with FP64 atomics
    for name, m in (("big", big), ("small", small)):  # the second call has the smaller input
        g = gs[name]
        print("case", name, opts, flush=True)
        d_xs = _dev(torch, m["perturb"](m["pts"]))
        d_xv = torch.empty(3 * g.nVol, dtype=torch.float64, device="cuda:0")
        d_dxv = _dev(torch, synthetic.random_dxv(g.nVol, seed=5, frac=0.3))
        ref = torch.empty(3 * g.nSurf, dtype=torch.float64, device="cuda:0")
        got = torch.full_like(ref, float("nan"))
        # no warp yet: refused
        assert L.idw_warp_deriv_dev_noalloc(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(got), stream) == _lib.IDW_ESTATE
        _lib.check(L.idw_warp_mesh_dev(g.ctx, _vp(d_xs), _vp(d_xv), stream))
        # the allocating form first: it launches the same kernels, so the
        # runtime has set up their per-lane private memory before the check
        _lib.check(L.idw_warp_deriv_dev(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(ref), stream))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _lib.check(L.idw_warp_deriv_dev_noalloc(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(got), stream))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        r = _rel(got.cpu().numpy(), ref.cpu().numpy())
        print(name, "noalloc vs allocating reverse: rel %.2e; device free memory %d -> %d" % (r, free0, free1),
              flush=True)
        assert np.isfinite(got.cpu().numpy()).all() and r <= tol
        assert free1 == free0  # nothing allocated
    # again on the big mesh after the small one has used the scratch array
    g, m = gs["big"], big
    d_xs = _dev(torch, m["perturb"](m["pts"]))
    d_dxv = _dev(torch, synthetic.random_dxv(g.nVol, seed=6, frac=1.0))
    d_xv = torch.empty(3 * g.nVol, dtype=torch.float64, device="cuda:0")
    _lib.check(L.idw_warp_mesh_dev(g.ctx, _vp(d_xs), _vp(d_xv), stream))
    a = torch.empty(3 * g.nSurf, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    _lib.check(L.idw_warp_deriv_dev_noalloc(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(a), stream))
    _lib.check(L.idw_warp_deriv_dev(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(b), stream))
    torch.cuda.synchronize()
    assert _rel(a.cpu().numpy(), b.cpu().numpy()) <= tol


def test_noalloc_reverse_refuses_surfaces_beyond_its_scratch():
    """Exact mode over 351,386 unique surface nodes needs 12 * 8 bytes per
    node and volume chunk, more than the 32 MiB array holds."""
    import torch

    L = _lib.load()
    m = synthetic.star_shells(n=243, nshell=1)
    g = _mesh(m, evalMode="exact")
    d_xs = _dev(torch, m["pts"])
    d_dxv = torch.zeros(3 * g.nVol, dtype=torch.float64, device="cuda:0")
    out = torch.empty(3 * g.nSurf, dtype=torch.float64, device="cuda:0")
    rc = L.idw_warp_deriv_dev_noalloc(g.ctx, _vp(d_xs), _vp(d_dxv), _vp(out), ctypes.c_void_p(0))
    assert rc == _lib.IDW_EINVAL and b"349525" in L.idw_last_error()
