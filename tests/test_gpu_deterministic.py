"""The deterministic fast reverse (option deterministic; needs an MI355X).

The reference sums each adjoint record serially (kd_tree.F90:942-1051), so
its warpDeriv gives the same bits on every run.  The default fast reverse adds
with FP64 atomics: equal to rounding, in an order that changes from run to
run.  With deterministic on, each plan entry's 12 sums go to a slot of their
own and a per-record reduction adds them in plan order, so:

  * two reverse products of one mesh, two instances of one mesh, a refilled
    plan, a recorded batch and the allocation-free entry point all give the
    same bits (FP64 and mixed precision; no, one and two symmetry planes);
  * the result equals the oracle's within the precision's bar (FP64:
    adjoint relative error <= 1e-10; mixed: <= 1e-6), and the atomic
    reverse's within the same bar: in FP64 the two differ only in the order
    of their sums, which moves the result by up to ~3e-12 relative where the
    mirror passes cancel (one symmetry plane through a sphere).
"""
import ctypes

import numpy as np
import pytest

from idwarp_amd import USMesh, _lib, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _case(name):
    if name == "star":
        return synthetic.star_shells(n=21, nshell=30), []
    half = synthetic.shifted(synthetic.sphere_omesh(17, 11, 11), (0.3, 2.5, 3.2))
    planes = [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0]), ([0.0, 0.0, 1.5], [0.0, 0.0, 1.0])]
    return half, planes[: {"sphere-1plane": 1, "sphere-2planes": 2}[name]]


def _mesh(m, planes, **opts):
    if planes:
        opts["symmetryPlanes"] = [[list(p), list(n)] for p, n in planes]
    g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4, **opts))
    g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    return g


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["star", "sphere-1plane", "sphere-2planes"])
def test_deterministic_reverse_is_bit_reproducible(name, prec):
    m, planes = _case(name)
    print("case", name, prec, len(m["xv0"]), "volume nodes", flush=True)
    xs = m["perturb"](m["pts"])
    v = synthetic.random_dxv(len(m["xv0"]), seed=11, frac=0.6)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4,
                       symmetryPlanes=planes)
    ora.setSurfaceCoordinates(xs)
    ora.warpMesh()
    dref = ora.warpDeriv(v)
    runs = []
    for inst in range(2):  # two instances of the same mesh
        g = _mesh(m, planes, precision=prec, deterministic=True)
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        for _ in range(2):
            g.warpDeriv(v, solverVec=False)
            runs.append(g.getdXs().ravel().copy())
        g.warpMesh()  # reusePlan off: the plan is filled and transposed again
        g.warpDeriv(v, solverVec=False)
        runs.append(g.getdXs().ravel().copy())
    for r in runs[1:]:
        assert np.array_equal(_bits(r), _bits(runs[0]))
    atomic = _mesh(m, planes, precision=prec)
    atomic.setSurfaceCoordinates(xs)
    atomic.warpMesh()
    atomic.warpDeriv(v, solverVec=False)
    ra, ro = _rel(runs[0], atomic.getdXs().ravel()), _rel(runs[0], dref)
    print("deterministic vs atomic reverse rel %.2e, vs oracle rel %.2e" % (ra, ro), flush=True)
    assert ro <= (1e-10 if prec == "fp64" else 1e-6)
    # the mixed atomic reverse takes FP32 far weights, the deterministic one FP64
    assert ra <= (1e-10 if prec == "fp64" else 1e-6)


def test_deterministic_reverse_in_a_batch_and_without_allocation():
    """The recorded batch and idw_warp_deriv_dev_noalloc give the eager
    deterministic reverse's bits, on every replay; the allocation-free call
    before any deterministic reverse of the plan is a state error.  Neither
    the batch nor the autotune option searches the tile parameters of a
    deterministic mesh: the forward's chunking orders its sums, so a choice
    made by timing could change the bits from run to run."""
    import torch

    from idwarp_amd.batch import Batch

    m = synthetic.star_shells(n=17, nshell=20)
    g = _mesh(m, [], precision="mixed", deterministic=True, autotune=True)
    xs = m["perturb"](m["pts"])
    v = synthetic.random_dxv(len(m["xv0"]), seed=12, frac=1.0)
    L = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    d_xs = torch.tensor(xs.ravel(), dtype=torch.float64, device="cuda:0")
    d_v = torch.tensor(v, dtype=torch.float64, device="cuda:0")
    d_xv = torch.empty(3 * g.nVol, dtype=torch.float64, device="cuda:0")
    got = torch.full((3 * g.nSurf,), float("nan"), dtype=torch.float64, device="cuda:0")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(L.idw_warp_mesh_dev(g.ctx, vp(d_xs), vp(d_xv), stream))
    assert L.idw_warp_deriv_dev_noalloc(g.ctx, vp(d_xs), vp(d_v), vp(got), stream) == _lib.IDW_ESTATE
    g.setSurfaceCoordinates(xs)
    g.warpMesh()
    g.warpDeriv(v, solverVec=False)
    ref = g.getdXs().ravel().copy()
    _lib.check(L.idw_warp_mesh_dev(g.ctx, vp(d_xs), vp(d_xv), stream))
    _lib.check(L.idw_warp_deriv_dev(g.ctx, vp(d_xs), vp(d_v), vp(got), stream))
    torch.cuda.synchronize()
    eager = got.cpu().numpy().copy()
    assert np.array_equal(_bits(eager), _bits(ref))
    got.fill_(float("nan"))
    _lib.check(L.idw_warp_deriv_dev_noalloc(g.ctx, vp(d_xs), vp(d_v), vp(got), stream))
    torch.cuda.synchronize()
    assert np.array_equal(_bits(got.cpu().numpy()), _bits(ref))
    b = Batch()
    outs = [torch.empty(3 * g.nSurf, dtype=torch.float64, device="cuda:0") for _ in range(2)]
    for o in outs:
        b.add_warp_deriv(g, d_xs, d_v, o)
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        b.run()
        torch.cuda.synchronize()
        for o in outs:
            assert np.array_equal(_bits(o.cpu().numpy()), _bits(ref)), rep
    assert b.stats()[0] == 1
    # neither the autotune option nor the batch searched the tile parameters
    assert g.getTuning()[2] == "untuned"
