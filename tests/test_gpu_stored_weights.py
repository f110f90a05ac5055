"""Exact mode with stored weights (option storeWeights; needs an MI355X).

The reference recomputes every pair's weight in the reverse product
(evalNodeExact_b, kd_tree.F90:984-1013).  The weights follow from the
undeformed surface and volume and the options alone, so with storeWeights
the warp keeps each pair's weight, as it used it, and the reverse products
after it read them.  On the CPU-sized sphere O-mesh 49x33x33 (53,361 volume
nodes, 1,617 unique surface nodes), and on a half sphere with one symmetry
plane (two mirror passes):

  * the warp is the same as without the option, and the reverse product
    equals the recomputing one (FP64: relative 1e-11, the two differing in
    the last bits of the weights; this adjoint is itself 2e-13 from the
    oracle's; mixed: the stored
    weights are the forward's, whose FP32/FP64 choice is made per volume
    wave, where the recomputing reverse chooses per surface wave: 1e-6) and
    the oracle's (FP64 1e-10, mixed 1e-6), for dXv ~ N(0,1) on a random 10%
    of the nodes (seed 5, as the 3.2M-node adjoint configuration seeds it);
  * the weights stay valid across warps of new surfaces and are rewritten
    when an option that shapes them changes;
  * the device memory grows by exactly passes x n_unique x ld x 8 bytes
    (ld: n_vol rounded up to 4).
"""
import numpy as np
import pytest

from idwarp_amd import USMesh, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(planes):
    if planes == 0:
        return synthetic.sphere_omesh(49, 33, 33), []
    half = synthetic.shifted(synthetic.sphere_omesh(25, 17, 17), (0.3, 2.5, 3.2))
    return half, [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])]


def _mesh(m, planes, **opts):
    if planes:
        opts["symmetryPlanes"] = [[list(p), list(n)] for p, n in planes]
    g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="exact", **opts))
    g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    return g


def _unique(g):
    import ctypes

    n = ctypes.c_int64(0)
    g._L.idw_get_sizes(g.ctx, None, None, ctypes.byref(n), None)
    return n.value


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("planes", [0, 1])
def test_stored_weights_equal_recompute(planes, prec):
    m, pl = _case(planes)
    nvol = len(m["xv0"])
    print("case", prec, "planes", planes, nvol, "volume nodes", flush=True)
    tol_w = 1e-12 if prec == "fp64" else 1e-6
    tol_adj = 1e-10 if prec == "fp64" else 1e-6
    g0 = _mesh(m, pl, precision=prec)
    g1 = _mesh(m, pl, precision=prec, storeWeights=True)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="exact", symmetryPlanes=pl)
    base = m["pts"]
    delta = m["perturb"](base) - base
    v = synthetic.random_dxv(nvol, seed=5, frac=0.1)
    mem0 = g1.getDeviceMemory()
    for p, a in enumerate((1.0, 0.6)):  # two surfaces: the stored weights serve both
        xs = base + a * delta
        ora.setSurfaceCoordinates(xs)
        ref = ora.warpMesh()
        dref = ora.warpDeriv(v)
        outs = []
        for g in (g0, g1):
            g.setSurfaceCoordinates(xs)
            g.warpMesh()
            grid = g.getWarpGrid().copy()
            g.warpDeriv(v, solverVec=False)
            outs.append((grid, g.getdXs().ravel().copy(), g.getTimings()[4]))
        (x0, d0, t0), (x1, d1, t1) = outs
        e1 = np.abs(x1 - ref).max() / ora.Ldef0
        r01, r1 = _rel(d1, d0), _rel(d1, dref)
        print("surface %d: stored vs recompute adjoint rel %.2e, vs oracle rel %.2e, warp err %.2e Ldef; "
              "reverse kernels %.3f ms recomputing, %.3f ms stored" % (p, r01, r1, e1, t0, t1), flush=True)
        assert np.abs(x1 - x0).max() <= tol_w * ora.Ldef0 and e1 <= tol_w
        assert r01 <= (1e-11 if prec == "fp64" else 1e-6) and r1 <= tol_adj
    passes = 2 ** len(pl)
    grown = g1.getDeviceMemory() - g0.getDeviceMemory()
    ld = (nvol + 3) // 4 * 4
    print("stored weights: %d bytes (%d passes x %d x %d x 8); context %d bytes recomputing, %d storing"
          % (grown, passes, _unique(g1), ld, g0.getDeviceMemory(), g1.getDeviceMemory()), flush=True)
    assert grown == passes * ld * _unique(g1) * 8
    assert g1.getDeviceMemory() > mem0
    # an option that shapes the weights: the next reverse warps again first
    for g in (g0, g1):
        g.setOption("LdefFact", 1.5)
        g.warpDeriv(v, solverVec=False)
    assert _rel(g1.getdXs().ravel(), g0.getdXs().ravel()) <= (1e-11 if prec == "fp64" else 1e-6)
    # the option off gives the memory back at the next warp
    g1.setOption("storeWeights", False)
    g1.warpMesh()
    assert g1.getDeviceMemory() == g0.getDeviceMemory()
