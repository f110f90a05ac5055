"""The measured configurations against the oracle on the device (needs an
MI355X), and the library's failure paths.

  * batched solves: four meshes of uneven size and mode, several surfaces and
    reverse seeds each, recorded into one HIP graph; every replay against the
    oracle (forward on sampled volume nodes, reverse for seeds carried by the
    sample), and the graph reused after eager calls;
  * the 20M-node configuration's recipe at 864k nodes in mixed precision at
    errTol 1e-3 and 1e-5 (and FP64), sampled against the oracle;
  * mixed precision far from the body (the ellipsoid O-mesh's outer field):
    the waves there keep FP64 terms and match the oracle to the FP64 bar;
  * the 3.2M-node ellipsoid O-mesh in exact mode, FP64 and mixed, sampled
    against the oracle, and its adjoint against central differences along 20
    random unit surface directions (the reference's check,
    verifyWarpDeriv.F90:119-141);
  * the surface whose root split depends on the 201-entry window;
  * mixed precision at the near-wall floor with volume nodes on the wall;
  * failure paths: calls out of order (IDW_ESTATE), non-finite input, an
    injected device-allocation failure, options the mixed mode rejects.

Tolerances: FP64 max|dXv| <= 1e-12 Ldef and adjoint relative error <= 1e-10;
mixed max|dXv| <= 1e-6 Ldef and adjoint relative error <= 1e-6.
"""
import ctypes
import os

import numpy as np
import pytest

from idwarp_amd import USMesh, _lib, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL, MIXED_TOL = 1e-12, 1e-6
O.set_threads(min(16, len(os.sched_getaffinity(0))))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _usmesh(m, **opts):
    g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", **opts))
    g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    return g


def _oracle(m, idx, **opts):
    o = {k: v for k, v in opts.items() if k not in ("precision", "nearWallFactor", "reusePlan")}
    return O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"][idx], **o)


def _sample_seed(nvol, idx, seed):
    """dXv ~ N(0,1) on the sampled nodes only: the oracle's reverse over the
    sample is then the whole reverse product."""
    g = np.random.default_rng(seed).standard_normal((len(idx), 3))
    full = np.zeros((nvol, 3))
    full[idx] = g
    return full.ravel(), g.ravel()


def test_c5_batch_of_uneven_meshes_against_oracle():
    """Four meshes (7,225 to 192,080 volume nodes; fast FP64, fast mixed,
    exact FP64), four surfaces Xs0 + a (Xs1 - Xs0) (a ~ U(0.5, 1.5), seeds
    100+p) and three reverse seeds (seeds 200+s) each, in one graph."""
    import torch

    from idwarp_amd.batch import Batch

    dev = torch.device("cuda:0")
    specs = [("sphere 25x17x17", synthetic.sphere_omesh(25, 17, 17), dict(evalMode="fast", errTol=1e-4)),
             ("ellipsoid 65x33x25", synthetic.ellipsoid_omesh(65, 33, 25),
              dict(evalMode="fast", errTol=1e-4, precision="mixed")),
             ("star n=41, 20 shells", synthetic.star_shells(41, 20),
              dict(evalMode="fast", errTol=1e-4, precision="mixed")),
             ("ellipsoid 33x17x13 exact", synthetic.ellipsoid_omesh(33, 17, 13), dict(evalMode="exact"))]
    cases = []
    b = Batch()
    for name, m, opts in specs:
        nvol = len(m["xv0"])
        idx = np.sort(np.random.default_rng(7).choice(nvol, min(nvol, 1500), replace=False))
        g = _usmesh(m, **opts)
        ora = _oracle(m, idx, **opts)
        base = m["pts"]
        delta = m["perturb"](base) - base
        surf = [base + np.random.default_rng(100 + p).uniform(0.5, 1.5) * delta for p in range(4)]
        seeds = [_sample_seed(nvol, idx, 200 + s) for s in range(3)]
        ref_xv = []
        for xs in surf:
            ora.setSurfaceCoordinates(xs)
            ref_xv.append(ora.warpMesh().reshape(-1, 3))
        ora.setSurfaceCoordinates(surf[0])
        ora.warpMesh()
        ref_dxs = [ora.warpDeriv(sd) for _, sd in seeds]
        d_xs = [torch.tensor(x.ravel(), dtype=torch.float64, device=dev) for x in surf]
        xv = [torch.empty(3 * nvol, dtype=torch.float64, device=dev) for _ in surf]
        dxs = [torch.empty(3 * len(base), dtype=torch.float64, device=dev) for _ in seeds]
        for s, o in zip(d_xs, xv):
            b.add_warp(g, s, o)
        for (full, _), o in zip(seeds, dxs):
            b.add_warp_deriv(g, d_xs[0], torch.tensor(full, dtype=torch.float64, device=dev), o)
        tol = MIXED_TOL if opts.get("precision") == "mixed" else FP64_TOL
        atol = MIXED_TOL if opts.get("precision") == "mixed" else 1e-10
        cases.append((name, g, idx, ref_xv, ref_dxs, xv, dxs, tol, atol, ora.Ldef0))
    for rep in range(3):
        for c in cases:
            for t in c[5] + c[6]:
                t.fill_(float("nan"))
        b.run()
        torch.cuda.synchronize()
        for name, g, idx, ref_xv, ref_dxs, xv, dxs, tol, atol, ldef in cases:
            worst = max(np.abs(t.view(-1, 3).cpu().numpy()[idx] - r).max() / ldef for r, t in zip(ref_xv, xv))
            radj = max(_rel(t.cpu().numpy(), r) for r, t in zip(ref_dxs, dxs))
            print("replay %d %s: max |dXv|/Ldef %.2e, adjoint rel %.2e" % (rep, name, worst, radj), flush=True)
            assert worst <= tol and radj <= atol, (rep, name)
        # eager work between replays leaves the recorded graph current
        g0 = cases[0][1]
        g0.setSurfaceCoordinates(specs[0][1]["pts"])
        g0.warpMesh()
        torch.cuda.synchronize()
    assert b.stats() == (1, 3)


def test_batch_graph_survives_eager_calls_and_option_pushes():
    """An eager warp (which pushes the mesh's options, reusePlan off) before
    and after a batch run does not make the batch record its graph again."""
    import torch

    from idwarp_amd.batch import Batch

    m = synthetic.star_shells(17, 6)
    g = _usmesh(m, evalMode="fast", errTol=1e-4)
    xs = m["perturb"](m["pts"])
    g.setSurfaceCoordinates(xs)
    g.warpMesh()
    ref = g.getWarpGrid().copy()
    b = Batch()
    d_xs = torch.tensor(xs.ravel(), dtype=torch.float64, device="cuda")
    out = torch.empty(3 * g.nVol, dtype=torch.float64, device="cuda")
    b.add_warp(g, d_xs, out)
    b.run()
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)
    for _ in range(2):
        g.warpMesh()
        assert not g.getOption("reusePlan")
        b.run()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref)
    assert b.stats() == (1, 3)
    g.setOption("errTol", 1e-3)  # a real change records again
    g._setMeshOptions()
    b.run()
    torch.cuda.synchronize()
    assert b.stats()[0] == 2


@pytest.mark.parametrize("tol", [1e-3, 1e-5])
def test_c3_recipe_mixed_errtol_sweep_against_oracle(tol):
    """The star body of the 20M-node configuration at 61-point cube-sphere
    edges with its 80 shells (1,728,160 volume nodes; the outer shells reach
    0.2, where the error test rejects clusters and errTol moves the
    decisions): mixed precision and FP64 against the oracle, same tree and
    errTol, on 2000 sampled nodes."""
    m = synthetic.star_shells(n=61, nshell=80)
    nvol = len(m["xv0"])
    idx = np.sort(np.random.default_rng(7).choice(nvol, 2000, replace=False))
    xs = m["perturb"](m["pts"])
    ora = _oracle(m, idx, evalMode="fast", errTol=tol)
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh().reshape(-1, 3)
    full, sd = _sample_seed(nvol, idx, 5)
    dref = ora.warpDeriv(sd)
    c = ora.counts()
    assert (c[:, 3] > c[:, 1]).sum() > 500  # many sampled nodes reject some cluster at this tolerance
    for prec in ("fp64", "mixed"):
        print("case C3 recipe errTol", tol, prec, flush=True)
        g = _usmesh(m, evalMode="fast", errTol=tol, precision=prec)
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        err = np.abs(g.getWarpGrid().reshape(-1, 3)[idx] - ref).max() / ora.Ldef0
        g.warpDeriv(full, solverVec=False)
        r = _rel(g.getdXs().ravel(), dref)
        print(prec, "sampled max |dXv|/Ldef %.2e, adjoint rel %.2e" % (err, r), flush=True)
        assert err <= (FP64_TOL if prec == "fp64" else MIXED_TOL)
        assert r <= (1e-10 if prec == "fp64" else MIXED_TOL)
        del g


def test_mixed_far_field_waves_evaluate_in_fp64():
    """The 10:1:1 ellipsoid O-mesh at 65x33x25 (its outer boundary 30
    semi-major lengths out), fast mode, mixed precision, four surfaces.
    Far from the body the terms' displacements at a node grow with the
    distance (rotations times up to 30) while their weighted mean stays
    small: FP32 weights and sums left errors near 1e-6 Ldef there.  A wave
    with a node more than 2 Ldef outside the surface's bounding sphere now
    keeps every term in FP64, so every sampled node that far out matches the
    oracle to the FP64 bar, and the whole sample stays far inside the mixed
    bar."""
    m = synthetic.ellipsoid_omesh(65, 33, 25)
    nvol = len(m["xv0"])
    idx = np.sort(np.random.default_rng(7).choice(nvol, 3000, replace=False))
    opts = dict(evalMode="fast", errTol=1e-4)
    ora = _oracle(m, idx, **opts)
    st = ora.surface_state()
    centre = st["Xu0"].mean(0)
    R = np.linalg.norm(st["Xu0"] - centre, axis=1).max()
    far = np.linalg.norm(m["xv0"][idx] - centre, axis=1) > R + 2.0 * ora.Ldef0 * (1 + 1e-9)
    assert 500 < far.sum() < len(idx) - 500
    g = _usmesh(m, precision="mixed", **opts)
    base = m["pts"]
    delta = m["perturb"](base) - base
    worst_far = worst = 0.0
    for p in range(4):
        xs = base + np.random.default_rng(100 + p).uniform(0.5, 1.5) * delta
        ora.setSurfaceCoordinates(xs)
        ref = ora.warpMesh().reshape(-1, 3)
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        e = np.abs(g.getWarpGrid().reshape(-1, 3)[idx] - ref).max(1) / ora.Ldef0
        worst_far, worst = max(worst_far, e[far].max()), max(worst, e.max())
    print("far nodes %d: max |dXv|/Ldef %.2e; all %d: %.2e" % (far.sum(), worst_far, len(idx), worst), flush=True)
    assert worst_far <= FP64_TOL
    assert worst <= 0.2 * MIXED_TOL


def test_c2_ellipsoid_exact_full_size_and_c4_finite_differences():
    """10:1:1 ellipsoid O-mesh 257x129x97 (3,215,841 volume nodes, 32,514
    unique surface nodes), exact mode.  FP64 and mixed forward on 2000
    sampled nodes and reverse for a seed carried by them, against the oracle;
    then the FP64 adjoint for dXv ~ N(0,1) on a random 10% of the nodes
    (seed 5) against central differences (h = 1e-6) along 20 random unit
    surface directions (seed 6), as verifyWarpDeriv.F90:119-141 does, and
    against the tangent through the dot-product identity.

    The FD quotient is formed from the difference of the two warped grids,
    whose entries (up to 30 in size) are rounded to ulp/2: that moves it by
    sigma = sqrt(sum dXv_i^2 (ulp(Xv+_i)^2 + ulp(Xv-_i)^2) / 12) / 2h, about
    3e-8 of it.  With the rotations on, the quotient also sees the rotation
    angle acos(n0 . n) (vectorUtils.f90:7-103) of nodes that barely turn:
    for an angle near 1e-6 the argument lies within 5e-13 of 1, so the
    angle is resolved to about 1e-4 of itself, a fine staircase that a 1e-6
    step samples and the derivative does not.  The h = 1e-6 quotient is held
    to 2e-5 relative there; with useRotations off (weights and translations
    only, the same kernels) it is held to 1e-6 relative plus 6 sigma."""
    m = synthetic.ellipsoid_omesh(257, 129, 97)
    nvol = len(m["xv0"])
    assert nvol == 3215841
    idx = np.sort(np.random.default_rng(7).choice(nvol, 2000, replace=False))
    xs = m["perturb"](m["pts"])
    ora = _oracle(m, idx, evalMode="exact")
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh().reshape(-1, 3)
    full, sd = _sample_seed(nvol, idx, 5)
    dref = ora.warpDeriv(sd)
    g = _usmesh(m, evalMode="exact")
    for prec in ("mixed", "fp64"):
        g.setOption("precision", prec)
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        err = np.abs(g.getWarpGrid().reshape(-1, 3)[idx] - ref).max() / ora.Ldef0
        g.warpDeriv(full, solverVec=False)
        r = _rel(g.getdXs().ravel(), dref)
        print("C2", prec, "sampled max |dXv|/Ldef %.2e, adjoint rel %.2e" % (err, r), flush=True)
        assert err <= (FP64_TOL if prec == "fp64" else MIXED_TOL)
        assert r <= (1e-10 if prec == "fp64" else MIXED_TOL)
    dxv = synthetic.random_dxv(nvol, seed=5, frac=0.1)
    h = 1e-6
    for rot in (True, False):
        g.setOption("useRotations", rot)
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        g.warpDeriv(dxv, solverVec=False)
        dxs = g.getdXs().ravel().copy()
        rng = np.random.default_rng(6)
        worst, worst_sig = 0.0, 0.0
        for k in range(20):
            p = rng.standard_normal(xs.size)
            p /= np.linalg.norm(p)
            grids = []
            for sgn in (1.0, -1.0):
                g.setSurfaceCoordinates((xs.ravel() + sgn * h * p).reshape(-1, 3))
                g.warpMesh()
                grids.append(g.getWarpGrid().copy())
            fd = dxv @ (grids[0] - grids[1]) / (2 * h)
            sigma = np.sqrt(np.sum(dxv ** 2 * (np.spacing(np.abs(grids[0])) ** 2 +
                                               np.spacing(np.abs(grids[1])) ** 2) / 12.0)) / (2 * h)
            ad = dxs @ p
            worst = max(worst, abs(fd - ad) / abs(ad))
            worst_sig = max(worst_sig, abs(fd - ad) / sigma)
            bound = 2e-5 * abs(ad) if rot else 1e-6 * abs(ad) + 6.0 * sigma
            assert abs(fd - ad) <= bound, (rot, k, fd, ad, sigma)
        print("C4 FD, rotations %s (20 directions, h=1e-6): max rel diff %.2e, max |diff| / sigma %.1f"
              % ("on" if rot else "off", worst, worst_sig), flush=True)
        # the exact identity <dXv, J p> = <J^T dXv, p> with the tangent kernels
        g.setSurfaceCoordinates(xs)
        p = rng.standard_normal(xs.size)
        jp = g.warpDerivFwd(p.reshape(-1, 3), solverVec=False)
        lhs, rhs = dxv @ jp, dxs @ p
        print("C4 dot product, rotations %s: rel %.2e" % ("on" if rot else "off", abs(lhs - rhs) / abs(lhs)),
              flush=True)
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # stored weights at this size would take 3,215,841 x 32,514 x 8 bytes
    # (836 GB): the warp reports the failed allocation, and the mesh warps
    # again once the option is off
    g.setOption("useRotations", True)
    g.setOption("storeWeights", True)
    g.setSurfaceCoordinates(xs)
    with pytest.raises(_lib.IdwError) as e:
        g.warpMesh()
    assert e.value.code == _lib.IDW_ENOMEM, e.value
    g.setOption("storeWeights", False)
    g.warpMesh()
    g.warpDeriv(full, solverVec=False)
    assert _rel(g.getdXs().ravel(), dref) <= 1e-10


def test_split_window_surface_fast_parity():
    """The surface whose root split takes the 201-entry window (x) rather than
    the whole range (z): device fast mode equals the oracle."""
    from tests.test_oracle_pins import _grid_faces
    from scipy.spatial import Delaunay

    i = np.arange(1, 201)
    win = np.stack([0.5 * ((37 * i) % 200) / 199, 0.75 * ((91 * i) % 200) / 199, i * 1e-9], 1)
    win = np.concatenate([win, [[1.0, 0.3, 201e-9], [0.2, 2.0, 202e-9]]])
    tri = Delaunay(win[:, :2]).simplices
    far = np.array([[0.1 * a, 0.5, 10.0 + b] for b in range(10) for a in range(10)])
    cf, sf = _grid_faces(10, 10, lambda a, b: len(win) + 10 * b + a, lambda a, b: 0)
    pts = np.concatenate([win, far])
    conn = np.concatenate([tri.astype(np.int32).ravel(), cf])
    sizes = np.concatenate([np.full(len(tri), 3, np.int32), sf])
    rng = np.random.default_rng(4)
    xv0 = np.concatenate([rng.uniform([-0.5, -0.5, 0.01], [1.5, 2.5, 20.0], (3000, 3)), pts])
    m = dict(pts=pts, conn=conn, faceSizes=sizes, xv0=xv0)
    xs = pts + 0.01 * np.sin(pts[:, [1, 2, 0]] * 3.0)
    ora = O.OracleMesh(pts, conn, sizes, xv0, evalMode="fast", errTol=1e-4)
    assert ora.tree()[0][0, 0] == 1
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    v = synthetic.random_dxv(len(xv0), seed=3, frac=0.5)
    dref = ora.warpDeriv(v)
    g = _usmesh(m, evalMode="fast", errTol=1e-4)
    g.setSurfaceCoordinates(xs)
    g.warpMesh()
    assert np.abs(g.getWarpGrid() - ref).max() <= FP64_TOL * ora.Ldef0
    g.warpDeriv(v, solverVec=False)
    assert _rel(g.getdXs().ravel(), dref) <= 1e-10


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_mixed_at_the_near_wall_floor_is_finite(mode):
    """Mixed precision with nearWallFactor at its floor (9.95e-8 for the
    default exponents) on a grid whose first layer is the wall: volume nodes
    coincide with surface points (d = 0, FP64 terms) and the nearest FP32
    pairs sit just above the floor, where q^5 ~ 1e35.  Every output is
    finite and within the mixed bar of the oracle."""
    m = dict(synthetic.sphere_omesh(25, 17, 9))
    # volume nodes on every surface point and 1.5e-7 Ldef off it (Ldef0 = 1)
    rng = np.random.default_rng(8)
    d = rng.standard_normal(m["pts"].shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    m["xv0"] = np.concatenate([m["xv0"], m["pts"] + 1.5e-7 * d])
    floor = _lib.load().idw_mixed_near_wall_floor(3.0, 5.0, 0.25)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode=mode, errTol=1e-4)
    g = _usmesh(m, evalMode=mode, errTol=1e-4, precision="mixed", nearWallFactor=floor)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    g.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    g.warpMesh()
    got = g.getWarpGrid()
    assert np.isfinite(got).all()
    err = np.abs(got - ref).max() / ora.Ldef0
    v = synthetic.random_dxv(len(m["xv0"]), seed=2, frac=1.0)
    g.warpDeriv(v, solverVec=False)
    dxs = g.getdXs().ravel()
    assert np.isfinite(dxs).all()
    r = _rel(dxs, ora.warpDeriv(v))
    print(mode, "floor %.3g: max |dXv|/Ldef %.2e, adjoint rel %.2e" % (floor, err, r), flush=True)
    assert err <= MIXED_TOL and r <= MIXED_TOL
    # below the floor the options are refused, by USMesh and by the library
    from idwarp_amd import Error

    with pytest.raises(Error):
        g.setOption("nearWallFactor", 0.5 * floor)
    o = g._copts()
    o.nearWallFactor = 0.0
    assert g._L.idw_set_options(g.ctx, ctypes.byref(o)) == _lib.IDW_EINVAL
    assert b"floor" in g._L.idw_last_error()


def test_calls_out_of_order_are_state_errors():
    """The reference aborts (EChk.F90:12-25) or reads a stale denominator
    (warpDeriv.F90:76); here every out-of-order call returns IDW_ESTATE and
    leaves the context usable."""
    L = _lib.load()
    o = _lib.IdwOptions()
    L.idw_default_options(ctypes.byref(o))
    ctx = ctypes.c_void_p(0)
    _lib.check(L.idw_create(ctypes.byref(ctx), 0, ctypes.byref(o)))
    try:
        buf = np.zeros(30)
        assert L.idw_warp_mesh(ctx) == _lib.IDW_ESTATE
        assert L.idw_warp_deriv(ctx, _lib.ptr(buf), 0) == _lib.IDW_ESTATE
        assert L.idw_get_dxs(ctx, _lib.ptr(buf), 30) == _lib.IDW_ESTATE
        assert L.idw_set_surface_coordinates(ctx, _lib.ptr(buf), 30) == _lib.IDW_ESTATE
        assert L.idw_warp_deriv_fwd(ctx, _lib.ptr(buf), 30, _lib.ptr(buf), 30) == _lib.IDW_ESTATE
        assert L.idw_get_denominator0(ctx, _lib.ptr(buf), 10) == _lib.IDW_ESTATE
        assert L.idw_get_ldef0(ctx, ctypes.byref(ctypes.c_double())) == _lib.IDW_ESTATE
        m = synthetic.sphere_omesh(13, 9, 5)
        xv0 = _lib.f64(m["xv0"])
        _lib.check(L.idw_set_volume_nodes(ctx, _lib.ptr(xv0), xv0.size))
        assert L.idw_warp_mesh(ctx) == _lib.IDW_ESTATE          # no surface yet
        dxv = np.zeros(xv0.size)
        assert L.idw_warp_deriv(ctx, _lib.ptr(dxv), dxv.size) == _lib.IDW_ESTATE
        assert b"initialize" in L.idw_last_error()
        dxs = np.zeros(3 * len(m["pts"]))
        assert L.idw_get_dxs(ctx, _lib.ptr(dxs), dxs.size) == _lib.IDW_ESTATE
    finally:
        L.idw_destroy(ctx)
    # a USMesh used in the same order raises and still works afterwards
    from idwarp_amd import Error

    g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast"))
    with pytest.raises(Exception):
        g.warpMesh()
    g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    g.setSurfaceCoordinates(m["pts"])
    g.warpMesh()
    assert np.array_equal(g.getWarpGrid(), m["xv0"].ravel())
    with pytest.raises(Error):
        g.setSurfaceCoordinates(m["pts"][:-1])


def test_non_finite_coordinates():
    """Host coordinates with NaN or inf are refused (IDW_EINVAL) at the
    volume nodes, the surface definition and the surface coordinates; the
    context keeps its previous state.  Device coordinates are not checked: a
    NaN surface point through the *_dev form gives NaN where it has weight,
    without a fault, and the next finite surface gives the oracle's grid."""
    import torch

    m = synthetic.sphere_omesh(13, 9, 5)
    g = _usmesh(m, evalMode="fast", errTol=1e-4)
    L, ctx = g._L, g.ctx
    bad = m["xv0"].copy()
    bad[5, 1] = np.nan
    assert L.idw_set_volume_nodes(ctx, _lib.ptr(_lib.f64(bad)), bad.size) == _lib.IDW_EINVAL
    assert b"not finite" in L.idw_last_error()
    xs = m["perturb"](m["pts"])
    xbad = xs.copy()
    xbad[3, 2] = np.inf
    with pytest.raises(Exception):
        g.setSurfaceCoordinates(xbad)
    g.setSurfaceCoordinates(xs)
    g.warpMesh()
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4)
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    assert np.abs(g.getWarpGrid() - ref).max() <= FP64_TOL * ora.Ldef0
    dev = torch.device("cuda:0")
    xn = xs.copy()
    xn[::7] = np.nan  # only the first copy of a duplicated point is read (kd_tree.F90:1488-1498)
    d_xs = torch.tensor(xn.ravel(), dtype=torch.float64, device=dev)
    d_xv = torch.empty(3 * g.nVol, dtype=torch.float64, device=dev)
    s = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(L.idw_warp_mesh_dev(ctx, ctypes.c_void_p(d_xs.data_ptr()), ctypes.c_void_p(d_xv.data_ptr()), s))
    torch.cuda.synchronize()
    assert torch.isnan(d_xv).any()
    d_xs.copy_(torch.tensor(xs.ravel(), dtype=torch.float64))
    _lib.check(L.idw_warp_mesh_dev(ctx, ctypes.c_void_p(d_xs.data_ptr()), ctypes.c_void_p(d_xv.data_ptr()), s))
    torch.cuda.synchronize()
    assert np.abs(d_xv.cpu().numpy() - ref).max() <= FP64_TOL * ora.Ldef0


@pytest.mark.parametrize("after", [3, 12, 30])
def test_device_allocation_failure_is_reported(after, monkeypatch):
    """IDW_TEST_FAIL_ALLOC_AFTER=n makes the context's (n+1)-th device
    allocation fail: set-up reports IDW_ENOMEM with a message instead of
    aborting, the context can be destroyed, and a fresh context without the
    hook works."""
    m = synthetic.sphere_omesh(13, 9, 5)
    monkeypatch.setenv("IDW_TEST_FAIL_ALLOC_AFTER", str(after))
    with pytest.raises(_lib.IdwError) as e:
        g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast"))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        g.setSurfaceCoordinates(m["pts"])
        g.warpMesh()
    assert e.value.code == _lib.IDW_ENOMEM and "injected" in str(e.value)
    monkeypatch.delenv("IDW_TEST_FAIL_ALLOC_AFTER")
    g = _usmesh(m, evalMode="fast")
    g.setSurfaceCoordinates(m["pts"])
    g.warpMesh()
    assert np.array_equal(g.getWarpGrid(), m["xv0"].ravel())


@pytest.mark.parametrize("case", ["fast-mixed", "fast-fp64", "exact-mixed", "exact-fp64"])
def test_autotuner_candidates_and_choice(case):
    """Every candidate of every tile parameter gives the oracle's warp and
    reverse within the precision's bar (star body with 80 shells, 9-point
    cube-sphere edges: 29,120 volume nodes; ellipsoid 33x17x13 in exact
    mode); the search times each applicable candidate, keeps the fastest,
    leaves its results within the bar and is kept for the same options."""
    mode, prec = case.split("-")
    m = synthetic.star_shells(n=9, nshell=80) if mode == "fast" else synthetic.ellipsoid_omesh(33, 17, 13)
    opts = dict(evalMode=mode, errTol=1e-4, precision=prec)
    ora = _oracle(m, np.arange(len(m["xv0"])), **opts)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    v = synthetic.random_dxv(len(m["xv0"]), seed=4, frac=0.5)
    dref = ora.warpDeriv(v)
    tol, atol = (MIXED_TOL, MIXED_TOL) if prec == "mixed" else (FP64_TOL, 1e-10)
    g = _usmesh(m, **opts)
    g.setSurfaceCoordinates(xs)
    cands = dict(USMesh.TUNING_SLOTS)
    for name, values in cands.items():
        for val in values:
            g.setTuning(**{name: val})
            g.warpMesh()
            err = np.abs(g.getWarpGrid() - ref).max() / ora.Ldef0
            g.warpDeriv(v, solverVec=False)
            r = _rel(g.getdXs().ravel(), dref)
            assert err <= tol and r <= atol, (name, val, err, r)
        g.setTuning(**{name: dict(forwardChunk=32, mergeWindow=32, walkBuffer=128, exactGroup=64,
                                  exactChunkShift=0)[name]})
    g.tune()
    chosen, table, state = g.getTuning()
    print(case, "tuning", chosen, table, flush=True)
    assert state == "current"
    timed = {"fast": ["walkBuffer", "forwardChunk"] + (["mergeWindow"] if prec == "mixed" else []),
             "exact": (["exactGroup"] if prec == "mixed" else []) + ["exactChunkShift"]}[mode]
    for name in timed:
        assert sorted(table[name]) == sorted(cands[name]) and min(table[name].values()) > 0
        assert chosen[name] == min(table[name], key=table[name].get)
    for name in set(cands) - set(timed):
        assert not table[name]
    g.warpMesh()
    assert np.abs(g.getWarpGrid() - ref).max() <= tol * ora.Ldef0
    g.warpDeriv(v, solverVec=False)
    assert _rel(g.getdXs().ravel(), dref) <= atol
    g.setOption("autotune", True)
    g.warpMesh()  # the choice is current: no new search
    assert g.getTuning()[1] == table
    g.setOption("errTol", 1e-3)
    g.warpMesh()  # other options: searched again
    assert g.getTuning()[2] == "current" and (mode == "exact" or g.getTuning()[1] != table)
