"""HIP warp vs the CPU oracle on identical inputs (needs an MI355X).

Tolerances: FP64 mode max|dXv| <= 1e-12 Ldef and adjoint
relative error <= 1e-10 (the oracle and the device differ only in pow vs
repeated products, FMA contraction and summation order); mixed mode
max|dXv| <= 1e-6 Ldef and adjoint relative error <= 1e-6.
"""
import numpy as np
import pytest

from idwarp_amd import USMesh, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13
MIXED_TOL = 1e-6


def _meshes():
    return {
        "sphere": synthetic.sphere_omesh(13, 9, 9),
        "ellipsoid": synthetic.ellipsoid_omesh(33, 17, 13),
        "star": synthetic.star_shells(n=9, nshell=6),
    }


MESHES = _meshes()


def _pair(m, **opts):
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], **{k: v for k, v in opts.items()
                                                                         if k not in ("precision", "nearWallFactor")})
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", **opts))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    return ora, gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_forward_and_adjoint_fp64(name, mode):
    m = MESHES[name]
    print("case", name, mode, flush=True)
    ora, gpu = _pair(m, evalMode=mode, errTol=1e-4)
    ldef = ora.Ldef0
    assert abs(gpu.getLdef0() - ldef) <= 1e-15 * ldef
    np.testing.assert_allclose(gpu.getDenominator0(), ora.den0, rtol=1e-12)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    got = gpu.getWarpGrid()
    err = np.abs(got - ref).max() / ldef
    print("max |dXv|/Ldef", err)
    assert err <= FP64_TOL
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=5, frac=0.3)
    dref = ora.warpDeriv(dxv)
    gpu.warpDeriv(dxv, solverVec=False)
    dgot = gpu.getdXs().ravel()
    r = _rel(dgot, dref)
    print("adjoint rel err", r)
    assert r <= 1e-10


@pytest.mark.parametrize("name", ["sphere", "star"])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_mixed_precision(name, mode):
    m = MESHES[name]
    ora, gpu = _pair(m, evalMode=mode, errTol=1e-4, precision="mixed", nearWallFactor=1e-3)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    err = np.abs(gpu.getWarpGrid() - ref).max() / ora.Ldef0
    print("mixed max |dXv|/Ldef", err)
    assert err <= MIXED_TOL
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=7, frac=0.5)
    dref = ora.warpDeriv(dxv)
    gpu.warpDeriv(dxv, solverVec=False)
    r = _rel(gpu.getdXs().ravel(), dref)
    print("mixed adjoint rel err", r)
    assert r <= 1e-6


def test_nondefault_options():
    """Options of the reference's second ONERA M6 instance (test_MultiUSMesh.py:161-172)."""
    m = MESHES["sphere"]
    opts = dict(aExp=4.0, bExp=6.0, LdefFact=100.0, alpha=0.3, errTol=1e-5, evalMode="fast", useRotations=False,
                zeroCornerRotations=False, cornerAngle=45.0, bucketSize=4)
    ora, gpu = _pair(m, **opts)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    assert np.abs(gpu.getWarpGrid() - ref).max() <= FP64_TOL * ora.Ldef0
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=11, frac=1.0)
    dref = ora.warpDeriv(dxv)
    gpu.warpDeriv(dxv, solverVec=False)
    assert _rel(gpu.getdXs().ravel(), dref) <= 1e-10


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_closed_forms(mode):
    """Zero perturbation keeps the grid bit-exact; a rigid translation moves
    every volume node by the same vector."""
    m = MESHES["star"]
    gpu = USMesh(options=dict(gridFile=m["xv0"], evalMode=mode, errTol=1e-4))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    gpu.setSurfaceCoordinates(m["pts"])
    gpu.warpMesh()
    assert np.array_equal(gpu.getWarpGrid(), m["xv0"].ravel())
    t = np.array([0.013, -0.02, 0.031])
    gpu.setSurfaceCoordinates(m["pts"] + t)
    gpu.warpMesh()
    d = gpu.getWarpGrid().reshape(-1, 3) - m["xv0"]
    assert np.abs(d - t).max() <= 1e-14


def test_solver_grid_and_errors():
    m = MESHES["sphere"]
    gpu = USMesh(options=dict(gridFile=m["xv0"], evalMode="fast"))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    nvol = len(m["xv0"])
    rng = np.random.default_rng(0)
    ind = rng.integers(0, 3 * nvol, size=3 * nvol + 17).astype(np.int32)  # duplicates included
    gpu.setExternalMeshIndices(ind)
    gpu.setSurfaceCoordinates(m["perturb"](m["pts"]))
    gpu.warpMesh()
    warp = gpu.getWarpGrid()
    np.testing.assert_array_equal(gpu.getSolverGrid(), warp[ind])
    dsol = rng.standard_normal(len(ind))
    gpu.warpDeriv(dsol, solverVec=True)
    a = gpu.getdXs().copy()
    dwarp = np.zeros(3 * nvol)
    np.add.at(dwarp, ind, dsol)
    gpu.warpDeriv(dwarp, solverVec=False)
    # the fast reverse sums cluster/point adjoints with FP64 atomics, so two
    # runs agree to rounding (in norm), not bit for bit
    assert _rel(gpu.getdXs().ravel(), a.ravel()) <= 1e-10
    from idwarp_amd import Error
    with pytest.raises(Error):
        gpu.setSurfaceCoordinates(m["pts"][:-1])


def test_plan_reuse_and_errtol_change():
    """reusePlan keeps the walk's plan only while errTol is unchanged; results
    equal the per-call walk and the oracle at each tolerance."""
    m = MESHES["star"]
    ora, gpu = _pair(m, evalMode="fast", errTol=1e-4, reusePlan=True)
    xs = m["perturb"](m["pts"])
    for tol in (1e-4, 1e-3, 1e-5, 1e-4):
        ora.setOptions(errTol=tol)
        gpu.setOption("errTol", tol)
        for shift in (0.0, 0.01):
            x = xs + shift
            ora.setSurfaceCoordinates(x)
            gpu.setSurfaceCoordinates(x)
            ref = ora.warpMesh()
            gpu.warpMesh()
            assert np.abs(gpu.getWarpGrid() - ref).max() <= FP64_TOL * ora.Ldef0, (tol, shift)
            dxv = synthetic.random_dxv(len(m["xv0"]), seed=3, frac=0.5)
            dref = ora.warpDeriv(dxv)
            gpu.warpDeriv(dxv, solverVec=False)
            assert _rel(gpu.getdXs().ravel(), dref) <= 1e-10, (tol, shift)


def test_autograd_wrapper_matches_warp_and_deriv():
    """torch.autograd.Function: forward equals warpMesh, backward equals
    warpDeriv/getdXs, and gradcheck-style central differences agree."""
    import torch

    from idwarp_amd.autograd import warp

    m = MESHES["sphere"]
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    xs_np = m["perturb"](m["pts"])
    gpu.setSurfaceCoordinates(xs_np)
    gpu.warpMesh()
    ref = gpu.getWarpGrid()
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=9, frac=1.0)
    gpu.warpDeriv(dxv, solverVec=False)
    gref = gpu.getdXs().ravel()
    xs = torch.tensor(xs_np.ravel(), dtype=torch.float64, device="cuda", requires_grad=True)
    xv = warp(gpu, xs)
    np.testing.assert_array_equal(xv.detach().cpu().numpy(), ref)
    xv.backward(torch.tensor(dxv, dtype=torch.float64, device="cuda"))
    # FP64 atomics make the reverse's summation order vary run to run
    assert _rel(xs.grad.cpu().numpy(), gref) <= 1e-10
    # central differences along two random directions
    rng = np.random.default_rng(4)
    for _ in range(2):
        p = torch.tensor(rng.standard_normal(xs.numel()), dtype=torch.float64, device="cuda")
        h = 1e-6
        with torch.no_grad():
            fp = warp(gpu, xs + h * p)
            fm = warp(gpu, xs - h * p)
        g = torch.tensor(dxv, dtype=torch.float64, device="cuda")
        fd = float(torch.dot(g, (fp - fm) / (2 * h)))
        ad = float(torch.dot(xs.grad, p))
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_libidwarp_compat_follows_reference_call_sequence():
    """The f2py-shaped binding driven by the reference's own call sequence
    (UnstructuredMesh.py:196-262 setSurfaceDefinition, :1051-1069 options,
    :379-426 warpMesh/warpDeriv) gives the oracle's results."""
    from idwarp_amd.libidwarp_compat import LibIdwarp

    m = MESHES["ellipsoid"]
    w = LibIdwarp()
    w.initpetsc(0)
    w.gridinput.errtol = 1e-4
    w.gridinput.evalmode = w.gridinput.eval_fast
    w.kd_tree.bucket_size = 8
    w.createcommongrid(m["xv0"].T)
    w.setexternalmeshindices(np.arange(m["xv0"].size))
    allPts = m["pts"]
    uniquePts, link, nUnique = w.pointreduce(allPts.T, 1e-12)
    uniquePts = uniquePts.T[0:nUnique]
    ind = np.lexsort(uniquePts.T)
    uniquePts = uniquePts[ind]
    inv = np.zeros_like(ind)
    inv[ind] = np.arange(len(ind))
    link = inv[link - 1] + 1
    w.setsymmetryplanes(np.zeros((3, 0)), np.zeros((3, 0)))
    w.initializewarping(np.ravel(allPts), uniquePts.T, link, m["faceSizes"], np.ravel(m["conn"]), "")
    w.deallocatepatchio()
    xs = m["perturb"](m["pts"])
    w.setsurfacecoordinates(np.ravel(xs))
    w.warpmesh()
    xv = w.getvolumecoordinates(w.griddata.warpmeshdof)
    solver = np.zeros(w.griddata.solvermeshdof)
    w.warp_to_solver_grid(xv, solver)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4)
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    assert np.abs(xv - ref).max() <= FP64_TOL * ora.Ldef0
    assert np.array_equal(solver, xv)
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=5, frac=0.1)
    dxvw = np.zeros(w.griddata.warpmeshdof)
    w.solver_to_warp_grid(dxv, dxvw)
    w.warpderiv(dxvw)
    dxs = np.zeros(allPts.size)
    w.getdxs(dxs)
    assert _rel(dxs, ora.warpDeriv(dxv)) <= 1e-10


@pytest.mark.parametrize("name", ["sphere", "star"])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_tangent_fp64_and_dot_product(name, mode):
    """warpDerivFwd against the oracle's (warpDerivFwd.F90:62-99), and the
    device dot-product identity <v, J p> == <J^T v, p> between the tangent
    and the reverse kernels."""
    m = MESHES[name]
    print("case", name, mode, flush=True)
    ora, gpu = _pair(m, evalMode=mode, errTol=1e-4)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ora.warpMesh()
    rng = np.random.default_rng(12)
    p = rng.standard_normal(xs.size)
    ref = ora.warpDerivFwd(p)
    got = gpu.warpDerivFwd(p.reshape(-1, 3), solverVec=False)
    r = _rel(got, ref)
    print("tangent rel err", r, "max", np.abs(got - ref).max() / np.abs(ref).max())
    assert r <= 1e-10
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    v = rng.standard_normal(got.size)
    gpu.warpDeriv(v, solverVec=False)
    jtv = gpu.getdXs().ravel()
    assert abs(v @ got - jtv @ p) <= 1e-9 * abs(v @ got)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_tangent_mixed(mode):
    m = MESHES["star"]
    ora, gpu = _pair(m, evalMode=mode, errTol=1e-4, precision="mixed", nearWallFactor=1e-3)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ora.warpMesh()
    p = np.random.default_rng(13).standard_normal(xs.size)
    ref = ora.warpDerivFwd(p)
    got = gpu.warpDerivFwd(p, solverVec=False)
    assert _rel(got, ref) <= 1e-6


def test_autograd_forward_mode():
    """torch forward-mode AD through WarpFunction.jvp equals warpDerivFwd."""
    import torch
    import torch.autograd.forward_ad as fwAD

    from idwarp_amd.autograd import warp

    m = MESHES["ellipsoid"]
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    xs_np = m["perturb"](m["pts"])
    gpu.setSurfaceCoordinates(xs_np)
    p = np.random.default_rng(14).standard_normal(xs_np.size)
    ref = gpu.warpDerivFwd(p, solverVec=False)
    xs = torch.tensor(xs_np.ravel(), dtype=torch.float64, device="cuda")
    with fwAD.dual_level():
        dual = fwAD.make_dual(xs, torch.tensor(p, dtype=torch.float64, device="cuda"))
        out = warp(gpu, dual)
        primal, tangent = fwAD.unpack_dual(out)
    np.testing.assert_array_equal(primal.cpu().numpy(), gpu.getWarpGrid())
    assert _rel(tangent.cpu().numpy(), ref) <= 1e-12


def _symmetric(nplanes):
    half = synthetic.shifted(synthetic.sphere_omesh(13, 9, 9), (0.3, 2.5, 3.2))
    planes = [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0]), ([0.0, 0.0, 1.5], [0.0, 0.0, 1.0])][:nplanes]
    return half, planes


@pytest.mark.parametrize("nplanes", [1, 2])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_symmetry_planes_fp64(mode, nplanes):
    """Symmetry planes (initializeWarping.F90:511-659): denominator
    estimate, warp, reverse and tangent over the 2^n mirror passes."""
    m, planes = _symmetric(nplanes)
    print("case", mode, nplanes, flush=True)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode=mode, errTol=1e-4,
                       symmetryPlanes=planes)
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode=mode, errTol=1e-4,
                              symmetryPlanes=[[list(p), list(n)] for p, n in planes]))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    np.testing.assert_allclose(gpu.getDenominator0(), ora.den0, rtol=1e-12)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    err = np.abs(gpu.getWarpGrid() - ref).max() / ora.Ldef0
    print("max |dXv|/Ldef", err)
    assert err <= FP64_TOL
    v = synthetic.random_dxv(len(m["xv0"]), seed=6, frac=0.5)
    gpu.warpDeriv(v, solverVec=False)
    r = _rel(gpu.getdXs().ravel(), ora.warpDeriv(v))
    print("adjoint rel err", r)
    assert r <= 1e-10
    p = np.random.default_rng(15).standard_normal(xs.size)
    t = _rel(gpu.warpDerivFwd(p, solverVec=False), ora.warpDerivFwd(p))
    print("tangent rel err", t)
    assert t <= 1e-10


def test_symmetry_planes_mixed_and_removal():
    m, planes = _symmetric(1)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4,
                       symmetryPlanes=planes)
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4, precision="mixed",
                              nearWallFactor=1e-3, symmetryPlanes=[[list(p), list(n)] for p, n in planes]))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    assert np.abs(gpu.getWarpGrid() - ref).max() <= MIXED_TOL * ora.Ldef0
    # removing the planes through the C ABI gives the plain warp again
    plain = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4)
    plain.setSurfaceCoordinates(xs)
    ref0 = plain.warpMesh()
    from idwarp_amd import _lib

    _lib.check(gpu._L.idw_set_symmetry_planes(gpu.ctx, None, None, 0))
    gpu.setOption("precision", "fp64")
    gpu.warpMesh()
    assert np.abs(gpu.getWarpGrid() - ref0).max() <= FP64_TOL * plain.Ldef0
    np.testing.assert_allclose(gpu.getDenominator0(), plain.den0, rtol=1e-12)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_c1_sphere_omesh_49x33x33(mode):
    """The reference's CPU-runnable configuration: spherical O-mesh 49x33x33
    (53,361 volume nodes, 1,617 wall points), normal bump plus 2 degree
    rotation; exact, then fast with errTol 1e-4; FP64 and mixed against the
    oracle."""
    O.set_threads(8)
    m = synthetic.sphere_omesh(49, 33, 33)
    assert len(m["xv0"]) == 53361 and len(m["pts"]) == 1617
    print("case C1", mode, flush=True)
    ora, gpu = _pair(m, evalMode=mode, errTol=1e-4)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    err = np.abs(gpu.getWarpGrid() - ref).max() / ora.Ldef0
    print("fp64 max |dXv|/Ldef", err)
    assert err <= FP64_TOL
    dxv = synthetic.random_dxv(len(m["xv0"]), seed=5, frac=0.1)
    dref = ora.warpDeriv(dxv)
    gpu.warpDeriv(dxv, solverVec=False)
    r = _rel(gpu.getdXs().ravel(), dref)
    print("fp64 adjoint rel err", r)
    assert r <= 1e-10
    gpu.setOption("precision", "mixed")
    gpu.warpMesh()
    errm = np.abs(gpu.getWarpGrid() - ref).max() / ora.Ldef0
    gpu.warpDeriv(dxv, solverVec=False)
    rm = _rel(gpu.getdXs().ravel(), dref)
    print("mixed max |dXv|/Ldef", errm, "adjoint", rm)
    assert errm <= MIXED_TOL and rm <= MIXED_TOL


def test_batch_graph_replays_match_eager():
    """Graph-recorded batch over two meshes (several surfaces each, reverse
    products): every replay, with eager work between replays, equals the
    eager per-call results."""
    import torch

    from idwarp_amd.batch import Batch

    dev = torch.device("cuda:0")
    cases = []
    for name, opts in (("sphere", dict(evalMode="fast", errTol=1e-4)),
                       ("star", dict(evalMode="fast", errTol=1e-4, precision="mixed", nearWallFactor=1e-3))):
        m = MESHES[name]
        g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", **opts))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        base = m["pts"]
        delta = m["perturb"](base) - base
        surfaces = [base + a * delta for a in (0.5, 1.0, 1.5)]
        seeds = [synthetic.random_dxv(len(m["xv0"]), seed=s, frac=0.2) for s in (1, 2)]
        ref_xv, ref_dxs = [], []
        for xs in surfaces:
            g.setSurfaceCoordinates(xs)
            g.warpMesh()
            ref_xv.append(g.getWarpGrid())
        for sd in seeds:
            g.setSurfaceCoordinates(surfaces[1])
            g.warpMesh()
            g.warpDeriv(sd, solverVec=False)
            ref_dxs.append(g.getdXs().ravel())
        cases.append((g, surfaces, seeds, ref_xv, ref_dxs))
    b = Batch()
    outs = []
    for g, surfaces, seeds, _, _ in cases:
        d_xs = [torch.tensor(x.ravel(), dtype=torch.float64, device=dev) for x in surfaces]
        xv = [torch.empty(3 * g.nVol, dtype=torch.float64, device=dev) for _ in surfaces]
        dxs = [torch.empty(3 * g.nSurf, dtype=torch.float64, device=dev) for _ in seeds]
        for s, o in zip(d_xs, xv):
            b.add_warp(g, s, o)
        for sd, o in zip(seeds, dxs):
            b.add_warp_deriv(g, d_xs[1], torch.tensor(sd, dtype=torch.float64, device=dev), o)
        outs.append((xv, dxs))
    for rep in range(3):
        for xv, dxs in outs:
            for t in xv + dxs:
                t.fill_(float("nan"))
        b.run()
        torch.cuda.synchronize()
        for (g, _, _, ref_xv, ref_dxs), (xv, dxs) in zip(cases, outs):
            for r, t in zip(ref_xv, xv):
                assert np.array_equal(t.cpu().numpy(), r), rep
            for r, t in zip(ref_dxs, dxs):
                # FP64 atomics make the reverse's summation order vary run to run
                assert _rel(t.cpu().numpy(), r) <= 1e-10, rep
        # eager work between replays
        g0 = cases[0][0]
        g0.setSurfaceCoordinates(cases[0][1][0])
        g0.warpMesh()
        torch.cuda.synchronize()
    layout, ms = b.info()
    print("batch layout", layout, "replay ms", ms)
    assert layout in (0, 1) and min(ms) > 0


def test_restart_file_roundtrip(tmp_path):
    """initializeWarping.F90:98-312 restart file: written when missing (count,
    estimate, costs), loaded when it matches, rewritten when its count or its
    estimate disagree."""
    m = MESHES["star"]
    path = str(tmp_path / "restart.bin")
    opts = dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4, restartFile=path)
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4)
    g1 = USMesh(options=opts)
    g1.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    status, costs = g1.getRestartInfo()
    n = len(m["xv0"])
    raw = open(path, "rb").read()
    assert status == "saved" and len(raw) == 4 + 16 * n
    assert np.frombuffer(raw[:4], np.int32)[0] == n
    fden = np.frombuffer(raw[4:4 + 8 * n], np.float64)
    fcost = np.frombuffer(raw[4 + 8 * n:], np.float64)
    assert np.array_equal(fden, g1.getDenominator0()) and np.array_equal(fcost, costs)
    np.testing.assert_allclose(fcost, ora.costs(), rtol=1e-12)
    xs = m["perturb"](m["pts"])
    g1.setSurfaceCoordinates(xs)
    g1.warpMesh()
    # loaded
    g2 = USMesh(options=opts)
    g2.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    assert g2.getRestartInfo()[0] == "loaded"
    assert np.array_equal(g2.getDenominator0(), fden)
    g2.setSurfaceCoordinates(xs)
    g2.warpMesh()
    assert np.array_equal(g2.getWarpGrid(), g1.getWarpGrid())
    # an estimate off by 1e-6 is rejected and the file rewritten
    bad = np.frombuffer(raw, np.uint8).copy()
    den_bad = (fden * (1 + 1e-6)).astype(np.float64)
    bad[4:4 + 8 * n] = np.frombuffer(den_bad.tobytes(), np.uint8)
    open(path, "wb").write(bad.tobytes())
    g3 = USMesh(options=opts)
    g3.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    assert g3.getRestartInfo()[0] == "saved"
    assert np.array_equal(np.frombuffer(open(path, "rb").read()[4:4 + 8 * n], np.float64), fden)
    # a file for another mesh size is rewritten
    open(path, "wb").write(np.int32(n + 1).tobytes() + bytes(16 * (n + 1)))
    g4 = USMesh(options=opts)
    g4.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    assert g4.getRestartInfo()[0] == "saved"


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_ahmed_body_openfoam_with_symmetry_plane(tmp_path, mode):
    """The reference's Ahmed-body example (examples/AhmedBodyCoarse: an
    OpenFOAM case read through fileType='OpenFOAM', warped about its patch,
    wall and slip faces with a y=0 symmetry plane and the example's options)
    against the oracle."""
    from tests.test_openfoam_cpu import ahmed_case

    from idwarp_amd import openfoam

    case = ahmed_case(tmp_path)
    opts = dict(gridFile=case, fileType="OpenFOAM", symmetryPlanes=[[[0, 0, 0], [0, 1, 0]]], aExp=3.0, bExp=5.0,
                alpha=1.0, LdefFact=0.20, evalMode=mode)
    gpu = USMesh(options=opts)
    coords0 = gpu.getSurfaceCoordinates()  # builds the internal OpenFOAM surface
    of = openfoam.read_polymesh(case)
    pts, conn, fs = openfoam.warp_surface(of)
    assert np.array_equal(coords0, pts)
    ora = O.OracleMesh(pts, conn, fs, of["x0"], evalMode=mode, aExp=3.0, bExp=5.0, alpha=1.0, LdefFact=0.20,
                       symmetryPlanes=[([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])])
    lo, hi = of["x0"].min(0), of["x0"].max(0)
    on_box = (np.isclose(pts[:, 0], lo[0]) | np.isclose(pts[:, 0], hi[0]) | np.isclose(pts[:, 2], lo[2]) |
              np.isclose(pts[:, 2], hi[2]) | np.isclose(pts[:, 1], hi[1]))
    xs = pts.copy()
    xs[~on_box, 0] += 0.005 * np.sin(3 * pts[~on_box, 0] + 1)
    xs[~on_box, 2] += 0.01 * np.cos(2 * pts[~on_box, 0])
    ora.setSurfaceCoordinates(xs)
    gpu.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    gpu.warpMesh()
    err = np.abs(gpu.getWarpGrid() - ref).max() / ora.Ldef0
    print("Ahmed", mode, "max |dXv|/Ldef", err, "moved surface points", int((~on_box).sum()))
    assert err <= FP64_TOL
    assert np.abs(ref - of["x0"].ravel()).max() > 1e-3
    v = synthetic.random_dxv(len(of["x0"]), seed=11, frac=0.5)
    gpu.warpDeriv(v, solverVec=False)
    r = _rel(gpu.getdXs().ravel(), ora.warpDeriv(v))
    print("adjoint rel err", r)
    assert r <= 1e-10


def test_plan_writing_warp_equals_plan_replay():
    """A warp that writes its plan (the walk, then the forward over it) and a
    warp that replays the stored plan (reusePlan) give the same volume
    nodes bit for bit and the same dXs to 1e-12 (star-shaped body, 61-point
    cube-sphere edges, 40 shells, 864,080 volume nodes), and the warp
    matches the oracle on a sample of 2000 volume nodes (FP64), with the
    mixed result within the mixed bar."""
    O.set_threads(8)
    m = synthetic.star_shells(n=61, nshell=40)
    nvol = len(m["xv0"])
    xs = m["perturb"](m["pts"])
    dxv = synthetic.random_dxv(nvol, seed=5, frac=0.1)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(nvol, 2000, replace=False))
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"][idx], evalMode="fast", errTol=1e-4)
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    ldef = ora.Ldef0
    for prec in ("fp64", "mixed"):
        print("case plan writing vs replay", prec, flush=True)
        gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4,
                                  precision=prec))
        gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        gpu.setSurfaceCoordinates(xs)
        gpu.warpMesh()  # walk + forward
        x1 = gpu.getWarpGrid().copy()
        gpu.warpDeriv(dxv, solverVec=False)
        d1 = gpu.getdXs().ravel().copy()
        gpu.setOption("reusePlan", True)
        gpu.warpMesh()  # the same plan, one forward launch
        x2 = gpu.getWarpGrid().copy()
        gpu.warpDeriv(dxv, solverVec=False)
        d2 = gpu.getdXs().ravel().copy()
        assert np.array_equal(x1, x2)
        assert _rel(d1, d2) <= 1e-12
        err = np.abs(x1.reshape(-1, 3)[idx] - ref.reshape(-1, 3)).max() / ldef
        print(prec, "sampled max |dXv|/Ldef", err)
        assert err <= (FP64_TOL if prec == "fp64" else MIXED_TOL)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_plot3d_grid_and_surface_files(tmp_path, mode):
    """USMesh driven by PLOT3D files as the reference's PLOT3D users drive
    it (fileType='PLOT3D', setSurfaceDefinitionFromFile, setSurfaceFromFile,
    writeGrid): a 13x9x9 sphere O-mesh in two volume zones that share their
    k=4 layer, its wall in two surface zones that share a row.  The warp must
    equal an in-memory USMesh on the same nodes bit for bit and the oracle
    within the FP64 bar; the written grid reads back as the warped grid."""
    from idwarp_amd import plot3d

    ni, nj, nk = 13, 9, 9
    m = synthetic.sphere_omesh(ni, nj, nk)
    lay = ni * nj
    grid = str(tmp_path / "grid.xyz")
    plot3d.write_grid(grid, [[ni, nj, 5], [ni, nj, 5]], np.concatenate([m["xv0"][:5 * lay], m["xv0"][4 * lay:]]))

    def wall_file(path, p):
        w = p.reshape(nj, ni, 3)
        plot3d.write_surface(path, [[ni, 5], [ni, 5]], np.concatenate([w[:5].reshape(-1, 3), w[4:].reshape(-1, 3)]))

    surf0, surf1 = str(tmp_path / "wall0.xyz"), str(tmp_path / "wall1.xyz")
    wall_file(surf0, m["pts"])
    wall_file(surf1, m["perturb"](m["pts"]))

    gpu = USMesh(options=dict(gridFile=grid, fileType="PLOT3D", evalMode=mode))
    gpu.setSurfaceDefinitionFromFile(surf0)
    gpu.setSurfaceFromFile(surf1)
    gpu.warpMesh()

    _, xvol = plot3d.read_grid(grid)
    pts, conn, _ = plot3d.read_surface(surf0)
    fs = np.full(len(conn), 4, np.int32)
    xs = plot3d.read_surface(surf1)[0]
    mem = USMesh(options=dict(gridFile=xvol, fileType="memory", evalMode=mode))
    mem.setSurfaceDefinition(pts, conn, fs)
    mem.setSurfaceCoordinates(xs)
    mem.warpMesh()
    ora = O.OracleMesh(pts, conn.ravel(), fs, xvol, evalMode=mode)
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()

    got = gpu.getWarpGrid()
    assert np.array_equal(got, mem.getWarpGrid())
    err = np.abs(got - ref).max() / ora.Ldef0
    print("PLOT3D", mode, "max |dXv|/Ldef", err, flush=True)
    assert err <= FP64_TOL
    assert np.abs(got - xvol.ravel()).max() > 1e-3
    out = str(tmp_path / "warped.xyz")
    gpu.writeGrid(out)
    sizes, back = plot3d.read_grid(out)
    assert sizes.tolist() == [[ni, nj, 5], [ni, nj, 5]]
    assert np.array_equal(back, got.reshape(-1, 3))
    # the shared layer moves as one (fast mode sums about a per-wave origin, so to rounding)
    assert np.abs(back[4 * lay:5 * lay] - back[5 * lay:6 * lay]).max() <= FP64_TOL * ora.Ldef0
    with pytest.raises(Exception):
        gpu.writeGrid()


@pytest.mark.parametrize("planes", [0, 1])
def test_mixed_adjoint_of_wave_pairs(planes):
    """The mixed-precision reverse merges the FP32 plans of neighbouring walk
    waves and adds each shared term once.  Star-shaped body with 59,422 volume
    nodes (929 walk waves: the last pair has one wave), with and without a
    symmetry plane: the adjoint must equal the FP64 reverse (per-wave path,
    same plan) and the oracle within the mixed-precision bar, and a tangent
    dot product must hold."""
    m = synthetic.star_shells(n=31, nshell=11)
    assert (len(m["xv0"]) + 63) // 64 == 929
    sym = [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])] if planes else None
    opts = dict(evalMode="fast", errTol=1e-4)  # default near-wall threshold: most terms are FP32
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], symmetryPlanes=sym, evalMode="fast", errTol=1e-4)
    res = {}
    for prec in ("fp64", "mixed"):
        gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", precision=prec, symmetryPlanes=sym, **opts))
        gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        xs = m["perturb"](m["pts"])
        gpu.setSurfaceCoordinates(xs)
        gpu.warpMesh()
        res[prec + "_xv"] = gpu.getWarpGrid().copy()
        dxv = synthetic.random_dxv(len(m["xv0"]), seed=9, frac=0.7)
        gpu.warpDeriv(dxv, solverVec=False)
        res[prec] = (gpu, dxv, gpu.getdXs().ravel().copy())
    ora.setSurfaceCoordinates(m["perturb"](m["pts"]))
    ora.warpMesh()
    dref = ora.warpDeriv(res["mixed"][1])
    r64 = _rel(res["fp64"][2], dref)
    rmix = _rel(res["mixed"][2], res["fp64"][2])
    rora = _rel(res["mixed"][2], dref)
    print("planes", planes, "fp64 vs oracle", r64, "mixed vs fp64", rmix, "mixed vs oracle", rora, flush=True)
    fwd = np.abs(res["mixed_xv"] - res["fp64_xv"]).max() / gpu.getLdef0()
    print("planes", planes, "forward mixed vs fp64 max |dXv|/Ldef", fwd, flush=True)
    assert fwd <= 1e-6
    assert r64 <= 1e-10
    assert rmix <= 1e-6 and rora <= 1e-6
    assert rmix >= 1e-12  # the FP32 path (and so the pair merge) did run
    gpu, dxv, jtv = res["mixed"]
    p = np.random.default_rng(10).standard_normal(jtv.size)
    jp = gpu.warpDerivFwd(p.reshape(-1, 3), solverVec=False)
    assert abs(dxv @ jp - jtv @ p) <= 1e-6 * abs(dxv @ jp)


@pytest.mark.parametrize("planes", [False, True])
def test_mixed_adjoint_sparse_and_dense_seed_groups(planes):
    """The mixed reverse compacts away the volume nodes whose seed is zero and
    takes groups of four walk waves with at most 12 seeds per 64 nodes in the
    group reverse (one atomic set per term of the four waves' union), the
    others in the pair reverse.  Same star body as above: seeds dense where
    x > 0.2 and on 5% of the nodes elsewhere, so one launch runs both; then
    a seed on 1% of the nodes, and a zero seed.  The adjoint must equal the
    FP64 reverse and the oracle within the mixed bar, and the tangent's dot
    product must hold."""
    m = synthetic.star_shells(n=31, nshell=11)
    xv0 = m["xv0"]
    sym = [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])] if planes else None
    opts = dict(evalMode="fast", errTol=1e-4)
    rng = np.random.default_rng(11)
    dense = xv0[:, 0] > 0.2
    mixed_seed = rng.standard_normal(xv0.shape)
    mixed_seed[~dense & (rng.random(len(xv0)) >= 0.05)] = 0.0
    assert dense.sum() > 5000 and (~dense).sum() > 5000
    seeds = {"half dense": mixed_seed.ravel(), "1%": synthetic.random_dxv(len(xv0), seed=12, frac=0.01),
             "zero": np.zeros(3 * len(xv0))}
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], xv0, symmetryPlanes=sym, **opts)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    ora.warpMesh()
    gpus = {}
    for prec in ("fp64", "mixed"):
        g = USMesh(options=dict(gridFile=xv0, fileType="memory", precision=prec, symmetryPlanes=sym, **opts))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        gpus[prec] = g
    for name, dxv in seeds.items():
        dref = ora.warpDeriv(dxv)
        got = {}
        for prec, g in gpus.items():
            g.warpDeriv(dxv, solverVec=False)
            got[prec] = g.getdXs().ravel().copy()
        if name == "zero":
            assert not np.any(got["mixed"]) and not np.any(got["fp64"]) and not np.any(dref)
            continue
        r64, rmix, rora = _rel(got["fp64"], dref), _rel(got["mixed"], got["fp64"]), _rel(got["mixed"], dref)
        print("planes", planes, name, "fp64 vs oracle %.2e, mixed vs fp64 %.2e, mixed vs oracle %.2e" %
              (r64, rmix, rora), flush=True)
        assert r64 <= 1e-10
        assert rmix <= 1e-6 and rora <= 1e-6
        assert rmix >= 1e-12  # the FP32 paths ran
        p = np.random.default_rng(13).standard_normal(got["mixed"].size)
        jp = gpus["mixed"].warpDerivFwd(p.reshape(-1, 3), solverVec=False)
        assert abs(dxv @ jp - got["mixed"] @ p) <= 1e-6 * abs(dxv @ jp)


@pytest.mark.parametrize("nplanes", [0, 1, 2])
def test_exact_mixed_warp_on_the_matrix_cores(nplanes):
    """The mixed exact warp sums every surface point on the matrix cores, a
    chunk at a time, and re-scans a chunk holding a pair nearer than the
    near-wall distance to add that pair in FP64.  Volume nodes are placed on
    every surface point (d = 0: near pairs in every pass) and 2e-6 off them,
    beyond the default near-wall distance (FP32 pairs whose r is tiny against
    the wave's extent).  For 0, 1 and 2 symmetry planes and each chunk size
    the warp and the reverse must stay within the mixed bar of the oracle."""
    m, planes = _symmetric(nplanes)
    m = dict(m)
    rng = np.random.default_rng(41)
    d = rng.standard_normal(m["pts"].shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    m["xv0"] = np.concatenate([m["xv0"], m["pts"], m["pts"] + 2.0e-6 * d])
    sym = [[list(p), list(n)] for p, n in planes] or None
    ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="exact", symmetryPlanes=planes or None)
    xs = m["perturb"](m["pts"])
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    v = synthetic.random_dxv(len(m["xv0"]), seed=42, frac=0.5)
    dref = ora.warpDeriv(v)
    gpu = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="exact", precision="mixed",
                              symmetryPlanes=sym))
    gpu.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    gpu.setSurfaceCoordinates(xs)
    for ch in (16, 32, 64):
        gpu.setTuning(exactGroup=ch)
        gpu.warpMesh()
        got = gpu.getWarpGrid()
        err = np.abs(got - ref).max() / ora.Ldef0
        gpu.warpDeriv(v, solverVec=False)
        r = _rel(gpu.getdXs().ravel(), dref)
        print("planes", nplanes, "chunk", ch, "max |dXv|/Ldef %.2e, adjoint rel %.2e" % (err, r), flush=True)
        assert np.isfinite(got).all() and err <= MIXED_TOL and r <= MIXED_TOL


@pytest.mark.parametrize("planes", [False, True])
def test_exact_adjoint_over_packed_seeded_nodes(planes):
    """The exact reverse runs over the volume nodes with a nonzero seed only,
    packed in traversal order by a count and an exclusive scan per block of
    256 nodes.  On a 309k-node ellipsoid (1,207 blocks: the scan takes two
    rounds of 1,024) a seed carried by 1,500 sampled nodes must give the
    oracle's reverse over those nodes (the other nodes add nothing), in FP64
    and mixed precision; so must a seed on the last node alone.  A zero seed
    gives zero.  (Dense seeds: the exact cases of the tests above.)"""
    m = synthetic.ellipsoid_omesh(97, 65, 49)
    xv0 = m["xv0"]
    n = len(xv0)
    assert (n + 255) // 256 > 1024
    sym = [([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])] if planes else None
    xs = m["perturb"](m["pts"])
    rng = np.random.default_rng(31)
    cases = {"sampled": np.sort(rng.choice(n, size=1500, replace=False)), "last node": np.array([n - 1])}
    gpus = {}
    for prec in ("fp64", "mixed"):
        g = USMesh(options=dict(gridFile=xv0, fileType="memory", evalMode="exact", precision=prec,
                                symmetryPlanes=sym))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        g.setSurfaceCoordinates(xs)
        g.warpMesh()
        gpus[prec] = g
    for name, idx in cases.items():
        ora = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], xv0[idx], symmetryPlanes=sym, evalMode="exact")
        ora.setSurfaceCoordinates(xs)
        ora.warpMesh()
        g = np.random.default_rng(32).standard_normal((len(idx), 3))
        dref = ora.warpDeriv(g.ravel())
        seed = np.zeros((n, 3))
        seed[idx] = g
        for prec, gpu in gpus.items():
            gpu.warpDeriv(seed.ravel(), solverVec=False)
            r = _rel(gpu.getdXs().ravel(), dref)
            print("planes", planes, name, prec, "adjoint rel err vs oracle %.2e" % r, flush=True)
            assert r <= (1e-10 if prec == "fp64" else 1e-6)
    for gpu in gpus.values():
        gpu.warpDeriv(np.zeros(3 * n), solverVec=False)
        assert not np.any(gpu.getdXs())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_bucket_tree(prec):
    """A surface of 8 points (a cube, 6 quads) is one bucket: the tree has no
    internal node, the walk visits the root bucket only and the cluster data
    has nothing to compute.  Fast mode then equals the exact sum, so the
    device's fast warp and reverse must match the oracle's exact ones."""
    c = np.array([[x, y, z] for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0)]) - 0.5
    quads = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]]
    conn = np.array(quads, np.int32).ravel()
    fs = np.full(6, 4, np.int32)
    rng = np.random.default_rng(21)
    d = rng.standard_normal((3000, 3))
    xv0 = (d / np.linalg.norm(d, axis=1, keepdims=True)) * rng.uniform(0.9, 3.0, (3000, 1))
    xs = c + 0.02 * rng.standard_normal(c.shape)
    ora = O.OracleMesh(c, conn, fs, xv0, evalMode="exact")
    ora.setSurfaceCoordinates(xs)
    ref = ora.warpMesh()
    v = synthetic.random_dxv(len(xv0), seed=3, frac=0.5)
    refb = ora.warpDeriv(v)
    gpu = USMesh(options=dict(gridFile=xv0, fileType="memory", evalMode="fast", errTol=1e-4, precision=prec))
    gpu.setSurfaceDefinition(c, conn, fs)
    gpu.setSurfaceCoordinates(xs)
    gpu.warpMesh()
    print("case single bucket", prec, flush=True)
    err = np.abs(gpu.getWarpGrid() - ref.ravel()).max() / ora.Ldef0
    gpu.warpDeriv(v, solverVec=False)
    rb = _rel(gpu.getdXs().ravel(), refb)
    print(prec, "max |dXv|/Ldef", err, "adjoint rel", rb)
    assert err <= (FP64_TOL if prec == "fp64" else MIXED_TOL)
    assert rb <= (1e-10 if prec == "fp64" else 1e-6)
