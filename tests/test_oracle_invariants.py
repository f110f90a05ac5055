"""Pin the oracle's IDW sums, kd-tree, error tables and reverse by what the
mathematics fixes (the reference's kd_tree.F90 needs PETSc and cannot be
built here, and its regression inputs are download-only):

  * closed forms: zero perturbation, rigid translation, a flat plate rotated
    about an in-plane axis (every Rodrigues rotation equals the rigid one);
  * the fast mode with errTol=0 reduces to the exact sum bit for bit;
  * a NumPy brute force of the exact sum with SciPy rotations;
  * kd-tree structure (split sizes, bucket rule, partition, widest axis,
    centroids, radii, area sums);
  * error tables against an independent NumPy evaluation at the reference
    sphere points; the denominator estimate against the exact denominator;
  * the reverse against central differences, the dot-product identity and
    the all-ones identity of the reference's complex-step test.
"""
import numpy as np
import pytest

from idwarp_amd import synthetic
from oracle import oracle as O

O.set_threads(4)


def _plate(n=9, L=1.0):
    """Flat square plate in z=0, quads, with a volume block above it."""
    x = np.linspace(-L, L, n)
    X, Y = np.meshgrid(x, x, indexing="xy")
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    p0 = (j * n + i).ravel()
    conn = np.stack([p0, p0 + 1, p0 + n + 1, p0 + n], 1).ravel().astype(np.int32)
    fs = np.full(len(p0), 4, np.int32)
    g = np.linspace(-1.5, 1.5, 7)
    zz = np.array([0.05, 0.2, 0.7, 1.5])
    V = np.stack(np.meshgrid(g, g, zz, indexing="ij"), -1).reshape(-1, 3)
    return pts, conn, fs, V


@pytest.fixture(scope="module")
def sphere():
    return synthetic.sphere_omesh(13, 9, 9)


@pytest.fixture(scope="module")
def star():
    return synthetic.star_shells(n=9, nshell=5)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_zero_perturbation_and_translation(star, mode):
    m = star
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode=mode, errTol=1e-4)
    om.setSurfaceCoordinates(m["pts"])
    assert np.array_equal(om.warpMesh(), m["xv0"].ravel())
    t = np.array([0.02, -0.01, 0.035])
    om.setSurfaceCoordinates(m["pts"] + t)
    d = om.warpMesh().reshape(-1, 3) - m["xv0"]
    assert np.abs(d - t).max() < 1e-14


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_plate_rigid_rotation(mode):
    """Normals (0,0,1) rotated about the x axis: the minimal rotation taking
    n0 to R n0 is R itself, b = Xu - R Xu0 = 0, so Xv = R Xv0 exactly."""
    pts, conn, fs, V = _plate()
    om = O.OracleMesh(pts, conn, fs, V, evalMode=mode, errTol=1e-4, zeroCornerRotations=False)
    a = 0.2
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    om.setSurfaceCoordinates(pts @ R.T)
    xv = om.warpMesh().reshape(-1, 3)
    assert np.abs(xv - V @ R.T).max() < 1e-13


def test_fast_with_zero_tolerance_is_exact(sphere):
    m = sphere
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="exact")
    xs = m["perturb"](m["pts"])
    om.setSurfaceCoordinates(xs)
    xe = om.warpMesh()
    om.setOptions(evalMode="fast", errTol=0.0)
    assert np.array_equal(om.warpMesh(), xe)
    om.setOptions(evalMode="fast", errTol=1e-3)
    xf = om.warpMesh()
    assert 0 < np.abs(xf - xe).max() < 0.05 * om.Ldef0


def _rodrigues_scipy(n0, n1):
    from scipy.spatial.transform import Rotation

    a = np.cross(n0, n1)
    s = np.linalg.norm(a)
    if s < 1.4901161193847656e-08:
        return np.eye(3)
    ang = np.arccos(min(1.0, np.dot(n0 / np.linalg.norm(n0), n1 / np.linalg.norm(n1))))
    return Rotation.from_rotvec(a / s * ang).as_matrix()


def test_exact_sum_brute_force():
    """Eq. of kd_tree.F90:680-689 evaluated with NumPy on the oracle's own
    nodal state, rotations from SciPy's rotation-vector map."""
    m = synthetic.sphere_omesh(9, 7, 5)
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="exact")
    st0 = om.surface_state()
    xs = m["perturb"](m["pts"])
    om.setSurfaceCoordinates(xs)
    xv = om.warpMesh().reshape(-1, 3)
    st = om.surface_state()
    Xu0, A = st0["Xu0"], st0["Ai"]
    Xu = xs[st["XuInd"] - 1]
    # normals of the perturbed surface from the oracle state (kept identical),
    # rotations recomputed independently
    Ms = np.array([_rodrigues_scipy(st0["normals0"][i], (st["Mi"][i] @ st0["normals0"][i]))
                   if not st0["isCorner"][i] else np.eye(3) for i in range(len(A))])
    assert np.allclose(Ms, st["Mi"], atol=1e-12)
    B = Xu - np.einsum("nij,nj->ni", Ms, Xu0)
    Ldef = om.Ldef0
    r = m["xv0"]
    d = np.sqrt(((r[:, None, :] - Xu0[None]) ** 2).sum(-1) + 1e-15)
    q = Ldef / d
    w = A[None] * (q ** 3 + 0.25 ** 5 * q ** 5)
    S = np.einsum("nij,mj->mni", Ms, r) + B[None] - r[:, None, :]
    ref = r + (w[..., None] * S).sum(1) / w.sum(1)[:, None]
    assert np.abs(ref - xv).max() < 1e-12 * Ldef


def test_tree_structure(star):
    m = star
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], bucketSize=8)
    ints, dbl = om.tree()
    st = om.surface_state()
    Xu0, A = st["Xu0"], st["Ai"]
    assert ints[0, 1] == 1 and ints[0, 2] == om.n
    for k, (dnum, l, u, n, left, right) in enumerate(ints):
        assert n == u - l + 1
        pts = Xu0[l - 1:u]
        X, orad, Ai = dbl[k, :3], dbl[k, 3], dbl[k, 4]
        assert np.allclose(X, pts.mean(0), atol=1e-15)
        assert abs(1 / orad - np.sqrt(((pts - X) ** 2).sum(1)).max()) < 1e-14
        assert abs(Ai - A[l - 1:u].sum()) <= 1e-14 * Ai
        if dnum == 0:
            assert u - l <= 8 and left == -1 and right == -1
            continue
        assert u - l > 8
        mid = (l + u) // 2
        assert tuple(ints[left, 1:3]) == (l, mid) and tuple(ints[right, 1:3]) == (mid + 1, u)
        assert left == k + 1  # depth-first pre-order
        c = dnum - 1
        assert Xu0[l - 1:mid, c].max() <= Xu0[mid:u, c].min()
        if n <= 201:  # spread over the whole range: order-independent
            spread = pts.max(0) - pts.min(0)
            assert c == int(np.argmax(spread))


def test_error_tables_and_denominator(star):
    m = star
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast")
    ints, dbl = om.tree()
    st = om.surface_state()
    Xu0, A, Ldef = st["Xu0"], st["Ai"], om.Ldef0
    rstar = [2, 3, 4, 5, 7.5, 10, 20, 40, 80, 160, 320, 640]

    def w(d):
        q = Ldef / d
        return q ** 3 + 0.25 ** 5 * q ** 5

    for k in np.flatnonzero(ints[:, 0] != 0)[:12]:
        l, u = ints[k, 1], ints[k, 2]
        X, orad, Ai, err = dbl[k, :3], dbl[k, 3], dbl[k, 4], dbl[k, 5:]
        for j, rs in enumerate(rstar):
            vp, _ = O.sphere_pts(X, rs / orad)
            ex = np.array([(A[l - 1:u] * w(np.linalg.norm(v - Xu0[l - 1:u], axis=1))).sum() for v in vp])
            ap = Ai * w(np.linalg.norm(vp - X, axis=1))
            assert abs(np.sqrt(((ex - ap) ** 2).mean()) - err[j]) <= 1e-10 * max(err[j], 1e-300)
    # far-field estimate close to the exact denominator
    d = np.sqrt(((m["xv0"][:, None, :] - Xu0[None]) ** 2).sum(-1) + 1e-16)
    den = (A[None] * w(d)).sum(1)
    rel = np.abs(om.den0 - den) / den
    assert rel.max() < 2e-2 and np.median(rel) < 1e-3


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_reverse_fd_dot_and_all_ones(sphere, mode):
    m = sphere
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode=mode, errTol=1e-4)
    xs = m["perturb"](m["pts"]).ravel()
    om.setSurfaceCoordinates(xs)
    om.warpMesh()
    rng = np.random.default_rng(6)
    dxv = rng.standard_normal(3 * len(m["xv0"]))
    dxs = om.warpDeriv(dxv)
    h = 1e-6
    for p in [rng.standard_normal(xs.size) for _ in range(3)] + [np.ones(xs.size)]:
        om.setSurfaceCoordinates(xs + h * p)
        xp = om.warpMesh()
        om.setSurfaceCoordinates(xs - h * p)
        xm = om.warpMesh()
        fd = dxv @ (xp - xm) / (2 * h)
        assert abs(fd - dxs @ p) <= 1e-6 * abs(fd)
    om.setSurfaceCoordinates(xs)


def test_only_first_duplicate_gets_sensitivity(sphere):
    """kd_tree.F90:1155-1159: XsPtrb(XuInd(i)) += Xub(i)."""
    m = sphere
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="exact")
    om.setSurfaceCoordinates(m["perturb"](m["pts"]))
    om.warpMesh()
    dxs = om.warpDeriv(np.ones(3 * len(m["xv0"]))).reshape(-1, 3)
    first = set(om.surface_state()["XuInd"] - 1)
    dup = [i for i in range(len(dxs)) if i not in first]
    assert dup and np.all(dxs[dup] == 0.0)


def test_point_reduce_semantics():
    rng = np.random.default_rng(1)
    base = rng.standard_normal((40, 3))
    pts = np.concatenate([base, base[[3, 7, 7, 20]], base[[5]] + 5e-13])
    uniq, link, nu = O.point_reduce(pts, 1e-12)
    assert nu == 40
    assert np.array_equal(uniq, base)
    assert list(link[40:]) == [4, 8, 8, 21, 6]


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_tangent_fd_and_dot_product(sphere, mode):
    """warpDerivFwd (warpDerivFwd.F90:62-99) against central differences of
    warpMesh, and the dot-product identity with warpDeriv,
    <v, J p> == <J^T v, p> (examples/structured/dotprod.py:38-49)."""
    m = sphere
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode=mode, errTol=1e-4)
    xs = m["perturb"](m["pts"]).ravel()
    om.setSurfaceCoordinates(xs)
    om.warpMesh()
    rng = np.random.default_rng(8)
    h = 1e-6
    v = rng.standard_normal(3 * len(m["xv0"]))
    jtv = om.warpDeriv(v)
    for _ in range(3):
        p = rng.standard_normal(xs.size)
        jp = om.warpDerivFwd(p)
        om.setSurfaceCoordinates(xs + h * p)
        xp = om.warpMesh()
        om.setSurfaceCoordinates(xs - h * p)
        xm = om.warpMesh()
        om.setSurfaceCoordinates(xs)
        om.warpMesh()
        fd = (xp - xm) / (2 * h)
        assert np.linalg.norm(jp - fd) <= 1e-6 * np.linalg.norm(fd)
        # both sides sum terms like K r and K Xu0 that cancel: rounding ~1e-11
        assert abs(v @ jp - jtv @ p) <= 1e-9 * abs(v @ jp)


def test_tangent_of_rigid_translation():
    """A translation tangent t at every surface point moves every volume
    node by t (M and the normals do not change)."""
    pts, conn, fs, V = _plate()
    for mode in ("exact", "fast"):
        om = O.OracleMesh(pts, conn, fs, V, evalMode=mode, errTol=1e-4)
        om.setSurfaceCoordinates(pts)
        om.warpMesh()
        t = np.array([0.3, -0.2, 0.7])
        jp = om.warpDerivFwd(np.tile(t, len(pts))).reshape(-1, 3)
        assert np.abs(jp - t).max() < 1e-13


def _symmetric_case():
    half = synthetic.shifted(synthetic.sphere_omesh(13, 9, 5), (0.3, 2.5, -0.2))
    plane = ([0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    return half, plane, synthetic.mirrored_body(half, *plane)


def test_symmetry_plane_equals_mirrored_body():
    """getMirrorPt / getMirrorNumerator (initializeWarping.F90:533-600): the
    body with a symmetry plane warps the volume as the body plus its mirror
    image does without one (exact mode, Ldef matched)."""
    half, plane, full = _symmetric_case()
    oh = O.OracleMesh(half["pts"], half["conn"], half["faceSizes"], half["xv0"], evalMode="exact",
                      symmetryPlanes=[plane])
    of = O.OracleMesh(full["pts"], full["conn"], full["faceSizes"], full["xv0"], evalMode="exact")
    of.setOptions(LdefFact=oh.Ldef0 / of.Ldef0)
    xs = half["perturb"](half["pts"])
    oh.setSurfaceCoordinates(xs)
    of.setSurfaceCoordinates(full["perturb"](full["pts"]))
    a, b = oh.warpMesh(), of.warpMesh()
    assert np.abs(a - b).max() <= 1e-12 * oh.Ldef0
    assert np.abs(a - half["xv0"].ravel()).max() > 1e-3  # the warp is not trivial
    # the reverse: dXs of the half body equals the full body's restricted to
    # its first copy plus the mirrored sensitivities of the second copy
    v = synthetic.random_dxv(len(half["xv0"]), seed=3, frac=1.0)
    ga = oh.warpDeriv(v).reshape(-1, 3)
    gb = of.warpDeriv(v).reshape(-1, 3)
    n = len(half["pts"])
    R = np.eye(3) - 2.0 * np.outer(plane[1], plane[1])
    assert np.abs(ga - (gb[:n] + gb[n:] @ R)).max() <= 1e-10 * np.abs(ga).max()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_symmetry_reverse_and_tangent_fd(mode):
    half, plane, _ = _symmetric_case()
    om = O.OracleMesh(half["pts"], half["conn"], half["faceSizes"], half["xv0"], evalMode=mode, errTol=1e-4,
                      symmetryPlanes=[plane, ([0.0, 0.0, -3.0], [0.0, 0.0, 1.0])])
    xs = half["perturb"](half["pts"]).ravel()
    om.setSurfaceCoordinates(xs)
    om.warpMesh()
    rng = np.random.default_rng(21)
    v = rng.standard_normal(3 * len(half["xv0"]))
    jtv = om.warpDeriv(v)
    h = 1e-6
    for _ in range(2):
        p = rng.standard_normal(xs.size)
        jp = om.warpDerivFwd(p)
        om.setSurfaceCoordinates(xs + h * p)
        xp = om.warpMesh()
        om.setSurfaceCoordinates(xs - h * p)
        xm = om.warpMesh()
        om.setSurfaceCoordinates(xs)
        om.warpMesh()
        fd = (xp - xm) / (2 * h)
        assert np.linalg.norm(jp - fd) <= 1e-6 * np.linalg.norm(fd)
        assert abs(v @ fd - jtv @ p) <= 1e-6 * abs(v @ fd)
        assert abs(v @ jp - jtv @ p) <= 1e-9 * abs(v @ jp)


def test_symmetry_fast_with_zero_tolerance_is_exact():
    half, plane, _ = _symmetric_case()
    kw = dict(symmetryPlanes=[plane])
    oe = O.OracleMesh(half["pts"], half["conn"], half["faceSizes"], half["xv0"], evalMode="exact", **kw)
    of = O.OracleMesh(half["pts"], half["conn"], half["faceSizes"], half["xv0"], evalMode="fast", errTol=0.0, **kw)
    xs = half["perturb"](half["pts"])
    oe.setSurfaceCoordinates(xs)
    of.setSurfaceCoordinates(xs)
    assert np.array_equal(oe.warpMesh(), of.warpMesh())


def test_load_balancing_costs_follow_the_work_counts(star):
    """dryRun (kd_tree.F90:789-830) counts n per bucket, 0.5 per internal
    visit and 1.1 per accepted cluster; its distance carries a 1e-16 eps that
    the warp's does not, which may only move a decision at a tie."""
    m = star
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"], evalMode="fast", errTol=1e-4)
    costs = om.costs()
    om.setSurfaceCoordinates(m["pts"])
    c = om.counts()
    expect = (c[:, 2] + 0.5 * c[:, 0] + 1.1 * c[:, 1]) / om.n
    assert np.allclose(costs, expect, rtol=1e-12, atol=0)
    assert costs.min() > 0
