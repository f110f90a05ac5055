"""Synthetic meshes for the volume-warp workloads (no file I/O, NumPy only).

Each generator returns a dict with
  ``xv0``       (N, 3) volume nodes, the "common grid" a CGNS reader would give,
  ``pts``       (Ns, 3) wall-surface points with duplicates (seams, poles, edges),
  ``conn``      flat 0-based connectivity into ``pts``,
  ``faceSizes`` points per face,
  ``perturb``   callable(pts, seed) -> perturbed (Ns, 3) surface; duplicated
                points always receive identical coordinates.

Families:
  * ``sphere_omesh``   structured spherical O-mesh, geometric radial growth
                       (49x33x33, growth 1.15 to r=20 is the CPU-sized case);
  * ``ellipsoid_omesh`` 10:1:1 ellipsoid O-mesh with a 1e-5 first spacing and a
                       seeded random Fourier surface displacement;
  * ``star_shells``    cube-sphere star-shaped body (spherical-harmonic radius)
                       with stretched shells of volume nodes; n=205 and 80
                       shells gives 19,975,840 volume nodes on 249,698 unique
                       surface nodes.
"""
import math

import numpy as np


# --------------------------------------------------------------------------
# spherical O-mesh
# --------------------------------------------------------------------------
def _structured_wall(ni, nj):
    """Quads of an ni x nj structured wall (i fastest), outward orientation."""
    i, j = np.meshgrid(np.arange(ni - 1), np.arange(nj - 1), indexing="xy")
    p0 = (j * ni + i).ravel()
    conn = np.stack([p0, p0 + ni, p0 + ni + 1, p0 + 1], axis=1)
    return conn.ravel().astype(np.int32), np.full(len(p0), 4, np.int32)


def sphere_omesh(ni=49, nj=33, nk=33, growth=1.15, r_outer=20.0):
    u = np.arange(ni) / (ni - 1)
    v = np.arange(nj) / (nj - 1)
    phi, theta = 2 * np.pi * u, np.pi * v
    d0 = (r_outer - 1.0) * (growth - 1.0) / (growth ** (nk - 1) - 1.0)
    radii = np.concatenate([[1.0], 1.0 + d0 * (growth ** np.arange(1, nk) - 1.0) / (growth - 1.0)])
    T, P = np.meshgrid(theta, phi, indexing="ij")  # (nj, ni), i fastest
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1).reshape(-1, 3)
    dirs[np.abs(dirs) < 1e-300] = 0.0
    xv0 = (radii[:, None, None] * dirs[None, :, :]).reshape(-1, 3)
    pts = dirs.copy()
    conn, fs = _structured_wall(ni, nj)
    uu, vv = np.meshgrid(u, v, indexing="xy")
    uu, vv = uu.ravel(), vv.ravel()

    def perturb(p, seed=0, amp=0.05, rot_deg=2.0):
        bump = amp * np.sin(np.pi * uu) * np.sin(np.pi * vv)
        q = p + bump[:, None] * dirs
        a = math.radians(rot_deg)
        R = np.array([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]])
        return q @ R.T

    return dict(xv0=xv0, pts=pts, conn=conn, faceSizes=fs, perturb=perturb, name="sphere_omesh_%dx%dx%d" % (ni, nj, nk))


# --------------------------------------------------------------------------
# ellipsoid O-mesh
# --------------------------------------------------------------------------
def _solve_ratio(first, total, nint):
    lo, hi = 1.0 + 1e-9, 3.0
    for _ in range(200):
        q = 0.5 * (lo + hi)
        s = first * (q ** nint - 1.0) / (q - 1.0)
        if s > total:
            hi = q
        else:
            lo = q
    return 0.5 * (lo + hi)


def ellipsoid_omesh(ni=257, nj=129, nk=97, axes=(1.0, 0.1, 0.1), first=1e-5, outer=30.0):
    a, b, c = axes
    u = np.arange(ni) / (ni - 1)
    v = np.arange(nj) / (nj - 1)
    phi, theta = 2 * np.pi * u, np.pi * v
    q = _solve_ratio(first, outer - a, nk - 1)
    s = np.concatenate([[0.0], first * (q ** np.arange(1, nk) - 1.0) / (q - 1.0)])
    T, P = np.meshgrid(theta, phi, indexing="ij")
    ct, st, cp, sp = np.cos(T).ravel(), np.sin(T).ravel(), np.cos(P).ravel(), np.sin(P).ravel()
    st[np.abs(st) < 1e-15] = 0.0
    X = np.stack([(a + s[:, None]) * ct[None], (b + s[:, None]) * (st * cp)[None], (c + s[:, None]) * (st * sp)[None]],
                 axis=-1)
    xv0 = X.reshape(-1, 3)
    pts = X[0].copy()
    conn, fs = _structured_wall(ni, nj)
    uu, vv = np.meshgrid(u, v, indexing="xy")
    uu, vv = uu.ravel(), vv.ravel()

    def perturb(p, seed=2, peak=0.04, kmax=8):
        rng = np.random.default_rng(seed)
        d = np.zeros_like(p)
        for comp in range(3):
            A = rng.standard_normal((kmax, kmax))
            B = rng.standard_normal((kmax, kmax))
            f = np.zeros(len(p))
            for k in range(1, kmax + 1):
                ck, sk = np.cos(2 * np.pi * k * uu), np.sin(2 * np.pi * k * uu)
                for l in range(1, kmax + 1):
                    f += (A[k - 1, l - 1] * ck + B[k - 1, l - 1] * sk) * np.sin(l * np.pi * vv) / (k * k + l * l)
            d[:, comp] = f
        mag = np.sqrt((d ** 2).sum(1)).max()
        return p + d * (peak / mag)

    return dict(xv0=xv0, pts=pts, conn=conn, faceSizes=fs, perturb=perturb, name="ellipsoid_omesh_%dx%dx%d" % (ni, nj, nk))


# --------------------------------------------------------------------------
# star-shaped cube-sphere body with volume shells
# --------------------------------------------------------------------------
def _real_sh(lmax, x, y, z):
    """Real spherical harmonics up to lmax at unit vectors, by the associated
    Legendre three-term recursion (orthonormal up to a constant per l)."""
    ct = np.clip(z, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = np.arctan2(y, x)
    out = {}
    P = {(0, 0): np.ones_like(ct)}
    for m in range(1, lmax + 1):
        P[(m, m)] = -(2 * m - 1) * st * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = (2 * m + 1) * ct * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * ct * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m > 0:
                out[(l, m)] = math.sqrt(2) * norm * P[(l, am)] * np.cos(am * ph)
            elif m < 0:
                out[(l, m)] = math.sqrt(2) * norm * P[(l, am)] * np.sin(am * ph)
            else:
                out[(l, m)] = norm * P[(l, 0)]
    return out


def _cube_sphere(n):
    """Unit-sphere cube-sphere directions (6 n^2 points incl. edge duplicates)
    and 2 triangles per cell."""
    t = np.tan(np.linspace(-np.pi / 4, np.pi / 4, n))  # equiangular
    A, B = np.meshgrid(t, t, indexing="xy")
    A, B = A.ravel(), B.ravel()
    one = np.ones_like(A)
    faces = [(one, A, B), (-one, -A, B), (-A, one, B), (A, -one, B), (A, B, one), (-A, B, -one)]
    dirs = np.concatenate([np.stack(f, 1) for f in faces])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    p0 = (j * n + i).ravel()
    tri = []
    for f in range(6):
        base = f * n * n + p0
        tri.append(np.stack([base, base + 1, base + n + 1], 1))
        tri.append(np.stack([base, base + n + 1, base + n], 1))
    conn = np.concatenate(tri).ravel().astype(np.int32)
    return dirs, conn, np.full(len(conn) // 3, 3, np.int32)


def _dedupe_dirs(dirs):
    """Map each direction to a canonical copy so duplicated edge points are
    bit-identical."""
    key = np.round(dirs * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return dirs[first][inv.ravel()], first, inv.ravel()


def star_shells(n=205, nshell=80, first=1e-5, ratio=1.1, amp=0.1, seed=3, lmax=6):
    dirs, conn, fs = _cube_sphere(n)
    dirs, first_idx, inv = _dedupe_dirs(dirs)
    ud = dirs[first_idx]
    rng = np.random.default_rng(seed)
    Y = _real_sh(lmax, ud[:, 0], ud[:, 1], ud[:, 2])
    f = np.zeros(len(ud))
    for key in sorted(Y):
        f += rng.standard_normal() * Y[key]
    rad = 1.0 + amp * f / np.abs(f).max()
    pts = (rad[:, None] * ud)[inv]
    s = np.concatenate([[0.0], first * (ratio ** np.arange(1, nshell) - 1.0) / (ratio - 1.0)])
    xv0 = ((rad[None, :, None] + s[:, None, None]) * ud[None, :, :]).reshape(-1, 3)
    ymax = np.abs(pts[:, 1]).max()

    def perturb(p, seed=4, peak=0.01, twist=0.05, lmax_p=4):
        r2 = np.random.default_rng(seed)
        nrm = np.linalg.norm(p, axis=1, keepdims=True)
        w = p / nrm
        Yp = _real_sh(lmax_p, w[:, 0], w[:, 1], w[:, 2])
        g = np.zeros(len(p))
        for key in sorted(Yp):
            g += r2.standard_normal() * Yp[key]
        q = p + (peak * g / np.abs(g).max())[:, None] * w
        ang = twist * q[:, 1] / ymax
        c, s_ = np.cos(ang), np.sin(ang)
        x, z = q[:, 0].copy(), q[:, 2].copy()
        q[:, 0] = c * x + s_ * z
        q[:, 2] = -s_ * x + c * z
        return q

    return dict(xv0=xv0, pts=pts, conn=conn, faceSizes=fs, perturb=perturb, name="star_shells_n%d_s%d" % (n, nshell))


def random_dxv(nvol, seed=5, frac=0.1):
    """Adjoint seed: N(0,1) per component, masked to a random fraction of nodes."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nvol, 3))
    d[rng.random(nvol) >= frac] = 0.0
    return d.ravel()


def shifted(m, offset):
    """The mesh m moved rigidly by offset (volume, surface and its
    perturbation)."""
    off = np.asarray(offset, np.float64)
    base = m["perturb"]
    return dict(m, xv0=m["xv0"] + off, pts=m["pts"] + off, perturb=lambda p: base(p - off) + off,
                name=m["name"] + "_shifted")


def reflect(x, point, normal):
    """Reflection of points x (N, 3) about the plane (point, unit normal)."""
    n = np.asarray(normal, np.float64)
    return x - 2.0 * np.outer((x - np.asarray(point, np.float64)) @ n, n)


def mirrored_body(m, point, normal):
    """Surface of m plus its mirror image about the plane (same face lists,
    offset), with the perturbation mirrored too; volume nodes are m's.  A
    warp of this surface without symmetry planes equals the warp of m with
    the plane, when m does not touch the plane."""
    npts = len(m["pts"])
    base = m["perturb"]

    def perturb(p):
        return np.concatenate([base(p[:npts]), reflect(base(reflect(p[npts:], point, normal)), point, normal)])

    return dict(xv0=m["xv0"], pts=np.concatenate([m["pts"], reflect(m["pts"], point, normal)]),
                conn=np.concatenate([m["conn"], m["conn"] + npts]), faceSizes=np.concatenate([m["faceSizes"]] * 2),
                perturb=perturb, name=m["name"] + "_mirrored")
