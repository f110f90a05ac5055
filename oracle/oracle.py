"""ctypes wrapper of the CPU oracle (oracle/idw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the idwarp_amd package.

``OracleMesh`` mirrors the reference's call sequence:
``USMesh.setSurfaceDefinition`` (UnstructuredMesh.py:196-262: pointReduce,
lexsort, relink, initializeWarping), then ``warpMesh`` (warpMesh.F90) and
``warpDeriv`` (warpDeriv.F90), with the denominator estimate of
initializeWarping.F90:240-247.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Options(ctypes.Structure):
    _fields_ = [
        ("aExp", ctypes.c_double),
        ("bExp", ctypes.c_double),
        ("LdefFact", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("errTol", ctypes.c_double),
        ("cornerAngle", ctypes.c_double),
        ("evalMode", ctypes.c_int32),
        ("useRotations", ctypes.c_int32),
        ("zeroCornerRotations", ctypes.c_int32),
        ("bucketSize", ctypes.c_int32),
    ]


DEFAULTS = dict(aExp=3.0, bExp=5.0, LdefFact=1.0, alpha=0.25, errTol=5e-4, evalMode="fast",
                useRotations=True, zeroCornerRotations=True, cornerAngle=30.0, bucketSize=8)


def make_options(**kw):
    o = dict(DEFAULTS)
    o.update(kw)
    return Options(o["aExp"], o["bExp"], o["LdefFact"], o["alpha"], o["errTol"], o["cornerAngle"],
                   1 if o["evalMode"] == "fast" else 0, int(bool(o["useRotations"])),
                   int(bool(o["zeroCornerRotations"])), int(o["bucketSize"]))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        P, I, D, L64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.POINTER(Options), dp, I, dp, I, ip, ip, I, ip, I]
        L.orc_free.argtypes = [P]
        L.orc_set_options.argtypes = [P, ctypes.POINTER(Options)]
        L.orc_set_threads.argtypes = [I]
        L.orc_den0.argtypes = [P, dp, L64, dp]
        L.orc_warp.argtypes = [P, dp, dp, L64, dp, dp, dp]
        L.orc_warp_deriv.argtypes = [P, dp, dp, L64, dp, dp, dp, dp]
        L.orc_warp_deriv_fwd.argtypes = [P, dp, dp, dp, L64, dp, dp, dp]
        L.orc_counts.argtypes = [P, dp, dp, L64, dp, dp]
        L.orc_costs.argtypes = [P, dp, L64, dp, dp]
        L.orc_set_symmetry_planes.restype = I
        L.orc_set_symmetry_planes.argtypes = [P, dp, dp, I]
        L.orc_point_reduce.restype = I
        L.orc_point_reduce.argtypes = [dp, I, D, dp, ip]
        L.orc_num_nodes.restype = I
        L.orc_num_nodes.argtypes = [P]
        L.orc_num_unique.restype = I
        L.orc_num_unique.argtypes = [P]
        L.orc_ldef0.restype = D
        L.orc_ldef0.argtypes = [P]
        L.orc_get_surface.argtypes = [P] + [P] * 8
        L.orc_get_tree.argtypes = [P, ip, dp]
        L.orc_element_props.argtypes = [dp, I, dp, dp]
        L.orc_element_props_b.argtypes = [dp, I, D, dp, dp]
        L.orc_rotation_matrix.argtypes = [dp, dp, dp]
        L.orc_rotation_matrix_b.argtypes = [dp, dp, dp, dp]
        L.orc_rotation_matrix_d.argtypes = [dp, dp, dp, dp, dp]
        L.orc_element_props_d.argtypes = [dp, dp, I, dp, dp, dp, dp]
        L.orc_sphere_pts.argtypes = [dp, D, dp, dp]
        L.orc_last_times.argtypes = [dp]
        _LIB = L
    return _LIB


def set_threads(n):
    lib().orc_set_threads(int(n))


def last_times():
    """(set-up seconds, volume-loop seconds) of the last warp or warpDeriv."""
    t = np.zeros(2)
    lib().orc_last_times(t)
    return t[0], t[1]


def point_reduce(pts, tol=1e-12):
    """pointReduce.F90: returns (uniquePts (nU,3), link (1-based), nU)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    uniq = np.zeros_like(pts)
    link = np.zeros(len(pts), np.int32)
    nu = lib().orc_point_reduce(pts.ravel(), len(pts), tol, uniq.ravel(), link)
    return uniq[:nu].copy(), link, nu


def surface_definition(pts, conn, faceSizes):
    """UnstructuredMesh.py:239-256: point reduce, lexsort by (z, y, x), relink."""
    uniq, link, nu = point_reduce(pts, 1e-12)
    ind = np.lexsort(uniq.T)
    uniq = uniq[ind]
    inv = np.zeros_like(ind)
    inv[ind] = np.arange(len(ind))
    link = (inv[link - 1] + 1).astype(np.int32)
    return np.ascontiguousarray(uniq), link


class OracleMesh:
    """One surface + volume mesh instance on the CPU oracle."""

    def __init__(self, pts, conn, faceSizes, xv0, symmetryPlanes=None, **opts):
        L = lib()
        self.opts = dict(DEFAULTS)
        self.opts.update(opts)
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, np.int32).ravel()
        faceSizes = np.ascontiguousarray(faceSizes, np.int32).ravel()
        self.npts = len(pts)
        uniq, link = surface_definition(pts, conn, faceSizes)
        self.link = link
        o = make_options(**self.opts)
        self.h = L.orc_create(ctypes.byref(o), pts.ravel(), len(pts), uniq.ravel(), len(uniq), link, faceSizes,
                              len(faceSizes), conn, len(conn))
        self.n = L.orc_num_unique(self.h)
        if symmetryPlanes:  # [[point, normal], ...] as USMesh's 'symmetryPlanes' option
            sp = np.ascontiguousarray([p for p, _ in symmetryPlanes], np.float64).ravel()
            sn = np.ascontiguousarray([n for _, n in symmetryPlanes], np.float64).ravel()
            if L.orc_set_symmetry_planes(self.h, sp, sn, len(symmetryPlanes)) != 0:
                raise ValueError("at most 6 symmetry planes")
        self.xv0 = np.ascontiguousarray(xv0, np.float64).reshape(-1, 3)
        self.nvol = len(self.xv0)
        self.den0 = np.zeros(self.nvol)
        L.orc_den0(self.h, self.xv0.ravel(), self.nvol, self.den0)
        self.den = None
        self.xs = pts.ravel().copy()

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_free(self.h)
            self.h = None

    def setOptions(self, **kw):
        self.opts.update(kw)
        lib().orc_set_options(self.h, ctypes.byref(make_options(**self.opts)))

    @property
    def Ldef0(self):
        return lib().orc_ldef0(self.h)

    def setSurfaceCoordinates(self, xs):
        xs = np.ascontiguousarray(xs, np.float64).ravel()
        assert xs.size == 3 * self.npts
        self.xs = xs.copy()

    def warpMesh(self):
        xv = np.zeros(3 * self.nvol)
        self.den = np.zeros(self.nvol)
        lib().orc_warp(self.h, self.xs, self.xv0.ravel(), self.nvol, self.den0, xv, self.den)
        return xv

    def warpDeriv(self, dxv):
        if self.den is None:
            raise RuntimeError("warpDeriv before warpMesh")
        dxs = np.zeros(3 * self.npts)
        lib().orc_warp_deriv(self.h, self.xs, self.xv0.ravel(), self.nvol, self.den0, self.den,
                             np.ascontiguousarray(dxv, np.float64).ravel(), dxs)
        return dxs

    def warpDerivFwd(self, dxs):
        """warpDerivFwd.F90: dXv = (dXv/dXs) dxs (3*npts), warp order, with
        the denominator of the last warp."""
        if self.den is None:
            raise RuntimeError("warpDerivFwd before warpMesh")
        dxv = np.zeros(3 * self.nvol)
        lib().orc_warp_deriv_fwd(self.h, self.xs, np.ascontiguousarray(dxs, np.float64).ravel(), self.xv0.ravel(),
                                 self.nvol, self.den0, self.den, dxv)
        return dxv

    def costs(self):
        """Load-balancing costs of the restart file (dryRun per volume node)."""
        out = np.zeros(self.nvol)
        lib().orc_costs(self.h, self.xv0.ravel(), self.nvol, self.den0, out)
        return out

    def counts(self, xv0=None, den0=None):
        xv0 = self.xv0 if xv0 is None else np.ascontiguousarray(xv0, np.float64).reshape(-1, 3)
        den0 = self.den0 if den0 is None else np.ascontiguousarray(den0, np.float64)
        out = np.zeros(4 * len(xv0))
        lib().orc_counts(self.h, self.xs, xv0.ravel(), len(xv0), den0, out)
        return out.reshape(-1, 4)

    def surface_state(self):
        n = self.n
        Xu0 = np.zeros(3 * n); Ai = np.zeros(n); n0 = np.zeros(3 * n); Mi = np.zeros(9 * n); Bi = np.zeros(3 * n)
        corner = np.zeros(n, np.int32); idx = np.zeros(n, np.int32); xuind = np.zeros(n, np.int32)
        lib().orc_get_surface(self.h, *[a.ctypes.data for a in (Xu0, Ai, n0, Mi, Bi, corner, idx, xuind)])
        return dict(Xu0=Xu0.reshape(n, 3), Ai=Ai, normals0=n0.reshape(n, 3), Mi=Mi.reshape(n, 3, 3).transpose(0, 2, 1),
                    Bi=Bi.reshape(n, 3), isCorner=corner, indexes=idx, XuInd=xuind)

    def tree(self):
        nn = lib().orc_num_nodes(self.h)
        ints = np.zeros(6 * nn, np.int32)
        dbl = np.zeros(17 * nn)
        lib().orc_get_tree(self.h, ints, dbl)
        return ints.reshape(nn, 6), dbl.reshape(nn, 17)


def element_props(points):
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    a = np.zeros(1); nrm = np.zeros(3)
    lib().orc_element_props(p.ravel(), len(p), a, nrm)
    return a[0], nrm


def element_props_b(points, areab, normalb):
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    pb = np.zeros(p.size)
    lib().orc_element_props_b(p.ravel(), len(p), float(areab), np.ascontiguousarray(normalb, np.float64), pb)
    return pb.reshape(-1, 3)


def rotation_matrix(v1, v2):
    M = np.zeros(9)
    lib().orc_rotation_matrix(np.ascontiguousarray(v1, np.float64), np.ascontiguousarray(v2, np.float64), M)
    return M.reshape(3, 3).T  # column-major storage -> M[r, c]


def rotation_matrix_b(v1, v2, Mb):
    v2b = np.zeros(3)
    lib().orc_rotation_matrix_b(np.ascontiguousarray(v1, np.float64), np.ascontiguousarray(v2, np.float64),
                                np.ascontiguousarray(np.asarray(Mb, np.float64).T).ravel(), v2b)
    return v2b


def element_props_d(points, pointsd):
    """GETELEMENTPROPS_D: (area, normal, area tangent, normal tangent)."""
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    pd = np.ascontiguousarray(pointsd, np.float64).reshape(-1, 3)
    a = np.zeros(1); ad = np.zeros(1); nrm = np.zeros(3); nd = np.zeros(3)
    lib().orc_element_props_d(p.ravel(), pd.ravel(), len(p), a, ad, nrm, nd)
    return a[0], nrm, ad[0], nd


def rotation_matrix_d(v1, v2, v2d):
    """GETROTATIONMATRIX3D_D: (M, Md) as [r, c] arrays."""
    M = np.zeros(9); Md = np.zeros(9)
    lib().orc_rotation_matrix_d(np.ascontiguousarray(v1, np.float64), np.ascontiguousarray(v2, np.float64),
                                np.ascontiguousarray(v2d, np.float64), M, Md)
    return M.reshape(3, 3).T, Md.reshape(3, 3).T


def sphere_pts(x0, radius):
    v = np.zeros(60); f = np.zeros(36)
    lib().orc_sphere_pts(np.ascontiguousarray(x0, np.float64), float(radius), v, f)
    return v.reshape(20, 3), f.reshape(12, 3)
