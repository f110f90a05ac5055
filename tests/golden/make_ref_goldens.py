"""Write tests/golden/ref_elements.npz from the reference's own Fortran.

The routines come from oracle/_ref/libidwref.so, which oracle/Makefile
(`make -C oracle ref`) compiles in place from the reference sources
(getElementProps.F90, vectorUtils.f90, getSpherePts.F90, the Tapenade
reverse routines getElementProps_b.f90 / vectorUtils_b.f90 and the tangent
routines getElementProps_d.f90 / vectorUtils_d.f90, with the reference's
-fdefault-real-8).  Inputs are seeded random faces and vector
pairs plus the branch edge cases of the rotation (no rotation below the
1.49e-8 axis tolerance, acos argument clipped at 1, opposite normals).

Run: python tests/golden/make_ref_goldens.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libidwref.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Ref:
    def __init__(self, path=LIB):
        self.L = ctypes.CDLL(os.path.abspath(path))

    def element_props(self, pts):
        pts = np.asfortranarray(np.asarray(pts, np.float64).T)  # (3, n)
        n = ctypes.c_int32(pts.shape[1])
        a = np.zeros(1)
        nrm = np.zeros(3)
        self.L.getelementprops_(_p(pts), ctypes.byref(n), _p(a), _p(nrm))
        return a[0], nrm

    def element_props_b(self, pts, areab, normalb):
        pts = np.asfortranarray(np.asarray(pts, np.float64).T)
        n = ctypes.c_int32(pts.shape[1])
        pb = np.zeros_like(pts, order="F")
        area, nrm = self.element_props(pts.T)
        a = np.array([area])
        ab = np.array([float(areab)])
        nb = np.array(normalb, np.float64).copy()
        self.L.getelementprops_b_(_p(pts), _p(pb), ctypes.byref(n), _p(a), _p(ab), _p(nrm.copy()), _p(nb))
        return np.ascontiguousarray(pb.T)

    def rotation(self, v1, v2):
        M = np.zeros((3, 3), order="F")
        self.L.getrotationmatrix3d_(_p(np.array(v1, np.float64)), _p(np.array(v2, np.float64)), _p(M))
        return np.array(M)

    def rotation_b(self, v1, v2, Mb):
        v2b = np.zeros(3)
        M = self.rotation(v1, v2)
        Mbf = np.asfortranarray(np.array(Mb, np.float64))
        self.L.getrotationmatrix3d_b_(_p(np.array(v1, np.float64)), _p(np.array(v2, np.float64)), _p(v2b),
                                      _p(np.asfortranarray(M)), _p(Mbf))
        return v2b

    def element_props_d(self, pts, ptsd):
        p = np.asfortranarray(np.asarray(pts, np.float64).T)
        pdot = np.asfortranarray(np.asarray(ptsd, np.float64).T)
        n = ctypes.c_int32(p.shape[1])
        a, ad, nrm, nd = np.zeros(1), np.zeros(1), np.zeros(3), np.zeros(3)
        self.L.getelementprops_d_(_p(p), _p(pdot), ctypes.byref(n), _p(a), _p(ad), _p(nrm), _p(nd))
        return a[0], nrm, ad[0], nd

    def rotation_d(self, v1, v2, v2d):
        M = np.zeros((3, 3), order="F")
        Md = np.zeros((3, 3), order="F")
        self.L.getrotationmatrix3d_d_(_p(np.array(v1, np.float64)), _p(np.array(v2, np.float64)),
                                      _p(np.array(v2d, np.float64)), _p(M), _p(Md))
        return np.array(M), np.array(Md)

    def sphere(self, x0, radius):
        v = np.zeros((3, 20), order="F")
        f = np.zeros((3, 12), order="F")
        r = np.array([float(radius)])
        self.L.getspherepts_(_p(np.array(x0, np.float64)), _p(r), _p(v), _p(f))
        return np.ascontiguousarray(v.T), np.ascontiguousarray(f.T)


def cases(seed=1234):
    rng = np.random.default_rng(seed)
    faces = []
    for npts in (3, 4, 4, 5, 6, 8):
        for _ in range(4):
            c = rng.standard_normal(3)
            t = np.sort(rng.uniform(0, 2 * np.pi, npts))
            e1, e2 = np.linalg.qr(rng.standard_normal((3, 3)))[0][:, :2].T
            p = c + np.outer(np.cos(t), e1) + np.outer(np.sin(t), e2) + 0.05 * rng.standard_normal((npts, 3))
            faces.append(p)
    # degenerate quad (collapsed edge, as at an O-mesh pole)
    q = rng.standard_normal((4, 3))
    q[1] = q[0]
    faces.append(q)
    vecs = []
    for _ in range(24):
        v1 = rng.standard_normal(3)
        vecs.append((v1, v1 + rng.uniform(0.01, 1.0) * rng.standard_normal(3)))
    v = rng.standard_normal(3)
    vecs.append((v, v.copy()))                                    # identical: no-rotation branch
    vecs.append((v, v + 1e-10 * rng.standard_normal(3)))          # below the axis tolerance
    u = v / np.linalg.norm(v)
    perp = np.cross(u, rng.standard_normal(3))
    vecs.append((u, u + 3e-8 * perp / np.linalg.norm(perp)))      # acos argument near/clipped at 1
    vecs.append((v, -v))                                          # opposite
    vecs.append((v, 2.5 * v + 1e-3 * perp))                       # non-unit inputs
    return faces, vecs, rng


def main():
    R = Ref()
    faces, vecs, rng = cases()
    out = {}
    for k, p in enumerate(faces):
        a, n = R.element_props(p)
        ab = rng.standard_normal()
        nb = rng.standard_normal(3)
        out["face%d_pts" % k] = p
        out["face%d_area" % k] = np.array([a])
        out["face%d_normal" % k] = n
        out["face%d_areab" % k] = np.array([ab])
        out["face%d_normalb" % k] = nb
        out["face%d_ptsb" % k] = R.element_props_b(p, ab, nb)
    for k, (v1, v2) in enumerate(vecs):
        Mb = rng.standard_normal((3, 3))
        out["rot%d_v1" % k] = v1
        out["rot%d_v2" % k] = v2
        out["rot%d_M" % k] = R.rotation(v1, v2)
        out["rot%d_Mb" % k] = Mb
        out["rot%d_v2b" % k] = R.rotation_b(v1, v2, Mb)
    for k in range(4):
        x0 = rng.standard_normal(3)
        r = rng.uniform(0.01, 10)
        v, f = R.sphere(x0, r)
        out["sph%d_x0" % k] = x0
        out["sph%d_r" % k] = np.array([r])
        out["sph%d_v" % k] = v
        out["sph%d_f" % k] = f
    # tangents (forward mode) from a separate stream so the arrays above do
    # not depend on them
    rd = np.random.default_rng(4321)
    for k, p in enumerate(faces):
        pd = rd.standard_normal(p.shape)
        a, n, ad, nd = R.element_props_d(p, pd)
        out["face%d_ptsd" % k] = pd
        out["face%d_aread" % k] = np.array([ad])
        out["face%d_normald" % k] = nd
    for k, (v1, v2) in enumerate(vecs):
        v2d = rd.standard_normal(3)
        M, Md = R.rotation_d(v1, v2, v2d)
        out["rot%d_v2d" % k] = v2d
        out["rot%d_Md" % k] = Md
        out["rot%d_Mfwd" % k] = M
    np.savez_compressed(os.path.join(HERE, "ref_elements.npz"), nfaces=len(faces), nrot=len(vecs), nsph=4, **out)
    print("wrote %d arrays" % len(out))


if __name__ == "__main__":
    main()
