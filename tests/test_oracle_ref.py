"""Pin the oracle's element-level routines to the reference's own Fortran.

Fixtures in tests/golden/ref_elements.npz were written by
tests/golden/make_ref_goldens.py from oracle/_ref/libidwref.so (the
reference's getElementProps.F90, vectorUtils.f90, getSpherePts.F90, the
Tapenade reverse getElementProps_b.f90 / vectorUtils_b.f90 and tangent
getElementProps_d.f90 / vectorUtils_d.f90, compiled in place).  The
oracle restates those routines with the same operation order, so the
comparison is bit for bit.  Where the reference build is present, fresh
random cases are compared as well.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_elements.npz"))
REF_SO = os.path.join(HERE, "..", "oracle", "_ref", "libidwref.so")


def test_element_props_bitwise():
    for k in range(int(GOLD["nfaces"])):
        p = GOLD["face%d_pts" % k]
        a, n = O.element_props(p)
        assert a == GOLD["face%d_area" % k][0]
        assert np.array_equal(n, GOLD["face%d_normal" % k])


def test_element_props_reverse_bitwise():
    for k in range(int(GOLD["nfaces"])):
        pb = O.element_props_b(GOLD["face%d_pts" % k], GOLD["face%d_areab" % k][0], GOLD["face%d_normalb" % k])
        assert np.array_equal(pb, GOLD["face%d_ptsb" % k]), k


def test_rotation_matrix_bitwise():
    for k in range(int(GOLD["nrot"])):
        M = O.rotation_matrix(GOLD["rot%d_v1" % k], GOLD["rot%d_v2" % k])
        assert np.array_equal(M, GOLD["rot%d_M" % k]), k


def test_rotation_reverse_bitwise():
    """Covers the no-rotation branch (axis below 1.49e-8), the acos argument
    clipped at 1 and opposite normals."""
    for k in range(int(GOLD["nrot"])):
        v2b = O.rotation_matrix_b(GOLD["rot%d_v1" % k], GOLD["rot%d_v2" % k], GOLD["rot%d_Mb" % k])
        assert np.array_equal(v2b, GOLD["rot%d_v2b" % k]), k


def test_element_props_tangent_bitwise():
    for k in range(int(GOLD["nfaces"])):
        a, n, ad, nd = O.element_props_d(GOLD["face%d_pts" % k], GOLD["face%d_ptsd" % k])
        assert a == GOLD["face%d_area" % k][0] and np.array_equal(n, GOLD["face%d_normal" % k]), k
        assert ad == GOLD["face%d_aread" % k][0], k
        assert np.array_equal(nd, GOLD["face%d_normald" % k]), k


def test_rotation_tangent_bitwise():
    """Same branch cases as the reverse: no rotation, clipped acos argument,
    opposite and non-unit normals."""
    for k in range(int(GOLD["nrot"])):
        M, Md = O.rotation_matrix_d(GOLD["rot%d_v1" % k], GOLD["rot%d_v2" % k], GOLD["rot%d_v2d" % k])
        assert np.array_equal(M, GOLD["rot%d_Mfwd" % k]), k
        assert np.array_equal(Md, GOLD["rot%d_Md" % k]), k


def test_tangent_and_reverse_are_transposes():
    """<J_d v2d, Mb> == <v2d, J_b Mb> for the rotation and the element
    routines: the tangent and the reverse restatements are consistent."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        v1 = rng.standard_normal(3)
        v2 = v1 + rng.standard_normal(3)
        v2d, Mb = rng.standard_normal(3), rng.standard_normal((3, 3))
        _, Md = O.rotation_matrix_d(v1, v2, v2d)
        assert abs(np.sum(Md * Mb) - O.rotation_matrix_b(v1, v2, Mb) @ v2d) <= 1e-12 * (1 + np.abs(Md).sum())
    for npts in (3, 4, 5, 8):
        p, pd = rng.standard_normal((npts, 3)), rng.standard_normal((npts, 3))
        ab, nb = rng.standard_normal(), rng.standard_normal(3)
        _, _, ad, nd = O.element_props_d(p, pd)
        lhs = ab * ad + nb @ nd
        assert abs(lhs - np.sum(O.element_props_b(p, ab, nb) * pd)) <= 1e-12 * (1 + abs(lhs))


def test_sphere_points_bitwise():
    for k in range(int(GOLD["nsph"])):
        v, f = O.sphere_pts(GOLD["sph%d_x0" % k], GOLD["sph%d_r" % k][0])
        assert np.array_equal(v, GOLD["sph%d_v" % k])
        assert np.array_equal(f, GOLD["sph%d_f" % k])


def test_rotation_is_a_rotation():
    """Independent of the reference: M is orthogonal with det 1 and maps the
    first normal's direction onto the second's."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        v1 = rng.standard_normal(3)
        v2 = v1 + rng.uniform(0.05, 1.5) * rng.standard_normal(3)
        M = O.rotation_matrix(v1, v2)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(M) - 1) < 1e-14
        assert np.allclose(M @ (v1 / np.linalg.norm(v1)), v2 / np.linalg.norm(v2), atol=1e-13)


def test_reverse_routines_are_adjoints():
    """<Jp, seed> == <p, J^T seed> by central differences of the forward
    routines (the Tapenade reverse is the exact derivative)."""
    rng = np.random.default_rng(4)
    h = 1e-6
    for npts in (3, 4, 6):
        p = rng.standard_normal((npts, 3))
        ab, nb = rng.standard_normal(), rng.standard_normal(3)
        pb = O.element_props_b(p, ab, nb)
        d = rng.standard_normal((npts, 3))
        ap, npl = O.element_props(p + h * d)
        am, nmi = O.element_props(p - h * d)
        fd = ab * (ap - am) / (2 * h) + nb @ (npl - nmi) / (2 * h)
        assert abs(fd - np.sum(pb * d)) <= 1e-7 * (1 + abs(fd))
    for _ in range(10):
        v1 = rng.standard_normal(3)
        v2 = v1 + rng.standard_normal(3)
        Mb = rng.standard_normal((3, 3))
        v2b = O.rotation_matrix_b(v1, v2, Mb)
        d = rng.standard_normal(3)
        fd = np.sum(Mb * (O.rotation_matrix(v1, v2 + h * d) - O.rotation_matrix(v1, v2 - h * d))) / (2 * h)
        assert abs(fd - v2b @ d) <= 1e-7 * (1 + abs(fd))


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build oracle/_ref not present")
def test_fresh_cases_against_reference_build():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mkgold", os.path.join(HERE, "golden", "make_ref_goldens.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    R = mk.Ref(REF_SO)
    faces, vecs, rng = mk.cases(seed=987)
    for p in faces:
        a, n = R.element_props(p)
        a2, n2 = O.element_props(p)
        assert a == a2 and np.array_equal(n, n2)
        ab, nb = rng.standard_normal(), rng.standard_normal(3)
        assert np.array_equal(R.element_props_b(p, ab, nb), O.element_props_b(p, ab, nb))
    for v1, v2 in vecs:
        Mb = rng.standard_normal((3, 3))
        assert np.array_equal(R.rotation(v1, v2), O.rotation_matrix(v1, v2))
        assert np.array_equal(R.rotation_b(v1, v2, Mb), O.rotation_matrix_b(v1, v2, Mb))
        v2d = rng.standard_normal(3)
        Mr, Mdr = R.rotation_d(v1, v2, v2d)
        Mo, Mdo = O.rotation_matrix_d(v1, v2, v2d)
        assert np.array_equal(Mr, Mo) and np.array_equal(Mdr, Mdo)
    for p in faces:
        pd = rng.standard_normal(p.shape)
        r = R.element_props_d(p, pd)
        o = O.element_props_d(p, pd)
        assert r[0] == o[0] and r[2] == o[2] and np.array_equal(r[1], o[1]) and np.array_equal(r[3], o[3])
