"""The library's host C++ on the CPU, built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/native/host_checks, compiled by
idwarp_amd.build.build_host_checks):

  * its kd-tree and surface maps (idw_tree.cpp: build_tree_for_range,
    setCenters, setAi, the XuInd and Ldef0 set-up of myCreateTree,
    kd_tree.F90:123-300, 536-593, 1365-1512) equal the oracle's bit for bit
    on meshes that exercise the 201-entry split window, several bucket sizes
    and duplicated points;
  * its pointReduce equals the oracle's;
  * every C-ABI entry point returns an error code, without a sanitizer
    report, for a null context, bad options and a machine without a GPU;
  * the walk's header-offset limit and the mixed-precision floor.
"""
import os
import subprocess

import numpy as np
import pytest

from idwarp_amd import build, synthetic
from oracle import oracle as O

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")


@pytest.fixture(scope="module")
def checker():
    srcs = [os.path.join(build.CSRC, f) for f in ("idw_api.hip", "idw_kernels.hip", "idw_tree.cpp",
                                                  "idw_internal.h", "idw_kernels.h")]
    srcs.append(os.path.join(build.ROOT, "tests", "native", "host_checks.cpp"))
    exe = build.HOST_CHECKS
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(s) for s in srcs):
        build.build_host_checks()
    return exe


def _run(exe, *args):
    r = subprocess.run([exe] + list(args), env=ENV, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    return r.stdout


def _read(path, dtypes):
    out = []
    with open(path, "rb") as f:
        for dt in dtypes:
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            out.append(np.frombuffer(f.read(n * np.dtype(dt).itemsize), dt))
    return out


def _lib_tree(exe, tmp, uniq, link, npts, sizes, conn, bucket, Ai):
    inp, out = str(tmp / "tree.in"), str(tmp / "tree.out")
    with open(inp, "wb") as f:
        for n, a in ((len(uniq), uniq.ravel()), (npts, link), (len(sizes), sizes), (len(conn), conn),
                     (bucket, None), (len(Ai), Ai)):
            f.write(np.int64(n).tobytes())
            if a is not None:
                f.write(np.ascontiguousarray(a).tobytes())
    _run(exe, "tree", inp, out)
    ints, dbl, xu0, xuind, fconn, n2es, n2e, ldef = _read(out, [np.int32, np.float64, np.float64, np.int32, np.int32,
                                                                  np.int32, np.int32, np.float64])
    return dict(ints=ints.reshape(-1, 5), dbl=dbl.reshape(-1, 5), xu0=xu0.reshape(-1, 3), xuInd=xuind,
                faceConn=fconn, n2eStart=n2es, n2e=n2e, ldef0=ldef[0])


def _window_surface():
    from scipy.spatial import Delaunay

    i = np.arange(1, 401)
    win = np.stack([0.5 * ((37 * i) % 400) / 399, 0.75 * ((91 * i) % 400) / 399, i * 1e-9], 1)
    tri = Delaunay(win[:, :2]).simplices.astype(np.int32)
    far = np.array([[0.1 * a, 0.5, 10.0 + b] for b in range(12) for a in range(12)])
    quads = [[len(win) + 12 * b + a, len(win) + 12 * b + a + 1, len(win) + 12 * (b + 1) + a + 1,
              len(win) + 12 * (b + 1) + a] for b in range(11) for a in range(11)]
    conn = np.concatenate([tri.ravel(), np.array(quads, np.int32).ravel()])
    sizes = np.concatenate([np.full(len(tri), 3, np.int32), np.full(len(quads), 4, np.int32)])
    return dict(pts=np.concatenate([win, far]), conn=conn, faceSizes=sizes, xv0=np.array([[0.5, 0.5, 3.0]]))


CASES = {
    "sphere": (lambda: synthetic.sphere_omesh(13, 9, 9), 8),
    "ellipsoid": (lambda: synthetic.ellipsoid_omesh(33, 17, 5), 8),
    "star-b1": (lambda: synthetic.star_shells(n=17, nshell=1), 1),
    "star-b4": (lambda: synthetic.star_shells(n=17, nshell=1), 4),
    "star-b18": (lambda: synthetic.star_shells(n=33, nshell=1), 18),
    "window": (_window_surface, 8),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_tree_equals_oracle(checker, tmp_path, name):
    make, bucket = CASES[name]
    m = make()
    om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"][:4], evalMode="fast", bucketSize=bucket)
    uniq, link = O.surface_definition(m["pts"], m["conn"], m["faceSizes"])
    st = om.surface_state()
    L = _lib_tree(checker, tmp_path, uniq, link, len(m["pts"]), np.asarray(m["faceSizes"], np.int32),
                  np.asarray(m["conn"], np.int32).ravel(), bucket, st["Ai"])
    ints, dbl = om.tree()
    assert len(L["ints"]) == len(ints)
    l, u, skip, depth, leaf = L["ints"].T
    assert np.array_equal(l + 1, ints[:, 1]) and np.array_equal(u + 1, ints[:, 2])
    assert np.array_equal(leaf == 1, ints[:, 0] == 0)
    inner = np.flatnonzero(leaf == 0)
    assert np.array_equal(ints[inner, 4], inner + 1)          # left child follows its parent
    assert np.array_equal(ints[inner, 5], skip[inner + 1])     # right child follows the left subtree
    assert np.array_equal(L["dbl"][:, :3], dbl[:, :3])          # centroids, bit for bit
    assert np.array_equal(L["dbl"][:, 3], dbl[:, 3])            # 1 / radius
    assert np.array_equal(L["dbl"][:, 4], dbl[:, 4])            # setAi
    assert np.array_equal(L["xu0"], st["Xu0"])
    assert np.array_equal(L["xuInd"] + 1, st["XuInd"])
    assert L["ldef0"] == om.Ldef0
    # nodeToElem keeps a face once per occurrence of the node in it, in face order
    assert L["n2eStart"][-1] == len(m["conn"])


def test_point_reduce_equals_oracle(checker, tmp_path):
    rng = np.random.default_rng(2)
    base = rng.standard_normal((400, 3))
    pts = np.concatenate([base, base[rng.integers(0, 400, 150)], base[:10] + 3e-13, base[10:20] + 2e-12])
    pts = pts[rng.permutation(len(pts))]
    inp, out = str(tmp_path / "r.in"), str(tmp_path / "r.out")
    with open(inp, "wb") as f:
        f.write(np.int64(len(pts)).tobytes() + pts.tobytes() + np.float64(1e-12).tobytes())
    _run(checker, "reduce", inp, out)
    uniq, link = _read(out, [np.float64, np.int32])
    u2, l2, n2 = O.point_reduce(pts, 1e-12)
    assert len(uniq) == 3 * n2 == 3 * 410
    assert np.array_equal(uniq.reshape(-1, 3), u2) and np.array_equal(link, l2)


def test_api_error_paths_without_a_device(checker):
    assert "ok api" in _run(checker, "api")


def test_walk_offset_limit_and_mixed_floor(checker):
    assert "ok limits" in _run(checker, "limits")
