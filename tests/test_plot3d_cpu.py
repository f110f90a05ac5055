"""Multiblock PLOT3D I/O (idwarp_amd/plot3d.py) against the reference's file
layout: Fortran unformatted sequential records, int32 sizes, float64
coordinates, zones in order, i fastest (readPlot3d.F90:27-95,
readPlot3dSurface :164-262, writePlot3d.F90)."""
import numpy as np
import pytest

from idwarp_amd import plot3d


def _rec(b):
    n = np.array([len(b)], "<i4").tobytes()
    return n + b + n


def test_grid_read_matches_a_hand_built_file(tmp_path):
    # two zones, 2x1x1 and 1x2x2; X, Y, Z of a zone in one record
    z1 = np.array([[0.0, 1.0], [2.0, 3.0], [4.0, 5.0]])  # X(2), Y(2), Z(2)
    z2 = np.arange(12, dtype=np.float64).reshape(3, 4) + 10.0
    raw = (_rec(np.array([2], "<i4").tobytes()) + _rec(np.array([2, 1, 1, 1, 2, 2], "<i4").tobytes()) +
           _rec(z1.astype("<f8").tobytes()) + _rec(z2.astype("<f8").tobytes()))
    f = tmp_path / "g.xyz"
    f.write_bytes(raw)
    sizes, nodes = plot3d.read_grid(str(f))
    assert sizes.tolist() == [[2, 1, 1], [1, 2, 2]]
    want = np.concatenate([z1.T, z2.T])
    assert np.array_equal(nodes, want)


def test_grid_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    sizes = np.array([[4, 3, 2], [2, 2, 5]])
    nodes = rng.standard_normal((int(np.prod(sizes, axis=1).sum()), 3))
    f = tmp_path / "g.xyz"
    plot3d.write_grid(str(f), sizes, nodes)
    s2, n2 = plot3d.read_grid(str(f))
    assert np.array_equal(s2, sizes) and np.array_equal(n2, nodes)


def test_surface_quads_follow_the_reference_rule(tmp_path):
    # zone 1: 3 x 2 points, zone 2: 2 x 3 points
    pts = np.arange(12 * 3, dtype=np.float64).reshape(12, 3)
    f = tmp_path / "s.xyz"
    plot3d.write_surface(str(f), [[3, 2], [2, 3]], pts)
    p2, conn, sizes = plot3d.read_surface(str(f))
    assert np.array_equal(p2, pts)
    assert sizes.tolist() == [[3, 2], [2, 3]]
    # n1 = (i, j), n2 = (i+1, j), n3 = (i+1, j+1), n4 = (i, j+1), i fastest, zone offsets 0 and 6
    assert conn.tolist() == [[0, 1, 4, 3], [1, 2, 5, 4], [6, 7, 9, 8], [8, 9, 11, 10]]


def test_bad_files_raise(tmp_path):
    f = tmp_path / "bad.xyz"
    f.write_bytes(_rec(np.array([1], "<i4").tobytes()) + b"\x10\x00\x00\x00abc")
    with pytest.raises(ValueError):
        plot3d.read_grid(str(f))
    g = tmp_path / "short.xyz"
    g.write_bytes(_rec(np.array([1], "<i4").tobytes()) + _rec(np.array([2, 2, 2], "<i4").tobytes()) +
                  _rec(np.zeros(5).tobytes()))
    with pytest.raises(ValueError):
        plot3d.read_grid(str(g))
    with pytest.raises(ValueError):
        plot3d.write_grid(str(tmp_path / "x.xyz"), [[2, 2, 2]], np.zeros((7, 3)))
