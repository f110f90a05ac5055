"""Multiblock PLOT3D grids and surfaces for USMesh.

The files are Fortran unformatted sequential files as the reference reads
and writes them (readPlot3d / readPlot3dSurface, src/IO/readPlot3d.F90;
writePlot3d, src/IO/writePlot3d.F90): every record is framed by its byte
length as a 4-byte integer, integers are 4 bytes and coordinates 8-byte
reals, multiblock, no iblank.

* grid file: [nZones] [il jl kl per zone] then per zone one record holding
  X(i,j,k), Y(i,j,k), Z(i,j,k), i fastest;
* surface file: [nZones] [il jl per zone] then per zone X, Y, Z over (i, j).

Volume nodes are numbered zone by zone, i fastest, then j, then k, as the
reference interlaces them.  A surface zone of il x jl points gives
(il-1)(jl-1) quadrilaterals n1 = (i, j), n2 = (i+1, j), n3 = (i+1, j+1),
n4 = (i, j+1), i fastest (readPlot3d.F90:226-233).
"""
import numpy as np


def _records(path):
    """The records of a Fortran unformatted sequential file, as bytes."""
    data = np.fromfile(path, dtype=np.uint8)
    out, pos = [], 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError("%s: truncated record marker" % path)
        n = int(data[pos:pos + 4].view("<i4")[0])
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data) or int(data[end:end + 4].view("<i4")[0]) != n:
            raise ValueError("%s: not a Fortran unformatted sequential file (bad record at byte %d)" % (path, pos))
        out.append(data[pos + 4:end].tobytes())
        pos = end + 4
    return out


def _write_records(path, records):
    with open(path, "wb") as f:
        for r in records:
            b = r.tobytes() if isinstance(r, np.ndarray) else bytes(r)
            n = np.array([len(b)], dtype="<i4").tobytes()
            f.write(n)
            f.write(b)
            f.write(n)


def _header(recs, path, ndim):
    if len(recs) < 2:
        raise ValueError("%s: missing header records" % path)
    nz = int(np.frombuffer(recs[0], dtype="<i4")[0])
    sizes = np.frombuffer(recs[1], dtype="<i4").reshape(-1, ndim)
    if len(sizes) != nz or len(recs) != 2 + nz:
        raise ValueError("%s: %d zones declared, %d size entries and %d zone records" %
                         (path, nz, len(sizes), len(recs) - 2))
    return sizes.astype(np.int64)


def read_grid(path):
    """(sizes (nZones, 3), nodes (N, 3)): a multiblock volume grid, nodes in
    the reference's order (zone, k, j, i)."""
    recs = _records(path)
    sizes = _header(recs, path, 3)
    nodes = []
    for z, (il, jl, kl) in enumerate(sizes):
        n = int(il * jl * kl)
        v = np.frombuffer(recs[2 + z], dtype="<f8")
        if v.size != 3 * n:
            raise ValueError("%s: zone %d holds %d values, expected %d" % (path, z + 1, v.size, 3 * n))
        nodes.append(v.reshape(3, n).T)
    return sizes, (np.concatenate(nodes) if nodes else np.zeros((0, 3)))


def write_grid(path, sizes, nodes):
    """Write a multiblock volume grid (writePlot3d): nodes (N, 3) in the
    order read_grid returns."""
    sizes = np.asarray(sizes, dtype=np.int64).reshape(-1, 3)
    nodes = np.asarray(nodes, dtype=np.float64).reshape(-1, 3)
    if int(np.prod(sizes, axis=1).sum()) != len(nodes):
        raise ValueError("zone sizes hold %d nodes, %d given" % (int(np.prod(sizes, axis=1).sum()), len(nodes)))
    recs = [np.array([len(sizes)], dtype="<i4"), sizes.astype("<i4").ravel()]
    off = 0
    for il, jl, kl in sizes:
        n = int(il * jl * kl)
        recs.append(np.ascontiguousarray(nodes[off:off + n].T).astype("<f8").ravel())
        off += n
    _write_records(path, recs)


def read_surface(path):
    """(pts (N, 3), conn (nElem, 4) zero-based, sizes (nZones, 2)): a
    multiblock surface and its quadrilaterals (readPlot3dSurface)."""
    recs = _records(path)
    sizes = _header(recs, path, 2)
    pts, conn, off = [], [], 0
    for z, (il, jl) in enumerate(sizes):
        n = int(il * jl)
        v = np.frombuffer(recs[2 + z], dtype="<f8")
        if v.size != 3 * n:
            raise ValueError("%s: zone %d holds %d values, expected %d" % (path, z + 1, v.size, 3 * n))
        pts.append(v.reshape(3, n).T)
        i, j = np.meshgrid(np.arange(il - 1), np.arange(jl - 1), indexing="xy")  # i fastest
        i, j = i.ravel(), j.ravel()
        base = off + j * il + i
        conn.append(np.stack([base, base + 1, base + il + 1, base + il], axis=1))
        off += n
    pts = np.concatenate(pts) if pts else np.zeros((0, 3))
    conn = np.concatenate(conn).astype(np.int32) if conn else np.zeros((0, 4), np.int32)
    return pts, conn, sizes


def write_surface(path, sizes, pts):
    """Write a multiblock surface (the layout read_surface reads)."""
    sizes = np.asarray(sizes, dtype=np.int64).reshape(-1, 2)
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    if int(np.prod(sizes, axis=1).sum()) != len(pts):
        raise ValueError("zone sizes hold %d points, %d given" % (int(np.prod(sizes, axis=1).sum()), len(pts)))
    recs = [np.array([len(sizes)], dtype="<i4"), sizes.astype("<i4").ravel()]
    off = 0
    for il, jl in sizes:
        n = int(il * jl)
        recs.append(np.ascontiguousarray(pts[off:off + n].T).astype("<f8").ravel())
        off += n
    _write_records(path, recs)
