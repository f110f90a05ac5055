"""MultiUSMesh on the GPU, after the reference's ONERA M6 overset test
(tests/test_MultiUSMesh.py:24-105, options :150-177): two nearfield zones,
one with default options and one with the test's non-default set, and a
background zone between them in CGNS order; the solver's volume in a random
node order; the surface sheared by x += 0.25 y.  Every instance must equal a
standalone USMesh on its zone bit for bit and the oracle within the FP64 bar;
the assembled solver grid, reverse and forward products must follow."""
import numpy as np
import pytest

from idwarp_amd import MultiUSMesh, USMesh, synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13
FIN_OPTS = dict(aExp=4.0, bExp=6.0, LdefFact=100.0, alpha=0.3, errTol=1e-5, useRotations=False,
                zeroCornerRotations=False, cornerAngle=45.0, bucketSize=4)


def _case():
    ni, nj, nk = 13, 9, 9
    body = synthetic.sphere_omesh(ni, nj, nk)
    fin = synthetic.shifted(synthetic.ellipsoid_omesh(17, 9, 9, axes=(1.0, 0.3, 0.2)), (60.0, 0.0, 0.0))
    g = np.linspace(-30.0, 90.0, 6)
    bg = np.stack(np.meshgrid(g, g[:4] - 30.0, g[:3] - 30.0, indexing="ij"), -1).transpose(2, 1, 0, 3)
    bv = body["xv0"].reshape(nk, nj, ni, 3)
    blocks = [("body.00000", bv[:5]), ("body.00001", bv[4:]), ("background.00000", bg), ("fin.00000", fin["xv0"])]
    npb = len(body["pts"])
    # solver surface: body wall (block 0), one background quad (block 2), fin wall (block 3)
    bq = bg[0, 0, :2, :].reshape(-1, 3)[[0, 1]]
    bq = np.concatenate([bq, bg[0, 1, :2, :].reshape(-1, 3)[[1, 0]]])
    pts = np.concatenate([body["pts"], bq, fin["pts"]])
    conn = np.concatenate([body["conn"], npb + np.arange(4), npb + 4 + fin["conn"]])
    fs = np.concatenate([body["faceSizes"], [4], fin["faceSizes"]])
    ids = np.concatenate([np.zeros(len(body["faceSizes"]), int), [2], np.full(len(fin["faceSizes"]), 3)])
    return blocks, body, fin, pts, conn, fs, ids


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_multi_instance_overset(mode):
    print("case MultiUSMesh", mode, flush=True)
    blocks, body, fin, pts, conn, fs, ids = _case()
    multi = MultiUSMesh(blocks, {"body": dict(evalMode=mode), "fin": dict(FIN_OPTS, evalMode=mode)})
    xcgns = np.concatenate([np.asarray(b, np.float64).ravel() for _, b in blocks])
    nnodes = xcgns.size // 3
    perm = np.random.default_rng(7).permutation(nnodes)
    ind = (3 * perm[:, None] + np.arange(3)).ravel()
    multi.setExternalMeshIndices(ind)
    multi.setSurfaceDefinition(pts, conn, fs, ids)
    zones = {"body": np.concatenate([np.asarray(b).reshape(-1, 3) for _, b in blocks[:2]]), "fin": fin["xv0"]}
    vols, ncoor = multi.getWarpGrid()
    assert ncoor == 3 * (len(zones["body"]) + len(zones["fin"]))
    assert [v.sum() for v in vols] == [zones["body"].sum(), zones["fin"].sum()]
    assert np.array_equal(multi.getSolverGrid(), xcgns[ind])

    coords0 = multi.getSurfaceCoordinates()
    npb = len(body["pts"])
    assert np.array_equal(coords0[:npb], body["pts"]) and np.array_equal(coords0[npb + 4:], fin["pts"])
    assert not coords0[npb:npb + 4].any()
    new = coords0.copy()
    new[:, 0] += 0.25 * coords0[:, 1]
    multi.setSurfaceCoordinates(new)
    multi.warpMesh()

    solver = multi.getSolverGrid()
    dXv = np.linspace(0.0, 1.0, len(ind))
    dXs = multi.warpDeriv(dXv)
    starts = {"body": 0, "fin": 3 * (nnodes - len(fin["xv0"]))}
    rows = {"body": slice(0, npb), "fin": slice(npb + 4, len(pts))}
    zopts = {"body": {}, "fin": FIN_OPTS}
    covered = np.zeros(len(ind), bool)
    for ii, z in enumerate(("body", "fin")):
        src = body if z == "body" else fin
        xs = new[rows[z]]
        alone = USMesh(options=dict(gridFile=zones[z], fileType="memory", evalMode=mode, **zopts[z]))
        alone.setSurfaceDefinition(src["pts"], src["conn"], src["faceSizes"])
        alone.setSurfaceCoordinates(xs)
        alone.warpMesh()
        assert np.array_equal(multi.meshes[ii].getCommonGrid(), alone.getCommonGrid()), z
        ora = O.OracleMesh(src["pts"], src["conn"], src["faceSizes"], zones[z], evalMode=mode, **zopts[z])
        ora.setSurfaceCoordinates(xs)
        ref = ora.warpMesh()
        err = np.abs(alone.getCommonGrid() - ref).max() / ora.Ldef0
        print(z, mode, "max |dXv|/Ldef", err, flush=True)
        assert err <= FP64_TOL
        mask = (ind >= starts[z]) & (ind < starts[z] + zones[z].size)
        covered |= mask
        assert np.array_equal(solver[mask], alone.getCommonGrid()[ind[mask] - starts[z]])
        v = np.zeros(zones[z].size)
        np.add.at(v, ind[mask] - starts[z], dXv[mask])
        want = ora.warpDeriv(v)
        got = dXs[rows[z]].ravel()
        r = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(z, mode, "adjoint rel err", r, flush=True)
        assert r <= 1e-10
    assert np.array_equal(solver[~covered], xcgns[ind[~covered]])
    assert np.abs(solver - xcgns[ind]).max() > 1e-2
    assert not dXs[npb:npb + 4].any()

    s = np.random.default_rng(8).standard_normal(new.shape)
    fwd = multi.warpDerivFwd(s)
    assert not fwd[~covered].any()
    assert abs(dXv @ fwd - dXs.ravel() @ s.ravel()) <= 1e-9 * abs(dXv @ fwd)
