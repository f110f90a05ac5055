"""OpenFOAM polyMesh input (UnstructuredMesh.py:967-1049) without a GPU:
the reader against the committed Ahmed-body fixture, a write/read round
trip, and the warping surface of _computeOFConn."""
import json
import os

import numpy as np
import pytest

from idwarp_amd import openfoam

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ahmed_polymesh.npz")
CASE = "/root/reference/examples/AhmedBodyCoarse"


def ahmed_case(tmp):
    """Write the fixture as an OpenFOAM case (boundary faces only) and
    return its directory."""
    z = np.load(FIX)
    bnd = [(n, b) for n, b in json.loads(str(z["boundaries"]))]
    openfoam.write_polymesh(str(tmp), z["x0"], z["faceSizes"], z["faceConn"], bnd)
    return str(tmp)


def test_roundtrip_and_surface(tmp_path):
    case = ahmed_case(tmp_path)
    of = openfoam.read_polymesh(case)
    z = np.load(FIX)
    assert np.array_equal(of["x0"], z["x0"])
    assert np.array_equal(of["faceSizes"], z["faceSizes"]) and np.array_equal(of["faceConn"], z["faceConn"])
    names = [n for n, _ in of["boundaries"]]
    assert names == ["slip", "sym", "inlet", "outlet", "lowerWall", "wall"]
    pts, conn, fs = openfoam.warp_surface(of)
    # patch, wall and slip faces; the symmetry patch is not part of the surface
    assert len(fs) == 504 + 36 + 36 + 1170 + 1708
    assert len(pts) == fs.sum() == len(conn) and np.array_equal(conn, np.arange(len(conn)))
    assert np.all(np.isin(pts, of["x0"]))


@pytest.mark.skipif(not os.path.isdir(CASE), reason="reference example not present")
def test_reader_matches_fixture_from_reference_case():
    of = openfoam.read_polymesh(CASE)
    z = np.load(FIX)
    first = int(z["firstBoundaryFace"])
    s = of["faceStart"]
    assert np.array_equal(of["x0"], z["x0"])
    assert np.array_equal(of["faceConn"][s[first]:], z["faceConn"])
    assert len(of["faceSizes"]) == 56734


def test_rejects_binary(tmp_path):
    p = tmp_path / "points"
    p.write_text("FoamFile\n{\n    format      binary;\n}\n3\n(\n)\n")
    with pytest.raises(ValueError):
        openfoam.read_points(str(p))
