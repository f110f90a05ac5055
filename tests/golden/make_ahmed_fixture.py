"""Write tests/golden/ahmed_polymesh.npz from the reference's example input
examples/AhmedBodyCoarse/constant/polyMesh (an OpenFOAM ASCII mesh, data
only): all 21,706 points, the boundary faces (renumbered from 0, internal
faces dropped) and the boundary table.  Read with idwarp_amd.openfoam.

Run: python tests/golden/make_ahmed_fixture.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
CASE = "/root/reference/examples/AhmedBodyCoarse"


def main():
    from idwarp_amd import openfoam

    of = openfoam.read_polymesh(CASE)
    first = min(b["startFace"] for _, b in of["boundaries"])
    last = max(b["startFace"] + b["nFaces"] for _, b in of["boundaries"])
    s = of["faceStart"]
    sizes = of["faceSizes"][first:last].astype(np.int32)
    conn = of["faceConn"][s[first]:s[last]].astype(np.int32)
    bnd = [(n, dict(b, startFace=b["startFace"] - first)) for n, b in of["boundaries"]]
    np.savez_compressed(os.path.join(HERE, "ahmed_polymesh.npz"), x0=of["x0"], faceSizes=sizes, faceConn=conn,
                        boundaries=np.array(json.dumps(bnd)), firstBoundaryFace=np.int64(first))
    print("points %d, boundary faces %d" % (len(of["x0"]), len(sizes)))


if __name__ == "__main__":
    main()
