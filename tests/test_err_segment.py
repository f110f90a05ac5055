"""The walk's error-table segment (idwarp_amd/csrc/idw_err_segment.h) equals
the reference's linear search over rstar (getError, kd_tree.F90:882-906,
radii :34-35): first i with dOvrR < rstar(i+1), the last entry beyond 640.

The header is compiled for the host with g++ and evaluated on every radius,
its floating-point neighbours and a log-uniform sample of 200,000 values,
directly and through the walk's table of binade eighths (err_cell)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "idwarp_amd", "csrc")
RSTAR = [2.0, 3.0, 4.0, 5.0, 7.5, 10.0, 20.0, 40.0, 80.0, 160.0, 320.0, 640.0]

PROGRAM = r"""
#include <cstdio>
#include "idw_err_segment.h"
int main() {
    double d;
    while (std::scanf("%la", &d) == 1)
        std::printf("%d %d\n", idw::err_segment(d), idw::err_segment(idw::err_cell_low(idw::err_cell(d))));
    return 0;
}
"""


def linear_search(d):
    for i in range(11):
        if d < RSTAR[i + 1]:
            return i
    return 11


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_err_segment_matches_linear_search(tmp_path):
    src = tmp_path / "seg.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "seg"
    subprocess.run(["g++", "-O2", "-DIDW_HOST_ONLY", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    vals = []
    for r in RSTAR:
        vals += [r, np.nextafter(r, 0.0), np.nextafter(r, np.inf)]
    rng = np.random.default_rng(11)
    vals += list(np.exp(rng.uniform(np.log(2.0), np.log(1e5), 200000)))
    vals += [2.0, 1e300, 1e3, 1279.9999999999998, 1280.0]
    vals = [v for v in vals if v >= 2.0]
    out = subprocess.run([str(exe)], input="\n".join(float(v).hex() for v in vals), capture_output=True, text=True,
                         check=True).stdout.split()
    got = np.array(out, dtype=int).reshape(-1, 2)
    want = np.array([linear_search(v) for v in vals])
    bad = np.nonzero(got[:, 0] != want)[0]
    assert bad.size == 0, [(vals[i], got[i, 0], want[i]) for i in bad[:5]]
    # the walk's table: the segment of a value's cell, taken at the cell's lowest value
    bad = np.nonzero(got[:, 1] != want)[0]
    assert bad.size == 0, [(vals[i], got[i, 1], want[i]) for i in bad[:5]]
