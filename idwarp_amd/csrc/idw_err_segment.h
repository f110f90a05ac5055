// Segment of the error-table interpolation (getError, kd_tree.F90:882-906).
// Plain C++ apart from the HIP qualifiers, so that host code (a test
// program built with g++) can include it with IDW_HOST_ONLY defined.
#pragma once
#include <cmath>

#ifdef IDW_HOST_ONLY
#define IDW_HD inline
#else
#define IDW_HD __host__ __device__ __forceinline__
#endif

namespace idw {

// Segment i of the error table, rstar(i) <= d < rstar(i+1), for d >= 2
// (11 = NERR-1 from the last radius, 640, on).  The radii 2, 3, 4, 5, 7.5
// and 10 * 2^k are exact thresholds on the binary exponent e and mantissa m
// of d = m 2^e, m in [0.5, 1): no rounding is involved.
IDW_HD int err_segment(double d) {
    int e;
    const double m = frexp(d, &e);
    int i;
    if (e >= 4) i = e + 1 - (m < 0.625 ? 1 : 0);
    else if (e == 3) i = 2 + (m >= 0.625 ? 1 : 0) + (m >= 0.9375 ? 1 : 0);
    else i = m >= 0.75 ? 1 : 0;
    return i < 11 ? i : 11;
}

// Cell of d >= 2 on a grid of eighths of each binade: the exponent and the
// three leading mantissa bits of d, (bits >> 49) - (1024 << 3).  Every radius
// lies on a cell boundary, so all of a cell's values share one segment; the
// cells from ERR_CELLS - 1 on (d >= 1024) all lie in the last segment.
constexpr int ERR_CELLS = 73;
IDW_HD int err_cell(double d) {
    unsigned long long bits;
    __builtin_memcpy(&bits, &d, sizeof bits);
    const int c = (int)(bits >> 49) - 8192;
    return c < ERR_CELLS - 1 ? c : ERR_CELLS - 1;
}

// Lowest value of cell c.
IDW_HD double err_cell_low(int c) { return ldexp(1.0 + (c & 7) * 0.125, 1 + (c >> 3)); }

}  // namespace idw
