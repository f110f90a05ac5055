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

}  // namespace idw
