// Device arithmetic of the fast-mode walk: getError (kd_tree.F90:882-906)
// over a table of binade cells, and the square root of its distance test.
// Both give the reference's values bit for bit on the arguments the walk
// passes them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "idw_err_segment.h"
#include "idw_internal.h"

namespace idw {

// One cell of err_cell: {rstar(i), orstar(i), i} of its segment i, and the
// byte offset of err(i) in a NodeRec.  The last segment is {640, 0, 11}, whose
// interpolation (1 - 0) err(11) + 0 * Ai returns err(11) exactly (Ai, the word
// after the table, is finite).
struct SegCell {
    double r, o;
    int i;
    unsigned int off;
    int pad[2];
};
static_assert(sizeof(SegCell) == 32, "SegCell layout");
static_assert(offsetof(NodeRec, Ai) == offsetof(NodeRec, err) + NERR * sizeof(double), "Ai follows the error table");

// rseg: {rstar(i), orstar(i)} for i < NERR-1; rlast = rstar(NERR-1)
__device__ __forceinline__ SegCell seg_cell(int c, const double2 *rseg, double rlast) {
    const int i = err_segment(err_cell_low(c));
    SegCell t;
    t.r = i < NERR - 1 ? rseg[i].x : rlast;
    t.o = i < NERR - 1 ? rseg[i].y : 0.0;
    t.i = i;
    t.off = (unsigned int)(offsetof(NodeRec, err) + i * sizeof(double));
    t.pad[0] = t.pad[1] = 0;
    return t;
}

// getError at dOvrR >= 2 (finite): (1 - fact) err(i) + fact err(i+1) with
// fact = (dOvrR - rstar(i)) orstar(i), unfused.  Any other argument (below 2,
// NaN) reads the first or the last cell and gives a value callers discard.
__device__ __forceinline__ double get_error_cells(const NodeRec *nd, double dOvrR, const SegCell *cells) {
#pragma clang fp contract(off)
    // err_cell(dOvrR) from the high word, clamped to the table at both ends
    const int c = min(max((int)((unsigned)__double2hiint(dOvrR) >> 17) - 8192, 0), ERR_CELLS - 1);
    const SegCell t = cells[c];
    // err(i), err(i+1) as words of the record; err(12) is followed by Ai
    const double *w = reinterpret_cast<const double *>(reinterpret_cast<const char *>(nd) + t.off);
    const double e0 = w[0], e1 = w[1];
    const double fact = (dOvrR - t.r) * t.o;
    return (1.0 - fact) * e0 + fact * e1;
}

// sqrt(x) for normal x >= 2^-767, bit for bit: the hardware reciprocal square
// root refined by Goldschmidt's coupled iteration (g -> sqrt(x), h -> 1/(2
// sqrt(x))) and two Newton corrections of g with fused residuals, the sequence
// the compiler emits for sqrt(double) minus its rescaling of tiny arguments and
// its pass-through of zero and infinity, which never apply here.
__device__ __forceinline__ double sqrt_normal(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

}  // namespace idw
