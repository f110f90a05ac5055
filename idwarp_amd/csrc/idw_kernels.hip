// HIP kernels of the MI355X volume warp (gfx950, wave64).
//
// Work split
//  * surface kernels: one thread per face or per unique surface node;
//  * exact mode: one lane per volume node streaming every surface point
//    (forward), one lane per surface node streaming a chunk of volume nodes
//    (reverse, deterministic chunk partials);
//  * fast mode: one lane per volume node, 64 spatially sorted volume nodes per
//    wave walking the kd-tree together.  The walk is wave-uniform: every wave
//    visits the union of the nodes its lanes need, in depth-first order, and a
//    lane only acts on a node if it descended through all of the node's
//    ancestors.  That per-lane condition is kept in one integer (the depth at
//    which the lane stopped), so no stack is needed and node and point records
//    are fetched with scalar loads.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "idw_err_segment.h"
#include "idw_internal.h"
#include "idw_kernels.h"
#include "idw_walk_math.h"

namespace idw {

__constant__ double c_rstar[NERR] = {2.0, 3.0, 4.0, 5.0, 7.5, 10.0, 20.0, 40.0, 80.0, 160.0, 320.0, 640.0};
// {rstar(i), orstar(i)} pairs; orstar(i) = 1/(rstar(i+1) - rstar(i)) as
// myCreateTree computes it (kd_tree.F90:1539-1541).
__constant__ double2 c_rseg[NERR - 1] = {
    {2.0, 1.0 / (3.0 - 2.0)},     {3.0, 1.0 / (4.0 - 3.0)},       {4.0, 1.0 / (5.0 - 4.0)},
    {5.0, 1.0 / (7.5 - 5.0)},     {7.5, 1.0 / (10.0 - 7.5)},      {10.0, 1.0 / (20.0 - 10.0)},
    {20.0, 1.0 / (40.0 - 20.0)},  {40.0, 1.0 / (80.0 - 40.0)},    {80.0, 1.0 / (160.0 - 80.0)},
    {160.0, 1.0 / (320.0 - 160.0)}, {320.0, 1.0 / (640.0 - 320.0)}};

constexpr double EPS_W = 1e-15;    // constants.F90:33
constexpr double EPS_DEN0 = 1e-16; // getWiEstimate, kd_tree.F90:770,774
constexpr double ROT_TOL = 1.4901161193847656e-08;  // vectorUtils.f90:44

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_block(int nblocks) {
    // Workgroups are dealt round-robin over the 8 XCDs; give each XCD a
    // contiguous range of work so its L2 sees spatially close volume nodes.
    const int b = blockIdx.x;
    const int per = (nblocks + 7) >> 3;
    return (b & 7) * per + (b >> 3);
}

// 1/sqrt(x) for normal x > 0: hardware estimate (v_rsq_f64) refined by two
// Newton steps y += y (1 - x y^2) / 2, within a few ulp of the rounded
// value; replaces sqrt + IEEE division in the weights q = Ldef / d.
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

template <typename T>
__device__ __forceinline__ T ipow_u(T q, int e) {
    T r = T(1);
    while (e > 0) {
        if (e & 1) r *= q;
        q *= q;
        e >>= 1;
    }
    return r;
}

// A_i * (q^aExp + alpha^bExp q^bExp) without the A factor.  EXPM = EXP_35
// compiles the default exponents (3, 5) as products; otherwise the mode is
// read from P at run time.
template <int EXPM = EXP_INT, typename T>
__device__ __forceinline__ T wfun(T q, const WParams &P) {
    if (EXPM == EXP_35 || P.expMode == EXP_35) {
        const T q2 = q * q, q3 = q2 * q, q5 = q3 * q2;
        return q3 + T(P.a5) * q5;
    } else if (P.expMode == EXP_INT) {
        return ipow_u(q, P.ia) + T(P.a5) * ipow_u(q, P.ib);
    } else {
        return pow(q, T(P.aExp)) + T(P.a5) * pow(q, T(P.bExp));
    }
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
#pragma clang fp contract(off)
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Symmetry planes.  getMirrorPt (initializeWarping.F90:533-577): reflect the
// volume node about each plane selected by the pass mask, in plane order.
// Evaluated without FMA contraction, as the reference rounds it: a mirror
// image that lands near the wall turns position ulps into weight errors
// (q^5 at small distance).
__device__ __forceinline__ void mirror_pt(const Mirror &M, double &x, double &y, double &z) {
#pragma clang fp contract(off)
    for (int j = 0; j < M.n; ++j) {
        if (!((M.mask >> j) & 1)) continue;
        const double *n = M.nrm[j], *p = M.p[j];
        const double dp = (x - p[0]) * n[0] + (y - p[1]) * n[1] + (z - p[2]) * n[2];
        x -= 2.0 * n[0] * dp;
        y -= 2.0 * n[1] * dp;
        z -= 2.0 * n[2] * dp;
    }
}

// getMirrorNumerator (:579-600): reflect the pass's numerator back.
__device__ __forceinline__ void mirror_vec(const Mirror &M, double &a, double &b, double &c) {
#pragma clang fp contract(off)
    for (int j = 0; j < M.n; ++j) {
        if (!((M.mask >> j) & 1)) continue;
        const double *n = M.nrm[j];
        const double dp = a * n[0] + b * n[1] + c * n[2];
        a -= 2.0 * dp * n[0];
        b -= 2.0 * dp * n[1];
        c -= 2.0 * dp * n[2];
    }
}

// getMirrorNumerator_b (:602-634): the transpose, planes in reverse order.
__device__ __forceinline__ void mirror_vec_b(const Mirror &M, double &a, double &b, double &c) {
#pragma clang fp contract(off)
    for (int j = M.n - 1; j >= 0; --j) {
        if (!((M.mask >> j) & 1)) continue;
        const double *n = M.nrm[j];
        const double dpb = -2.0 * (a * n[0] + b * n[1] + c * n[2]);
        a += n[0] * dpb;
        b += n[1] * dpb;
        c += n[2] * dpb;
    }
}

// End of a forward launch (warpMesh.F90:56-70): reflect the pass's
// numerator back, combine it with the other passes' (numerator and
// denominator are summed over the passes in pass order), and on the last
// pass write Xv = Xv0 + num / den, or num / den for the tangent
// (warpDerivFwd.F90:92-97).
__device__ __forceinline__ void finish_fwd(const WParams &P, const Mirror &M, int k, int64_t j, const double4 &v0,
                                           double n0, double n1, double n2, double dd, double *__restrict__ xv,
                                           double *__restrict__ den, double4 *__restrict__ acc) {
    if (M.mask) mirror_vec(M, n0, n1, n2);
    if (P.pass == PASS_FIRST) {
        acc[k] = make_double4(n0, n1, n2, dd);
        return;
    }
    if (P.pass != PASS_ONLY) {
        const double4 a = acc[k];
        n0 = a.x + n0;
        n1 = a.y + n1;
        n2 = a.z + n2;
        dd = a.w + dd;
        if (P.pass == PASS_MID) {
            acc[k] = make_double4(n0, n1, n2, dd);
            return;
        }
    }
    const double oden = 1.0 / dd;
    const bool tg = P.tangent != 0;
    xv[3 * j] = (tg ? 0.0 : v0.x) + n0 * oden;
    xv[3 * j + 1] = (tg ? 0.0 : v0.y) + n1 * oden;
    xv[3 * j + 2] = (tg ? 0.0 : v0.z) + n2 * oden;
    if (!tg) den[k] = dd;
}

// ---------------------------------------------------------------------------
// Surface: face props, nodal props (computeNodalProperties, kd_tree.F90:1580)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_xu(const double *xs, const int32_t *xuInd, int32_t node, double *p) {
    const int64_t j = xuInd[node];
    p[0] = xs[3 * j];
    p[1] = xs[3 * j + 1];
    p[2] = xs[3 * j + 2];
}

// The surface kernels below (face properties, nodal normals, rotations,
// corner flags) evaluate the reference's expressions in its operation order
// without FMA contraction, so that they round as the reference does: the
// rotation angle acos(n0 . n) is ill-conditioned for small rotations
// (d angle = d dot / sin(angle)), and an ulp in the dot product of a nearly
// unrotated node becomes a visible rotation difference that far volume
// nodes multiply by their distance.

// getElementProps.F90: fan of cross products about the face centroid.
__global__ void k_face_props(int nface, const int32_t *__restrict__ facePtr, const int32_t *__restrict__ faceConn,
                             const int32_t *__restrict__ xuInd, const double *__restrict__ xs,
                             double4 *__restrict__ fprops) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int p0 = facePtr[f], np = facePtr[f + 1] - p0;
    double c[3] = {0.0, 0.0, 0.0}, p[3];
    for (int i = 0; i < np; ++i) {
        load_xu(xs, xuInd, faceConn[p0 + i], p);
        c[0] += p[0]; c[1] += p[1]; c[2] += p[2];
    }
    c[0] /= np; c[1] /= np; c[2] /= np;
    double first[3], prev[3], nrm[3] = {0.0, 0.0, 0.0}, area = 0.0, cr[3];
    load_xu(xs, xuInd, faceConn[p0], first);
    for (int k = 0; k < 3; ++k) first[k] -= c[k];
    prev[0] = first[0]; prev[1] = first[1]; prev[2] = first[2];
    for (int i = 0; i < np; ++i) {
        double nx[3];
        if (i + 1 < np) {
            load_xu(xs, xuInd, faceConn[p0 + i + 1], nx);
            for (int k = 0; k < 3; ++k) nx[k] -= c[k];
        } else {
            nx[0] = first[0]; nx[1] = first[1]; nx[2] = first[2];
        }
        cross3(prev, nx, cr);
        const double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + 1e-15);
        area += 0.5 * cn;
        nrm[0] += cr[0] / cn; nrm[1] += cr[1] / cn; nrm[2] += cr[2] / cn;
        prev[0] = nx[0]; prev[1] = nx[1]; prev[2] = nx[2];
    }
    const double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    fprops[f] = make_double4(area / np, nrm[0] / nn, nrm[1] / nn, nrm[2] / nn);
}

// getRotationMatrix3d (vectorUtils.f90:31-103), row-major M.
__device__ void rodrigues(const double *v1, const double *v2, double *M) {
#pragma clang fp contract(off)
    double axis[3];
    cross3(v1, v2, axis);
    const double axisMag = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2] + 1e-30);
    double s = 0.0, oc = 0.0;
    if (axisMag < ROT_TOL) {
        axis[0] = 1.0; axis[1] = 0.0; axis[2] = 0.0;
    } else {
        const double m1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2] + 1e-30);
        const double m2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2] + 1e-30);
        for (int k = 0; k < 3; ++k) axis[k] /= axisMag;
        const double dot = (v1[0] / m1) * (v2[0] / m2) + (v1[1] / m1) * (v2[1] / m2) + (v1[2] / m1) * (v2[2] / m2);
        const double ang = acos(fmin(1.0, dot));
        s = sin(ang);
        oc = 1.0 - cos(ang);
    }
    // M = I + sin(a) A + (1 - cos(a)) A^2 with A the skew matrix of the axis
    const double A[9] = {0.0, -axis[2], axis[1], axis[2], 0.0, -axis[0], -axis[1], axis[0], 0.0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double C = A[3 * r] * A[c] + A[3 * r + 1] * A[3 + c] + A[3 * r + 2] * A[6 + c];
            M[3 * r + c] = (r == c ? 1.0 : 0.0) + s * A[3 * r + c] + oc * C;
        }
}

// Per node: area-weighted normal over incident faces; at init it becomes
// normals0 and A_i, later the rotation M_i (or I), stored as K = M - I and
// d = Xu - Xu0.  Zero-area nodes get M = I, b = 0 (kd_tree.F90:1651-1663).
__global__ void k_node_props(int n, int initial, int useRot, const int32_t *__restrict__ n2eStart,
                             const int32_t *__restrict__ n2e, const double4 *__restrict__ fprops,
                             const int32_t *__restrict__ xuInd, const int32_t *__restrict__ corner,
                             const double *__restrict__ xs, double4 *__restrict__ ps, double *__restrict__ normals0,
                             double4 *__restrict__ nprops, Rec64 *__restrict__ pd, Rec32 *__restrict__ pd32) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sa = 0.0, sn[3] = {0.0, 0.0, 0.0};
    for (int e = n2eStart[i]; e < n2eStart[i + 1]; ++e) {
        const double4 fp = fprops[n2e[e]];
        sa += fp.x;
        sn[0] += fp.x * fp.y; sn[1] += fp.x * fp.z; sn[2] += fp.x * fp.w;
    }
    const double nrm[3] = {sn[0] / sa, sn[1] / sa, sn[2] / sa};
    nprops[i] = make_double4(sa, sn[0], sn[1], sn[2]);
    if (initial) {
        normals0[3 * i] = nrm[0]; normals0[3 * i + 1] = nrm[1]; normals0[3 * i + 2] = nrm[2];
        double4 p = ps[i];
        p.w = sa;
        ps[i] = p;
        return;
    }
    const double4 p = ps[i];
    Rec64 out;
    out.x[0] = p.x; out.x[1] = p.y; out.x[2] = p.z;
    out.A = p.w;
    double M[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    double xu[3];
    load_xu(xs, xuInd, i, xu);
    if (p.w == 0.0) {
        for (int k = 0; k < 9; ++k) out.K[k] = 0.0;
        out.d[0] = out.d[1] = out.d[2] = 0.0;
    } else {
        if (useRot && !corner[i]) {
            const double n0[3] = {normals0[3 * i], normals0[3 * i + 1], normals0[3 * i + 2]};
            rodrigues(n0, nrm, M);
        }
        for (int k = 0; k < 9; ++k) out.K[k] = M[k] - ((k == 0 || k == 4 || k == 8) ? 1.0 : 0.0);
        out.d[0] = xu[0] - p.x; out.d[1] = xu[1] - p.y; out.d[2] = xu[2] - p.z;
    }
    pd[i] = out;
    Rec32 o32;
    o32.x[0] = p.x; o32.x[1] = p.y; o32.x[2] = p.z;
    o32.A = (float)p.w;
    for (int k = 0; k < 9; ++k) o32.K[k] = (float)out.K[k];
    for (int k = 0; k < 3; ++k) o32.d[k] = (float)out.d[k];
    pd32[i] = o32;
}

// ---------------------------------------------------------------------------
// Tangent of the nodal properties (COMPUTENODALPROPERTIES_D,
// kd_tree.F90:1275-1360) for a surface tangent dxs
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(const double *a, const double *b) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Face area / point count and unit normal with their tangents
// (GETELEMENTPROPS_D, getElementProps_d.f90:7-94).
__global__ void k_face_props_d(int nface, const int32_t *__restrict__ facePtr, const int32_t *__restrict__ faceConn,
                               const int32_t *__restrict__ xuInd, const double *__restrict__ xs,
                               const double *__restrict__ dxs, double4 *__restrict__ fprops,
                               double4 *__restrict__ fpropsd) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int p0 = facePtr[f], np = facePtr[f + 1] - p0;
    double c[3] = {0.0, 0.0, 0.0}, cd[3] = {0.0, 0.0, 0.0}, p[3], pd[3];
    for (int i = 0; i < np; ++i) {
        load_xu(xs, xuInd, faceConn[p0 + i], p);
        load_xu(dxs, xuInd, faceConn[p0 + i], pd);
        for (int k = 0; k < 3; ++k) { c[k] += p[k]; cd[k] += pd[k]; }
    }
    for (int k = 0; k < 3; ++k) { c[k] /= np; cd[k] /= np; }
    double first[3], firstd[3], prev[3], prevd[3];
    load_xu(xs, xuInd, faceConn[p0], first);
    load_xu(dxs, xuInd, faceConn[p0], firstd);
    for (int k = 0; k < 3; ++k) {
        first[k] -= c[k]; firstd[k] -= cd[k];
        prev[k] = first[k]; prevd[k] = firstd[k];
    }
    double nrm[3] = {0.0, 0.0, 0.0}, nrmd[3] = {0.0, 0.0, 0.0}, area = 0.0, aread = 0.0;
    for (int i = 0; i < np; ++i) {
        double nx[3], nxd[3];
        if (i + 1 < np) {
            load_xu(xs, xuInd, faceConn[p0 + i + 1], nx);
            load_xu(dxs, xuInd, faceConn[p0 + i + 1], nxd);
            for (int k = 0; k < 3; ++k) { nx[k] -= c[k]; nxd[k] -= cd[k]; }
        } else {
            for (int k = 0; k < 3; ++k) { nx[k] = first[k]; nxd[k] = firstd[k]; }
        }
        double cr[3], t1[3], t2[3], crd[3];
        cross3(prev, nx, cr);
        cross3(prevd, nx, t1);
        cross3(prev, nxd, t2);
        for (int k = 0; k < 3; ++k) crd[k] = t1[k] + t2[k];
        const double cn = sqrt(dot3(cr, cr) + 1e-15);
        const double cnd = dot3(cr, crd) / cn;
        area += 0.5 * cn;
        aread += 0.5 * cnd;
        for (int k = 0; k < 3; ++k) {
            nrmd[k] += (crd[k] - cr[k] * cnd / cn) / cn;
            nrm[k] += cr[k] / cn;
        }
        for (int k = 0; k < 3; ++k) { prev[k] = nx[k]; prevd[k] = nxd[k]; }
    }
    const double nn = sqrt(dot3(nrm, nrm));
    const double nnd = dot3(nrm, nrmd) / nn;
    fprops[f] = make_double4(area / np, nrm[0] / nn, nrm[1] / nn, nrm[2] / nn);
    fpropsd[f] = make_double4(aread / np, (nrmd[0] - nrm[0] * nnd / nn) / nn, (nrmd[1] - nrm[1] * nnd / nn) / nn,
                              (nrmd[2] - nrm[2] * nnd / nn) / nn);
}

// Tangent of the rotation taking v1 to v2 for a tangent v2d of v2
// (GETROTATIONMATRIX3D_D, vectorUtils_d.f90:7-129), row-major: with
// M = I + sin(a) A + (1 - cos(a)) A^2, Md = (cos(a) A + sin(a) A^2) a' +
// sin(a) A' + (1 - cos(a)) (A' A + A A').  Zero below the axis tolerance.
__device__ void rodrigues_d(const double *v1, const double *v2, const double *v2d, double *Md) {
#pragma clang fp contract(off)
    for (int k = 0; k < 9; ++k) Md[k] = 0.0;
    double ax[3], axd[3];
    cross3(v1, v2, ax);
    cross3(v1, v2d, axd);
    const double am = sqrt(dot3(ax, ax) + 1e-30);
    if (am < ROT_TOL) return;
    const double amd = dot3(ax, axd) / am;
    double u[3], ud[3];
    for (int k = 0; k < 3; ++k) {
        ud[k] = (axd[k] - ax[k] * amd / am) / am;
        u[k] = ax[k] / am;
    }
    const double m1 = sqrt(dot3(v1, v1) + 1e-30), m2 = sqrt(dot3(v2, v2) + 1e-30);
    const double m2d = dot3(v2, v2d) / m2;
    double vv1[3], vv2[3], vv2d[3];
    for (int k = 0; k < 3; ++k) {
        vv1[k] = v1[k] / m1;
        vv2[k] = v2[k] / m2;
        vv2d[k] = (v2d[k] - v2[k] * m2d / m2) / m2;
    }
    const double dot = dot3(vv1, vv2);
    double arg = 1.0, argd = 0.0;
    if (1.0 > dot) {
        arg = dot;
        argd = dot3(vv1, vv2d);
    }
    const double angd = (arg == 1.0 || arg == -1.0) ? 0.0 : -(argd / sqrt(1.0 - arg * arg));
    const double ang = acos(arg), sn = sin(ang), cs = cos(ang), oc = 1.0 - cs;
    const double A[9] = {0.0, -u[2], u[1], u[2], 0.0, -u[0], -u[1], u[0], 0.0};
    const double Ad[9] = {0.0, -ud[2], ud[1], ud[2], 0.0, -ud[0], -ud[1], ud[0], 0.0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double C = 0.0, Cd = 0.0;
            for (int q = 0; q < 3; ++q) {
                C += A[3 * r + q] * A[3 * q + c];
                Cd += Ad[3 * r + q] * A[3 * q + c] + A[3 * r + q] * Ad[3 * q + c];
            }
            Md[3 * r + c] = (A[3 * r + c] * cs + C * sn) * angd + sn * Ad[3 * r + c] + oc * Cd;
        }
}

// Tangent point records: the forward kernels evaluate
// sum w (K (r - Xu0) + d) / sum w, and the tangent of the displacement
// M r + b - r with b = Xu - M Xu0 is Md (r - Xu0) + Xud, so the record holds
// K = Md and d = Xud (evalNodeExact_d, kd_tree.F90:1205-1228).  Zero-area
// nodes carry zero weight; their tangent record is set to zero.
__global__ void k_node_props_d(int n, int useRot, const int32_t *__restrict__ n2eStart,
                               const int32_t *__restrict__ n2e, const double4 *__restrict__ fprops,
                               const double4 *__restrict__ fpropsd, const int32_t *__restrict__ xuInd,
                               const int32_t *__restrict__ corner, const double *__restrict__ dxs,
                               const double4 *__restrict__ ps, const double *__restrict__ normals0,
                               Rec64 *__restrict__ pd, Rec32 *__restrict__ pd32) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sa = 0.0, sad = 0.0, sn[3] = {0.0, 0.0, 0.0}, snd[3] = {0.0, 0.0, 0.0};
    for (int e = n2eStart[i]; e < n2eStart[i + 1]; ++e) {
        const double4 fp = fprops[n2e[e]], fd = fpropsd[n2e[e]];
        sa += fp.x;
        sad += fd.x;
        const double fn[3] = {fp.y, fp.z, fp.w}, fnd[3] = {fd.y, fd.z, fd.w};
        for (int k = 0; k < 3; ++k) {
            snd[k] += fd.x * fn[k] + fp.x * fnd[k];
            sn[k] += fp.x * fn[k];
        }
    }
    const double4 p = ps[i];
    Rec64 out;
    out.x[0] = p.x; out.x[1] = p.y; out.x[2] = p.z;
    out.A = p.w;
    double Md[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double xud[3];
    load_xu(dxs, xuInd, i, xud);
    if (p.w == 0.0 || sa == 0.0) {
        xud[0] = xud[1] = xud[2] = 0.0;
    } else if (useRot && !corner[i]) {
        double nrm[3], nrmd[3];
        for (int k = 0; k < 3; ++k) {
            nrm[k] = sn[k] / sa;
            nrmd[k] = (snd[k] * sa - sn[k] * sad) / (sa * sa);
        }
        const double n0[3] = {normals0[3 * i], normals0[3 * i + 1], normals0[3 * i + 2]};
        rodrigues_d(n0, nrm, nrmd, Md);
    }
    for (int k = 0; k < 9; ++k) out.K[k] = Md[k];
    for (int k = 0; k < 3; ++k) out.d[k] = xud[k];
    pd[i] = out;
    Rec32 o32;
    o32.x[0] = p.x; o32.x[1] = p.y; o32.x[2] = p.z;
    o32.A = (float)p.w;
    for (int k = 0; k < 9; ++k) o32.K[k] = (float)out.K[k];
    for (int k = 0; k < 3; ++k) o32.d[k] = (float)out.d[k];
    pd32[i] = o32;
}

// determineCorners (kd_tree.F90:1666-1715) against the initial surface.
__global__ void k_corners(int n, double threshold, const int32_t *__restrict__ n2eStart, const int32_t *__restrict__ n2e,
                          const double4 *__restrict__ fprops, const double *__restrict__ normals0,
                          int32_t *__restrict__ corner) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    for (int e = n2eStart[i]; e < n2eStart[i + 1]; ++e) {
        const double4 fp = fprops[n2e[e]];
        const double dp = fp.y * normals0[3 * i] + fp.z * normals0[3 * i + 1] + fp.w * normals0[3 * i + 2];
        if (fabs(dp) < threshold) c = 1;
    }
    corner[i] = c;
}

// setData (kd_tree.F90:595-617) bottom up: per node the sums over its
// points of K (9), d (3) and K (X - p) (3, X the node's centroid), first for
// every bucket from its points, then for each internal node from its two
// children (node + 1 and that child's skip), shifting the children's
// K (X_c - p) sums to X with K_c (X - X_c).  An internal node's cluster
// record is K = mean(K), d = mean(d) + mean(K (X - p)), the form of
// d = K X + mean(b) with b = Xu - M Xu0 (the same for point and tangent
// records, which enter linearly).
__device__ __forceinline__ void leaf_sums(int node, const NodeRec *__restrict__ nodes,
                                          const double4 *__restrict__ ps, const Rec64 *__restrict__ pd,
                                          double *csum) {
    const NodeRec &r = nodes[node];
    double acc[15];
    for (int k = 0; k < 15; ++k) acc[k] = 0.0;
    for (int i = r.l; i <= r.u; ++i) {
        const Rec64 &q = pd[i];
        const double4 p = ps[i];
        const double dx = r.X[0] - p.x, dy = r.X[1] - p.y, dz = r.X[2] - p.z;
        for (int k = 0; k < 9; ++k) acc[k] += q.K[k];
        for (int k = 0; k < 3; ++k) {
            acc[9 + k] += q.d[k];
            acc[12 + k] += q.K[3 * k] * dx + q.K[3 * k + 1] * dy + q.K[3 * k + 2] * dz;
        }
    }
    double *o = csum + (int64_t)node * CSUM;
    for (int k = 0; k < 15; ++k) o[k] = acc[k];
}

__device__ __forceinline__ void node_sums(int node, const NodeRec *__restrict__ nodes, double *csum,
                                          Rec64 *__restrict__ rec, Rec32 *__restrict__ rec32) {
    const NodeRec &r = nodes[node];
    const int ch[2] = {node + 1, nodes[node + 1].skip};
    double acc[15];
    for (int k = 0; k < 15; ++k) acc[k] = 0.0;
    for (int c = 0; c < 2; ++c) {
        const double *cs = csum + (int64_t)ch[c] * CSUM;
        const NodeRec &rc = nodes[ch[c]];
        const double sx = r.X[0] - rc.X[0], sy = r.X[1] - rc.X[1], sz = r.X[2] - rc.X[2];
        for (int k = 0; k < 12; ++k) acc[k] += cs[k];
        for (int k = 0; k < 3; ++k) acc[12 + k] += cs[12 + k] + (cs[3 * k] * sx + cs[3 * k + 1] * sy + cs[3 * k + 2] * sz);
    }
    double *o = csum + (int64_t)node * CSUM;
    for (int k = 0; k < 15; ++k) o[k] = acc[k];
    Rec64 e;
    Rec32 e32;
    for (int k = 0; k < 3; ++k) e.x[k] = e32.x[k] = r.X[k];
    e.A = r.Ai;
    e32.A = (float)r.Ai;
    for (int k = 0; k < 9; ++k) {
        e.K[k] = acc[k] / r.n;
        e32.K[k] = (float)e.K[k];
    }
    for (int k = 0; k < 3; ++k) {
        e.d[k] = acc[9 + k] / r.n + acc[12 + k] / r.n;
        e32.d[k] = (float)e.d[k];
    }
    rec[node] = e;
    rec32[node] = e32;
}

// The lower tree, one block per cut root (a node at the cut depth, or a
// bucket above it): its subtree, the node range [root, skip), one depth at a
// time from the subtree's deepest (cut[2 b + 1]) to the root's, a barrier
// between depths.  Every sum a node reads is its children's, written by this
// block one depth earlier.
__global__ void __launch_bounds__(256) k_subtree_sums(const int32_t *__restrict__ cut,
                                                      const NodeRec *__restrict__ nodes,
                                                      const double4 *__restrict__ ps, const Rec64 *__restrict__ pd,
                                                      double *csum, Rec64 *__restrict__ rec,
                                                      Rec32 *__restrict__ rec32) {
    const int root = cut[2 * blockIdx.x], deepest = cut[2 * blockIdx.x + 1];
    const int end = nodes[root].skip, top = nodes[root].depth;
    for (int d = deepest; d >= top; --d) {
        for (int k = root + (int)threadIdx.x; k < end; k += blockDim.x) {
            if (nodes[k].depth != d) continue;
            if (nodes[k].leaf) leaf_sums(k, nodes, ps, pd, csum);
            else node_sums(k, nodes, csum, rec, rec32);
        }
        __syncthreads();
    }
}

// The internal nodes above the cut, in one block: level i is
// list[T.start[i] .. T.start[i] + T.count[i]), from the deepest to the root.
__global__ void __launch_bounds__(1024) k_top_sums(TopLevels T, const int32_t *__restrict__ list,
                                                   const NodeRec *__restrict__ nodes, double *csum,
                                                   Rec64 *__restrict__ rec, Rec32 *__restrict__ rec32) {
    for (int i = 0; i < T.n; ++i) {
        for (int t = threadIdx.x; t < T.count[i]; t += blockDim.x) node_sums(list[T.start[i] + t], nodes, csum, rec, rec32);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Exact mode (evalNodeExact on the root, kd_tree.F90:664-690)
// ---------------------------------------------------------------------------
// The VALU form, one lane per volume node: the FP64 forward, and in mixed
// precision the stored-weight forward (the mixed forward otherwise runs on
// the matrix cores, k_exact_fwd_mc).  G: surface points per FP32 partial sum
// in mixed precision.  STORE: every pair's weight, as used, is also written to
// wout[i ld + k] (surface point i, sorted volume node k; rows padded to
// ld = wrow(nvol), a multiple of 4) for the reverse.
template <int PREC, int EXPM, int G, bool STORE = false>
__global__ void __launch_bounds__(256) k_exact_fwd(int nvol, const double4 *__restrict__ vol,
                                                   const int32_t *__restrict__ perm, int n,
                                                   const double4 *__restrict__ ps, const Rec64 *__restrict__ pd,
                                                   const Rec32 *__restrict__ pd32, WParams P, Mirror M,
                                                   double *__restrict__ xv, double *__restrict__ den,
                                                   double4 *__restrict__ acc, double *__restrict__ wout = nullptr) {
    const int nb = (nvol + 255) / 256;
    const int blk = xcd_block(nb);
    if (blk >= nb) return;
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    const double4 v0 = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    double4 v = v0;
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    double n0 = 0.0, n1 = 0.0, n2 = 0.0, dd = 0.0;
    if (PREC == IDW_PREC_FP64) {
        for (int i = 0; i < n; ++i) {
            const double4 s = ps[i];
            const Rec64 &q = pd[i];
            const double rx = v.x - s.x, ry = v.y - s.y, rz = v.z - s.z;
            const double lo = P.Ldef * rsqrt_nr(rx * rx + ry * ry + rz * rz + EPS_W);
            const double w = s.w * wfun<EXPM>(lo, P);
            if (STORE && valid) wout[(int64_t)i * wrow(nvol) + k] = w;
            n0 += w * (q.K[0] * rx + q.K[1] * ry + q.K[2] * rz + q.d[0]);
            n1 += w * (q.K[3] * rx + q.K[4] * ry + q.K[5] * rz + q.d[1]);
            n2 += w * (q.K[6] * rx + q.K[7] * ry + q.K[8] * rz + q.d[2]);
            dd += w;
        }
    } else {
        // FP32 weight and displacement on the FP64 difference r - Xu0; pairs
        // nearer than nearWallFactor*Ldef for any lane of the wave stay FP64.
        const float Lf = (float)P.Ldef;
        const double T2 = P.nearT2;
        for (int i0 = 0; i0 < n; i0 += G) {
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, ad = 0.f;
            const int ie = min(i0 + G, n);
            for (int i = i0; i < ie; ++i) {
                const double4 s = ps[i];
                const double rx = v.x - s.x, ry = v.y - s.y, rz = v.z - s.z;
                const double d2 = rx * rx + ry * ry + rz * rz;
                if (__any(valid && d2 < T2)) {
                    const Rec64 &q = pd[i];
                    const double lo = P.Ldef * rsqrt_nr(d2 + EPS_W);
                    const double w = s.w * wfun<EXPM>(lo, P);
                    if (STORE && valid) wout[(int64_t)i * wrow(nvol) + k] = w;
                    n0 += w * (q.K[0] * rx + q.K[1] * ry + q.K[2] * rz + q.d[0]);
                    n1 += w * (q.K[3] * rx + q.K[4] * ry + q.K[5] * rz + q.d[1]);
                    n2 += w * (q.K[6] * rx + q.K[7] * ry + q.K[8] * rz + q.d[2]);
                    dd += w;
                } else {
                    const Rec32 &q = pd32[i];
                    const float fx = (float)rx, fy = (float)ry, fz = (float)rz;
                    // |r|^2 + eps (evalNodeExact's eps) in FP64, rounded once
                    const float lo = Lf * __frsqrt_rn((float)(d2 + EPS_W));
                    const float w = (float)s.w * wfun<EXPM>(lo, P);
                    if (STORE && valid) wout[(int64_t)i * wrow(nvol) + k] = (double)w;
                    a0 += w * (q.K[0] * fx + q.K[1] * fy + q.K[2] * fz + q.d[0]);
                    a1 += w * (q.K[3] * fx + q.K[4] * fy + q.K[5] * fz + q.d[1]);
                    a2 += w * (q.K[6] * fx + q.K[7] * fy + q.K[8] * fz + q.d[2]);
                    ad += w;
                }
            }
            n0 += a0; n1 += a1; n2 += a2; dd += ad;
        }
    }
    if (valid) finish_fwd(P, M, k, perm[k], v0, n0, n1, n2, dd, xv, den, acc);
}

// Reverse seed numb = dXv / den (warpDeriv.F90:76-78), sorted order; for a
// mirror pass the seed is reflected (getMirrorNumerator_b) and the mirrored
// positions are written to volm.
__global__ void k_prep_numb(int nvol, const int32_t *__restrict__ perm, const double *__restrict__ dxv,
                            const double *__restrict__ den, Mirror M, const double4 *__restrict__ vol,
                            double4 *__restrict__ numb, double4 *__restrict__ volm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvol) return;
    const int64_t j = perm[k];
    const double oden = 1.0 / den[k];
    double a = dxv[3 * j] * oden, b = dxv[3 * j + 1] * oden, c = dxv[3 * j + 2] * oden;
    if (M.mask) {
        mirror_vec_b(M, a, b, c);
        double4 v = vol[k];
        mirror_pt(M, v.x, v.y, v.z);
        volm[k] = v;
    }
    numb[k] = make_double4(a, b, c, 0.0);
}

// A seed: a nonzero component of dXv (node k in traversal order).
__device__ __forceinline__ bool seeded(int nvol, int64_t k, const int32_t *__restrict__ perm,
                                       const double *__restrict__ dxv) {
    if (k >= nvol) return false;
    const int64_t j = perm[k];
    return dxv[3 * j] != 0.0 || dxv[3 * j + 1] != 0.0 || dxv[3 * j + 2] != 0.0;
}

// The exact reverse runs over the seeded nodes only: a node whose seed is
// zero adds exactly zero to every sum.  They are packed in traversal order,
// so each surface node's sums keep their order: per block of 256 nodes the
// seeded count (k_seed_count), an exclusive scan of the counts
// (k_scan_counts, in place, the total after the last), then the packed
// seeds dXv/den and (mirrored) positions (k_pack_seeds).
__global__ void __launch_bounds__(256) k_seed_count(int nvol, const int32_t *__restrict__ perm,
                                                    const double *__restrict__ dxv, int32_t *__restrict__ cnt) {
    __shared__ int wc[4];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long b = __ballot(seeded(nvol, k, perm, dxv));
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Exclusive scan of v[0..n) in place, the total to v[n]; one workgroup of 1024.
__global__ void __launch_bounds__(1024) k_scan_counts(int n, int32_t *__restrict__ v) {
    __shared__ int ws[16];
    __shared__ int carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + t;
        const int x = i < n ? v[i] : 0;
        int s = x;  // inclusive scan over the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(s, d, 64);
            if (lane >= d) s += y;
        }
        if (lane == 63) ws[w] = s;
        __syncthreads();
        if (w == 0) {  // inclusive scan of the 16 wave totals
            int u = lane < 16 ? ws[lane] : 0;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int y = __shfl_up(u, d, 64);
                if (lane >= d) u += y;
            }
            if (lane < 16) ws[lane] = u;
        }
        __syncthreads();
        const int before = carry + (w > 0 ? ws[w - 1] : 0) + s - x;
        if (i < n) v[i] = before;
        __syncthreads();  // carry and ws are read
        if (t == 1023) carry = before + x;
        __syncthreads();
    }
    if (t == 0) v[n] = carry;
}

__global__ void __launch_bounds__(256) k_pack_seeds(int nvol, const int32_t *__restrict__ perm,
                                                    const double *__restrict__ dxv, const double *__restrict__ den,
                                                    Mirror M, const double4 *__restrict__ vol,
                                                    const int32_t *__restrict__ off, double4 *__restrict__ numb,
                                                    double4 *__restrict__ pos) {
    __shared__ int wc[4];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool sd = seeded(nvol, k, perm, dxv);
    const unsigned long long b = __ballot(sd);
    if (lane == 0) wc[w] = __popcll(b);
    __syncthreads();
    if (!sd) return;
    int r = off[blockIdx.x] + __popcll(b & ((1ull << lane) - 1ull));
    for (int u = 0; u < w; ++u) r += wc[u];
    const int64_t j = perm[k];
    const double oden = 1.0 / den[k];
    double a = dxv[3 * j] * oden, bb = dxv[3 * j + 1] * oden, cc = dxv[3 * j + 2] * oden;
    double4 v = vol[k];
    if (M.mask) {
        mirror_vec_b(M, a, bb, cc);
        mirror_pt(M, v.x, v.y, v.z);
    }
    numb[r] = make_double4(a, bb, cc, 0.0);
    pos[r] = v;
}

// evalNodeExact_b (kd_tree.F90:984-1013) transposed: one lane per surface
// node, a chunk of the packed seeded nodes per block row (the nchunk chunks
// split the *count nodes evenly).  Accumulates
// Mb'(r,c) = sum s_r (x_c - Xu0_c) (the reverse's Mib - Xu0 (x) Bib of
// COMPUTENODALPROPERTIES_B, kd_tree.F90:1115-1118) and Bb = sum s.
template <int PREC, int EXPM>
__global__ void __launch_bounds__(256) k_exact_rev(int nchunk, const int32_t *__restrict__ count,
                                                   const double4 *__restrict__ vol,
                                                   const double4 *__restrict__ numb, int n,
                                                   const double4 *__restrict__ ps, WParams P,
                                                   double *__restrict__ partial) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    const int nv = *count, chunk = (nv + nchunk - 1) / nchunk;
    const int j0 = c * chunk, j1 = min(nv, j0 + chunk);
    const bool valid = i < n;
    const double4 s = valid ? ps[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    double acc[12];
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double4 v = vol[j];
        const double4 nb = numb[j];
        const double rx = v.x - s.x, ry = v.y - s.y, rz = v.z - s.z;
        double w;
        if (PREC == IDW_PREC_FP64) {
            const double lo = P.Ldef * rsqrt_nr(rx * rx + ry * ry + rz * rz + EPS_W);
            w = s.w * wfun<EXPM>(lo, P);
        } else {
            const double d2 = rx * rx + ry * ry + rz * rz;
            if (__any(valid && d2 < P.nearT2)) {
                const double lo = P.Ldef * rsqrt_nr(d2 + EPS_W);
                w = s.w * wfun<EXPM>(lo, P);
            } else {
                const float lo = (float)P.Ldef * __frsqrt_rn((float)(d2 + EPS_W));
                w = (double)((float)s.w * wfun<EXPM>(lo, P));
            }
        }
        const double s0 = w * nb.x, s1 = w * nb.y, s2 = w * nb.z;
        acc[0] += s0 * rx; acc[1] += s0 * ry; acc[2] += s0 * rz;
        acc[3] += s1 * rx; acc[4] += s1 * ry; acc[5] += s1 * rz;
        acc[6] += s2 * rx; acc[7] += s2 * ry; acc[8] += s2 * rz;
        acc[9] += s0; acc[10] += s1; acc[11] += s2;
    }
    if (valid) {
        double *o = partial + ((int64_t)c * n + i) * 12;
        for (int k = 0; k < 12; ++k) o[k] = acc[k];
    }
}

// k_exact_rev with the forward's stored weights (w[i ld + j]) in place of
// recomputed ones: lane i streams its own row of the weights, four at a time
// (chunk is a multiple of 4, so every load is 32-byte aligned; the row's
// padding is loaded, never used).
__global__ void __launch_bounds__(256) k_exact_rev_w(int nvol, int chunk, const double4 *__restrict__ vol,
                                                     const double4 *__restrict__ numb, int n,
                                                     const double4 *__restrict__ ps, const double *__restrict__ wst,
                                                     double *__restrict__ partial) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    const int j0 = c * chunk, j1 = min(nvol, j0 + chunk);
    const bool valid = i < n;
    const double4 s = valid ? ps[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double *wr = wst + (int64_t)(valid ? i : 0) * wrow(nvol);
    double acc[12];
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    for (int j4 = j0; j4 < j1; j4 += 4) {
        const double4 w4 = *reinterpret_cast<const double4 *>(wr + j4);
        const double ws[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = j4 + q;
            if (j >= j1) break;  // uniform: the chunk's end
            const double4 v = vol[j];
            const double4 nb = numb[j];
            const double rx = v.x - s.x, ry = v.y - s.y, rz = v.z - s.z;
            const double w = ws[q];
            const double s0 = w * nb.x, s1 = w * nb.y, s2 = w * nb.z;
            acc[0] += s0 * rx; acc[1] += s0 * ry; acc[2] += s0 * rz;
            acc[3] += s1 * rx; acc[4] += s1 * ry; acc[5] += s1 * rz;
            acc[6] += s2 * rx; acc[7] += s2 * ry; acc[8] += s2 * rz;
            acc[9] += s0; acc[10] += s1; acc[11] += s2;
        }
    }
    if (valid) {
        double *o = partial + ((int64_t)c * n + i) * 12;
        for (int k = 0; k < 12; ++k) o[k] = acc[k];
    }
}

__global__ void k_exact_rev_reduce(int n, int nchunk, const double *__restrict__ partial, double *__restrict__ padj,
                                   int add) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 12) return;
    const int i = t / 12, k = t % 12;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[((int64_t)c * n + i) * 12 + k];
    double *o = padj + (int64_t)i * ADJ_STRIDE + k;
    *o = add ? *o + s : s;
}

// ---------------------------------------------------------------------------
// Fast mode: wave-uniform kd-tree walk
// ---------------------------------------------------------------------------
// getError (kd_tree.F90:882-906): linear interpolation of the node's error
// table at dOvrR = dist * oradius >= 2; rseg holds {rstar(i), orstar(i)}.
static_assert(NERR == 12, "err_segment covers the 12 reference radii");
__device__ __forceinline__ double get_error(const NodeRec *nd, double dOvrR, const double2 *rseg) {
#pragma clang fp contract(off)
    const int i = err_segment(dOvrR);
    if (i == NERR - 1) return nd->err[NERR - 1];
    const double2 seg = rseg[i];
    const double fact = (dOvrR - seg.x) * seg.y;
    return (1.0 - fact) * nd->err[i] + fact * nd->err[i + 1];
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// This lane's bit of a wave-uniform lane mask, as a lane predicate.
__device__ __forceinline__ bool lane_in(unsigned long long m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// Wait for every outstanding vector-memory operation (vmcnt(0)).  The
// counter retires in order, so a wait for a load issued after stores or
// atomics also waits for those: the kernels place it where that costs least.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// LDS written by one lane and read by another lane of the same wave: the
// wave's LDS operations execute in order, the barrier keeps the compiler
// from moving them.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WBUF: plan entries of each kind a wave gathers in LDS before writing them
// out (64, 128 or 256, a tuning parameter).

// Wave-level record of the walk: one entry per term (accepted cluster, or
// point of an evaluated bucket) with the mask of the lanes it serves, in
// depth-first order.  The wave owns slots [off, end) sized by the counting
// pass; FP32 terms fill them from the front, FP64 terms (mixed precision
// near the wall) from the back, and the split is stored per wave.  Entries of
// both kinds are gathered in LDS and written out WBUF at a time, so the walk's
// vector-memory queue (which retires in order) is rarely held up by stores
// in front of its error-table loads.  Writes outside the range are dropped
// and a count that disagrees with the counting pass is reported in
// *mismatch.
template <int WBUF>
struct PlanWriter {
    PlanEntry *plan;
    int64_t off, end, nf = 0, nb = 0;
    unsigned int *mismatch;
    PlanEntry *buf, *buf64;  // this wave's LDS buffers (WBUF entries each)
    int cf = 0, cb = 0;      // entries in buf, buf64 (wave-uniform)
    __device__ __forceinline__ void flush() {
        const int lane = threadIdx.x & 63;
        wave_sync();
        for (int j = lane; j < cf; j += 64) {
            const int64_t at = off + nf + j;
            if (at < end) plan[at] = buf[j];
        }
        nf += cf;
        cf = 0;
        wave_sync();
    }
    __device__ __forceinline__ void flush64() {
        const int lane = threadIdx.x & 63;
        wave_sync();
        for (int j = lane; j < cb; j += 64) {
            const int64_t at = end - 1 - (nb + j);
            if (at >= off) plan[at] = buf64[j];
        }
        nb += cb;
        cb = 0;
        wave_sync();
    }
    // terms base .. base+n-1 for the lanes of mask, found at tree node node
    __device__ __forceinline__ void put(int base, int n, unsigned long long mask, bool fp64, int node) {
        const int lane = threadIdx.x & 63;
        if (n > WBUF / 2) {
            if (fp64) flush64();
            else flush();
            for (int i = lane; i < n; i += 64) {
                const int64_t at = fp64 ? end - 1 - (nb + i) : off + nf + i;
                if (at >= off && at < end) {
                    PlanEntry e;
                    e.mask = mask;
                    e.rec = base + i;
                    e.node = node;
                    plan[at] = e;
                }
            }
            if (fp64) nb += n;
            else nf += n;
            return;
        }
        int &c = fp64 ? cb : cf;
        if (c + n > WBUF) {
            if (fp64) flush64();
            else flush();
        }
        if (lane < n) {
            PlanEntry e;
            e.mask = mask;
            e.rec = base + lane;
            e.node = node;
            (fp64 ? buf64 : buf)[c + lane] = e;
        }
        c += n;
    }
    // allFp64: the wave's FP32 terms are evaluated in FP64 as well (its
    // split moves to the front of its range)
    __device__ __forceinline__ void finish(int64_t *split, int wave, bool allFp64) {
        flush();
        flush64();
        if ((threadIdx.x & 63) == 0) {
            split[wave] = allFp64 ? off : off + nf;
            if (nf + nb != end - off) atomicAdd(mismatch, 1u);
        }
    }
};

__device__ __forceinline__ NodeHdr load_hdr(const NodeRec *nodes, int node) {
    return *reinterpret_cast<const NodeHdr *>(nodes + node);
}

// Walk header of the node at byte offset off (a multiple of sizeof(WalkHdr)).
__device__ __forceinline__ WalkHdr load_whdr(const WalkHdr *w, unsigned off) {
    return *reinterpret_cast<const WalkHdr *>(reinterpret_cast<const char *>(w) + off);
}

// One visit of the walk below at node (walk header h):
// the lane decisions, the plan entries, and the wave's next node.
template <bool COUNT, int WBUF>
__device__ __forceinline__ int walk_visit(const WalkHdr &h, int node, const double4 &v, double tol,
                                          int &stop, PlanWriter<WBUF> &pw, int64_t &cnt, const NodeRec *nodes,
                                          int nnodes, const SegCell *cells, double nearT2, double nearT) {
#pragma clang fp contract(off)
    const double cx = v.x - h.X[0], cy = v.y - h.X[1], cz = v.z - h.X[2];
    // active lanes: not below the depth where they accepted a cluster
    const unsigned long long mAct = __ballot(h.depth <= stop);
    int next;
    if (h.skip == node + 1) {  // a bucket
        if (COUNT) {
            cnt += h.n;
        } else {
            bool near = false;
            if (nearT2 > 0.0) {
                // distance to the bounding sphere below nearT
                const double rn = h.r + nearT;
                near = (__ballot(cx * cx + cy * cy + cz * cz < rn * rn) & mAct) != 0;
            }
            pw.put(nnodes + h.l, h.n, mAct, near, node);
        }
        next = h.skip;
    } else {
        const double d2 = cx * cx + cy * cy + cz * cz;
        // The decisions are kept as lane masks (compare results) so that
        // no lane predicate goes through a vector register.  dOvrR =
        // sqrt(d2) orad < 2 is certain without the square root when
        // d2 orad^2 < 4 (1 - 2.5e-12): the roundings move either side by a
        // few 1e-16.
        const unsigned long long mSure = __ballot((d2 * h.r) * h.r < 3.99999999999);
        const unsigned long long mNeed = mAct & ~mSure;
        unsigned long long mAcc = 0;
        if (mNeed) {
            // every lane evaluates the test, the lanes of mNeed use it
            // (d2 >= 3.99 rad^2 for them, and the tree build keeps every
            // internal radius above 1e-100)
            const double dOvrR = sqrt_normal(d2) * h.r;
            const double err = get_error_cells(nodes + node, dOvrR, cells);
            mAcc = mNeed & __ballot(dOvrR >= 2.0) & __ballot(err < tol);
        }
        const unsigned long long mDesc = mAct & ~mAcc;
        // descending lanes stop nowhere, accepting lanes stop here
        stop = lane_in(mDesc) ? INT32_MAX : stop;
        stop = lane_in(mAcc) ? h.depth : stop;
        if (mAcc) {
            if (COUNT) ++cnt;
            else pw.put(node, 1, mAcc, false, node);
        }
        next = mDesc ? node + 1 : h.skip;
    }
    return uniform(next);
}

// The walk of evalNode (kd_tree.F90:619-662) for 64 volume nodes per wave:
// records every point of every bucket the wave evaluates (evalNodeExact)
// and every cluster some lane accepts (evalNodeApprox, :692-713), with the
// lane masks, in depth-first order.  The decisions use only init-time data
// (centroids, radii, error tables, denominator0) and errTol.  COUNT: terms
// per wave only.  Mixed precision (nearT2 > 0) keeps some terms in FP64:
// the points of a bucket when an active lane lies within sqrt(nearT2) of
// the bucket's bounding sphere, and every term of a wave with a node more
// than sqrt(farT2) outside the root's bounding sphere (the whole surface).
// Far from the body the terms' displacements at the node, K (v - x) + d,
// grow with the distance while their weighted mean stays small, so FP32
// weights and sums would lose against the cancellation; such waves
// evaluate few terms.
template <bool COUNT, int WBUF>
__global__ void __launch_bounds__(256) k_plan_walk(int nvol, int blk0, int nbc, const double4 *__restrict__ vol,
                                                   int nnodes, const WalkHdr *__restrict__ whdr,
                                                   const NodeRec *__restrict__ nodes, double errTol, double nearT2,
                                                   double farT2, int64_t *__restrict__ waveCount,
                                                   PlanEntry *__restrict__ plan,
                                                   const int64_t *__restrict__ waveOff,
                                                   int64_t *__restrict__ waveSplit,
                                                   unsigned int *__restrict__ mismatch, Mirror M) {
#pragma clang fp contract(off)
    __shared__ SegCell cells[ERR_CELLS];
    __shared__ PlanEntry wbuf[COUNT ? 1 : 4][COUNT ? 1 : WBUF], wbuf64[COUNT ? 1 : 4][COUNT ? 1 : WBUF];
    const int lb = COUNT ? (int)blockIdx.x : xcd_block(nbc);
    if (lb >= nbc) return;  // block-uniform
    const int blk = blk0 + lb;
    if (threadIdx.x < ERR_CELLS) cells[threadIdx.x] = seg_cell(threadIdx.x, c_rseg, c_rstar[NERR - 1]);
    __syncthreads();
    const double nearT = sqrt(nearT2);
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    const double tol = errTol * v.w;  // errTol * denominator0
    int stop = valid ? INT32_MAX : -1;  // depth at which this lane stopped descending
    const int wave = uniform(blk * 4 + (threadIdx.x >> 6));
    // the node's coordinates are in before the loop, so that the loop need
    // not wait for the plan's stores in flight to learn it
    wait_vmem();
    PlanWriter<WBUF> pw{plan, COUNT ? 0 : waveOff[wave], COUNT ? 0 : waveOff[wave + 1], 0, 0, mismatch,
                  wbuf[COUNT ? 0 : threadIdx.x >> 6], wbuf64[COUNT ? 0 : threadIdx.x >> 6]};
    int64_t cnt = 0;
    // Before each visit the walk touches a word of both possible successors'
    // headers (node + 1 when a lane descends, skip otherwise), which brings
    // their cache lines into the scalar cache; after the visit it loads the
    // chosen one, a cache hit.  Loading both headers whole and selecting
    // between them took twelve scalar selects per visit on a
    // scalar-issue-bound loop.  Header nnodes is a zeroed pad: neither
    // successor needs a clamp.
    constexpr unsigned HB = sizeof(WalkHdr);
    int node = 0;
    WalkHdr h = load_whdr(whdr, 0);
    // mixed precision: a wave with a node more than sqrt(farT2) outside the
    // root's bounding sphere keeps every term in FP64
    bool farWave = false;
    if (!COUNT && farT2 > 0.0) {
        const double R = h.skip == 1 ? h.r : 1.0 / h.r;  // the root's radius (bucket: rad, else 1 / orad)
        const double cx = v.x - h.X[0], cy = v.y - h.X[1], cz = v.z - h.X[2], t = R + sqrt(farT2);
        farWave = __ballot(valid && cx * cx + cy * cy + cz * cz > t * t) != 0ull;
    }
    unsigned touched = 0;
    for (;;) {
        const unsigned *pa =
            reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(whdr) + (unsigned)(node + 1) * HB);
        const unsigned *pb = reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(whdr) + (unsigned)h.skip * HB);
        const unsigned ta = pa[0], tb = pb[0];
        const int next = walk_visit<COUNT, WBUF>(h, node, v, tol, stop, pw, cnt, nodes, nnodes, cells, nearT2, nearT);
        if (next >= nnodes) break;
        touched += ta ^ tb;
        h = load_whdr(whdr, (unsigned)next * HB);
        node = next;
    }
    // The touches have no other use.  This adds zero to an output under a
    // condition the compiler cannot rule out, so they are not dropped.
    if (touched == 0x9e3779b9u && (threadIdx.x & 63) == 0) {
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(waveCount + wave), 0ull);
        else atomicAdd(mismatch, 0u);
    }
    if (COUNT) {
        if ((threadIdx.x & 63) == 0) waveCount[wave] = cnt;
    } else {
        pw.finish(waveSplit, wave, farWave);
    }
}

// Mixed precision keeps FP64 for buckets that an active lane approaches
// within sqrt(nearT2).  This walk flags those lanes (the wall layer and
// anything closer than nearT) so that they can be given waves of their
// own: key = near bit above the node's 63-bit Hilbert key.
// Bits 0..20 of x spread to every third bit.
__device__ __forceinline__ uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Index of the cell (x, y, z) (21 bits each) along a 3-D Hilbert curve, as a
// 63-bit key: Skilling's transform of the axes to the curve's transposed
// index ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), then
// its bits interleaved.  Consecutive keys are neighbouring cells, where a
// Morton order jumps between distant ones at its octant boundaries, so the
// 64 nodes of a wave stay closer together.
__device__ __forceinline__ uint64_t hilbert_key(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t X[3] = {x, y, z};
    constexpr uint32_t M = 1u << 20;
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;  // invert the low bits
            } else {        // exchange the low bits of X[0] and X[i]
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];  // Gray encode
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    return (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
}

__global__ void __launch_bounds__(256) k_near_keys(int nvol, const double4 *__restrict__ vol, int nnodes,
                                                   const NodeRec *__restrict__ nodes, double errTol, double nearT2,
                                                   double3 lo, double3 scale, uint64_t *__restrict__ keys,
                                                   int32_t *__restrict__ ids) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < nvol;
    const double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    const double tol = errTol * v.w;
    const double nearT = sqrt(nearT2);
    int stop = valid ? INT32_MAX : -1;
    bool near = false;
    int node = 0;
    NodeHdr h = load_hdr(nodes, 0);
    for (;;) {
        const NodeHdr hA = load_hdr(nodes, min(node + 1, nnodes - 1));
        const NodeHdr hB = load_hdr(nodes, min(h.skip, nnodes - 1));
        const bool active = h.depth <= stop;
        const double cx = v.x - h.X[0], cy = v.y - h.X[1], cz = v.z - h.X[2];
        int next;
        if (h.leaf) {
            const double rn = h.rad + nearT;
            near = near || (active && cx * cx + cy * cy + cz * cz < rn * rn);
            next = h.skip;
        } else {
            bool descend = false;
            if (active) {
                const double d2 = cx * cx + cy * cy + cz * cz;
                if ((d2 * h.orad) * h.orad < 3.99999999999) {
                    descend = true;
                } else {
                    const double dOvrR = sqrt(d2) * h.orad;
                    if (dOvrR < 2.0) descend = true;
                    else if (!(get_error(nodes + node, dOvrR, c_rseg) < tol)) descend = true;
                }
                stop = descend ? INT32_MAX : h.depth;
            }
            next = __any(descend) ? node + 1 : h.skip;
        }
        next = uniform(next);
        if (next >= nnodes) break;
        h = (next == node + 1) ? hA : hB;
        node = next;
    }
    if (!valid) return;
    auto q = [](double t) { return (uint32_t)fmin(fmax(t, 0.0), 2097151.0); };
    keys[k] = ((uint64_t)near << 63) |
              hilbert_key(q((v.x - lo.x) * scale.x), q((v.y - lo.y) * scale.y), q((v.z - lo.z) * scale.z));
    ids[k] = k;
}

__global__ void k_permute_vol(int64_t nvol, const int32_t *__restrict__ order, const double4 *__restrict__ vin,
                              const int32_t *__restrict__ pin, double4 *__restrict__ vout, int32_t *__restrict__ pout) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvol) return;
    const int32_t o = order[k];
    vout[k] = vin[o];
    pout[k] = pin[o];
}

typedef double dvec4 __attribute__((ext_vector_type(4)));

// wave-uniform copy of lane 0's value
__device__ __forceinline__ double first_lane(double x) {
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                            __builtin_amdgcn_readfirstlane(__double2loint(x)));
}

// ---------------------------------------------------------------------------
// Fast-mode sums over the plan on the matrix cores
// ---------------------------------------------------------------------------
// The wave's volume nodes are taken relative to the wave origin o (lane 0's
// node, g = v - o).  A term at x contributes w (K (v - x) + d) = w (K g + c)
// with c = d - K (x - o), so the wave's numerators and denominators are
//   S(node, :) = sum over terms of w(node, term) [K (9), c (3), 1],
// a (64 nodes x terms) by (terms x 16) product evaluated with
// v_mfma_f64_16x16x4f64.  The weights are computed directly in the
// A-operand layout: for node block t and term group q, lane l computes the
// weight of node 16t + l%16 with term 4q + l/16; the 16-wide B rows of the
// terms are staged in LDS.  The reverse reduces over the wave's nodes the
// same way (terms x nodes by nodes x 16).
//
// MFMA lane layout (16x16x4 f64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; register j of D holds D[4j + l/16][l%16].

// Terms staged per chunk in the forward (FCH: 16, 32 or 64, a tuning
// parameter; 64 / FCH lanes stage each term).  FP32 terms (mixed precision)
// and FP64 terms have chunk layouts of their own, so that the FP32 one, which
// carries the bulk, holds no FP64 B rows: at 64 terms a wave's chunk takes
// 7.4 KB and four workgroups of four waves fit a CU's LDS alongside.
constexpr int BROW = 17;  // padded LDS row (doubles) of a staged B row

constexpr int BROWF = 17;  // padded LDS row (floats) of a staged FP32 B row

// Per-term weight factors.  The default exponents (3, 5) give
// w = A L^3 p^3 + A L^3 alpha^5 L^2 p^5 with p = 1/|r|, evaluated as
// p^3 (f0 + f1 p^2) from f0 = A L^3 and f1 = A L^3 alpha^5 L^2; other
// exponents use f0 = A and the general form A wfun(L p).
template <int EXPM>
__device__ __forceinline__ double2 weight_factors(double A, const WParams &P) {
    if (EXPM == EXP_35) {
        const double L3 = P.Ldef * P.Ldef * P.Ldef;
        return make_double2(A * L3, A * L3 * (P.a5 * P.Ldef * P.Ldef));
    }
    return make_double2(A, 0.0);
}

// FP32 weight of one pair from |r|^2 (mixed precision) and the term's factors.
template <int EXPM>
__device__ __forceinline__ float weight_f32(float d2, float2 f, float Lf, const WParams &P) {
    if (EXPM == EXP_35) {
        const float p = __frsqrt_rn(d2), p2 = p * p;
        return (p * p2) * fmaf(f.y, p2, f.x);
    }
    return f.x * wfun<EXPM>(Lf * __frsqrt_rn(d2), P);
}

// FP64 weight of one pair from |r|^2 and the term's factors.
template <int EXPM>
__device__ __forceinline__ double weight_f64(double d2, double2 f, const WParams &P) {
    if (EXPM == EXP_35) {
        const double p = rsqrt_nr(d2), p2 = p * p;
        return (p * p2) * fma(f.y, p2, f.x);
    }
    return f.x * wfun<EXPM>(P.Ldef * rsqrt_nr(d2), P);
}

// A chunk of staged terms: B rows [K (9), c (3), 1, 0, 0, 0], positions x
// with the eps added to |r|^2 (EPS_W for points, evalNodeExact; 0 for
// clusters, evalNodeApprox), the lane masks and the weight factors.
template <int FCH>
struct FwdChunk32 {
    float B[FCH][BROWF];
    double X[FCH][4];  // x (3), eps
    unsigned long long mask[FCH];
    float2 F[FCH];
};
template <int FCH>
struct FwdChunk64 {
    double B[FCH][BROW];
    double X[FCH][4];
    unsigned long long mask[FCH];
    double2 F[FCH];
};
// FP64 terms per chunk: FCH in FP64 precision; in mixed precision (a short
// tail of FP64 terms per near-wall wave) at most 32, within the FP32 chunk's
// footprint.
constexpr int fch64(int prec, int fch) { return prec == IDW_PREC_MIXED && fch > 32 ? 32 : fch; }
template <int PREC, int FCH>
union FwdLds {
    FwdChunk32<FCH> c32;
    FwdChunk64<fch64(PREC, FCH)> c64;
    double acc[32][BROW];  // half of the wave's 64 x 16 sums
};

// Entry of plan slot e for this lane, or an empty entry (no lanes, record
// -1) beyond the end of the range.
// The load is unconditional (a slot inside the range is read), so the
// loops that prefetch entries carry no branches around their loads.
// Requires end > 0.
__device__ __forceinline__ PlanEntry plan_entry(const PlanEntry *__restrict__ plan, int64_t e, int64_t end) {
    PlanEntry pe = plan[min(e, end - 1)];
    if (e >= end) {
        pe.mask = 0ull;
        pe.rec = -1;
        pe.node = -1;
    }
    return pe;
}

// Row of this lane's term in a chunk of 64 (one term per lane) ordered by the
// 16-node blocks its mask touches: terms that touch the same blocks sit side
// by side, so the sums over four rows at a time find more blocks that none of
// the four touches and skip them; empty entries (mask 0) go last.  A stable
// radix sort of the 4-bit block signature, one bit per pass, each pass a
// split of the 64 lanes moved with ds_permute; then each sorted position is
// sent back to the term's lane.  All 64 lanes must be active.
__device__ __forceinline__ int block_order_row(unsigned long long mask) {
    const int lane = threadIdx.x & 63;
    const unsigned lo = (unsigned)mask, hi = (unsigned)(mask >> 32);
    const unsigned sig = ((lo & 0xffffu) ? 1u : 0u) | ((lo >> 16) ? 2u : 0u) | ((hi & 0xffffu) ? 4u : 0u) |
                         ((hi >> 16) ? 8u : 0u);
    unsigned v = ((15u - sig) << 8) | (unsigned)lane;  // key above the term's lane
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const bool one = (v >> (8 + b)) & 1u;
        const unsigned long long ones = __ballot(one);
        const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(ones >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)ones, 0u));
        const int dst = one ? 64 - __popcll(ones) + below : lane - below;  // a permutation of 0..63
        v = (unsigned)__builtin_amdgcn_ds_permute(dst << 2, (int)v);
    }
    return __builtin_amdgcn_ds_permute((int)((v & 63u) << 2), lane);
}

// Stage FCH terms of the plan in LDS, 64 / FCH lanes per term: term r is
// staged by lanes r + FCH p.  Part p = 0 writes row 0 of K and c (all rows
// when one lane stages a term), x, eps, the mask and the weight factors; the
// other parts the remaining rows; the last part the constant column.  pe:
// this lane's entry (record -1: padding row).  W32 reads the FP32 records
// (mixed precision) into an FP32 chunk; with one lane per term its rows are
// ordered by block_order_row (the padding rows last) unless ORDER is off.
// FP64 chunks keep plan order, so that FP64 sums do not depend on the chunk
// size.
template <bool W32, int EXPM, int FCH, typename Chunk, bool ORDER = true>
__device__ __forceinline__ void stage_terms(Chunk &C, PlanEntry pe, int nnodes, int nrec,
                                            const Rec64 *__restrict__ rec, const Rec32 *__restrict__ rec32,
                                            const WParams &P, double ox, double oy, double oz) {
    static_assert(FCH == 16 || FCH == 32 || FCH == 64, "16, 32 or 64 terms per chunk");
    constexpr int LPT = 64 / FCH;  // lanes per term
    const int lane = threadIdx.x & 63, part = lane / FCH;
    const int r = W32 && ORDER && LPT == 1 ? block_order_row(pe.mask) : lane & (FCH - 1);
    // K rows [a0, a1) of this lane: 1 lane {0,1,2}; 2 lanes {0}, {1,2}; 4 lanes {0}, {1}, {2}, {}
    const int a0 = LPT == 1 ? 0 : (LPT == 2 ? (part == 0 ? 0 : 1) : min(part, 3));
    const int a1 = LPT == 1 ? 3 : (LPT == 2 ? (part == 0 ? 1 : 3) : min(part + 1, 3));
    const bool lo = part == 0, hi = part == LPT - 1;
    if (pe.rec >= 0) {
        pe.rec = min(pe.rec, nrec - 1);  // a plan that disagreed with its count stays in bounds
        double x[3];
        if (W32) for (int i = 0; i < 3; ++i) x[i] = rec32[pe.rec].x[i];
        else for (int i = 0; i < 3; ++i) x[i] = rec[pe.rec].x[i];
        const double y0 = x[0] - ox, y1 = x[1] - oy, y2 = x[2] - oz;
        for (int a = a0; a < a1; ++a) {
            double K0, K1, K2, d;
            if (W32) {
                const Rec32 &q = rec32[pe.rec];
                K0 = q.K[3 * a]; K1 = q.K[3 * a + 1]; K2 = q.K[3 * a + 2]; d = q.d[a];
            } else {
                const Rec64 &q = rec[pe.rec];
                K0 = q.K[3 * a]; K1 = q.K[3 * a + 1]; K2 = q.K[3 * a + 2]; d = q.d[a];
            }
            const double c = d - (K0 * y0 + K1 * y1 + K2 * y2);
            C.B[r][3 * a] = K0;
            C.B[r][3 * a + 1] = K1;
            C.B[r][3 * a + 2] = K2;
            C.B[r][9 + a] = c;
        }
        if (hi) {
            C.B[r][12] = 1.0f;
            C.B[r][13] = C.B[r][14] = C.B[r][15] = 0.0f;
        }
        if (lo) {
            const double eps = pe.rec < nnodes ? 0.0 : EPS_W;
            if constexpr (W32) {  // |v - x|^2 = |g|^2 + (|y|^2 + eps) - 2 g.y with g = v - o, y = x - o
                C.X[r][0] = -2.0 * y0;
                C.X[r][1] = -2.0 * y1;
                C.X[r][2] = -2.0 * y2;
                C.X[r][3] = fma(y0, y0, fma(y1, y1, fma(y2, y2, eps)));
            } else {
                for (int i = 0; i < 3; ++i) C.X[r][i] = x[i];
                C.X[r][3] = eps;
            }
            C.mask[r] = pe.mask;
            const double2 f = weight_factors<EXPM>(W32 ? (double)rec32[pe.rec].A : rec[pe.rec].A, P);
            if constexpr (W32) C.F[r] = make_float2((float)f.x, (float)f.y);
            else C.F[r] = f;
        }
    } else {
        constexpr int W = 16 / LPT;  // B columns each part clears
        for (int j = part * W; j < (part + 1) * W; ++j) C.B[r][j] = 0.0f;
        if (lo) {
            for (int i = 0; i < 3; ++i) C.X[r][i] = 0.0;
            C.X[r][3] = 1.0;
            C.mask[r] = 0ull;
            if constexpr (W32) C.F[r] = make_float2(0.0f, 0.0f);
            else C.F[r] = make_double2(0.0, 0.0);
        }
    }
}

typedef float fvec4 __attribute__((ext_vector_type(4)));

// Swap helpers over the lane rows (16 lanes each) of two registers:
// rows 2-3 of x with rows 0-1 of y, and odd rows of x with even rows of y.
__device__ __forceinline__ void swap_rows32(float &x, float &y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_rows16(float &x, float &y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}

// FP32 sums come in the f32 16x16x4 layout (register j of lane l: row
// 4(l/16) + j, column l%16); the FP64 sums they are added to are kept in the
// same layout while only FP32 chunks add to them, and moved to the f64
// layout (register j: row 4j + l/16) once, by transposing the 4 x 4 blocks
// over (lane row, register) with two rounds of row swaps.  Each sum then
// takes the same additions in the same order as a per-chunk transpose of
// the FP32 sums would give it.
__device__ __forceinline__ void add_f32_sums(dvec4 &D, fvec4 F) {
    D[0] += (double)F[0];
    D[1] += (double)F[1];
    D[2] += (double)F[2];
    D[3] += (double)F[3];
}
__device__ __forceinline__ void swap_rows32(double &x, double &y) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                     false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                     false);
    x = __hiloint2double((int)hi[0], (int)lo[0]);
    y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap_rows16(double &x, double &y) {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                     false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                     false);
    x = __hiloint2double((int)hi[0], (int)lo[0]);
    y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void f32_to_f64_layout(dvec4 &D) {
    double d0 = D[0], d1 = D[1], d2 = D[2], d3 = D[3];
    swap_rows32(d0, d2);
    swap_rows32(d1, d3);
    swap_rows16(d0, d1);
    swap_rows16(d2, d3);
    D = dvec4{d0, d1, d2, d3};
}

// Forward sums of the terms in plan entries [a, b) into D (node block t in
// D[t]); vx/vy/vz: positions of the lane's rows (nodes 16t + lane%16).
// The entries of the next chunk are requested while this one is summed.
// W32 (mixed precision, FP32 terms): FP32 weights and B rows on
// v_mfma_f32_16x16x4f32, each chunk's FP32 sums added to the FP64 sums, which
// W32 keeps in the f32 layout (the caller moves them to the f64 layout);
// vx/vy/vz are then the rows' positions g relative to the wave origin and
// g2 = |g|^2, and |r|^2 is formed in FP64 as g2 + (|y|^2 + eps) - 2 g.y
// (y = x - o staged with the term).  Its cancellation, about
// 1e-16 ((|g| + |y|) / |r|)^2 relative to |r|^2, stays below FP32's rounding
// while |r| > (|g| + |y|) / 4000.  The terms FP32 takes are accepted
// clusters (at least two radii away) and the points of buckets whose
// bounding spheres no active lane comes within the near-wall distance of, so
// a node at the wall keeps its close points in FP64; a point just beyond
// that distance from a node of a wide wave would lose digits here.
template <bool W32, int EXPM, int FCH, typename Chunk>
__device__ __forceinline__ void fwd_terms(int64_t a, int64_t b, const PlanEntry *__restrict__ plan, int nnodes,
                                          int nrec, const Rec64 *__restrict__ rec, const Rec32 *__restrict__ rec32,
                                          const WParams &P, Chunk &C, double ox, double oy, double oz,
                                          const double (&vx)[4], const double (&vy)[4], const double (&vz)[4],
                                          const double (&g2)[4], dvec4 (&D)[4]) {
    const int lane = threadIdx.x & 63, n = lane & 15, R = lane >> 4, r = lane & (FCH - 1);
    const float Lf = (float)P.Ldef;
    if (a >= b) return;
    PlanEntry pe = plan_entry(plan, a + r, b);
    for (int64_t c0 = a; c0 < b; c0 += FCH) {
        const int cnt = (int)min((int64_t)FCH, b - c0);
        wave_sync();  // the previous chunk is consumed
        stage_terms<W32, EXPM, FCH>(C, pe, nnodes, nrec, rec, rec32, P, ox, oy, oz);
        pe = plan_entry(plan, c0 + FCH + r, b);  // next chunk
        wave_sync();
        const int ng = (cnt + 3) >> 2;
        if constexpr (W32) {
            fvec4 F[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) F[t] = fvec4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int q = 0; q < ng; ++q) {
                const int tr = 4 * q + R;
                const float Bv = C.B[tr][n];
                const unsigned long long m = C.mask[tr] >> n;
                const double x0 = C.X[tr][0], x1 = C.X[tr][1], x2 = C.X[tr][2], eps = C.X[tr][3];
                const float2 f = C.F[tr];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const unsigned long long on = __ballot((m >> (16 * t)) & 1ull);
                    // node block t idle for all four terms: nothing to add
                    // (about one block in seven)
                    if (!on) continue;
                    // |r|^2 + eps in FP64 (x0..x2 = -2 y, eps = |y|^2 + eps), rounded once
                    const double d2 = fma(vx[t], x0, fma(vy[t], x1, fma(vz[t], x2, g2[t] + eps)));
                    const float w = weight_f32<EXPM>((float)d2, f, Lf, P);
                    F[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(lane_in(on) ? w : 0.0f, Bv, F[t], 0, 0, 0);
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) add_f32_sums(D[t], F[t]);
        } else {
#pragma unroll 1
            for (int q = 0; q < ng; ++q) {
                const int tr = 4 * q + R;
                const double Bv = C.B[tr][n];
                const unsigned long long m = C.mask[tr] >> n;
                const double x0 = C.X[tr][0], x1 = C.X[tr][1], x2 = C.X[tr][2], eps = C.X[tr][3];
                const double2 f = C.F[tr];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const unsigned long long on = __ballot((m >> (16 * t)) & 1ull);
                    if (!on) continue;  // node block t idle for all four terms
                    const double rx = vx[t] - x0, ry = vy[t] - x1, rz = vz[t] - x2;
                    const double w = weight_f64<EXPM>(rx * rx + ry * ry + rz * rz + eps, f, P);
                    D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(lane_in(on) ? w : 0.0, Bv, D[t], 0, 0, 0);
                }
            }
        }
    }
}

// Forward of the fast mode replaying the plan: accepted clusters add
// Ai w(dist) (Kbar (x - X) + e) (evalNodeApprox), bucket points their exact
// terms (evalNodeExact); then Xv = Xv0 + num / den (warpMesh.F90:65-70).
// Mixed precision: FP32 weights for the wave's FP32 terms, FP64 for the
// rest; the products are summed in FP64 either way.
template <int PREC, int EXPM, int FCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_fast_fwd(int nvol, int blk0, int nbc,
                                                  const double4 *__restrict__ vol,
                                                  const int32_t *__restrict__ perm, const PlanEntry *__restrict__ plan,
                                                  const int64_t *__restrict__ waveOff,
                                                  const int64_t *__restrict__ waveSplit, int nnodes, int nrec,
                                                  const Rec64 *__restrict__ rec, const Rec32 *__restrict__ rec32,
                                                  WParams P, Mirror M, double *__restrict__ xv,
                                                  double *__restrict__ den, double4 *__restrict__ acc) {
    __shared__ FwdLds<PREC, FCH> lds[4];
    const int lb = xcd_block(nbc);
    if (lb >= nbc) return;
    const int blk = blk0 + lb;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, n = lane & 15, R = lane >> 4;
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    const int wave = uniform(blk * 4 + wv);
    // the lane's (mirrored) node, loaded again where needed rather than kept live
    auto node_pos = [&]() {
        double4 p = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
        if (M.mask) mirror_pt(M, p.x, p.y, p.z);
        return p;
    };
    // positions of the lane's rows: nodes 16t + lane%16
    auto rows_of = [&](const double4 &p, double (&px)[4], double (&py)[4], double (&pz)[4]) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int src = 16 * t + n;
            px[t] = __shfl(p.x, src, 64);
            py[t] = __shfl(p.y, src, 64);
            pz[t] = __shfl(p.z, src, 64);
        }
    };
    double4 v = node_pos();
    // wave origin: lane 0 is valid whenever any lane of the wave is
    const double ox = first_lane(v.x), oy = first_lane(v.y), oz = first_lane(v.z);
    dvec4 D[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) D[t] = dvec4{0.0, 0.0, 0.0, 0.0};
    FwdLds<PREC, FCH> &L = lds[wv];
    constexpr int F64 = fch64(PREC, FCH);
    const int64_t e0 = waveOff[wave], e1 = waveOff[wave + 1];
    double vx[4], vy[4], vz[4], g2[4];
    int64_t es = e0;
    if (PREC == IDW_PREC_MIXED) {
        es = min(max(waveSplit[wave], e0), e1);
        if (es > e0) {  // the FP32 terms, with the rows relative to the origin
            rows_of(v, vx, vy, vz);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                vx[t] -= ox;
                vy[t] -= oy;
                vz[t] -= oz;
                g2[t] = fma(vx[t], vx[t], fma(vy[t], vy[t], vz[t] * vz[t]));
            }
            fwd_terms<true, EXPM, FCH>(e0, es, plan, nnodes, nrec, rec, rec32, P, L.c32, ox, oy, oz, vx, vy, vz, g2,
                                       D);
#pragma unroll
            for (int t = 0; t < 4; ++t) f32_to_f64_layout(D[t]);  // D was zero before the FP32 terms
            v = node_pos();
        }
    }
    rows_of(v, vx, vy, vz);
    fwd_terms<false, EXPM, F64>(es, e1, plan, nnodes, nrec, rec, rec32, P, L.c64, ox, oy, oz, vx, vy, vz, g2, D);
    // D[t] register j: node 16t + 4j + R, column n.  Through LDS, 32 nodes
    // at a time, each lane gathers its node's 13 sums.  The lane's own
    // (mirrored) position is row R of its four rows.
    const double gx = (R == 0 ? vx[0] : R == 1 ? vx[1] : R == 2 ? vx[2] : vx[3]) - ox;
    const double gy = (R == 0 ? vy[0] : R == 1 ? vy[1] : R == 2 ? vy[2] : vy[3]) - oy;
    const double gz = (R == 0 ? vz[0] : R == 1 ? vz[1] : R == 2 ? vz[2] : vz[3]) - oz;
    double n0 = 0.0, n1 = 0.0, n2 = 0.0, dd = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        wave_sync();
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j) L.acc[16 * tt + 4 * j + R][n] = D[2 * h + tt][j];
        wave_sync();
        if ((lane >> 5) == h) {
            const double *s = L.acc[lane & 31];
            n0 = s[0] * gx + s[1] * gy + s[2] * gz + s[9];
            n1 = s[3] * gx + s[4] * gy + s[5] * gz + s[10];
            n2 = s[6] * gx + s[7] * gy + s[8] * gz + s[11];
            dd = s[12];
        }
    }
    if (valid) {
        // reloaded rather than kept live through the sums
        const double4 *vk = vol + k;
        const double4 vb = make_double4(__builtin_nontemporal_load(&vk->x), __builtin_nontemporal_load(&vk->y),
                                        __builtin_nontemporal_load(&vk->z), 0.0);
        finish_fwd(P, M, k, perm[k], vb, n0, n1, n2, dd, xv, den, acc);
    }
}

// Exact-mode forward in mixed precision on the matrix cores
// (evalNodeExact, kd_tree.F90:664-690, for every surface point): each wave
// takes the surface points as terms of the fast forward's form (x = Xu0, A,
// K = M - I, d), FCH at a time, with FP32 weights in the A-operand layout and
// v_mfma_f32_16x16x4f32 sums added to FP64 sums per chunk.  r = g - y is
// formed in FP64 from the origin-relative positions (y = x - o recovered
// exactly from the staged -2 y): the near-wall rule here is per pair, so an
// FP32 pair can lie just beyond the threshold while the wave spans far more,
// where the fast forward's dot-product form of |r|^2 would cancel.  A (node, point) pair
// nearer than the near-wall distance is left out of the FP32 sums; a chunk
// with such a pair is scanned again after its sums, each lane testing its
// own node against the chunk's points with the same arithmetic (so the same
// decisions), and adds those pairs in FP64 (rsqrt_nr, FP64 displacement) as
// the VALU kernel evaluates them.  Rows past the last node are summed like
// the others and never written.
template <int EXPM, int FCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_exact_fwd_mc(
    int nvol, const double4 *__restrict__ vol, const int32_t *__restrict__ perm, int n, int nnodes, int nrec,
    const Rec64 *__restrict__ rec, const Rec32 *__restrict__ rec32, const double4 *__restrict__ ps, WParams P,
    Mirror M, double *__restrict__ xv, double *__restrict__ den, double4 *__restrict__ acc) {
    __shared__ FwdLds<IDW_PREC_MIXED, FCH> lds[4];
    const int nb = (nvol + 255) / 256;
    const int blk = xcd_block(nb);
    if (blk >= nb) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nn = lane & 15, R = lane >> 4;
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    // wave origin: lane 0 is valid whenever any lane of the wave is
    const double ox = first_lane(v.x), oy = first_lane(v.y), oz = first_lane(v.z);
    // the lane's rows: nodes 16t + lane%16, relative to the origin
    double gx[4], gy[4], gz[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int src = 16 * t + nn;
        gx[t] = __shfl(v.x, src, 64) - ox;
        gy[t] = __shfl(v.y, src, 64) - oy;
        gz[t] = __shfl(v.z, src, 64) - oz;
    }
    // the lane's own node: row R
    const double ogx = R == 0 ? gx[0] : R == 1 ? gx[1] : R == 2 ? gx[2] : gx[3];
    const double ogy = R == 0 ? gy[0] : R == 1 ? gy[1] : R == 2 ? gy[2] : gy[3];
    const double ogz = R == 0 ? gz[0] : R == 1 ? gz[1] : R == 2 ? gz[2] : gz[3];
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, ed = 0.0;  // FP64 sums of the node's near pairs
    dvec4 D[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) D[t] = dvec4{0.0, 0.0, 0.0, 0.0};
    FwdLds<IDW_PREC_MIXED, FCH> &L = lds[wv];
    FwdChunk32<FCH> &C = L.c32;
    const float Lf = (float)P.Ldef;
    const double T2 = P.nearT2;
    const int r = lane & (FCH - 1);
    const unsigned long long vmask = __ballot(valid);
    for (int c0 = 0; c0 < n; c0 += FCH) {
        const int cnt = min(FCH, n - c0);
        PlanEntry pe;
        pe.mask = r < cnt ? vmask : 0ull;
        pe.rec = r < cnt ? nnodes + c0 + r : -1;
        pe.node = 0;
        wave_sync();  // the previous chunk is consumed
        stage_terms<true, EXPM, FCH, FwdChunk32<FCH>, false>(C, pe, nnodes, nrec, rec, rec32, P, ox, oy, oz);
        wave_sync();
        const int ng = (cnt + 3) >> 2;
        fvec4 F[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) F[t] = fvec4{0.0f, 0.0f, 0.0f, 0.0f};
        unsigned long long nearBits = 0ull;  // lanes with a near pair, over the chunk
#pragma unroll 1
        for (int q = 0; q < ng; ++q) {
            const int tr = 4 * q + R;
            const bool live = tr < cnt;  // a padding row's weight must stay 0 (its r may vanish)
            const float Bv = C.B[tr][nn];
            const double x0 = C.X[tr][0], x1 = C.X[tr][1], x2 = C.X[tr][2];  // -2 y
            const float2 f = C.F[tr];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                // r = g - y (one rounding each), |r|^2 + eps for a point; padding rows are at the origin
                const double rx = fma(x0, 0.5, gx[t]), ry = fma(x1, 0.5, gy[t]), rz = fma(x2, 0.5, gz[t]);
                const double d2 = fma(rz, rz, fma(ry, ry, fma(rx, rx, EPS_W)));
                const bool nr = live && d2 < T2;
                nearBits |= __ballot(nr);
                const float w = weight_f32<EXPM>((float)d2, f, Lf, P);
                F[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(live && !nr ? w : 0.0f, Bv, F[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) add_f32_sums(D[t], F[t]);
        if (nearBits) {
            // the chunk's near pairs in FP64, each lane for its own node
            double4 p = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
            if (M.mask) mirror_pt(M, p.x, p.y, p.z);
            for (int j = 0; j < cnt; ++j) {
                const double rx = fma(C.X[j][0], 0.5, ogx), ry = fma(C.X[j][1], 0.5, ogy),
                             rz = fma(C.X[j][2], 0.5, ogz);
                if (!(fma(rz, rz, fma(ry, ry, fma(rx, rx, EPS_W))) < T2)) continue;
                const int i = c0 + j;
                const double4 sp = ps[i];
                const Rec64 &pr = rec[nnodes + i];
                const double ax = p.x - sp.x, ay = p.y - sp.y, az = p.z - sp.z;
                const double lo = P.Ldef * rsqrt_nr(ax * ax + ay * ay + az * az + EPS_W);
                const double w = sp.w * wfun<EXPM>(lo, P);
                e0 += w * (pr.K[0] * ax + pr.K[1] * ay + pr.K[2] * az + pr.d[0]);
                e1 += w * (pr.K[3] * ax + pr.K[4] * ay + pr.K[5] * az + pr.d[1]);
                e2 += w * (pr.K[6] * ax + pr.K[7] * ay + pr.K[8] * az + pr.d[2]);
                ed += w;
            }
        }
    }
    // D[t] register j (the f32 layout: only FP32 chunks added to it): node
    // 16t + 4R + j, column n.  Through LDS, 32 nodes at a time, each lane
    // gathers its node's 13 sums.
    double n0 = 0.0, n1 = 0.0, n2 = 0.0, dd = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        wave_sync();
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j) L.acc[16 * tt + 4 * R + j][nn] = D[2 * h + tt][j];
        wave_sync();
        if ((lane >> 5) == h) {
            const double *sm = L.acc[lane & 31];
            n0 = sm[0] * ogx + sm[1] * ogy + sm[2] * ogz + sm[9];
            n1 = sm[3] * ogx + sm[4] * ogy + sm[5] * ogz + sm[10];
            n2 = sm[6] * ogx + sm[7] * ogy + sm[8] * ogz + sm[11];
            dd = sm[12];
        }
    }
    if (valid) {
        const double4 vb = vol[k];
        finish_fwd(P, M, k, perm[k], vb, n0 + e0, n1 + e1, n2 + e2, dd + ed, xv, den, acc);
    }
}


constexpr int RT = 16;  // terms per reverse tile
constexpr int GROUP_NS = 4;  // walk waves per group of the sparse-seed reverse (k_fast_rev_group)

// A volume node whose seed s = dXv / den is zero adds nothing to the reverse
// sums.  compact_nodes puts the wave's nodes with a nonzero seed first, in
// lane order, with their lane index in the low word of w (the bit of the
// plan masks that serves them), and pads them to a multiple of four with
// rows that add exactly zero: seed 0 at position (1e30, 1e30, 1e30), whose
// weights are finite (FP64) or 0 (FP32).  Returns the number of node quads
// the reverse sums over (0 when every seed is zero).  cs: the compacted
// seeds.
__device__ __forceinline__ int compact_nodes(bool valid, const double4 &v, const double (&sd)[3], double4 *cnode,
                                             double (*cs)[3]) {
    const int lane = threadIdx.x & 63;
    const bool nz = valid && (sd[0] != 0.0 || sd[1] != 0.0 || sd[2] != 0.0);
    const unsigned long long mnz = __ballot(nz);
    const int cnt = __popcll(mnz), cnt4 = (cnt + 3) & ~3;
    if (nz) {
        const int at = __popcll(mnz & ((1ull << lane) - 1ull));
        cnode[at] = make_double4(v.x, v.y, v.z, __hiloint2double(0, lane));
        for (int c = 0; c < 3; ++c) cs[at][c] = sd[c];
    }
    if (lane >= cnt && lane < cnt4) {
        cnode[lane] = make_double4(1e30, 1e30, 1e30, __hiloint2double(0, 0));
        for (int c = 0; c < 3; ++c) cs[lane][c] = 0.0;
    }
    return uniform(cnt4 >> 2);
}

// This lane's bit of a plan mask for a compacted node row (its lane index in
// the low word of w).
__device__ __forceinline__ bool serves(unsigned long long mask, const double4 &c) {
    return (mask >> (__double2loint(c.w) & 63)) & 1ull;
}

struct RevLds {
    double4 node[64];    // mirrored positions of the wave's nodes with a nonzero seed (compact_nodes)
    double s[64][3];     // their seeds
    double Y[RT][4];     // x - o of the tile's terms
    int32_t idx[RT];     // their records (-1: padding)
};

// Reverse sums of the terms in plan entries [a, b): per tile of 16 terms,
// P(term, :) = sum over the wave's nodes of w(node, term) G(node, :) with
// G = [s (x) g (9), s (3)] (B in registers), then
// Mb'(a, c) += P(3a + c) - P(9 + a) (x - o)_c and Bb(a) += P(9 + a) into
// the term's adjoint record.  nodes: the nq quads of compacted node rows
// (compact_nodes); B[q] their G rows.  Lane l computes the weight of term
// l%16 with node row 4q + l/16, in FP64.  The next tile's record and the
// entries of the tile after it are requested before this tile's sums and are
// in registers before its atomics are issued.
// With part (deterministic reverse) the 12 sums of the term at plan slot e
// are written to part[12 e ..] instead of being added to its record.
template <int EXPM>
__device__ __forceinline__ void rev_terms(int64_t a, int64_t b, const PlanEntry *__restrict__ plan, int nnodes,
                                          int nrec, const Rec64 *__restrict__ rec, const WParams &P,
                                          const double4 *nodes, int nq, double (*Y)[4], int32_t *idxs,
                                          const double (&B)[16], double ox, double oy, double oz,
                                          double *__restrict__ adj, double *__restrict__ part = nullptr) {
    const int lane = threadIdx.x & 63, m = lane & 15, R = lane >> 4;
    const int src = (lane & ~15) | (9 + (m < 9 ? m / 3 : 0));
    if (a >= b) return;
    auto record = [&](const PlanEntry &pe) {  // x, A of an entry (zero for padding); unconditional load
        const double4 v = *reinterpret_cast<const double4 *>(rec + min(max(pe.rec, 0), nrec - 1));
        return pe.rec >= 0 ? v : make_double4(0.0, 0.0, 0.0, 0.0);
    };
    PlanEntry pe = plan_entry(plan, a + m, b);
    double4 xa = record(pe);
    PlanEntry pn = plan_entry(plan, a + RT + m, b);
    for (int64_t t0 = a; t0 < b; t0 += RT) {
        const double4 xn = record(pn);
        const PlanEntry pnn = plan_entry(plan, t0 + 2 * RT + m, b);
        const int idx = pe.rec >= 0 ? min(pe.rec, nrec - 1) : -1;
        const double eps = (idx >= 0 && idx < nnodes) ? 0.0 : EPS_W;
        wave_sync();  // the previous tile's reads of Y / idx are done
        if (R == 0) {
            Y[m][0] = xa.x - ox;
            Y[m][1] = xa.y - oy;
            Y[m][2] = xa.z - oz;
            idxs[m] = idx;
        }
        const double2 f = weight_factors<EXPM>(xa.w, P);
        dvec4 D0 = {0.0, 0.0, 0.0, 0.0}, D1 = D0;
        auto sums = [&](auto dense) {
            // dense: all 64 seeds nonzero, the rows are the wave's nodes in
            // lane order and row 4q + R is bit 4q + R of the masks
            const unsigned long long mR = pe.mask >> R;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (!dense && q >= nq) continue;  // rows past 4 nq have zero seeds
                const double4 c = nodes[4 * q + R];
                const bool bit = dense ? ((mR >> (4 * q)) & 1ull) != 0 : serves(pe.mask, c);
                const unsigned long long on = __ballot(bit);
                if (!on) continue;  // node rows 4q..4q+3 idle for all the tile's terms
                const double rx = c.x - xa.x, ry = c.y - xa.y, rz = c.z - xa.z;
                double w = weight_f64<EXPM>(rx * rx + ry * ry + rz * rz + eps, f, P);
                w = bit ? w : 0.0;
                if (q & 1) D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(w, B[q], D1, 0, 0, 0);
                else D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(w, B[q], D0, 0, 0, 0);
            }
        };
        if (nq == 16) sums(std::true_type{});
        else sums(std::false_type{});
        const dvec4 Ds = D0 + D1;
        wait_vmem();  // the prefetches are in before this tile's atomics go out
        wave_sync();  // Y / idx of this tile are written
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int tr = 4 * t + R;
            double val = Ds[t];
            const double pb = __shfl(val, src, 64);
            if (m < 9) val -= pb * Y[tr][m % 3];
            const int id = idxs[tr];
            if (part) {
                if (id >= 0 && m < 12) part[(t0 + tr) * 12 + m] = val;
            } else if (id >= 0 && m < 12 && val != 0.0) {
                atomicAdd(adj + (int64_t)id * ADJ_STRIDE + m, val);
            }
        }
        pe = pn;
        xa = xn;
        pn = pnn;
    }
}

// Reverse of the fast mode (evalNode_b / evalNodeApprox_b, kd_tree.F90:
// 942-1031) replaying the forward's plan: for accepted cluster c,
// C(a,b) += sum s_a (x_b - X_b), B(a) += sum s_a; for bucket point i,
// Mb'(a,b) += sum s_a (x_b - Xu0_b), Bb(a) += sum s_a, with
// s = w * dXv / den (warpDeriv.F90:76-78).  The adjoint records are indexed
// like the term records (clusters, then points).  FP64 weights throughout:
// mixed precision takes k_fast_rev_pair, whose FP64 terms come back here
// through rev_terms.
template <int PREC, int EXPM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_fast_rev(int nvol, const double4 *__restrict__ vol,
                                                  const int32_t *__restrict__ perm, const double *__restrict__ dxv,
                                                  const double *__restrict__ den, const PlanEntry *__restrict__ plan,
                                                  const int64_t *__restrict__ waveOff,
                                                  const int64_t *__restrict__ waveSplit, int nnodes, int nrec,
                                                  const Rec64 *__restrict__ rec, WParams P, Mirror M,
                                                  double *__restrict__ adj, double *__restrict__ part) {
    static_assert(PREC == IDW_PREC_FP64, "mixed precision reverses through k_fast_rev_pair (or here, deterministic)");
    __shared__ RevLds lds[4];
    const int nb = (nvol + 255) / 256;
    const int blk = xcd_block(nb);
    if (blk >= nb) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, n = lane & 15, R = lane >> 4;
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    const int wave = uniform(blk * 4 + wv);
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (valid) {
        const int64_t j = perm[k];
        const double oden = 1.0 / den[k];
        s0 = dxv[3 * j] * oden;
        s1 = dxv[3 * j + 1] * oden;
        s2 = dxv[3 * j + 2] * oden;
    }
    if (M.mask) {
        mirror_pt(M, v.x, v.y, v.z);
        mirror_vec_b(M, s0, s1, s2);
    }
    // wave origin: lane 0 is valid whenever any lane of the wave is
    const double ox = first_lane(v.x), oy = first_lane(v.y), oz = first_lane(v.z);
    RevLds &L = lds[wv];
    const double sd[3] = {s0, s1, s2};
    const int nq = compact_nodes(valid, v, sd, L.node, L.s);
    // no seed in the wave: nothing to add (the deterministic reverse still
    // writes its slots' zeros)
    if (nq == 0 && !part) return;
    wave_sync();
    // B operand of node row group q: G(row 4q + R, column n)
    const int sa = n < 9 ? n / 3 : (n < 12 ? n - 9 : 0), gb = n % 3;
    const double og = gb == 0 ? ox : (gb == 1 ? oy : oz);
    double B[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const double sv = L.s[4 * q + R][sa], gv = reinterpret_cast<const double *>(&L.node[4 * q + R])[gb] - og;
        B[q] = q < nq ? (n < 9 ? sv * gv : (n < 12 ? sv : 0.0)) : 0.0;
    }
    const int64_t e0 = waveOff[wave], e1 = waveOff[wave + 1];
    rev_terms<EXPM>(e0, e1, plan, nnodes, nrec, rec, P, L.node, nq, L.Y, L.idx, B, ox, oy, oz, adj, part);
}

// ---------------------------------------------------------------------------
// Deterministic fast reverse: per-slot sums reduced per record in plan order
// ---------------------------------------------------------------------------
// The record of every plan slot and the slot itself, the keys and values of
// the stable sort that lists each record's slots in increasing order.  A slot
// without a record (never written by the reverse) takes key nrec, past every
// record's range.
__global__ void k_plan_recs(int64_t total, const PlanEntry *__restrict__ plan, unsigned *__restrict__ keys,
                            int32_t *__restrict__ vals, int nrec) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int r = plan[e].rec;
    keys[e] = (unsigned)(r < 0 ? nrec : min(r, nrec - 1));
    vals[e] = (int32_t)e;
}

// recStart[r] = first position of record r in the sorted keys (r <= nrec).
__global__ void k_rec_bounds(int nrec, int64_t total, const unsigned *__restrict__ keys, int64_t *__restrict__ recStart) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrec) return;
    int64_t lo = 0, hi = total;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < (unsigned)r) lo = mid + 1;
        else hi = mid;
    }
    recStart[r] = lo;
}

// Chunks of the records' sorted slot ranges: record r has
// max(1, ceil(n_r / DET_CHUNK)) of them, so that a cluster accepted by most
// waves is summed by many waves instead of one.
__global__ void k_rec_chunks(int nrec, const int64_t *__restrict__ recStart, int64_t *__restrict__ count) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int64_t n = recStart[r + 1] - recStart[r];
    count[r] = n > DET_CHUNK ? (n + DET_CHUNK - 1) / DET_CHUNK : 1;
}

__global__ void k_chunk_recs(int nrec, const int64_t *__restrict__ chunkOff, int32_t *__restrict__ chunkRec) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    for (int64_t c = chunkOff[r]; c < chunkOff[r + 1]; ++c) chunkRec[c] = r;
}

// One wave per chunk: lane group g (12 lanes, 5 groups) reads slots
// g, g + 5, ... of the chunk, each slot's 12 sums as one 96-byte access, and
// lane m of a group adds sum m; the groups are then added in order 0..4.
__global__ void __launch_bounds__(256) k_det_chunks(int64_t nchunk, const int32_t *__restrict__ chunkRec,
                                                    const int64_t *__restrict__ chunkOff,
                                                    const int64_t *__restrict__ recStart,
                                                    const int32_t *__restrict__ pos, const double *__restrict__ part,
                                                    double *__restrict__ chunkPart) {
    const int lane = threadIdx.x & 63, g = lane / 12, m = lane - 12 * g;
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nchunk) return;  // wave-uniform
    const int r = chunkRec[c];
    const int64_t s = recStart[r] + (c - chunkOff[r]) * DET_CHUNK;
    const int64_t e = min(recStart[r + 1], s + DET_CHUNK);
    double acc = 0.0;
    if (g < 5) {
        int64_t i = s + g;
        for (; i + 15 < e; i += 20) {  // four slots of the group in flight
            const double a0 = part[(int64_t)pos[i] * 12 + m], a1 = part[(int64_t)pos[i + 5] * 12 + m];
            const double a2 = part[(int64_t)pos[i + 10] * 12 + m], a3 = part[(int64_t)pos[i + 15] * 12 + m];
            acc += a0;
            acc += a1;
            acc += a2;
            acc += a3;
        }
        for (; i < e; i += 5) acc += part[(int64_t)pos[i] * 12 + m];
    }
    double tot = acc;
#pragma unroll
    for (int k = 1; k < 5; ++k) tot += __shfl(acc, 12 * k + (lane < 12 ? lane : 0), 64);
    if (lane < 12) chunkPart[c * 12 + lane] = tot;
}

// Per record and sum: its chunks' partials added in chunk order, stored (the
// first mirror pass) or added (later passes) into the adjoint record.  Every
// step has a fixed order, so the result is the same on every run.
__global__ void k_det_finish(int nrec, const int64_t *__restrict__ chunkOff, const double *__restrict__ chunkPart,
                             double *__restrict__ adj, int add) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrec * 12) return;
    const int r = (int)(t / 12), m = (int)(t - 12 * (int64_t)r);
    double acc = 0.0;
    for (int64_t c = chunkOff[r]; c < chunkOff[r + 1]; ++c) acc += chunkPart[c * 12 + m];
    double &o = adj[(int64_t)r * ADJ_STRIDE + m];
    o = add ? o + acc : acc;
}

// ---------------------------------------------------------------------------
// Reverse of a pair of walk waves (mixed precision)
// ---------------------------------------------------------------------------
// Neighbouring waves share most of their terms (at C3 a record appears in
// 1.71 of two neighbouring waves' plans on average), and the reverse is held
// back by its atomics.  So one wave takes the 128 volume nodes of walk waves A
// = 2p and B = 2p + 1: their FP32 plan ranges, both in depth-first order, are
// merged on the fly, and a term in both plans is summed over all 128 nodes
// and added to its adjoint record once.  Terms of one plan only are summed
// over that wave's 64 nodes, so no node is weighted against a term its wave
// does not use.  FP32 weights and G on v_mfma_f32_16x16x4f32; the FP64 terms
// (near the wall) take the per-wave path with FP64 sums.
// MQ: plan entries of each wave in the merge buffers (8, 16 or 32, a tuning
// parameter); a queue holds < RT items before a merge round adds at most MQ
// to it, so QCAP = RT + MQ.
template <int MQ>
struct PairMergeLds {
    static constexpr int QCAP = RT + MQ;
    unsigned long long keyA[MQ], keyB[MQ], mskA[MQ], mskB[MQ];
    // union terms by kind: 0 in both plans, 1 in A's only, 2 in B's only
    int32_t qrec[3][QCAP];
    unsigned long long qmA[2][QCAP];  // A masks of kinds 0 and 1
    unsigned long long qmB[2][QCAP];  // B masks of kinds 0 and 2
};

template <int MQ>
struct PairLds {
    double4 node[128];  // mirrored positions of the nodes with a nonzero seed: wave A's at 0, wave B's at 64
    union {
        PairMergeLds<MQ> q;  // during the merge
        double s[128][3];    // seeds, while the B operands are built
    };
    double Y[RT][4];  // x - o of the tile's terms
    int32_t idx[RT];  // their records (-1: padding)
};

// Depth-first position of a plan entry: its tree node, then its record (a
// bucket's points in order).  Strictly increasing along a wave's FP32 range.
__device__ __forceinline__ unsigned long long plan_key(const PlanEntry &e) {
    return ((unsigned long long)(unsigned)e.node << 32) | (unsigned)e.rec;
}

// Number of the first n keys of a sorted LDS array below k.
__device__ __forceinline__ int count_below(const unsigned long long *a, int n, unsigned long long k) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// A tile of up to RT union terms of one kind (per lane m = lane % 16: term m).
struct PairTile {
    int kind, rec;
    unsigned long long mA, mB;
    double4 xa;  // x, A of the record (zero for padding)
};

template <int EXPM, int MQ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_fast_rev_pair(int nvol, int nwave, const double4 *__restrict__ vol, const int32_t *__restrict__ perm,
                const double *__restrict__ dxv, const double *__restrict__ den, const PlanEntry *__restrict__ plan,
                const int64_t *__restrict__ waveOff, const int64_t *__restrict__ waveSplit, int nnodes, int nrec,
                const Rec64 *__restrict__ rec, WParams P, Mirror M, double *__restrict__ adj,
                const uint32_t *__restrict__ groupSparse) {
    static_assert(MQ == 8 || MQ == 16 || MQ == 32, "both merge buffers are filled by one load per lane");
    __shared__ PairLds<MQ> lds[4];
    const int npair = (nwave + 1) >> 1;
    const int nblk = (npair + 3) >> 2;
    const int blk = xcd_block(nblk);
    if (blk >= nblk) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, R = lane >> 4;
    const int pair = uniform(blk * 4 + wv);
    if (pair >= npair) return;  // wave-uniform
    // a group of walk waves with sparse seeds is the group reverse's
    if (groupSparse && groupSparse[pair / (GROUP_NS / 2)]) return;
    PairLds<MQ> &L = lds[wv];
    // the pair's nodes: lane j holds node j of wave A and node j of wave B
    auto node_seed = [&](int k, double4 &v, double (&sd)[3]) {  // mirrored position and seed s = dXv / den
        const bool valid = k < nvol;
        v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
        sd[0] = sd[1] = sd[2] = 0.0;
        if (valid) {
            const int64_t j = perm[k];
            const double oden = 1.0 / den[k];
            sd[0] = dxv[3 * j] * oden;
            sd[1] = dxv[3 * j + 1] * oden;
            sd[2] = dxv[3 * j + 2] * oden;
        }
        if (M.mask) {
            mirror_pt(M, v.x, v.y, v.z);
            mirror_vec_b(M, sd[0], sd[1], sd[2]);
        }
        v.w = 0.0;
    };
    // each wave's nodes with a nonzero seed, compacted: rows 64 h + i of the
    // node array, nqA / nqB quads of rows (the rest add nothing)
    double ox, oy, oz;
    int nq[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double4 v;
        double sd[3];
        const int k = pair * 128 + 64 * h + lane;
        node_seed(k, v, sd);
        if (h == 0) {
            // pair origin: lane 0's node of wave A (valid whenever any node of the pair is)
            ox = first_lane(v.x);
            oy = first_lane(v.y);
            oz = first_lane(v.z);
        }
        nq[h] = compact_nodes(k < nvol, v, sd, L.node + 64 * h, L.s + 64 * h);
    }
    const int nqA = nq[0], nqB = nq[1];
    if (nqA == 0 && nqB == 0) return;  // no seed in the pair: nothing to add
    wave_sync();
    // G(row nd, column m) = s_a g_b (9), s_a (3), 0 with g = v - o, from LDS
    const int sa = m < 9 ? m / 3 : (m < 12 ? m - 9 : 0), gb = m % 3;
    const double og = gb == 0 ? ox : (gb == 1 ? oy : oz);
    auto g_entry = [&](int nd) {
        const double sv = L.s[nd][sa];
        const double gv = reinterpret_cast<const double *>(&L.node[nd])[gb] - og;
        return m < 9 ? sv * gv : (m < 12 ? sv : 0.0);
    };
    float GA[16], GB[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        GA[q] = q < nqA ? (float)g_entry(4 * q + R) : 0.0f;
        GB[q] = q < nqB ? (float)g_entry(64 + 4 * q + R) : 0.0f;
        __builtin_amdgcn_sched_barrier(0);
    }
    const float Lf = (float)P.Ldef;
    const int src9 = (lane & ~15) | (9 + (m < 9 ? m / 3 : 0));
    const int wA = 2 * pair, wB = min(2 * pair + 1, nwave);  // wave B == nwave: an empty range
    const int64_t a0 = waveOff[wA], a1 = waveOff[wA + 1];
    const int64_t b0 = wB < nwave ? waveOff[wB] : 0, b1 = wB < nwave ? waveOff[wB + 1] : 0;
    const int64_t asp = min(max(waveSplit[wA], a0), a1);
    const int64_t bsp = wB < nwave ? min(max(waveSplit[wB], b0), b1) : 0;
    // merge state (wave-uniform): next unread entry and end of A's and B's FP32
    // ranges, entries in the merge buffers, items in the three queues
    int64_t ia = a0, ib = b0;
    const int64_t ea = asp, eb = bsp;
    int na = 0, nb = 0, q0 = 0, q1 = 0, q2 = 0;
    // Lane i < MQ tops up A's buffer from entry ia + i, lane MQ + i B's from
    // ib + i.  Those entries are requested one merge round ahead (when ia and
    // ib last moved), so a round does not wait for its loads.
    const bool mA = lane < MQ;
    const int mi = mA ? lane : lane - MQ;
    auto next_entry = [&](int64_t ia_, int64_t ib_) {
        const int64_t idx = (mA ? ia_ : ib_) + mi;
        PlanEntry e{};
        if (idx < (mA ? ea : eb)) e = plan[idx];
        return e;
    };
    PlanEntry pre = next_entry(ia, ib);
    auto record = [&](int r) {  // x, A of a record (zero for padding); unconditional load
        const double4 v = *reinterpret_cast<const double4 *>(rec + min(max(r, 0), nrec - 1));
        return r >= 0 ? v : make_double4(0.0, 0.0, 0.0, 0.0);
    };
    // Each pass: find the next tile (a full queue, or at the end what is left,
    // merging as needed) and request its records; then add the previous
    // tile's sums to the adjoint records; then sum the new tile.
    PairTile T;
    T.kind = -1;
    fvec4 F0 = {0.0f, 0.0f, 0.0f, 0.0f}, F1 = F0;
    for (;;) {
        PairTile Tn;
        Tn.kind = -1;
        for (;;) {
            int kind = q0 >= RT ? 0 : (q1 >= RT ? 1 : (q2 >= RT ? 2 : -1));
            const bool done = ia >= ea && ib >= eb && na == 0 && nb == 0;
            if (kind < 0 && done) kind = q0 > 0 ? 0 : (q1 > 0 ? 1 : (q2 > 0 ? 2 : -1));
            if (kind >= 0) {
                const int qk = kind == 0 ? q0 : (kind == 1 ? q1 : q2);
                const int cnt = min(qk, RT);
                const int base = qk - cnt;
                const bool ok = m < cnt;
                Tn.kind = kind;
                Tn.rec = ok ? L.q.qrec[kind][base + m] : -1;
                Tn.mA = (ok && kind != 2) ? L.q.qmA[kind][base + m] : 0ull;
                Tn.mB = (ok && kind != 1) ? L.q.qmB[kind == 0 ? 0 : 1][base + m] : 0ull;
                if (kind == 0) q0 = uniform(base);
                else if (kind == 1) q1 = uniform(base);
                else q2 = uniform(base);
                Tn.xa = record(Tn.rec);
                break;
            }
            if (done) break;
            // One merge round: top up both buffers to MQ entries, then move every entry
            // whose key is at most K = min(last key loaded of each plan not yet read to its
            // end) into the queues.  No entry past K can precede or equal a key still
            // unread, and at least one buffer empties, so the merge ends.  Lanes 0..31
            // hold A's buffer, lanes 32..63 B's.
            const bool isA = lane < MQ;
            const int i = isA ? lane : lane - MQ;
            const int la = (int)min((int64_t)(MQ - na), ea - ia);
            const int lb = (int)min((int64_t)(MQ - nb), eb - ib);
            wave_sync();
            if (isA ? i < la : i < lb) {
                const PlanEntry e = pre;
                const int at = (isA ? na : nb) + i;
                if (isA) {
                    L.q.keyA[at] = plan_key(e);
                    L.q.mskA[at] = e.mask;
                } else {
                    L.q.keyB[at] = plan_key(e);
                    L.q.mskB[at] = e.mask;
                }
            }
            ia += la;
            na = uniform(na + la);
            ib += lb;
            nb = uniform(nb + lb);
            pre = next_entry(ia, ib);
            wave_sync();
            const unsigned long long KA = ia < ea ? L.q.keyA[na - 1] : ~0ull;
            const unsigned long long KB = ib < eb ? L.q.keyB[nb - 1] : ~0ull;
            const unsigned long long K = KA < KB ? KA : KB;
            const int nown = isA ? na : nb, noth = isA ? nb : na;
            const unsigned long long *other = isA ? L.q.keyB : L.q.keyA;
            const bool have = i < nown;
            const unsigned long long k = have ? (isA ? L.q.keyA[i] : L.q.keyB[i]) : 0ull;
            const unsigned long long msk = have ? (isA ? L.q.mskA[i] : L.q.mskB[i]) : 0ull;
            const int c = have ? count_below(other, noth, k) : 0;
            const bool twin = have && c < noth && other[c] == k;
            const bool emit = have && k <= K;
            const unsigned long long m0 = __ballot(emit && isA && twin);
            const unsigned long long m1 = __ballot(emit && isA && !twin);
            const unsigned long long m2 = __ballot(emit && !isA && !twin);
            const unsigned long long below = (1ull << lane) - 1ull;
            if (emit && (isA || !twin)) {
                const int kind = isA ? (twin ? 0 : 1) : 2;
                const unsigned long long mk = kind == 0 ? m0 : (kind == 1 ? m1 : m2);
                const int slot = (kind == 0 ? q0 : (kind == 1 ? q1 : q2)) + __popcll(mk & below);
                L.q.qrec[kind][slot] = (int32_t)(unsigned)k;
                if (kind == 0) {
                    L.q.qmA[0][slot] = msk;
                    L.q.qmB[0][slot] = L.q.mskB[c];  // the twin in B's buffer
                } else if (kind == 1) {
                    L.q.qmA[1][slot] = msk;
                } else {
                    L.q.qmB[1][slot] = msk;
                }
            }
            q0 = uniform(q0 + __popcll(m0));
            q1 = uniform(q1 + __popcll(m1));
            q2 = uniform(q2 + __popcll(m2));
            // the emitted entries are a prefix of each buffer: move the rest to the front
            const int nae = __popcll(__ballot(emit && isA)), nbe = __popcll(__ballot(emit && !isA));
            const int e_own = isA ? nae : nbe;
            const bool mv = i + e_own < nown;
            const unsigned long long k2 = mv ? (isA ? L.q.keyA[i + e_own] : L.q.keyB[i + e_own]) : 0ull;
            const unsigned long long m2v = mv ? (isA ? L.q.mskA[i + e_own] : L.q.mskB[i + e_own]) : 0ull;
            wave_sync();
            if (mv) {
                if (isA) {
                    L.q.keyA[i] = k2;
                    L.q.mskA[i] = m2v;
                } else {
                    L.q.keyB[i] = k2;
                    L.q.mskB[i] = m2v;
                }
            }
            na = uniform(na - nae);
            nb = uniform(nb - nbe);
        }
        if (T.kind >= 0) {
            wait_vmem();  // the next tile's records are in before this tile's atomics go out
            wave_sync();  // Y / idx of this tile are written
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int tr = 4 * R + t;  // f32 layout: register t holds term row 4 (l/16) + t
                double val = (double)F0[t] + (double)F1[t];
                const double pb = __shfl(val, src9, 64);
                if (m < 9) val -= pb * L.Y[tr][m % 3];
                const int id = L.idx[tr];
                if (id >= 0 && m < 12 && val != 0.0) atomicAdd(adj + (int64_t)id * ADJ_STRIDE + m, val);
            }
        }
        if (Tn.kind < 0) break;
        T = Tn;
        const int idx = T.rec >= 0 ? min(T.rec, nrec - 1) : -1;
        wave_sync();  // the previous tile's reads of Y / idx are done
        if (R == 0) {
            L.Y[m][0] = T.xa.x - ox;
            L.Y[m][1] = T.xa.y - oy;
            L.Y[m][2] = T.xa.z - oz;
            L.idx[m] = idx;
        }
        const double2 fd = weight_factors<EXPM>(T.xa.w, P);
        const float2 f = make_float2((float)fd.x, (float)fd.y);
        const double epsT = T.rec >= nnodes ? EPS_W : 0.0;  // points carry evalNodeExact's eps, clusters none
        F0 = fvec4{0.0f, 0.0f, 0.0f, 0.0f};
        F1 = F0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 0 ? T.kind == 2 : T.kind == 1) continue;  // wave-uniform: the half has no lanes
            const int nqh = h == 0 ? nqA : nqB;
            const unsigned long long mh = h == 0 ? T.mA : T.mB;
            auto sums = [&](auto dense) {
                // dense: all 64 seeds nonzero, the rows are the wave's nodes
                // in lane order and row 4q + R is bit 4q + R of the masks
                const unsigned long long mR = mh >> R;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    if (dense || q < nqh) {  // rows past 4 nqh have zero seeds
                        const double4 c = L.node[64 * h + 4 * q + R];
                        const bool on = dense ? ((mR >> (4 * q)) & 1ull) != 0 : serves(mh, c);
                        const double rx = c.x - T.xa.x, ry = c.y - T.xa.y, rz = c.z - T.xa.z;
                        const float wf =
                            weight_f32<EXPM>((float)fma(rz, rz, fma(ry, ry, fma(rx, rx, epsT))), f, Lf, P);
                        const float w = on ? wf : 0.0f;
                        const float g = h == 0 ? GA[q] : GB[q];
                        if (q & 1) F1 = __builtin_amdgcn_mfma_f32_16x16x4f32(w, g, F1, 0, 0, 0);
                        else F0 = __builtin_amdgcn_mfma_f32_16x16x4f32(w, g, F0, 0, 0, 0);
                    }
                }
            };
            if (nqh == 16) sums(std::true_type{});
            else sums(std::false_type{});
        }
    }
    // the FP64 terms, per wave, with FP64 G (the merge buffers are free
    // again: the compacted seeds are written once more, to the same rows)
    if (asp >= a1 && bsp >= b1) return;
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double4 v;
        double sd[3];
        const int k = pair * 128 + 64 * h + lane;
        node_seed(k, v, sd);
        compact_nodes(k < nvol, v, sd, L.node + 64 * h, L.s + 64 * h);
    }
    wave_sync();
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const int64_t e0 = h == 0 ? asp : bsp, e1 = h == 0 ? a1 : b1;
        const int nqh = h == 0 ? nqA : nqB;
        if (e0 >= e1 || nqh == 0) continue;
        double B[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            B[q] = q < nqh ? g_entry(64 * h + 4 * q + R) : 0.0;
            __builtin_amdgcn_sched_barrier(0);
        }
        rev_terms<EXPM>(e0, e1, plan, nnodes, nrec, rec, P, L.node + 64 * h, nqh, L.Y, L.idx, B, ox, oy, oz, adj);
    }
}

// ---------------------------------------------------------------------------
// Reverse over groups of walk waves with sparse seeds (mixed precision)
// ---------------------------------------------------------------------------
// With few seeded nodes the pair reverse's time is its memory-side FP64
// atomics: one set of 12 per term of a pair's union, whatever the number of
// seeded nodes behind it.  A record appears in 1.71 of two neighbouring walk
// waves' plans, 2.88 of four and 8.0 of sixteen, so one set per term of the
// union of NS > 2 walk waves takes fewer atomics.  One wave merges the NS
// depth-first FP32 ranges on the fly (64 / NS lanes and buffered entries per
// range), queues each union term once with its NS lane masks, and sums a
// tile of 16 union terms over the seeded node quads of every walk wave that
// uses it.  A group goes here when at most 12 of each 64 of its nodes carry
// a seed, and the rows then fit a fixed LDS budget; this kernel, launched
// first, flags the groups it takes and the pair reverse leaves them.

template <int NS>
struct GroupLds {
    static constexpr int MQ = 64 / NS;             // buffered plan entries per walk wave
    static constexpr int MAXS = 12 * NS;           // seeded nodes of a sparse group
    static constexpr int ROWS = MAXS + 3 * NS;     // their rows, each walk wave's padded to quads
    static constexpr int QCAP = RT + 64;           // union queue: < RT items, then at most 64 per merge round
    double4 node[ROWS];    // mirrored positions (w: the node's lane in its walk wave, low word)
    float G[ROWS][12];     // G rows in FP32: s (x) g (9), s (3) with g = v - o; columns 12-15 are 0
    unsigned long long key[NS][MQ], msk[NS][MQ];
    int32_t qrec[QCAP];
    unsigned long long qm[NS][QCAP];
    double Y[RT][4];  // x - o of the tile's terms
    int32_t idx[RT];  // their records (-1: padding)
    int32_t nq[NS], base[NS];  // each walk wave's node quads and first row
};

template <int EXPM, int NS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_fast_rev_group(int nvol, int nwave, const double4 *__restrict__ vol,
                                                       const int32_t *__restrict__ perm, const double *__restrict__ dxv,
                                                       const double *__restrict__ den,
                                                       const PlanEntry *__restrict__ plan,
                                                       const int64_t *__restrict__ waveOff,
                                                       const int64_t *__restrict__ waveSplit, int nnodes, int nrec,
                                                       const Rec64 *__restrict__ rec, WParams P, Mirror M,
                                                       double *__restrict__ adj, uint32_t *__restrict__ groupSparse) {
    static_assert(NS == 2 || NS == 4 || NS == 8, "the lanes split evenly over the walk waves");
    using Lds = GroupLds<NS>;
    constexpr int MQ = Lds::MQ;
    __shared__ Lds L;
    const int ngroup = (nwave + NS - 1) / NS;
    const int grp = xcd_block(ngroup);
    if (grp >= ngroup) return;
    const int lane = threadIdx.x & 63, m = lane & 15, R = lane >> 4;
    const int w0 = grp * NS;
    const int64_t node0 = (int64_t)w0 * 64;
    // the group's seeded nodes per walk wave
    unsigned long long nzm[NS];
    int tot = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        nzm[s] = __ballot(seeded(nvol, node0 + 64 * s + lane, perm, dxv));
        tot += __popcll(nzm[s]);
    }
    const bool sparse = tot <= Lds::MAXS;
    if (lane == 0) groupSparse[grp] = sparse ? 1u : 0u;
    if (!sparse || tot == 0) return;  // the pair reverse takes a dense group; no seed: nothing to add
    // origin: the group's first node (valid: some node of the group is)
    double4 o = vol[node0];
    if (M.mask) mirror_pt(M, o.x, o.y, o.z);
    const double ox = first_lane(o.x), oy = first_lane(o.y), oz = first_lane(o.z);
    // rows: walk wave s's seeded nodes at base[s] .. base[s] + 4 nq[s], in lane order, padded with rows
    // that add exactly zero (seed 0 at 1e30)
    int nq[NS], base[NS];
    int at = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int cnt = __popcll(nzm[s]);
        nq[s] = (cnt + 3) >> 2;
        base[s] = at;
        const int64_t k = node0 + 64 * s + lane;
        if ((nzm[s] >> lane) & 1ull) {
            double4 v = vol[k];
            const int64_t j = perm[k];
            const double oden = 1.0 / den[k];
            double sd[3] = {dxv[3 * j] * oden, dxv[3 * j + 1] * oden, dxv[3 * j + 2] * oden};
            if (M.mask) {
                mirror_pt(M, v.x, v.y, v.z);
                mirror_vec_b(M, sd[0], sd[1], sd[2]);
            }
            const int row = at + __popcll(nzm[s] & ((1ull << lane) - 1ull));
            L.node[row] = make_double4(v.x, v.y, v.z, __hiloint2double(0, lane));
            const double g[3] = {v.x - ox, v.y - oy, v.z - oz};
            float *Gr = L.G[row];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
#pragma unroll
                for (int b = 0; b < 3; ++b) Gr[3 * a + b] = (float)(sd[a] * g[b]);
                Gr[9 + a] = (float)sd[a];
            }
        }
        if (lane >= cnt && lane < 4 * nq[s]) {
            const int row = at + lane;
            L.node[row] = make_double4(1e30, 1e30, 1e30, __hiloint2double(0, 0));
#pragma unroll
            for (int c = 0; c < 12; ++c) L.G[row][c] = 0.0f;
        }
        at += 4 * nq[s];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s)
        if (lane == s) {
            L.nq[s] = nq[s];
            L.base[s] = base[s];
        }
    // merge state of walk wave sid = lane / MQ, held by its MQ lanes: next
    // unread entry ia, end ea of its FP32 range, entries buffered na
    const int sid = lane / MQ, i = lane % MQ;
    const int ws = w0 + sid;
    int64_t ia = 0, ea = 0;
    if (ws < nwave) {
        ia = waveOff[ws];
        ea = min(max(waveSplit[ws], ia), waveOff[ws + 1]);
    }
    int na = 0, qn = 0;
    PlanEntry pre{};
    if (ia + i < ea) pre = plan[ia + i];
    auto record = [&](int r) {  // x, A of a record (zero for padding); unconditional load
        const double4 v = *reinterpret_cast<const double4 *>(rec + min(max(r, 0), nrec - 1));
        return r >= 0 ? v : make_double4(0.0, 0.0, 0.0, 0.0);
    };
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long mine = (MQ == 64 ? ~0ull : ((1ull << MQ) - 1ull)) << (sid * MQ);
    const float Lf = (float)P.Ldef;
    const int src9 = (lane & ~15) | (9 + (m < 9 ? m / 3 : 0));
    wave_sync();
    // tile state: record, x and A, and the NS masks of term m
    int trec = -1, nrec_t = -1;
    double4 txa = make_double4(0.0, 0.0, 0.0, 0.0), nxa = txa;
    unsigned long long tm[NS], nm[NS];
    bool have = false;
    fvec4 F0 = {0.0f, 0.0f, 0.0f, 0.0f}, F1 = F0;
    for (;;) {
        // next tile: a full queue, or at the end what is left, merging as needed
        bool next = false;
        for (;;) {
            const bool done = __ballot(ia < ea || na > 0) == 0ull;
            if (qn >= RT || (done && qn > 0)) {
                const int cnt = min(qn, RT), b = qn - cnt;
                const bool ok = m < cnt;
                nrec_t = ok ? L.qrec[b + m] : -1;
#pragma unroll
                for (int t = 0; t < NS; ++t) nm[t] = ok ? L.qm[t][b + m] : 0ull;
                qn = uniform(b);
                nxa = record(nrec_t);
                next = true;
                break;
            }
            if (done) break;
            // One merge round: top up every buffer to MQ entries, then move every entry whose key
            // is at most K = min(last key loaded of each range not read to its end) to the queue,
            // once per distinct key with the masks of every range holding it.  No key past K can
            // precede one still unread, and at least one buffer empties, so the merge ends.
            const int la = (int)min((int64_t)(MQ - na), ea - ia);
            wave_sync();
            if (i < la) {
                L.key[sid][na + i] = plan_key(pre);
                L.msk[sid][na + i] = pre.mask;
            }
            ia += la;
            na += la;
            pre = PlanEntry{};
            if (ia + i < ea) pre = plan[ia + i];
            wave_sync();
            const unsigned long long last = ia < ea ? L.key[sid][na - 1] : ~0ull;
            unsigned long long K = ~0ull;
            int n[NS];
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)last, t * MQ);
                const unsigned hi = __builtin_amdgcn_readlane((unsigned)(last >> 32), t * MQ);
                const unsigned long long kt = ((unsigned long long)hi << 32) | lo;
                K = kt < K ? kt : K;
                n[t] = __builtin_amdgcn_readlane(na, t * MQ);
            }
            const bool hv = i < na;
            const unsigned long long k = hv ? L.key[sid][i] : ~0ull;
            const unsigned long long mk = hv ? L.msk[sid][i] : 0ull;
            const bool emit = hv && k <= K;
            // the same key in the other buffers (a key is unique within a range)
            unsigned long long mt[NS];
            bool earlier = false;  // a lower-numbered range holds the key: it queues it
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const int c = count_below(L.key[t], n[t], k);
                const bool twin = emit && t != sid && c < n[t] && L.key[t][c] == k;
                mt[t] = t == sid ? mk : (twin ? L.msk[t][c] : 0ull);
                earlier = earlier || (twin && t < sid);
            }
            const bool owner = emit && !earlier;
            const unsigned long long mo = __ballot(owner);
            if (owner) {
                const int slot = qn + __popcll(mo & below);
                L.qrec[slot] = (int32_t)(unsigned)k;
#pragma unroll
                for (int t = 0; t < NS; ++t) L.qm[t][slot] = mt[t];
            }
            qn = uniform(qn + __popcll(mo));
            // the consumed entries are a prefix of each buffer: move the rest to the front
            const int e_own = __popcll(__ballot(emit) & mine);
            const bool mv = i + e_own < na;
            const unsigned long long k2 = mv ? L.key[sid][i + e_own] : 0ull;
            const unsigned long long m2 = mv ? L.msk[sid][i + e_own] : 0ull;
            wave_sync();
            if (mv) {
                L.key[sid][i] = k2;
                L.msk[sid][i] = m2;
            }
            na -= e_own;
        }
        if (have) {
            wave_sync();  // Y / idx of this tile are written
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int tr = 4 * R + t;  // f32 layout: register t holds term row 4 (l/16) + t
                double val = (double)F0[t] + (double)F1[t];
                const double pb = __shfl(val, src9, 64);
                if (m < 9) val -= pb * L.Y[tr][m % 3];
                const int id = L.idx[tr];
                if (id >= 0 && m < 12 && val != 0.0) atomicAdd(adj + (int64_t)id * ADJ_STRIDE + m, val);
            }
        }
        if (!next) break;
        have = true;
        trec = nrec_t;
        txa = nxa;
#pragma unroll
        for (int t = 0; t < NS; ++t) tm[t] = nm[t];
        const int idx = trec >= 0 ? min(trec, nrec - 1) : -1;
        wave_sync();  // the previous tile's reads of Y / idx are done
        if (R == 0) {
            L.Y[m][0] = txa.x - ox;
            L.Y[m][1] = txa.y - oy;
            L.Y[m][2] = txa.z - oz;
            L.idx[m] = idx;
        }
        const double2 fd = weight_factors<EXPM>(txa.w, P);
        const float2 f = make_float2((float)fd.x, (float)fd.y);
        const double epsT = trec >= nnodes ? EPS_W : 0.0;  // points carry evalNodeExact's eps, clusters none
        F0 = fvec4{0.0f, 0.0f, 0.0f, 0.0f};
        F1 = F0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            // walk wave s: its seeded node quads, when a term of the tile serves it
            if (nq[s] == 0 || __ballot(tm[s] != 0ull) == 0ull) continue;
            const unsigned long long ms = tm[s];
#pragma unroll 2
            for (int q = 0; q < nq[s]; ++q) {
                const int row = base[s] + 4 * q + R;
                const double4 c = L.node[row];
                const bool on = serves(ms, c);
                const double rx = c.x - txa.x, ry = c.y - txa.y, rz = c.z - txa.z;
                const float wf = weight_f32<EXPM>((float)fma(rz, rz, fma(ry, ry, fma(rx, rx, epsT))), f, Lf, P);
                const float w = on ? wf : 0.0f;
                const float g = m < 12 ? L.G[row][m] : 0.0f;
                if (q & 1) F1 = __builtin_amdgcn_mfma_f32_16x16x4f32(w, g, F1, 0, 0, 0);
                else F0 = __builtin_amdgcn_mfma_f32_16x16x4f32(w, g, F0, 0, 0, 0);
            }
        }
    }
    // the FP64 terms, per walk wave, with FP64 G rows (their seeds read
    // again: a padding row's lane holds no seeded node)
    const int sa = m < 9 ? m / 3 : (m < 12 ? m - 9 : 0), gb = m % 3;
    const double og = gb == 0 ? ox : (gb == 1 ? oy : oz);
#pragma unroll 1
    for (int s = 0; s < NS; ++s) {
        const int w = w0 + s, nqs = uniform(L.nq[s]), bs = uniform(L.base[s]);
        if (w >= nwave || nqs == 0) continue;
        const int64_t e1 = waveOff[w + 1], e0 = min(max(waveSplit[w], waveOff[w]), e1);
        if (e0 >= e1) continue;
        const unsigned long long nzs = __ballot(seeded(nvol, node0 + 64 * s + lane, perm, dxv));
        double B[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            double gv = 0.0;
            if (q < nqs) {
                const int row = bs + 4 * q + R;
                const double4 c = L.node[row];
                const int ln = __double2loint(c.w) & 63;
                if ((nzs >> ln) & 1ull && 4 * q + R < __popcll(nzs)) {
                    const int64_t kk = node0 + 64 * s + ln;
                    const int64_t j = perm[kk];
                    const double oden = 1.0 / den[kk];
                    double sd[3] = {dxv[3 * j] * oden, dxv[3 * j + 1] * oden, dxv[3 * j + 2] * oden};
                    if (M.mask) mirror_vec_b(M, sd[0], sd[1], sd[2]);
                    const double sv = sa == 0 ? sd[0] : (sa == 1 ? sd[1] : sd[2]);
                    const double pv = (gb == 0 ? c.x : (gb == 1 ? c.y : c.z)) - og;
                    gv = m < 9 ? sv * pv : (m < 12 ? sv : 0.0);
                }
            }
            B[q] = gv;
        }
        rev_terms<EXPM>(e0, e1, plan, nnodes, nrec, rec, P, L.node + bs, nqs, L.Y, L.idx, B, ox, oy, oz, adj);
    }
}

// Per-lane work of the fast walk, summed over all volume nodes: internal
// node visits V, accepted clusters A, leaf point pairs L and error checks E
// (the quantities of dryRun's cost model, kd_tree.F90:789-830).
__global__ void __launch_bounds__(256) k_fast_count(int nvol, const double4 *__restrict__ vol, int nnodes,
                                                    const NodeRec *__restrict__ nodes, double errTol,
                                                    unsigned long long *__restrict__ out, Mirror M) {
#pragma clang fp contract(off)
    __shared__ unsigned long long red[8];
    if (threadIdx.x < 8) red[threadIdx.x] = 0;
    __syncthreads();
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < nvol;
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    const double tol = errTol * v.w;
    int stop = valid ? INT32_MAX : -1;
    unsigned long long cV = 0, cA = 0, cL = 0, cE = 0;
    unsigned long long wV = 0, wA = 0, wL = 0, wP = 0;  // per wave (uniform)
    int node = 0;
    while (node < nnodes) {
        const NodeRec *r = nodes + node;
        const int depth = uniform(r->depth);
        const bool active = depth <= stop;
        if (uniform(r->leaf)) {
            if (active) cL += r->n;
            ++wL;
            wP += uniform(r->n);
            node = uniform(r->skip);
            continue;
        }
        ++wV;
        const double cx = v.x - r->X[0], cy = v.y - r->X[1], cz = v.z - r->X[2];
        const double dist = sqrt(cx * cx + cy * cy + cz * cz);
        bool descend = false;
        if (active) {
            ++cV;
            const double dOvrR = dist * r->orad;
            if (dOvrR < 2.0) descend = true;
            else {
                ++cE;
                if (get_error(r, dOvrR, c_rseg) < tol) ++cA;
                else descend = true;
            }
            stop = descend ? INT32_MAX : depth;
        }
        if (__any(active && !descend && dist * r->orad >= 2.0)) ++wA;
        node = __any(descend) ? node + 1 : uniform(r->skip);
    }
    atomicAdd(&red[0], cV);
    atomicAdd(&red[1], cA);
    atomicAdd(&red[2], cL);
    atomicAdd(&red[3], cE);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&red[4], wV);
        atomicAdd(&red[5], wA);
        atomicAdd(&red[6], wL);
        atomicAdd(&red[7], wP);
    }
    __syncthreads();
    if (threadIdx.x < 8) atomicAdd(out + threadIdx.x, red[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Nodal reverse (setData_b + COMPUTENODALPROPERTIES_B, kd_tree.F90:1033-1162)
// ---------------------------------------------------------------------------
// GETROTATIONMATRIX3D_B (vectorUtils_b.f90:7-154): d<Mb, M>/dv2, row-major Mb.
__device__ void rodrigues_b(const double *v1, const double *v2, const double *Mb, double *v2b) {
#pragma clang fp contract(off)
    double ax0[3];
    cross3(v1, v2, ax0);
    const double axmag = sqrt(ax0[0] * ax0[0] + ax0[1] * ax0[1] + ax0[2] * ax0[2] + 1e-30);
    v2b[0] = v2b[1] = v2b[2] = 0.0;
    if (axmag < ROT_TOL) return;  // M = I there: no dependence on v2
    const double m1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2] + 1e-30);
    const double m2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2] + 1e-30);
    const double ax[3] = {ax0[0] / axmag, ax0[1] / axmag, ax0[2] / axmag};
    const double u1[3] = {v1[0] / m1, v1[1] / m1, v1[2] / m1};
    const double u2[3] = {v2[0] / m2, v2[1] / m2, v2[2] / m2};
    const double dot = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
    const bool clipped = !(1.0 > dot);
    const double arg = clipped ? 1.0 : dot;
    const double ang = acos(arg), sn = sin(ang), cs = cos(ang);
    const double A[9] = {0.0, -ax[2], ax[1], ax[2], 0.0, -ax[0], -ax[1], ax[0], 0.0};
    // angle adjoint: d/da of sin(a) A + (1-cos a) A^2
    double sA = 0.0, sC = 0.0, Ab[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double C = A[3 * r] * A[c] + A[3 * r + 1] * A[3 + c] + A[3 * r + 2] * A[6 + c];
            sA += A[3 * r + c] * Mb[3 * r + c];
            sC += C * Mb[3 * r + c];
        }
    const double angb = cs * sA + sn * sC;
    // skew adjoint: sin(a) Mb + (1-cos a) (Mb A^T + A^T Mb)
    const double oc = 1.0 - cs;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
            for (int m = 0; m < 3; ++m) t += Mb[3 * r + m] * A[3 * c + m] + A[3 * m + r] * Mb[3 * m + c];
            Ab[3 * r + c] = sn * Mb[3 * r + c] + oc * t;
        }
    double axb[3] = {Ab[7] - Ab[5], Ab[2] - Ab[6], Ab[3] - Ab[1]};
    // angle = acos(u1 . u2)
    const double argb = (arg == 1.0 || arg == -1.0) ? 0.0 : -(angb / sqrt(1.0 - arg * arg));
    double u2b[3] = {0.0, 0.0, 0.0};
    if (!clipped) { u2b[0] = u1[0] * argb; u2b[1] = u1[1] * argb; u2b[2] = u1[2] * argb; }
    // u2 = v2 / |v2|
    const double m2b = -((v2[0] * u2b[0] + v2[1] * u2b[1] + v2[2] * u2b[2]) / (m2 * m2));
    for (int k = 0; k < 3; ++k) v2b[k] += u2b[k] / m2 + v2[k] * (m2b / m2);
    // axis = (v1 x v2) / |v1 x v2|
    const double axmagb = -((ax0[0] * axb[0] + ax0[1] * axb[1] + ax0[2] * axb[2]) / (axmag * axmag));
    double ax0b[3];
    for (int k = 0; k < 3; ++k) ax0b[k] = axb[k] / axmag + ax0[k] * (axmagb / axmag);
    // d(v1 x v2)/dv2: v2b += ax0b x v1
    v2b[0] += ax0b[1] * v1[2] - ax0b[2] * v1[1];
    v2b[1] += ax0b[2] * v1[0] - ax0b[0] * v1[2];
    v2b[2] += ax0b[0] * v1[1] - ax0b[1] * v1[0];
}

// Per node: complete Mb', Bb (leaf/exact part in padj plus the shares of the
// accepted clusters containing the point, setData_b), then the rotation
// backward and the backward of normal = sumNormal / sumArea.
__global__ void k_nodal_rev_a(int n, int fast, int useRot, const NodeRec *__restrict__ nodes,
                              const double *__restrict__ nadj, const double *__restrict__ padj,
                              const double4 *__restrict__ ps, const double4 *__restrict__ nprops,
                              const double *__restrict__ normals0, const int32_t *__restrict__ corner,
                              double *__restrict__ ntmp) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double Mb[9], Bb[3];
    const double *pa = padj + (int64_t)i * ADJ_STRIDE;
    for (int k = 0; k < 9; ++k) Mb[k] = pa[k];
    for (int k = 0; k < 3; ++k) Bb[k] = pa[9 + k];
    const double4 p = ps[i];
    if (fast) {
        int node = 0;
        while (!nodes[node].leaf) {
            const NodeRec &r = nodes[node];
            const double *na = nadj + (int64_t)node * ADJ_STRIDE;
            const double ovrN = 1.0 / r.n;
            const double dx = r.X[0] - p.x, dy = r.X[1] - p.y, dz = r.X[2] - p.z;
            for (int q = 0; q < 3; ++q) {
                const double b = na[9 + q];
                Mb[3 * q] += (na[3 * q] + b * dx) * ovrN;
                Mb[3 * q + 1] += (na[3 * q + 1] + b * dy) * ovrN;
                Mb[3 * q + 2] += (na[3 * q + 2] + b * dz) * ovrN;
                Bb[q] += b * ovrN;
            }
            const int left = node + 1;
            node = (i <= nodes[left].u) ? left : nodes[left].skip;
        }
    }
    double nb[3] = {0.0, 0.0, 0.0};
    const double4 np4 = nprops[i];
    const double sa = np4.x, sn[3] = {np4.y, np4.z, np4.w};
    if (useRot && !corner[i]) {
        const double n0[3] = {normals0[3 * i], normals0[3 * i + 1], normals0[3 * i + 2]};
        const double nrm[3] = {sn[0] / sa, sn[1] / sa, sn[2] / sa};
        rodrigues_b(n0, nrm, Mb, nb);
    }
    double *o = ntmp + (int64_t)i * 8;
    o[0] = Bb[0]; o[1] = Bb[1]; o[2] = Bb[2];
    o[3] = nb[0] / sa; o[4] = nb[1] / sa; o[5] = nb[2] / sa;
    o[6] = (-(sn[0] * nb[0] / sa) + -(sn[1] * nb[1] / sa) + -(sn[2] * nb[2] / sa)) / sa;
    o[7] = 0.0;
}

// GETELEMENTPROPS_B per face with the seeds of all its slots summed:
// area seed sum (sab + n_f . snb)/nPts, unit-normal seed dA_f * sum snb.
// The fan's points are reloaded (centroid-relative) where an edge needs
// them, and each point's seed is written once, to its face slot, then
// shifted by the centroid's share: no per-point arrays, so nothing goes to
// scratch memory.  The arithmetic and its order are the per-face form's
// (point j's seed sums edge j's term and edge j-1's, from zero, in the
// order edge 0 ... edge np-1 adds them).
__device__ __forceinline__ void fan_point(const double *xs, const int32_t *xuInd, const int32_t *conn, int j,
                                          const double *c, double *p) {
#pragma clang fp contract(off)
    load_xu(xs, xuInd, conn[j], p);
    p[0] -= c[0]; p[1] -= c[1]; p[2] -= c[2];
}

__global__ void k_nodal_rev_face(int nface, const int32_t *__restrict__ facePtr, const int32_t *__restrict__ faceConn,
                                 const int32_t *__restrict__ xuInd, const double *__restrict__ xs,
                                 const double4 *__restrict__ fprops, const double *__restrict__ ntmp,
                                 double *__restrict__ pb) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int p0 = facePtr[f], np = facePtr[f + 1] - p0;
    const int32_t *conn = faceConn + p0;
    const double4 fp = fprops[f];
    double areab = 0.0, nrb[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < np; ++j) {
        const double *t = ntmp + (int64_t)conn[j] * 8;
        areab += (t[6] + (fp.y * t[3] + fp.z * t[4] + fp.w * t[5])) / np;
        nrb[0] += t[3]; nrb[1] += t[4]; nrb[2] += t[5];
    }
    nrb[0] *= fp.x; nrb[1] *= fp.x; nrb[2] *= fp.x;
    // recompute the fan (radial vectors, cross products) of getElementProps
    double c[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < np; ++j) {
        double q[3];
        load_xu(xs, xuInd, conn[j], q);
        c[0] += q[0]; c[1] += q[1]; c[2] += q[2];
    }
    c[0] /= np; c[1] /= np; c[2] /= np;
    double first[3], a[3], b[3], nsum[3] = {0.0, 0.0, 0.0};
    fan_point(xs, xuInd, conn, 0, c, first);
    a[0] = first[0]; a[1] = first[1]; a[2] = first[2];
    for (int j = 0; j < np; ++j) {
        if (j + 1 < np) fan_point(xs, xuInd, conn, j + 1, c, b);
        else { b[0] = first[0]; b[1] = first[1]; b[2] = first[2]; }
        double cr[3];
        cross3(a, b, cr);
        const double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + 1e-15);
        nsum[0] += cr[0] / cn; nsum[1] += cr[1] / cn; nsum[2] += cr[2] / cn;
        a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
    }
    // normal = nsum / |nsum|
    const double t2 = nsum[0] * nsum[0] + nsum[1] * nsum[1] + nsum[2] * nsum[2], t = sqrt(t2);
    const double g = -((nsum[0] * nrb[0] + nsum[1] * nrb[1] + nsum[2] * nrb[2]) / (t2 * t));
    const double nsb[3] = {nrb[0] / t + nsum[0] * g, nrb[1] / t + nsum[1] * g, nrb[2] / t + nsum[2] * g};
    // seed of the cross product cr = a x b of the edge (a, b)
    auto edge_crb = [&](const double *ea, const double *eb, double *crb) {
        double cr[3];
        cross3(ea, eb, cr);
        const double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + 1e-15);
        // area += cn/2, nsum += cr/cn
        const double cnb = 0.5 * areab - (cr[0] * nsb[0] + cr[1] * nsb[1] + cr[2] * nsb[2]) / (cn * cn);
        for (int k = 0; k < 3; ++k) crb[k] = nsb[k] / cn + cr[k] * (cnb / cn);
    };
    // edge j-1 (from point j-1 to point j): its seed and its first point;
    // before point 0 that is edge np-1
    double prevA[3], prevCrb[3];
    fan_point(xs, xuInd, conn, np - 1, c, prevA);
    edge_crb(prevA, first, prevCrb);
    double cb[3] = {0.0, 0.0, 0.0};
    a[0] = first[0]; a[1] = first[1]; a[2] = first[2];
    for (int j = 0; j < np; ++j) {
        if (j + 1 < np) fan_point(xs, xuInd, conn, j + 1, c, b);
        else { b[0] = first[0]; b[1] = first[1]; b[2] = first[2]; }
        double crb[3];
        edge_crb(a, b, crb);
        // cr = a x b: ab += b x crb (edge j), bb += crb x a (edge j-1's a)
        const double xa[3] = {b[1] * crb[2] - b[2] * crb[1], b[2] * crb[0] - b[0] * crb[2],
                              b[0] * crb[1] - b[1] * crb[0]};
        const double xb[3] = {prevCrb[1] * prevA[2] - prevCrb[2] * prevA[1],
                              prevCrb[2] * prevA[0] - prevCrb[0] * prevA[2],
                              prevCrb[0] * prevA[1] - prevCrb[1] * prevA[0]};
        double rv[3];
        for (int k = 0; k < 3; ++k) {
            rv[k] = 0.0;
            if (j == 0) { rv[k] += xa[k]; rv[k] += xb[k]; }
            else { rv[k] += xb[k]; rv[k] += xa[k]; }
            cb[k] -= rv[k];
        }
        double *o = pb + 3 * (int64_t)(p0 + j);
        o[0] = rv[0]; o[1] = rv[1]; o[2] = rv[2];
        prevA[0] = a[0]; prevA[1] = a[1]; prevA[2] = a[2];
        prevCrb[0] = crb[0]; prevCrb[1] = crb[1]; prevCrb[2] = crb[2];
        a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
    }
    for (int j = 0; j < np; ++j) {
        double *o = pb + 3 * (int64_t)(p0 + j);
        for (int k = 0; k < 3; ++k) o[k] = o[k] + cb[k] / np;
    }
}

// dXs(XuInd(i)) = Bb_i + sum of the face-slot adjoints of node i
// (kd_tree.F90:1146-1159: only the first duplicate receives sensitivity).
__global__ void k_nodal_rev_gather(int n, const int32_t *__restrict__ slotStart, const int32_t *__restrict__ slots,
                                   const double *__restrict__ pb, const double *__restrict__ ntmp,
                                   const int32_t *__restrict__ xuInd, double *__restrict__ dxs) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *t = ntmp + (int64_t)i * 8;
    double x[3] = {t[0], t[1], t[2]};
    for (int e = slotStart[i]; e < slotStart[i + 1]; ++e) {
        const double *q = pb + 3 * (int64_t)slots[e];
        x[0] += q[0]; x[1] += q[1]; x[2] += q[2];
    }
    const int64_t j = xuInd[i];
    dxs[3 * j] = x[0]; dxs[3 * j + 1] = x[1]; dxs[3 * j + 2] = x[2];
}

__global__ void k_fill(int64_t n, double *__restrict__ a, double v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// ---------------------------------------------------------------------------
// Init-time kernels
// ---------------------------------------------------------------------------
__constant__ double c_sphere[20][3];

// computeErrors (kd_tree.F90:832-880): for internal node q and radius j,
// the RMS over the 20 sphere points of (exact - cluster) denominator.  One
// block per internal node, one thread per (j, sphere point).
__global__ void __launch_bounds__(256) k_errors(const int32_t *__restrict__ internal, NodeRec *__restrict__ nodes,
                                                const double4 *__restrict__ ps, WParams P) {
    __shared__ double e2[NERR][20];
    const int node = internal[blockIdx.x];
    const NodeRec &r = nodes[node];
    const int t = threadIdx.x;
    if (t < NERR * 20) {
        const int j = t / 20, ii = t % 20;
        const double radius = 1.0 / r.orad * c_rstar[j];
        const double s3 = sqrt(3.0);
        double vp[3];
        for (int k = 0; k < 3; ++k) vp[k] = c_sphere[ii][k] * radius / s3 + r.X[k];
        double ex = 0.0;
        for (int i = r.l; i <= r.u; ++i) {
            const double4 p = ps[i];
            const double rx = vp[0] - p.x, ry = vp[1] - p.y, rz = vp[2] - p.z;
            ex += p.w * wfun(P.Ldef / sqrt(rx * rx + ry * ry + rz * rz), P);
        }
        const double rx = vp[0] - r.X[0], ry = vp[1] - r.X[1], rz = vp[2] - r.X[2];
        const double ap = r.Ai * wfun(P.Ldef / sqrt(rx * rx + ry * ry + rz * rz), P);
        const double d = ex - ap;
        e2[j][ii] = d * d;
    }
    __syncthreads();
    if (t < NERR) {
        double s = 0.0;
        for (int ii = 0; ii < 20; ++ii) s += e2[t][ii];
        nodes[node].err[t] = sqrt(s / 20);
    }
}

// getWiEstimate (kd_tree.F90:751-787) for every volume node: far when
// dist*oradius > 5 with dist including 1e-16.  Writes vol[k].w.
__global__ void __launch_bounds__(256) k_den0(int nvol, double4 *__restrict__ vol, int nnodes,
                                              const NodeRec *__restrict__ nodes, const double4 *__restrict__ ps,
                                              WParams P, Mirror M) {
#pragma clang fp contract(off)
    const int nb = (nvol + 255) / 256;
    const int blk = xcd_block(nb);
    if (blk >= nb) return;
    const int k = blk * 256 + threadIdx.x;
    const bool valid = k < nvol;
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    int stop = valid ? INT32_MAX : -1;
    // mirror passes after the first add to the same running sum
    // (initializeWarping.F90:240-247)
    double dd = M.mask ? v.w : 0.0;
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    int node = 0;
    while (node < nnodes) {
        const NodeRec *r = nodes + node;
        const int depth = uniform(r->depth);
        const bool active = depth <= stop;
        if (uniform(r->leaf)) {
            const int l = uniform(r->l), u = uniform(r->u);
            for (int i = l; i <= u; ++i) {
                const double4 p = ps[i];
                const double rx = v.x - p.x, ry = v.y - p.y, rz = v.z - p.z;
                const double w = p.w * wfun(P.Ldef / sqrt(rx * rx + ry * ry + rz * rz + EPS_DEN0), P);
                if (active) dd += w;
            }
            node = uniform(r->skip);
            continue;
        }
        const double cx = v.x - r->X[0], cy = v.y - r->X[1], cz = v.z - r->X[2];
        const double dist = sqrt(cx * cx + cy * cy + cz * cz + EPS_DEN0);
        bool descend = false;
        if (active) {
            if (dist * r->orad > 5.0) dd += r->Ai * wfun(P.Ldef / dist, P);
            else descend = true;
            stop = descend ? INT32_MAX : depth;
        }
        node = __any(descend) ? node + 1 : uniform(r->skip);
    }
    if (valid) vol[k].w = dd;
}

// dryRun (kd_tree.F90:789-830) per volume node: n per bucket, 0.5 per
// internal node visited and 1.1 per accepted cluster, with the dry run's own
// distance eps; summed over the mirror passes after division by the unique
// surface node count (initializeWarping.F90:258-265).  These are the
// load-balancing costs the restart file stores.
__global__ void __launch_bounds__(256) k_dry_run(int nvol, const double4 *__restrict__ vol, int nnodes,
                                                 const NodeRec *__restrict__ nodes, double errTol, int nsurf,
                                                 Mirror M, double *__restrict__ cost) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < nvol;
    double4 v = valid ? vol[k] : make_double4(0.0, 0.0, 0.0, 1.0);
    if (M.mask) mirror_pt(M, v.x, v.y, v.z);
    const double tol = errTol * v.w;
    int stop = valid ? INT32_MAX : -1;
    double c = 0.0;
    int node = 0;
    while (node < nnodes) {
        const NodeRec *r = nodes + node;
        const int depth = uniform(r->depth);
        const bool active = depth <= stop;
        if (uniform(r->leaf)) {
            if (active) c += (double)uniform(r->n);
            node = uniform(r->skip);
            continue;
        }
        bool descend = false;
        if (active) {
            const double cx = v.x - r->X[0], cy = v.y - r->X[1], cz = v.z - r->X[2];
            const double dOvrR = sqrt(cx * cx + cy * cy + cz * cz + 1e-16) * r->orad;
            c += 0.5;
            if (dOvrR < 2.0) descend = true;
            else if (get_error(r, dOvrR, c_rseg) < tol) c += 1.1;
            else descend = true;
            stop = descend ? INT32_MAX : depth;
        }
        node = __any(descend) ? node + 1 : uniform(r->skip);
    }
    if (valid) cost[k] = (M.mask ? cost[k] : 0.0) + c / nsurf;
}

__global__ void k_curve_keys(int64_t nvol, const double *__restrict__ xv0, double3 lo, double3 scale,
                         uint64_t *__restrict__ keys, int32_t *__restrict__ ids) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvol) return;
    auto q = [](double t) { return (uint32_t)fmin(fmax(t, 0.0), 2097151.0); };
    keys[k] = hilbert_key(q((xv0[3 * k] - lo.x) * scale.x), q((xv0[3 * k + 1] - lo.y) * scale.y),
                          q((xv0[3 * k + 2] - lo.z) * scale.z));
    ids[k] = (int32_t)k;
}

__global__ void k_build_vol(int64_t nvol, const double *__restrict__ xv0, const int32_t *__restrict__ perm,
                            double4 *__restrict__ vol) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvol) return;
    const int64_t j = perm[k];
    vol[k] = make_double4(xv0[3 * j], xv0[3 * j + 1], xv0[3 * j + 2], 0.0);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
static inline unsigned xcd_blocks(int64_t nb) { return (unsigned)(((nb + 7) / 8) * 8); }
static inline unsigned xcd_grid(int64_t n) { return xcd_blocks((n + 255) / 256); }

void launch_face_props(hipStream_t st, const DevSurface &S, const double *xs) {
    if (S.nface) k_face_props<<<grid_for(S.nface, 256), 256, 0, st>>>(S.nface, S.facePtr, S.faceConn, S.xuInd, xs, S.fprops);
}

void launch_props_tangent(hipStream_t st, const DevSurface &S, const double *xs, const double *dxs, int useRot) {
    if (S.nface)
        k_face_props_d<<<grid_for(S.nface, 256), 256, 0, st>>>(S.nface, S.facePtr, S.faceConn, S.xuInd, xs, dxs,
                                                               S.fprops, S.fpropsd);
    k_node_props_d<<<grid_for(S.n, 256), 256, 0, st>>>(S.n, useRot, S.n2eStart, S.n2e, S.fprops, S.fpropsd, S.xuInd,
                                                       S.corner, dxs, S.ps, S.normals0, S.pd, S.pd32);
}

void launch_node_props(hipStream_t st, const DevSurface &S, const double *xs, int initial, int useRot) {
    k_node_props<<<grid_for(S.n, 256), 256, 0, st>>>(S.n, initial, useRot, S.n2eStart, S.n2e, S.fprops, S.xuInd,
                                                     S.corner, xs, S.ps, S.normals0, S.nprops, S.pd, S.pd32);
}

void launch_corners(hipStream_t st, const DevSurface &S, double threshold) {
    k_corners<<<grid_for(S.n, 256), 256, 0, st>>>(S.n, threshold, S.n2eStart, S.n2e, S.fprops, S.normals0, S.corner);
}

void launch_cluster_data(hipStream_t st, const DevSurface &S) {
    if (!S.nInternal) return;  // a single bucket: no cluster records
    k_subtree_sums<<<S.nCut, 256, 0, st>>>(S.cut, S.nodes, S.ps, S.pd, S.csum, S.rec, S.rec32);
    if (S.top.n) k_top_sums<<<1, 1024, 0, st>>>(S.top, S.topOrder, S.nodes, S.csum, S.rec, S.rec32);
}

void launch_errors(hipStream_t st, const DevSurface &S, const WParams &P, const double *sphere60) {
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sphere), sphere60, sizeof(double) * 60, 0, hipMemcpyHostToDevice, st);
    if (S.nInternal) k_errors<<<S.nInternal, 256, 0, st>>>(S.internal, S.nodes, S.ps, P);
}

void launch_den0(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M) {
    k_den0<<<xcd_grid(V.n), 256, 0, st>>>((int)V.n, V.vol, S.nnodes, S.nodes, S.ps, P, M);
}

void launch_dry_run(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                    double *cost) {
    k_dry_run<<<grid_for(V.n, 256), 256, 0, st>>>((int)V.n, V.vol, S.nnodes, S.nodes, errTol, S.n, M, cost);
}

void launch_curve_keys(hipStream_t st, int64_t n, const double *xv0, double lo[3], double scale[3], uint64_t *keys,
                   int32_t *ids) {
    k_curve_keys<<<grid_for(n, 256), 256, 0, st>>>(n, xv0, make_double3(lo[0], lo[1], lo[2]),
                                               make_double3(scale[0], scale[1], scale[2]), keys, ids);
}

void launch_near_keys(hipStream_t st, const DevSurface &S, const DevVolume &V, double errTol, double nearT2,
                      const double lo[3], const double scale[3], uint64_t *keys, int32_t *ids) {
    k_near_keys<<<grid_for(V.n, 256), 256, 0, st>>>((int)V.n, V.vol, S.nnodes, S.nodes, errTol, nearT2,
                                                    make_double3(lo[0], lo[1], lo[2]),
                                                    make_double3(scale[0], scale[1], scale[2]), keys, ids);
}

void launch_permute_vol(hipStream_t st, int64_t n, const int32_t *order, const double4 *vin, const int32_t *pin,
                        double4 *vout, int32_t *pout) {
    k_permute_vol<<<grid_for(n, 256), 256, 0, st>>>(n, order, vin, pin, vout, pout);
}

void launch_build_vol(hipStream_t st, int64_t n, const double *xv0, const int32_t *perm, double4 *vol) {
    k_build_vol<<<grid_for(n, 256), 256, 0, st>>>(n, xv0, perm, vol);
}

__global__ void k_zero_u32(int64_t n, unsigned int *__restrict__ a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0u;
}

void launch_zero_u32(hipStream_t st, unsigned int *a, int64_t n) {
    if (n > 0) k_zero_u32<<<grid_for(n, 256), 256, 0, st>>>(n, a);
}

void launch_fill(hipStream_t st, double *a, int64_t n, double v) {
    if (n > 0) k_fill<<<grid_for(n, 256), 256, 0, st>>>(n, a, v);
}

template <int PREC, int EXPM, int FCH>
static void fast_fwd_launch(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                            double *xv, const DevPlan *plan, int64_t blk0, int64_t nblk) {
    k_fast_fwd<PREC, EXPM, FCH><<<xcd_blocks(nblk), 256, 0, st>>>(
        (int)V.n, (int)blk0, (int)nblk, V.vol, V.perm, plan->entries, plan->waveOff, plan->waveSplit, S.nnodes,
        S.nnodes + S.n, S.rec, S.rec32, P, M, xv, V.den, V.acc);
}

template <int PREC, int EXPM>
static void fast_fwd_fch(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                         double *xv, const DevPlan *plan, int64_t blk0, int64_t nblk, int fch) {
    if (fch == 16) fast_fwd_launch<PREC, EXPM, 16>(st, S, V, P, M, xv, plan, blk0, nblk);
    else if (fch == 64) fast_fwd_launch<PREC, EXPM, 64>(st, S, V, P, M, xv, plan, blk0, nblk);
    else fast_fwd_launch<PREC, EXPM, 32>(st, S, V, P, M, xv, plan, blk0, nblk);
}

template <int EXPM>
static void exact_fwd_mc(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                         double *xv, int g) {
    const unsigned grid = xcd_grid(V.n);
    const int nrec = S.nnodes + S.n;
    if (g == 16)
        k_exact_fwd_mc<EXPM, 16><<<grid, 256, 0, st>>>((int)V.n, V.vol, V.perm, S.n, S.nnodes, nrec, S.rec, S.rec32,
                                                      S.ps, P, M, xv, V.den, V.acc);
    else if (g == 32)
        k_exact_fwd_mc<EXPM, 32><<<grid, 256, 0, st>>>((int)V.n, V.vol, V.perm, S.n, S.nnodes, nrec, S.rec, S.rec32,
                                                      S.ps, P, M, xv, V.den, V.acc);
    else
        k_exact_fwd_mc<EXPM, 64><<<grid, 256, 0, st>>>((int)V.n, V.vol, V.perm, S.n, S.nnodes, nrec, S.rec, S.rec32,
                                                      S.ps, P, M, xv, V.den, V.acc);
}

template <int EXPM>
static void forward_impl(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                         int fast, int prec, double *xv, const DevPlan *plan, int64_t blk0, int64_t nblk,
                         const Tuning &T, double *wout) {
    if (nblk < 0) nblk = grid_for(V.n, 256) - blk0;
    if (fast && nblk <= 0) return;
    if (!fast && wout) {
        // stored weights: G = 8 in mixed precision
        if (prec == IDW_PREC_FP64)
            k_exact_fwd<IDW_PREC_FP64, EXPM, 8, true><<<xcd_grid(V.n), 256, 0, st>>>(
                (int)V.n, V.vol, V.perm, S.n, S.ps, S.pd, S.pd32, P, M, xv, V.den, V.acc, wout);
        else
            k_exact_fwd<IDW_PREC_MIXED, EXPM, 8, true><<<xcd_grid(V.n), 256, 0, st>>>(
                (int)V.n, V.vol, V.perm, S.n, S.ps, S.pd, S.pd32, P, M, xv, V.den, V.acc, wout);
    } else if (!fast) {
        // G only shapes the mixed-precision kernel
        if (prec == IDW_PREC_FP64)
            k_exact_fwd<IDW_PREC_FP64, EXPM, 8><<<xcd_grid(V.n), 256, 0, st>>>((int)V.n, V.vol, V.perm, S.n, S.ps, S.pd,
                                                                              S.pd32, P, M, xv, V.den, V.acc);
        else exact_fwd_mc<EXPM>(st, S, V, P, M, xv, T.exactG);
    } else {
        if (prec == IDW_PREC_FP64) fast_fwd_fch<IDW_PREC_FP64, EXPM>(st, S, V, P, M, xv, plan, blk0, nblk, T.fch);
        else fast_fwd_fch<IDW_PREC_MIXED, EXPM>(st, S, V, P, M, xv, plan, blk0, nblk, T.fch);
    }
}

void launch_prep_numb(hipStream_t st, const DevVolume &V, const Mirror &M, const double *dxv) {
    k_prep_numb<<<grid_for(V.n, 256), 256, 0, st>>>((int)V.n, V.perm, dxv, V.den, M, V.vol, V.numb, V.volm);
}

void launch_plan_count(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                       int64_t *waveCount) {
    const int nb = (int)grid_for(V.n, 256);
    k_plan_walk<true, 128><<<nb, 256, 0, st>>>((int)V.n, 0, nb, V.vol, S.nnodes, S.whdr, S.nodes, errTol, 0.0, 0.0,
                                               waveCount, nullptr, nullptr, nullptr, nullptr, M);
}

template <int WBUF>
static void plan_fill_launch(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                             double nearT2, double farT2, const DevPlan &plan, int64_t blk0, int64_t nblk) {
    k_plan_walk<false, WBUF><<<xcd_blocks(nblk), 256, 0, st>>>((int)V.n, (int)blk0, (int)nblk, V.vol, S.nnodes, S.whdr,
                                                               S.nodes, errTol, nearT2, farT2, nullptr, plan.entries,
                                                               plan.waveOff, plan.waveSplit, plan.mismatch, M);
}

void launch_plan_fill(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                      double nearT2, double farT2, const DevPlan &plan, const Tuning &T, int64_t blk0, int64_t nblk) {
    if (nblk < 0) nblk = grid_for(V.n, 256) - blk0;
    if (nblk <= 0) return;
    if (T.wbuf == 64) plan_fill_launch<64>(st, S, V, M, errTol, nearT2, farT2, plan, blk0, nblk);
    else if (T.wbuf == 256) plan_fill_launch<256>(st, S, V, M, errTol, nearT2, farT2, plan, blk0, nblk);
    else plan_fill_launch<128>(st, S, V, M, errTol, nearT2, farT2, plan, blk0, nblk);
}

// Mixed-precision fast reverse: the groups of GROUP_NS walk waves with
// sparse seeds in k_fast_rev_group (which flags them), then the other
// groups' pairs of walk waves in k_fast_rev_pair.
template <int EXPM, int MQ>
static void rev_pair_launch(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                            const double *dxv, const DevPlan *plan) {
    const int64_t ngroup = (plan->nwave + GROUP_NS - 1) / GROUP_NS;
    k_fast_rev_group<EXPM, GROUP_NS><<<xcd_blocks(ngroup), 64, 0, st>>>(
        (int)V.n, (int)plan->nwave, V.vol, V.perm, dxv, V.den, plan->entries, plan->waveOff, plan->waveSplit, S.nnodes,
        S.nnodes + S.n, S.rec, P, M, S.adj, plan->groupSparse);
    k_fast_rev_pair<EXPM, MQ><<<xcd_blocks(((plan->nwave + 1) / 2 + 3) / 4), 256, 0, st>>>(
        (int)V.n, (int)plan->nwave, V.vol, V.perm, dxv, V.den, plan->entries, plan->waveOff, plan->waveSplit, S.nnodes,
        S.nnodes + S.n, S.rec, P, M, S.adj, plan->groupSparse);
}

template <int EXPM>
static void reverse_impl(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                         int fast, int prec, double *partial, int nchunk, const double *dxv, const DevPlan *plan,
                         const Tuning &T, double *detPart, const double *wst) {
    if (!fast) {
        dim3 g(grid_for(S.n, 256), nchunk);
        if (wst) {  // every node (its row of weights); chunks rounded to 4 nodes: the last may be short or empty
            launch_prep_numb(st, V, M, dxv);
            const double4 *pos = M.mask ? V.volm : V.vol;
            const int chunk = (int)((V.n + nchunk - 1) / nchunk);
            k_exact_rev_w<<<g, 256, 0, st>>>((int)V.n, (chunk + 3) & ~3, pos, V.numb, S.n, S.ps, wst, partial);
        } else {  // the seeded nodes, packed
            const int nb = (int)grid_for(V.n, 256);
            k_seed_count<<<nb, 256, 0, st>>>((int)V.n, V.perm, dxv, V.seedOff);
            k_scan_counts<<<1, 1024, 0, st>>>(nb, V.seedOff);
            k_pack_seeds<<<nb, 256, 0, st>>>((int)V.n, V.perm, dxv, V.den, M, V.vol, V.seedOff, V.numb, V.cpos);
            const int32_t *cnt = V.seedOff + nb;
            if (prec == IDW_PREC_FP64)
                k_exact_rev<IDW_PREC_FP64, EXPM><<<g, 256, 0, st>>>(nchunk, cnt, V.cpos, V.numb, S.n, S.ps, P, partial);
            else
                k_exact_rev<IDW_PREC_MIXED, EXPM><<<g, 256, 0, st>>>(nchunk, cnt, V.cpos, V.numb, S.n, S.ps, P,
                                                                     partial);
        }
        // the first pass stores, later passes add
        k_exact_rev_reduce<<<grid_for((int64_t)S.n * 12, 256), 256, 0, st>>>(S.n, nchunk, partial, S.padj,
                                                                              M.mask != 0);
    } else {
        const unsigned g = xcd_grid(V.n);
        // the deterministic reverse takes the per-wave path with FP64 weights in either precision
        if (prec == IDW_PREC_FP64 || detPart)
            k_fast_rev<IDW_PREC_FP64, EXPM><<<g, 256, 0, st>>>(
                (int)V.n, V.vol, V.perm, dxv, V.den, plan->entries, plan->waveOff, plan->waveSplit, S.nnodes,
                S.nnodes + S.n, S.rec, P, M, S.adj, detPart);
        else if (T.mq == 8)
            rev_pair_launch<EXPM, 8>(st, S, V, P, M, dxv, plan);
        else if (T.mq == 16)
            rev_pair_launch<EXPM, 16>(st, S, V, P, M, dxv, plan);
        else
            rev_pair_launch<EXPM, 32>(st, S, V, P, M, dxv, plan);
    }
}

void launch_reverse_prep(hipStream_t st, const DevSurface &S, int fast) {
    if (fast) launch_fill(st, S.adj, ((int64_t)S.nnodes + S.n) * ADJ_STRIDE, 0.0);
}

void launch_fast_count(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                       unsigned long long *out) {
    k_fast_count<<<grid_for(V.n, 256), 256, 0, st>>>((int)V.n, V.vol, S.nnodes, S.nodes, errTol, out, M);
}

void launch_forward(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                    int fast, int prec, double *xv, const DevPlan *plan, const Tuning &T, int64_t blk0, int64_t nblk,
                    double *wout) {
    if (P.expMode == EXP_35) forward_impl<EXP_35>(st, S, V, P, M, fast, prec, xv, plan, blk0, nblk, T, wout);
    else forward_impl<EXP_INT>(st, S, V, P, M, fast, prec, xv, plan, blk0, nblk, T, wout);
}

void launch_reverse(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                    int fast, int prec, double *partial, int nchunk, const double *dxv, const DevPlan *plan,
                    const Tuning &T, double *detPart, const double *wst) {
    if (P.expMode == EXP_35)
        reverse_impl<EXP_35>(st, S, V, P, M, fast, prec, partial, nchunk, dxv, plan, T, detPart, wst);
    else reverse_impl<EXP_INT>(st, S, V, P, M, fast, prec, partial, nchunk, dxv, plan, T, detPart, wst);
}

void launch_plan_recs(hipStream_t st, const DevPlan &plan, int nrec, unsigned *keys, int32_t *vals) {
    if (plan.total > 0) k_plan_recs<<<grid_for(plan.total, 256), 256, 0, st>>>(plan.total, plan.entries, keys, vals, nrec);
}

void launch_rec_bounds(hipStream_t st, int nrec, int64_t total, const unsigned *keys, int64_t *recStart) {
    k_rec_bounds<<<grid_for((int64_t)nrec + 1, 256), 256, 0, st>>>(nrec, total, keys, recStart);
}

void launch_rec_chunks(hipStream_t st, int nrec, const int64_t *recStart, int64_t *count) {
    k_rec_chunks<<<grid_for(nrec, 256), 256, 0, st>>>(nrec, recStart, count);
}

void launch_chunk_recs(hipStream_t st, int nrec, const int64_t *chunkOff, int32_t *chunkRec) {
    k_chunk_recs<<<grid_for(nrec, 256), 256, 0, st>>>(nrec, chunkOff, chunkRec);
}

void launch_det_reduce(hipStream_t st, const DevSurface &S, const DevPlan &plan, const double *part,
                       double *chunkPart, int add) {
    const int nrec = S.nnodes + S.n;
    if (plan.nchunk > 0)
        k_det_chunks<<<grid_for(plan.nchunk, 4), 256, 0, st>>>(plan.nchunk, plan.chunkRec, plan.chunkOff,
                                                              plan.recStart, plan.detPos, part, chunkPart);
    k_det_finish<<<grid_for((int64_t)nrec * 12, 256), 256, 0, st>>>(nrec, plan.chunkOff, chunkPart, S.adj, add);
}

void launch_nodal_reverse(hipStream_t st, const DevSurface &S, const double *xs, int fast, int useRot, double *dxs,
                          int64_t npts) {
    k_nodal_rev_a<<<grid_for(S.n, 256), 256, 0, st>>>(S.n, fast, useRot, S.nodes, S.nadj, S.padj, S.ps, S.nprops,
                                                      S.normals0, S.corner, S.ntmp);
    if (S.nface)
        k_nodal_rev_face<<<grid_for(S.nface, 128), 128, 0, st>>>(S.nface, S.facePtr, S.faceConn, S.xuInd, xs, S.fprops,
                                                                 S.ntmp, S.pb);
    launch_fill(st, dxs, 3 * npts, 0.0);
    k_nodal_rev_gather<<<grid_for(S.n, 256), 256, 0, st>>>(S.n, S.slotStart, S.slots, S.pb, S.ntmp, S.xuInd, dxs);
}

}  // namespace idw
