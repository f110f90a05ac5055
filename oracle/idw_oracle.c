/*
 * idw_oracle.c -- CPU restatement of IDWarp's volume-warp hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every routine restates the reference Fortran with the same operation order
 * (FP64, no FMA contraction: build with -ffp-contract=off) so that, compiled
 * with gcc, results agree with the reference's gfortran/-fdefault-real-8 build
 * to the last few ulps.  Index conventions follow the reference: tree ranges
 * l..u and faceConn/link/XuInd are 1-based; arrays are stored 0-based.
 * Mi is stored column-major per node, Mi[9*i + r + 3*c] == Mi(r+1, c+1, i+1).
 *
 * Pinning (see tests/test_oracle_*.py):
 *  - element props, Rodrigues rotation, their Tapenade reverse and the sphere
 *    points are compared with the reference's own Fortran compiled into
 *    oracle/_ref/ (oracle/Makefile);
 *  - the IDW sums, tree traversal, setData(_b), computeErrors and the
 *    denominator estimate are pinned by closed forms (rigid translation and
 *    rotation reproduce exactly), brute force on tiny meshes, the fast==exact
 *    identity at errTol=0, central finite differences and the dot-product
 *    identity; kd_tree.F90 itself needs PETSc and cannot be built here;
 *  - symmetry planes by the mirrored-body identity (a body plus its mirror
 *    image without planes equals the body with the plane, exact mode).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NERR 12
/* kd_tree.F90:34-35 */
static const double RSTAR[NERR] = {2.0, 3.0, 4.0, 5.0, 7.5, 10.0, 20.0, 40.0, 80.0, 160.0, 320.0, 640.0};
/* constants.F90:31-33 (built with -fdefault-real-8, so these are FP64 literals) */
static const double C_PI = 3.1415926535897931;
static const double C_PHI = 2.1180339887498949;
static const double C_EPS = 1e-15;

typedef struct {
    double aExp, bExp, LdefFact, alpha, errTol, cornerAngle;
    int32_t evalMode; /* 0 exact, 1 fast (gridInput.f90:16-17) */
    int32_t useRotations, zeroCornerRotations, bucketSize;
} orc_options;

typedef struct {
    int dnum;
    double val;
    int l, u, n;
    int left, right;
    double Ai, X[3], oradius, err[NERR];
    double Mi[9], Bi[3], Mib[9], Bib[3];
    double Mid[9], Bid[3]; /* tangent cluster means (setData_d) */
} onode;

typedef struct {
    orc_options opt; /* current gridInput values */
    int n, nface, nlink, nconn;
    int *facePtr;   /* 0..nface */
    int *faceConn;  /* 1-based tree index */
    int *n2eStart;  /* n+1, CSR of nodeToElem (face ids 1-based, duplicates kept) */
    int *n2e;
    int *XuInd;     /* 1-based point index into Xs */
    int *indexes;   /* tree permutation, 1-based into the_data */
    double *the_data, *Xu0, *Xu, *Ai, *normals0, *normals, *Mi, *Bi;
    double *Mib, *Bib, *Xub, *normalsb;
    double *Mid, *Bid, *Xud; /* tangent of Mi, Bi, Xu (COMPUTENODALPROPERTIES_D) */
    int *isCorner;
    double Ldef0, Ldef, alphaToBexp, aExp, bExp, errTol;
    double orstar[NERR - 1];
    onode *nodes;
    int nnodes, cap, bucket;
    int nplanes; /* symmetry planes (setSymmetryPlanes); nLoop = 2^nplanes */
    double symmPts[6][3], symmNormals[6][3];
} octx;

/* ------------------------------------------------------------------ */
/* Symmetry planes (initializeWarping.F90:511-659).  Loop index kk is  */
/* 1-based as in the reference; plane j (1-based) is applied when      */
/* mod(kk-1, 2^j) >= 2^(j-1).                                          */
/* ------------------------------------------------------------------ */
static int mirror_on(int kk, int j) { return (kk - 1) % (1 << j) >= (1 << (j - 1)); }

/* getMirrorPt (initializeWarping.F90:533-577) */
static void mirror_pt(const octx *t, const double *pt, double *r, int kk) {
    r[0] = pt[0]; r[1] = pt[1]; r[2] = pt[2];
    for (int j = 1; j <= t->nplanes; j++) {
        if (!mirror_on(kk, j)) continue;
        const double *p = t->symmPts[j - 1], *n = t->symmNormals[j - 1];
        double d[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
        double dp = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
        for (int k = 0; k < 3; k++) r[k] = r[k] - 2 * n[k] * dp;
    }
}

/* getMirrorNumerator (:579-600) */
static void mirror_num(const octx *t, double *num, int kk) {
    for (int j = 1; j <= t->nplanes; j++) {
        if (!mirror_on(kk, j)) continue;
        const double *n = t->symmNormals[j - 1];
        double dp = num[0] * n[0] + num[1] * n[1] + num[2] * n[2];
        for (int k = 0; k < 3; k++) num[k] = num[k] - 2.0 * dp * n[k];
    }
}

/* getMirrorNumerator_b (:602-634) */
static void mirror_num_b(const octx *t, double *numb, int kk) {
    for (int j = t->nplanes; j >= 1; j--) {
        if (!mirror_on(kk, j)) continue;
        const double *n = t->symmNormals[j - 1];
        double dpb = -(n[0] * 2.0 * numb[0] + n[1] * 2.0 * numb[1] + n[2] * 2.0 * numb[2]);
        for (int k = 0; k < 3; k++) numb[k] = numb[k] + n[k] * dpb;
    }
}

/* getMirrorNumerator_d (:636-659) */
static void mirror_num_d(const octx *t, double *numd, int kk) {
    for (int j = 1; j <= t->nplanes; j++) {
        if (!mirror_on(kk, j)) continue;
        const double *n = t->symmNormals[j - 1];
        double dpd = n[0] * numd[0] + n[1] * numd[1] + n[2] * numd[2];
        for (int k = 0; k < 3; k++) numd[k] = numd[k] - 2.0 * n[k] * dpd;
    }
}

/* ------------------------------------------------------------------ */
/* vectorUtils.f90:7-103                                              */
/* ------------------------------------------------------------------ */
static void cross3(const double *v1, const double *v2, double *c) {
    c[0] = (v1[1] * v2[2] - v1[2] * v2[1]);
    c[1] = (v1[2] * v2[0] - v1[0] * v2[2]);
    c[2] = (v1[0] * v2[1] - v1[1] * v2[0]);
}

static double getmag(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-30); }

static const double ROT_TOL = 1.4901161193847656e-08; /* vectorUtils.f90:44 */

/* Skew matrix A of the axis and C = A*A in the reference's explicit form. */
static void rot_AC(const double *axis, double A[3][3], double C[3][3]) {
    A[0][0] = 0.0;      A[0][1] = -axis[2]; A[0][2] = axis[1];
    A[1][0] = axis[2];  A[1][1] = 0.0;      A[1][2] = -axis[0];
    A[2][0] = -axis[1]; A[2][1] = axis[0];  A[2][2] = 0.0;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            C[r][c] = A[r][0] * A[0][c] + A[r][1] * A[1][c] + A[r][2] * A[2][c];
}

/* getRotationMatrix3d (vectorUtils.f90:31-103); M column-major */
void orc_rotation_matrix(const double *v1, const double *v2, double *M) {
    double magV1 = getmag(v1), magV2 = getmag(v2);
    double axis[3], A[3][3], C[3][3], angle;
    cross3(v1, v2, axis);
    double axisMag = getmag(axis);
    if (axisMag < ROT_TOL) {
        angle = 0.0;
        axis[0] = 1.0; axis[1] = 0.0; axis[2] = 0.0;
    } else {
        double vv1[3], vv2[3];
        for (int k = 0; k < 3; k++) axis[k] = axis[k] / axisMag;
        for (int k = 0; k < 3; k++) { vv1[k] = v1[k] / magV1; vv2[k] = v2[k] / magV2; }
        double dot = vv1[0] * vv2[0] + vv1[1] * vv2[1] + vv1[2] * vv2[2];
        double arg = dot < 1.0 ? dot : 1.0;
        angle = acos(arg);
    }
    rot_AC(axis, A, C);
    double s = sin(angle), oc = 1.0 - cos(angle);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            M[r + 3 * c] = (r == c ? 1.0 : 0.0) + s * A[r][c] + oc * C[r][c];
}

/* GETMAG_B (vectorUtils_b.f90:181-201) */
static void getmag_b(const double *v, double *vb, double magb) {
    double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-30;
    double tempb = (s == 0.0) ? 0.0 : magb / (2.0 * sqrt(s));
    vb[0] = vb[0] + 2 * v[0] * tempb;
    vb[1] = vb[1] + 2 * v[1] * tempb;
    vb[2] = vb[2] + 2 * v[2] * tempb;
}

/* CROSS_PRODUCT_3D_B (vectorUtils_b.f90:155-180) */
static void cross3_b(const double *v1, double *v1b, const double *v2, double *v2b, double *cb) {
    /* (a + b) - c, as the Fortran statements associate */
    v1b[0] = v1b[0] + v2[1] * cb[2] - v2[2] * cb[1];
    v2b[1] = v2b[1] + v1[0] * cb[2] - v1[2] * cb[0];
    v1b[1] = v1b[1] + v2[2] * cb[0] - v2[0] * cb[2];
    v2b[0] = v2b[0] + v1[2] * cb[1] - v1[1] * cb[2];
    v1b[2] = v1b[2] + v2[0] * cb[1] - v2[1] * cb[0];
    v2b[2] = v2b[2] + v1[1] * cb[0] - v1[0] * cb[1];
    cb[0] = cb[1] = cb[2] = 0.0;
}

/* GETROTATIONMATRIX3D_B (vectorUtils_b.f90:7-154): adds d<Mb,M>/dv2 into v2b. */
void orc_rotation_matrix_b(const double *v1, const double *v2, const double *Mb, double *v2b) {
    double magv1 = getmag(v1), magv2 = getmag(v2);
    double axis0[3], axis[3], A[3][3], C[3][3], angle = 0.0, arg = 0.0;
    double vv1[3] = {0, 0, 0}, vv2[3];
    int rotated, clipped = 0;
    cross3(v1, v2, axis0);
    double axismag = getmag(axis0);
    if (axismag < ROT_TOL) {
        rotated = 0;
        axis[0] = 1.0; axis[1] = 0.0; axis[2] = 0.0;
    } else {
        rotated = 1;
        for (int k = 0; k < 3; k++) axis[k] = axis0[k] / axismag;
        for (int k = 0; k < 3; k++) { vv1[k] = v1[k] / magv1; vv2[k] = v2[k] / magv2; }
        double dot = vv1[0] * vv2[0] + vv1[1] * vv2[1] + vv1[2] * vv2[2];
        if (1.0 > dot) { arg = dot; } else { arg = 1.0; clipped = 1; }
        angle = acos(arg);
    }
    rot_AC(axis, A, C);
    double ab[3][3], cb[3][3], mib[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) mib[r][c] = Mb[r + 3 * c];
    double sa = 0.0, sc = 0.0;
    /* SUM(a*mib), SUM(c*mib): Fortran array order (column-major) */
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) { sa += A[r][c] * mib[r][c]; sc += C[r][c] * mib[r][c]; }
    double angleb = cos(angle) * sa + sin(angle) * sc;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            ab[r][c] = sin(angle) * mib[r][c];
            cb[r][c] = (1.0 - cos(angle)) * mib[r][c];
        }
    /* reverse of C = A*A, in the reference's statement order (1-based (r,c) -> [r-1][c-1]) */
    ab[2][0] = ab[2][0] + A[0][2] * cb[2][2] + A[0][1] * cb[2][1] + (A[0][0] + A[2][2]) * cb[2][0] + A[1][2] * cb[1][0] + A[0][2] * cb[0][0];
    ab[0][2] = ab[0][2] + A[2][0] * cb[2][2] + A[1][0] * cb[1][2] + (A[0][0] + A[2][2]) * cb[0][2] + A[2][1] * cb[0][1] + A[2][0] * cb[0][0];
    ab[2][1] = ab[2][1] + A[1][2] * cb[2][2] + (A[1][1] + A[2][2]) * cb[2][1] + A[1][0] * cb[2][0] + A[1][2] * cb[1][1] + A[0][2] * cb[0][1];
    ab[1][2] = ab[1][2] + A[2][1] * cb[2][2] + (A[1][1] + A[2][2]) * cb[1][2] + A[2][1] * cb[1][1] + A[2][0] * cb[1][0] + A[0][1] * cb[0][2];
    ab[0][1] = ab[0][1] + A[2][0] * cb[2][1] + A[1][0] * cb[1][1] + A[1][2] * cb[0][2] + (A[0][0] + A[1][1]) * cb[0][1] + A[1][0] * cb[0][0];
    ab[0][0] = ab[0][0] + A[2][0] * cb[2][0] + A[1][0] * cb[1][0] + A[0][2] * cb[0][2] + A[0][1] * cb[0][1] + 2 * A[0][0] * cb[0][0];
    ab[1][0] = ab[1][0] + A[2][1] * cb[2][0] + A[0][2] * cb[1][2] + A[0][1] * cb[1][1] + (A[0][0] + A[1][1]) * cb[1][0] + A[0][1] * cb[0][0];
    double axisb[3] = {0, 0, 0};
    axisb[0] = axisb[0] + ab[2][1] - ab[1][2];
    axisb[1] = axisb[1] + ab[0][2] - ab[2][0];
    axisb[2] = axisb[2] + ab[1][0] - ab[0][1];
    double magv2b;
    if (!rotated) {
        magv2b = 0.0;
        axisb[0] = axisb[1] = axisb[2] = 0.0;
        getmag_b(axis0, axisb, 0.0);
        double v1b[3] = {0, 0, 0};
        cross3_b(v1, v1b, v2, v2b, axisb);
        getmag_b(v2, v2b, magv2b);
        return;
    }
    double argb;
    if (arg == 1.0 || arg == -1.0) argb = 0.0;
    else argb = -(angleb / sqrt(1.0 - arg * arg));
    double vv2b[3] = {0, 0, 0};
    if (!clipped) {
        vv2b[0] = vv2b[0] + vv1[0] * argb;
        vv2b[1] = vv2b[1] + vv1[1] * argb;
        vv2b[2] = vv2b[2] + vv1[2] * argb;
    }
    for (int k = 0; k < 3; k++) v2b[k] = v2b[k] + vv2b[k] / magv2;
    magv2b = -((v2[0] * vv2b[0] + v2[1] * vv2b[1] + v2[2] * vv2b[2]) / (magv2 * magv2));
    double axismagb = -((axis0[0] * axisb[0] + axis0[1] * axisb[1] + axis0[2] * axisb[2]) / (axismag * axismag));
    for (int k = 0; k < 3; k++) axisb[k] = axisb[k] / axismag;
    getmag_b(axis0, axisb, axismagb);
    double v1b[3] = {0, 0, 0};
    cross3_b(v1, v1b, v2, v2b, axisb);
    getmag_b(v2, v2b, magv2b);
}

/* CROSS_PRODUCT_3D_D (vectorUtils_d.f90:138-151), statement order kept */
static void cross3_d(const double *v1, const double *v1d, const double *v2, const double *v2d, double *c,
                     double *cd) {
    cd[0] = v2[2] * v1d[1] + v1[1] * v2d[2] - v2[1] * v1d[2] - v1[2] * v2d[1];
    c[0] = v1[1] * v2[2] - v1[2] * v2[1];
    cd[1] = v2[0] * v1d[2] + v1[2] * v2d[0] - v2[2] * v1d[0] - v1[0] * v2d[2];
    c[1] = v1[2] * v2[0] - v1[0] * v2[2];
    cd[2] = v2[1] * v1d[0] + v1[0] * v2d[1] - v2[0] * v1d[1] - v1[1] * v2d[0];
    c[2] = v1[0] * v2[1] - v1[1] * v2[0];
}

/* GETMAG_D (vectorUtils_d.f90:155-176) */
static double getmag_d(const double *v, const double *vd, double *magd) {
    double argd = 2 * v[0] * vd[0] + 2 * v[1] * vd[1] + 2 * v[2] * vd[2];
    double arg = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-30;
    double t = sqrt(arg);
    *magd = (arg == 0.0) ? 0.0 : argd / (2.0 * t);
    return t;
}

/* GETROTATIONMATRIX3D_D (vectorUtils_d.f90:7-129): M and its tangent Md
 * for a tangent v2d of v2 (v1 fixed); both column-major. */
void orc_rotation_matrix_d(const double *v1, const double *v2, const double *v2d, double *M, double *Md) {
    double magv1 = getmag(v1), magv2d, magv2 = getmag_d(v2, v2d, &magv2d);
    double v1d[3] = {0, 0, 0}, axis[3], axisd[3], axismagd;
    cross3_d(v1, v1d, v2, v2d, axis, axisd);
    double axismag = getmag_d(axis, axisd, &axismagd);
    double angle, angled;
    if (axismag < ROT_TOL) {
        angle = 0.0;
        axis[0] = 1.0; axis[1] = 0.0; axis[2] = 0.0;
        axisd[0] = axisd[1] = axisd[2] = 0.0;
        angled = 0.0;
    } else {
        double vv1[3], vv2[3], vv2d[3], arg, argd;
        for (int k = 0; k < 3; k++) axisd[k] = (axisd[k] - axis[k] * axismagd / axismag) / axismag;
        for (int k = 0; k < 3; k++) axis[k] = axis[k] / axismag;
        for (int k = 0; k < 3; k++) vv1[k] = v1[k] / magv1;
        for (int k = 0; k < 3; k++) vv2d[k] = (v2d[k] - v2[k] * magv2d / magv2) / magv2;
        for (int k = 0; k < 3; k++) vv2[k] = v2[k] / magv2;
        if (1.0 > vv1[0] * vv2[0] + vv1[1] * vv2[1] + vv1[2] * vv2[2]) {
            argd = vv1[0] * vv2d[0] + vv1[1] * vv2d[1] + vv1[2] * vv2d[2];
            arg = vv1[0] * vv2[0] + vv1[1] * vv2[1] + vv1[2] * vv2[2];
        } else {
            arg = 1.0;
            argd = 0.0;
        }
        if (arg == 1.0 || arg == -1.0) angled = 0.0;
        else angled = -(argd / sqrt(1.0 - arg * arg));
        angle = acos(arg);
    }
    double A[3][3], C[3][3], a[3][3];
    rot_AC(axis, A, C);
    /* skew matrix of the axis tangent; its square's tangent as the
     * reference's statements write it (1-based (r,c) -> [r-1][c-1]) */
    a[0][0] = 0.0;       a[0][1] = -axisd[2]; a[0][2] = axisd[1];
    a[1][0] = axisd[2];  a[1][1] = 0.0;       a[1][2] = -axisd[0];
    a[2][0] = -axisd[1]; a[2][1] = axisd[0];  a[2][2] = 0.0;
    double c[3][3];
    c[0][0] = 2 * A[0][0] * a[0][0] + A[1][0] * a[0][1] + A[0][1] * a[1][0] + A[2][0] * a[0][2] + A[0][2] * a[2][0];
    c[0][1] = A[0][1] * a[0][0] + (A[0][0] + A[1][1]) * a[0][1] + A[0][1] * a[1][1] + A[2][1] * a[0][2] + A[0][2] * a[2][1];
    c[0][2] = A[0][2] * a[0][0] + (A[0][0] + A[2][2]) * a[0][2] + A[1][2] * a[0][1] + A[0][1] * a[1][2] + A[0][2] * a[2][2];
    c[1][0] = (A[0][0] + A[1][1]) * a[1][0] + A[1][0] * a[0][0] + A[1][0] * a[1][1] + A[2][0] * a[1][2] + A[1][2] * a[2][0];
    c[1][1] = A[0][1] * a[1][0] + A[1][0] * a[0][1] + 2 * A[1][1] * a[1][1] + A[2][1] * a[1][2] + A[1][2] * a[2][1];
    c[1][2] = A[0][2] * a[1][0] + A[1][0] * a[0][2] + A[1][2] * a[1][1] + (A[1][1] + A[2][2]) * a[1][2] + A[1][2] * a[2][2];
    c[2][0] = (A[0][0] + A[2][2]) * a[2][0] + A[2][0] * a[0][0] + A[1][0] * a[2][1] + A[2][1] * a[1][0] + A[2][0] * a[2][2];
    c[2][1] = A[0][1] * a[2][0] + A[2][0] * a[0][1] + (A[1][1] + A[2][2]) * a[2][1] + A[2][1] * a[1][1] + A[2][1] * a[2][2];
    c[2][2] = A[0][2] * a[2][0] + A[2][0] * a[0][2] + A[1][2] * a[2][1] + A[2][1] * a[1][2] + 2 * A[2][2] * a[2][2];
    double sn = sin(angle), oc = 1.0 - cos(angle), cs = cos(angle);
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            Md[r + 3 * q] = (A[r][q] * cs + C[r][q] * sn) * angled + sn * a[r][q] + oc * c[r][q];
            M[r + 3 * q] = (r == q ? 1.0 : 0.0) + sn * A[r][q] + oc * C[r][q];
        }
}

/* ------------------------------------------------------------------ */
/* getElementProps.F90:1-46 and GETELEMENTPROPS_B (getElementProps_b.f90) */
/* points: 3*npts, xyz per point                                      */
/* ------------------------------------------------------------------ */
#define MAXFACEPTS 20
void orc_element_props(const double *pts, int npts, double *area, double *normal) {
    double center[3] = {0, 0, 0}, rv[MAXFACEPTS][3], cr[3];
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) center[k] = center[k] + pts[3 * i + k];
    for (int k = 0; k < 3; k++) center[k] = center[k] / npts;
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) rv[i][k] = pts[3 * i + k] - center[k];
    normal[0] = normal[1] = normal[2] = 0.0;
    *area = 0.0;
    for (int i = 0; i < npts; i++) {
        cross3(rv[i], rv[i < npts - 1 ? i + 1 : 0], cr);
        double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + 1e-15);
        *area = *area + 0.5 * cn;
        for (int k = 0; k < 3; k++) normal[k] = normal[k] + cr[k] / cn;
    }
    double nn = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    for (int k = 0; k < 3; k++) normal[k] = normal[k] / nn;
}

/* Adds the reverse of (area, unit normal) into pointsb (3*npts). */
void orc_element_props_b(const double *pts, int npts, double areab, const double *normalb_in, double *pointsb) {
    double center[3] = {0, 0, 0}, rv[MAXFACEPTS][3], rvb[MAXFACEPTS][3];
    double crs[MAXFACEPTS][3], cn[MAXFACEPTS], normal[3] = {0, 0, 0};
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) center[k] = center[k] + pts[3 * i + k];
    for (int k = 0; k < 3; k++) center[k] = center[k] / npts;
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) { rv[i][k] = pts[3 * i + k] - center[k]; rvb[i][k] = 0.0; }
    for (int i = 0; i < npts; i++) {
        cross3(rv[i], rv[i < npts - 1 ? i + 1 : 0], crs[i]);
        cn[i] = sqrt(crs[i][0] * crs[i][0] + crs[i][1] * crs[i][1] + crs[i][2] * crs[i][2] + 1e-15);
        for (int k = 0; k < 3; k++) normal[k] = normal[k] + crs[i][k] / cn[i];
    }
    double temp = normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2];
    double temp0 = sqrt(temp);
    double tmpb[3], normalb[3];
    for (int k = 0; k < 3; k++) { tmpb[k] = normalb_in[k]; normalb[k] = tmpb[k] / temp0; }
    double tempb = (temp == 0.0) ? 0.0 : -((normal[0] * tmpb[0] + normal[1] * tmpb[1] + normal[2] * tmpb[2]) / (2.0 * (temp0 * temp0 * temp0)));
    for (int k = 0; k < 3; k++) normalb[k] = normalb[k] + 2 * normal[k] * tempb;
    double crossb[3] = {0, 0, 0};
    for (int i = npts - 1; i >= 0; i--) {
        for (int k = 0; k < 3; k++) crossb[k] = crossb[k] + normalb[k] / cn[i];
        double cnb = 0.5 * areab - (crs[i][0] * normalb[0] + crs[i][1] * normalb[1] + crs[i][2] * normalb[2]) / (cn[i] * cn[i]);
        double s = crs[i][0] * crs[i][0] + crs[i][1] * crs[i][1] + crs[i][2] * crs[i][2] + 1e-15;
        double tb = (s == 0.0) ? 0.0 : cnb / (2.0 * sqrt(s));
        for (int k = 0; k < 3; k++) crossb[k] = crossb[k] + 2 * crs[i][k] * tb;
        int i2 = i < npts - 1 ? i + 1 : 0;
        cross3_b(rv[i], rvb[i], rv[i2], rvb[i2], crossb);
    }
    double centerb[3] = {0, 0, 0};
    for (int i = npts - 1; i >= 0; i--)
        for (int k = 0; k < 3; k++) {
            pointsb[3 * i + k] = pointsb[3 * i + k] + rvb[i][k];
            centerb[k] = centerb[k] - rvb[i][k];
        }
    for (int k = 0; k < 3; k++) centerb[k] = centerb[k] / npts;
    for (int i = npts - 1; i >= 0; i--)
        for (int k = 0; k < 3; k++) pointsb[3 * i + k] = pointsb[3 * i + k] + centerb[k];
}

/* GETELEMENTPROPS_D (getElementProps_d.f90:7-94): area, unit normal and
 * their tangents for point tangents ptsd (3*npts). */
void orc_element_props_d(const double *pts, const double *ptsd, int npts, double *area, double *aread,
                         double *normal, double *normald) {
    double center[3] = {0, 0, 0}, centerd[3] = {0, 0, 0};
    double rv[MAXFACEPTS][3], rvd[MAXFACEPTS][3], cr[3], crd[3];
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) {
            centerd[k] = centerd[k] + ptsd[3 * i + k];
            center[k] = center[k] + pts[3 * i + k];
        }
    for (int k = 0; k < 3; k++) { centerd[k] = centerd[k] / npts; center[k] = center[k] / npts; }
    for (int i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) {
            rvd[i][k] = ptsd[3 * i + k] - centerd[k];
            rv[i][k] = pts[3 * i + k] - center[k];
        }
    for (int k = 0; k < 3; k++) normal[k] = normald[k] = 0.0;
    *area = *aread = 0.0;
    for (int i = 0; i < npts; i++) {
        int i2 = i < npts - 1 ? i + 1 : 0;
        cross3_d(rv[i], rvd[i], rv[i2], rvd[i2], cr, crd);
        double arg1d = 2 * cr[0] * crd[0] + 2 * cr[1] * crd[1] + 2 * cr[2] * crd[2];
        double arg1 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + 1e-15;
        double temp = sqrt(arg1);
        double cnd = (arg1 == 0.0) ? 0.0 : arg1d / (2.0 * temp);
        double cn = temp;
        *aread = *aread + 0.5 * cnd;
        *area = *area + 0.5 * cn;
        for (int k = 0; k < 3; k++) normald[k] = normald[k] + (crd[k] - cr[k] * cnd / cn) / cn;
        for (int k = 0; k < 3; k++) normal[k] = normal[k] + cr[k] / cn;
    }
    double arg1d = 2 * normal[0] * normald[0] + 2 * normal[1] * normald[1] + 2 * normal[2] * normald[2];
    double arg1 = normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2];
    double temp = sqrt(arg1);
    double r1d = (arg1 == 0.0) ? 0.0 : arg1d / (2.0 * temp);
    double r1 = temp;
    for (int k = 0; k < 3; k++) normald[k] = (normald[k] - normal[k] * r1d / r1) / r1;
    for (int k = 0; k < 3; k++) normal[k] = normal[k] / r1;
}

/* getSpherePts.F90:1-71; vpts 3x20, fpts 3x12 (xyz per point) */
void orc_sphere_pts(const double *x0, double radius, double *vpts, double *fpts) {
    const double o = 1.0, z = 0.0, ip = 1.0 / C_PHI, p = C_PHI;
    const double nom[20][3] = {
        {-o, -o, -o}, {o, -o, -o}, {-o, o, -o}, {o, o, -o}, {-o, -o, o}, {o, -o, o}, {-o, o, o}, {o, o, o},
        {z, -ip, -p}, {z, ip, -p}, {z, -ip, p}, {z, ip, p},
        {-ip, -p, z}, {ip, -p, z}, {-ip, p, z}, {ip, p, z},
        {-p, z, -ip}, {-p, z, ip}, {p, z, -ip}, {p, z, ip}};
    static const int inds[12][5] = {
        {3, 4, 10, 15, 16}, {7, 8, 12, 15, 16}, {4, 8, 16, 19, 20}, {3, 7, 15, 17, 18},
        {1, 3, 9, 10, 17}, {2, 4, 9, 10, 19}, {5, 7, 11, 12, 18}, {6, 8, 11, 12, 20},
        {1, 2, 9, 13, 14}, {5, 6, 11, 13, 14}, {2, 6, 14, 19, 20}, {1, 5, 13, 17, 18}};
    double s3 = sqrt(3.0);
    for (int i = 0; i < 20; i++)
        for (int k = 0; k < 3; k++) vpts[3 * i + k] = nom[i][k] * radius / s3 + x0[k];
    if (!fpts) return;
    for (int i = 0; i < 12; i++) {
        for (int k = 0; k < 3; k++) fpts[3 * i + k] = 0.0;
        for (int j = 0; j < 5; j++)
            for (int k = 0; k < 3; k++) fpts[3 * i + k] = fpts[3 * i + k] + vpts[3 * (inds[i][j] - 1) + k];
    }
    double ovr5 = 1.0 / 5.0;
    for (int i = 0; i < 36; i++) fpts[i] = fpts[i] * ovr5;
}

/* ------------------------------------------------------------------ */
/* pointReduce.F90:1-83 (radius search inclusive, kd_tree2.f90:1537)  */
/* link is 1-based; returns nUnique                                   */
/* ------------------------------------------------------------------ */
static const double *g_sort_pts;
static int cmp_x(const void *a, const void *b) {
    double xa = g_sort_pts[3 * *(const int *)a], xb = g_sort_pts[3 * *(const int *)b];
    if (xa < xb) return -1;
    if (xa > xb) return 1;
    return (*(const int *)a) - (*(const int *)b);
}
int orc_point_reduce(const double *pts, int n, double tol, double *unique, int *link) {
    if (n == 0) return 0;
    int *ord = malloc(sizeof(int) * n), *pos = malloc(sizeof(int) * n), *found = malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) ord[i] = i;
    g_sort_pts = pts;
    qsort(ord, n, sizeof(int), cmp_x);
    for (int i = 0; i < n; i++) pos[ord[i]] = i;
    double tol2 = tol * tol;
    int nUnique = 0;
    for (int i = 0; i < n; i++) link[i] = 0;
    for (int i = 0; i < n; i++) {
        if (link[i] != 0) continue;
        int nf = 0;
        const double *p = pts + 3 * i;
        for (int dir = -1; dir <= 1; dir += 2) {
            for (int q = pos[i] + (dir > 0 ? 0 : -1); q >= 0 && q < n; q += dir) {
                int j = ord[q];
                double dx = pts[3 * j] - p[0];
                if (dx * dx > tol2) break; /* sd >= dx*dx, and dx grows along the sort */
                double dy = pts[3 * j + 1] - p[1], dz = pts[3 * j + 2] - p[2];
                if (dx * dx + dy * dy + dz * dz <= tol2) found[nf++] = j;
            }
        }
        nUnique++;
        for (int k = 0; k < 3; k++) unique[3 * (nUnique - 1) + k] = p[k];
        if (nf == 1) link[i] = nUnique;
        else
            for (int j = 0; j < nf; j++) link[found[j]] = nUnique;
    }
    free(ord); free(pos); free(found);
    return nUnique;
}

/* ------------------------------------------------------------------ */
/* kd-tree build (kd_tree.F90:123-300)                                */
/* ------------------------------------------------------------------ */
#define DAT(t, c, j) ((t)->the_data[3 * ((j) - 1) + (c) - 1]) /* the_data(c, j), 1-based */

static int new_node(octx *t) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = realloc(t->nodes, sizeof(onode) * t->cap);
    }
    memset(&t->nodes[t->nnodes], 0, sizeof(onode));
    t->nodes[t->nnodes].left = t->nodes[t->nnodes].right = -1;
    return t->nnodes++;
}

static double spread_in_coordinate(octx *t, int c, int l, int u) {
    int *ind = t->indexes - 1; /* 1-based view */
    double smin = DAT(t, c, ind[l]), smax = smin;
    int ulocal = u < l + 200 ? u : l + 200;
    for (int i = l + 1; i <= ulocal; i++) {
        double v = DAT(t, c, ind[i]);
        if (smin > v) smin = v;
        if (smax < v) smax = v;
    }
    return smax - smin;
}

static int most_spread_coordinate(octx *t, int l, int u) {
    int res = 0;
    double bsp = -1.0;
    for (int i = 1; i <= 3; i++) {
        double sp = spread_in_coordinate(t, i, l, u);
        if (sp > bsp) { res = i; bsp = sp; }
    }
    return res;
}

static void select_on_coordinate(octx *t, int c, int k, int li, int ui) {
    int *ind = t->indexes - 1;
    int l = li, u = ui;
    while (l < u) {
        int tt = ind[l], m = l;
        for (int i = l + 1; i <= u; i++) {
            if (DAT(t, c, ind[i]) < DAT(t, c, tt)) {
                m = m + 1;
                int s = ind[m]; ind[m] = ind[i]; ind[i] = s;
            }
        }
        int s = ind[l]; ind[l] = ind[m]; ind[m] = s;
        if (m <= k) l = m + 1;
        if (m >= k) u = m - 1;
    }
}

static int build_range(octx *t, int l, int u) {
    int id = new_node(t);
    if ((u - l) <= t->bucket) {
        onode *r = &t->nodes[id];
        r->dnum = 0; r->val = 0.0; r->l = l; r->u = u; r->n = u - l + 1;
    } else {
        int c = most_spread_coordinate(t, l, u);
        int m = (l + u) / 2;
        select_on_coordinate(t, c, m, l, u);
        t->nodes[id].dnum = c;
        t->nodes[id].val = DAT(t, c, t->indexes[m - 1]);
        t->nodes[id].l = l; t->nodes[id].u = u; t->nodes[id].n = u - l + 1;
        int left = build_range(t, l, m);
        int right = build_range(t, m + 1, u);
        t->nodes[id].left = left;
        t->nodes[id].right = right;
    }
    return id;
}

/* setCenters (kd_tree.F90:558-593) */
static void set_centers(octx *t, int id) {
    onode *np = &t->nodes[id];
    np->X[0] = np->X[1] = np->X[2] = 0.0;
    for (int i = np->l; i <= np->u; i++)
        for (int k = 0; k < 3; k++) np->X[k] = np->X[k] + DAT(t, k + 1, t->indexes[i - 1]);
    for (int k = 0; k < 3; k++) np->X[k] = np->X[k] / np->n;
    double radius = 0.0;
    for (int i = np->l; i <= np->u; i++) {
        double dx[3];
        for (int k = 0; k < 3; k++) dx[k] = DAT(t, k + 1, t->indexes[i - 1]) - np->X[k];
        double dist = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        if (dist > radius) radius = dist;
        np->oradius = 1.0 / radius;
    }
    if (np->dnum != 0) {
        set_centers(t, np->left);
        set_centers(t, t->nodes[id].right);
    }
}

/* setAi (kd_tree.F90:536-556) */
static void set_ai(octx *t, int id) {
    onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        np->Ai = 0.0;
        for (int i = np->l; i <= np->u; i++) np->Ai = np->Ai + t->Ai[i - 1];
    } else {
        set_ai(t, np->left);
        set_ai(t, np->right);
        np->Ai = (t->nodes[np->left].Ai + t->nodes[np->right].Ai);
    }
}

static inline double wfun(const octx *t, double ldefodist) {
    return pow(ldefodist, t->aExp) + t->alphaToBexp * pow(ldefodist, t->bExp);
}

/* computeErrors (kd_tree.F90:832-880) for one internal node; the
 * reference recurses over the tree, each node's table is independent. */
static void compute_errors_node(octx *t, int id) {
    onode *np = &t->nodes[id];
    if (np->dnum == 0) return;
    double vpts[60];
    for (int j = 0; j < NERR; j++) {
        double dExact[20] = {0}, dApprox[20] = {0};
        orc_sphere_pts(np->X, 1.0 / np->oradius * RSTAR[j], vpts, NULL);
        np->err[j] = 0.0;
        for (int ii = 0; ii < 20; ii++) {
            const double *v = vpts + 3 * ii;
            for (int i = np->l; i <= np->u; i++) {
                const double *x = t->Xu0 + 3 * (i - 1);
                double rr0 = v[0] - x[0], rr1 = v[1] - x[1], rr2 = v[2] - x[2];
                double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2);
                double lo = t->Ldef / dist;
                dExact[ii] = dExact[ii] + t->Ai[i - 1] * wfun(t, lo);
            }
            double rr0 = v[0] - np->X[0], rr1 = v[1] - np->X[1], rr2 = v[2] - np->X[2];
            double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2);
            double lo = t->Ldef / dist;
            dApprox[ii] = dApprox[ii] + np->Ai * wfun(t, lo);
            double e = dExact[ii] - dApprox[ii];
            np->err[j] = np->err[j] + e * e;
        }
        np->err[j] = sqrt(np->err[j] / 20);
    }
}

static int g_threads = 1;

static void compute_errors(octx *t) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int id = 0; id < t->nnodes; id++) compute_errors_node(t, id);
}

/* ------------------------------------------------------------------ */
/* computeNodalProperties (kd_tree.F90:1580-1664), determineCorners   */
/* (:1666-1715)                                                        */
/* ------------------------------------------------------------------ */
static int face_points(const octx *t, int ind, double *pts) {
    int npts = t->facePtr[ind] - t->facePtr[ind - 1];
    for (int kk = 1; kk <= npts; kk++) {
        int node = t->faceConn[t->facePtr[ind - 1] + kk - 1];
        for (int k = 0; k < 3; k++) pts[3 * (kk - 1) + k] = t->Xu[3 * (node - 1) + k];
    }
    return npts;
}

static void compute_nodal_properties(octx *t, const double *xs, int initial) {
    for (int i = 0; i < t->n; i++) {
        int j = t->XuInd[i];
        for (int k = 0; k < 3; k++) t->Xu[3 * i + k] = xs[3 * (j - 1) + k];
    }
    memset(t->normals, 0, sizeof(double) * 3 * t->n);
    memset(t->Mi, 0, sizeof(double) * 9 * t->n);
    memset(t->Bi, 0, sizeof(double) * 3 * t->n);
    double pts[3 * MAXFACEPTS];
    for (int i = 0; i < t->n; i++) {
        double sumArea = 0.0, sumNormal[3] = {0, 0, 0};
        for (int e = t->n2eStart[i]; e < t->n2eStart[i + 1]; e++) {
            int ind = t->n2e[e];
            int npts = face_points(t, ind, pts);
            double fa, fn[3];
            orc_element_props(pts, npts, &fa, fn);
            double dA = fa / (t->facePtr[ind] - t->facePtr[ind - 1]);
            sumArea = sumArea + dA;
            for (int k = 0; k < 3; k++) sumNormal[k] = sumNormal[k] + dA * fn[k];
        }
        for (int k = 0; k < 3; k++) t->normals[3 * i + k] = sumNormal[k] / sumArea;
        if (initial) {
            for (int k = 0; k < 3; k++) t->normals0[3 * i + k] = t->normals[3 * i + k];
            t->Ai[i] = sumArea;
        } else {
            double *M = t->Mi + 9 * i;
            if (t->opt.useRotations && !t->isCorner[i]) {
                orc_rotation_matrix(t->normals0 + 3 * i, t->normals + 3 * i, M);
            } else {
                memset(M, 0, sizeof(double) * 9);
                M[0] = M[4] = M[8] = 1.0;
            }
            const double *x0 = t->Xu0 + 3 * i;
            for (int r = 0; r < 3; r++)
                t->Bi[3 * i + r] = t->Xu[3 * i + r] - (M[r] * x0[0] + M[r + 3] * x0[1] + M[r + 6] * x0[2]);
        }
    }
    for (int i = 0; i < t->n; i++) {
        if (t->Ai[i] == 0.0) {
            double *M = t->Mi + 9 * i;
            t->normals[3 * i] = t->normals[3 * i + 1] = t->normals[3 * i + 2] = 0.0;
            t->Bi[3 * i] = t->Bi[3 * i + 1] = t->Bi[3 * i + 2] = 0.0;
            memset(M, 0, sizeof(double) * 9);
            M[0] = M[4] = M[8] = 1.0;
        }
    }
}

static void determine_corners(octx *t) {
    double threshold = cos(t->opt.cornerAngle * C_PI / 180.0);
    double pts[3 * MAXFACEPTS];
    for (int i = 0; i < t->n; i++) {
        t->isCorner[i] = 0;
        if (!t->opt.zeroCornerRotations) continue;
        for (int e = t->n2eStart[i]; e < t->n2eStart[i + 1]; e++) {
            int npts = face_points(t, t->n2e[e], pts);
            double fa, fn[3];
            orc_element_props(pts, npts, &fa, fn);
            const double *n0 = t->normals0 + 3 * i;
            double dp = fn[0] * n0[0] + fn[1] * n0[1] + fn[2] * n0[2];
            if (fabs(dp) < threshold) t->isCorner[i] = 1;
        }
    }
}

/* setData (kd_tree.F90:595-617) */
static void set_data(octx *t, int id) {
    onode *np = &t->nodes[id];
    if (np->dnum == 0) return;
    for (int k = 0; k < 3; k++) np->Bi[k] = 0.0;
    for (int k = 0; k < 9; k++) np->Mi[k] = 0.0;
    for (int i = np->l; i <= np->u; i++) {
        for (int k = 0; k < 3; k++) np->Bi[k] = np->Bi[k] + t->Bi[3 * (i - 1) + k];
        for (int k = 0; k < 9; k++) np->Mi[k] = np->Mi[k] + t->Mi[9 * (i - 1) + k];
    }
    for (int k = 0; k < 3; k++) np->Bi[k] = np->Bi[k] / np->n;
    for (int k = 0; k < 9; k++) np->Mi[k] = np->Mi[k] / np->n;
    set_data(t, np->left);
    set_data(t, t->nodes[id].right);
}

/* ------------------------------------------------------------------ */
/* Evaluators                                                         */
/* ------------------------------------------------------------------ */
typedef struct { double V, A, L, E; } ocount;

/* evalNodeExact (kd_tree.F90:664-690) */
static void eval_exact(const octx *t, const onode *np, const double *r, double *num, double *den) {
    for (int i = np->l; i <= np->u; i++) {
        const double *x0 = t->Xu0 + 3 * (i - 1), *M = t->Mi + 9 * (i - 1), *b = t->Bi + 3 * (i - 1);
        double rr0 = r[0] - x0[0], rr1 = r[1] - x0[1], rr2 = r[2] - x0[2];
        double lo = t->Ldef / sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + C_EPS);
        double Wi = t->Ai[i - 1] * wfun(t, lo);
        for (int k = 0; k < 3; k++) {
            double Si = M[k] * r[0] + M[k + 3] * r[1] + M[k + 6] * r[2] + b[k] - r[k];
            num[k] = num[k] + Wi * Si;
        }
        *den = *den + Wi;
    }
}

/* getError (kd_tree.F90:882-906) */
static double get_error(const octx *t, const onode *np, double dist) {
    double dOvrR = dist * np->oradius;
    double err = np->err[NERR - 1];
    for (int i = 0; i < 11; i++) {
        if (dOvrR < RSTAR[i + 1]) {
            double fact = (dOvrR - RSTAR[i]) * t->orstar[i];
            err = (1.0 - fact) * np->err[i] + fact * np->err[i + 1];
            break;
        }
    }
    return err;
}

/* evalNode + evalNodeApprox (kd_tree.F90:619-662, 692-713) */
static void eval_node(const octx *t, int id, const double *r, double *num, double *den, double approxDen, ocount *cnt) {
    const onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        if (cnt) cnt->L += np->n;
        eval_exact(t, np, r, num, den);
        return;
    }
    if (cnt) cnt->V += 1;
    double d0 = r[0] - np->X[0], d1 = r[1] - np->X[1], d2 = r[2] - np->X[2];
    double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (dist * np->oradius < 2.0) {
        eval_node(t, np->left, r, num, den, approxDen, cnt);
        eval_node(t, np->right, r, num, den, approxDen, cnt);
        return;
    }
    if (cnt) cnt->E += 1;
    double err = get_error(t, np, dist);
    if (err < t->errTol * approxDen) {
        if (cnt) cnt->A += 1;
        double lo = t->Ldef / dist;
        double Wi = np->Ai * wfun(t, lo);
        for (int k = 0; k < 3; k++) {
            double Si = np->Mi[k] * r[0] + np->Mi[k + 3] * r[1] + np->Mi[k + 6] * r[2] + np->Bi[k] - r[k];
            num[k] = num[k] + Wi * Si;
        }
        *den = *den + Wi;
    } else {
        eval_node(t, np->left, r, num, den, approxDen, cnt);
        eval_node(t, np->right, r, num, den, approxDen, cnt);
    }
}

/* dryRun (kd_tree.F90:789-830): operation-count estimate of one walk */
static void dry_run(const octx *t, int id, const double *r, double approxDen, double *c) {
    const onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        *c = *c + (double)np->n;
        return;
    }
    double rr0 = r[0] - np->X[0], rr1 = r[1] - np->X[1], rr2 = r[2] - np->X[2];
    double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + 1e-16);
    *c = *c + 0.5;
    if (dist * np->oradius < 2.0) {
        dry_run(t, np->left, r, approxDen, c);
        dry_run(t, np->right, r, approxDen, c);
        return;
    }
    double err = get_error(t, np, dist);
    if (err < t->errTol * approxDen) {
        *c = *c + 1.1;
    } else {
        dry_run(t, np->left, r, approxDen, c);
        dry_run(t, np->right, r, approxDen, c);
    }
}

/* getWiEstimate (kd_tree.F90:751-787) */
static void wi_estimate(const octx *t, int id, const double *r, double *den) {
    const onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        for (int i = np->l; i <= np->u; i++) {
            const double *x0 = t->Xu0 + 3 * (i - 1);
            double rr0 = r[0] - x0[0], rr1 = r[1] - x0[1], rr2 = r[2] - x0[2];
            double lo = t->Ldef / sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + 1e-16);
            *den = *den + t->Ai[i - 1] * wfun(t, lo);
        }
        return;
    }
    double rr0 = r[0] - np->X[0], rr1 = r[1] - np->X[1], rr2 = r[2] - np->X[2];
    double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + 1e-16);
    if (dist * np->oradius > 5.0) {
        double lo = t->Ldef / dist;
        *den = *den + np->Ai * wfun(t, lo);
    } else {
        wi_estimate(t, np->left, r, den);
        wi_estimate(t, np->right, r, den);
    }
}

/* Reverse evaluators (kd_tree.F90:942-1051).  Mib/Bib per point and per
 * node are passed explicitly so threads can privatise them. */
static void eval_exact_b(const octx *t, const onode *np, const double *r, const double *numb, double *Mib, double *Bib) {
    for (int i = np->l; i <= np->u; i++) {
        const double *x0 = t->Xu0 + 3 * (i - 1);
        double rr0 = r[0] - x0[0], rr1 = r[1] - x0[1], rr2 = r[2] - x0[2];
        double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + C_EPS);
        double lo = t->Ldef / dist;
        double wi = t->Ai[i - 1] * wfun(t, lo);
        double sib[3] = {wi * numb[0], wi * numb[1], wi * numb[2]};
        double *M = Mib + 9 * (i - 1), *B = Bib + 3 * (i - 1);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) M[k + 3 * c] = M[k + 3 * c] + r[c] * sib[k];
        for (int k = 0; k < 3; k++) B[k] = B[k] + sib[k];
    }
}

static void eval_node_b(const octx *t, int id, const double *r, const double *numb, double approxDen,
                        double *Mib, double *Bib, double *nMib, double *nBib) {
    const onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        eval_exact_b(t, np, r, numb, Mib, Bib);
        return;
    }
    double rr0 = r[0] - np->X[0], rr1 = r[1] - np->X[1], rr2 = r[2] - np->X[2];
    double dist = sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2);
    if (dist * np->oradius < 2.0) {
        eval_node_b(t, np->left, r, numb, approxDen, Mib, Bib, nMib, nBib);
        eval_node_b(t, np->right, r, numb, approxDen, Mib, Bib, nMib, nBib);
        return;
    }
    double err = get_error(t, np, dist);
    if (err < t->errTol * approxDen) {
        double lo = t->Ldef / dist;
        double wi = np->Ai * wfun(t, lo);
        double sib[3] = {wi * numb[0], wi * numb[1], wi * numb[2]};
        double *M = nMib + 9 * id, *B = nBib + 3 * id;
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) M[k + 3 * c] = M[k + 3 * c] + r[c] * sib[k];
        for (int k = 0; k < 3; k++) B[k] = B[k] + sib[k];
    } else {
        eval_node_b(t, np->left, r, numb, approxDen, Mib, Bib, nMib, nBib);
        eval_node_b(t, np->right, r, numb, approxDen, Mib, Bib, nMib, nBib);
    }
}

/* setData_b (kd_tree.F90:1033-1051) */
static void set_data_b(octx *t, int id) {
    onode *np = &t->nodes[id];
    if (np->dnum == 0) return;
    double ovrN = 1.0 / np->n;
    for (int i = np->l; i <= np->u; i++) {
        for (int k = 0; k < 3; k++) t->Bib[3 * (i - 1) + k] = t->Bib[3 * (i - 1) + k] + np->Bib[k] * ovrN;
        for (int k = 0; k < 9; k++) t->Mib[9 * (i - 1) + k] = t->Mib[9 * (i - 1) + k] + np->Mib[k] * ovrN;
    }
    set_data_b(t, np->left);
    set_data_b(t, t->nodes[id].right);
}

/* COMPUTENODALPROPERTIES_B (kd_tree.F90:1057-1162); adds into xsb. */
static void compute_nodal_properties_b(octx *t, const double *xs, double *xsb) {
    for (int i = 0; i < t->n; i++) {
        int j = t->XuInd[i];
        for (int k = 0; k < 3; k++) t->Xu[3 * i + k] = xs[3 * (j - 1) + k];
    }
    memset(t->normals, 0, sizeof(double) * 3 * t->n);
    memset(t->normalsb, 0, sizeof(double) * 3 * t->n);
    memset(t->Xub, 0, sizeof(double) * 3 * t->n);
    int maxe = 1;
    for (int i = 0; i < t->n; i++)
        if (t->n2eStart[i + 1] - t->n2eStart[i] > maxe) maxe = t->n2eStart[i + 1] - t->n2eStart[i];
    double (*pts)[3 * MAXFACEPTS] = malloc(sizeof(double) * 3 * MAXFACEPTS * maxe);
    double (*fnorm)[3] = malloc(sizeof(double) * 3 * maxe), *fda = malloc(sizeof(double) * maxe);
    int *fnp = malloc(sizeof(int) * maxe), *find = malloc(sizeof(int) * maxe);
    for (int i = 0; i < t->n; i++) {
        double sumarea = 0.0, sumnormal[3] = {0, 0, 0};
        int nelem = t->n2eStart[i + 1] - t->n2eStart[i];
        for (int jj = 0; jj < nelem; jj++) {
            int ind = t->n2e[t->n2eStart[i] + jj];
            int npts = face_points(t, ind, pts[jj]);
            double fa;
            orc_element_props(pts[jj], npts, &fa, fnorm[jj]);
            fda[jj] = fa / (t->facePtr[ind] - t->facePtr[ind - 1]);
            fnp[jj] = npts; find[jj] = ind;
            sumarea = sumarea + fda[jj];
            for (int k = 0; k < 3; k++) sumnormal[k] = sumnormal[k] + fda[jj] * fnorm[jj][k];
        }
        for (int k = 0; k < 3; k++) t->normals[3 * i + k] = sumnormal[k] / sumarea;
        double *xub = t->Xub + 3 * i, *bib = t->Bib + 3 * i, *mib = t->Mib + 9 * i;
        const double *x0 = t->Xu0 + 3 * i;
        for (int k = 0; k < 3; k++) xub[k] = xub[k] + bib[k];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) mib[k + 3 * c] = mib[k + 3 * c] - x0[c] * bib[k];
        bib[0] = bib[1] = bib[2] = 0.0;
        if (t->opt.useRotations && !t->isCorner[i])
            orc_rotation_matrix_b(t->normals0 + 3 * i, t->normals + 3 * i, mib, t->normalsb + 3 * i);
        memset(mib, 0, sizeof(double) * 9);
        double *nb = t->normalsb + 3 * i;
        double sumnormalb[3];
        for (int k = 0; k < 3; k++) sumnormalb[k] = nb[k] / sumarea;
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + (-(sumnormal[k] * nb[k] / sumarea));
        double sumareab = s / sumarea;
        nb[0] = nb[1] = nb[2] = 0.0;
        for (int jj = nelem - 1; jj >= 0; jj--) {
            double dab = sumareab + (fnorm[jj][0] * sumnormalb[0] + fnorm[jj][1] * sumnormalb[1] + fnorm[jj][2] * sumnormalb[2]);
            double fnb[3] = {fda[jj] * sumnormalb[0], fda[jj] * sumnormalb[1], fda[jj] * sumnormalb[2]};
            int ind = find[jj];
            double faceareab = dab / (t->facePtr[ind] - t->facePtr[ind - 1]);
            double pb[3 * MAXFACEPTS];
            memset(pb, 0, sizeof(pb));
            orc_element_props_b(pts[jj], fnp[jj], faceareab, fnb, pb);
            for (int kk = fnp[jj] - 1; kk >= 0; kk--) {
                int node = t->faceConn[t->facePtr[ind - 1] + kk];
                for (int k = 0; k < 3; k++) t->Xub[3 * (node - 1) + k] = t->Xub[3 * (node - 1) + k] + pb[3 * kk + k];
            }
        }
    }
    free(pts); free(fnorm); free(fda); free(fnp); free(find);
    for (int i = 0; i < t->n; i++) {
        int j = t->XuInd[i];
        for (int k = 0; k < 3; k++) xsb[3 * (j - 1) + k] = xsb[3 * (j - 1) + k] + t->Xub[3 * i + k];
    }
}

/* COMPUTENODALPROPERTIES_D (kd_tree.F90:1275-1360) for the non-initial
 * point: tangents Mid, Bid of Mi, Bi for a tangent xsd (3*npts) of the
 * surface; Xud = xsd at the first duplicate of each unique node. */
static void compute_nodal_properties_d(octx *t, const double *xs, const double *xsd) {
    for (int i = 0; i < t->n; i++) {
        int j = t->XuInd[i];
        for (int k = 0; k < 3; k++) {
            t->Xud[3 * i + k] = xsd[3 * (j - 1) + k];
            t->Xu[3 * i + k] = xs[3 * (j - 1) + k];
        }
    }
    double pts[3 * MAXFACEPTS], ptsd[3 * MAXFACEPTS];
    for (int i = 0; i < t->n; i++) {
        double sumarea = 0.0, sumnormal[3] = {0, 0, 0}, sumaread = 0.0, sumnormald[3] = {0, 0, 0};
        for (int e = t->n2eStart[i]; e < t->n2eStart[i + 1]; e++) {
            int ind = t->n2e[e];
            int npts = t->facePtr[ind] - t->facePtr[ind - 1];
            for (int kk = 0; kk < npts; kk++) {
                int node = t->faceConn[t->facePtr[ind - 1] + kk];
                for (int k = 0; k < 3; k++) {
                    ptsd[3 * kk + k] = t->Xud[3 * (node - 1) + k];
                    pts[3 * kk + k] = t->Xu[3 * (node - 1) + k];
                }
            }
            double fa, fad, fn[3], fnd[3];
            orc_element_props_d(pts, ptsd, npts, &fa, &fad, fn, fnd);
            double dad = fad / npts, da = fa / npts;
            sumaread = sumaread + dad;
            sumarea = sumarea + da;
            for (int k = 0; k < 3; k++) sumnormald[k] = sumnormald[k] + dad * fn[k] + da * fnd[k];
            for (int k = 0; k < 3; k++) sumnormal[k] = sumnormal[k] + da * fn[k];
        }
        double nd[3], nrm[3];
        for (int k = 0; k < 3; k++) nd[k] = (sumnormald[k] * sumarea - sumnormal[k] * sumaread) / (sumarea * sumarea);
        for (int k = 0; k < 3; k++) nrm[k] = sumnormal[k] / sumarea;
        double M[9], *Md = t->Mid + 9 * i;
        if (t->opt.useRotations && !t->isCorner[i]) {
            orc_rotation_matrix_d(t->normals0 + 3 * i, nrm, nd, M, Md);
        } else {
            memset(Md, 0, sizeof(double) * 9);
        }
        const double *x0 = t->Xu0 + 3 * i;
        for (int r = 0; r < 3; r++)
            t->Bid[3 * i + r] = t->Xud[3 * i + r] - x0[0] * Md[r] - x0[1] * Md[r + 3] - x0[2] * Md[r + 6];
    }
}

/* setData_d (kd_tree.F90:1247-1267) */
static void set_data_d(octx *t, int id) {
    onode *np = &t->nodes[id];
    if (np->dnum == 0) return;
    for (int k = 0; k < 3; k++) np->Bid[k] = 0.0;
    for (int k = 0; k < 9; k++) np->Mid[k] = 0.0;
    for (int i = np->l; i <= np->u; i++) {
        for (int k = 0; k < 3; k++) np->Bid[k] = np->Bid[k] + t->Bid[3 * (i - 1) + k];
        for (int k = 0; k < 9; k++) np->Mid[k] = np->Mid[k] + t->Mid[9 * (i - 1) + k];
    }
    for (int k = 0; k < 3; k++) np->Bid[k] = np->Bid[k] / np->n;
    for (int k = 0; k < 9; k++) np->Mid[k] = np->Mid[k] / np->n;
    set_data_d(t, np->left);
    set_data_d(t, t->nodes[id].right);
}

/* evalNodeExact_d (kd_tree.F90:1205-1228) */
static void eval_exact_d(const octx *t, const onode *np, const double *r, double *numd) {
    for (int i = np->l; i <= np->u; i++) {
        const double *x0 = t->Xu0 + 3 * (i - 1), *M = t->Mid + 9 * (i - 1), *b = t->Bid + 3 * (i - 1);
        double rr0 = r[0] - x0[0], rr1 = r[1] - x0[1], rr2 = r[2] - x0[2];
        double lo = t->Ldef / sqrt(rr0 * rr0 + rr1 * rr1 + rr2 * rr2 + C_EPS);
        double Wi = t->Ai[i - 1] * wfun(t, lo);
        for (int k = 0; k < 3; k++) {
            double Si = M[k] * r[0] + M[k + 3] * r[1] + M[k + 6] * r[2] + b[k];
            numd[k] = numd[k] + Wi * Si;
        }
    }
}

/* evalNode_d + evalNodeApprox_d (kd_tree.F90:1164-1203, 1230-1245) */
static void eval_node_d(const octx *t, int id, const double *r, double *numd, double approxDen) {
    const onode *np = &t->nodes[id];
    if (np->dnum == 0) {
        eval_exact_d(t, np, r, numd);
        return;
    }
    double d0 = r[0] - np->X[0], d1 = r[1] - np->X[1], d2 = r[2] - np->X[2];
    double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (dist * np->oradius < 2.0) {
        eval_node_d(t, np->left, r, numd, approxDen);
        eval_node_d(t, np->right, r, numd, approxDen);
        return;
    }
    double err = get_error(t, np, dist);
    if (err < t->errTol * approxDen) {
        double lo = t->Ldef / dist;
        double Wi = np->Ai * wfun(t, lo);
        for (int k = 0; k < 3; k++) {
            double Si = np->Mid[k] * r[0] + np->Mid[k + 3] * r[1] + np->Mid[k + 6] * r[2] + np->Bid[k];
            numd[k] = numd[k] + Wi * Si;
        }
    } else {
        eval_node_d(t, np->left, r, numd, approxDen);
        eval_node_d(t, np->right, r, numd, approxDen);
    }
}

/* ------------------------------------------------------------------ */
/* Public entry points                                                 */
/* ------------------------------------------------------------------ */
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* _setMeshOptions (UnstructuredMesh.py:1051-1069): only gridInput changes;
 * tree exponents, error tables and den0 keep their init-time values. */
void orc_set_options(void *h, const orc_options *o) { ((octx *)h)->opt = *o; }

void orc_free(void *h) {
    octx *t = h;
    if (!t) return;
    free(t->facePtr); free(t->faceConn); free(t->n2eStart); free(t->n2e); free(t->XuInd); free(t->indexes);
    free(t->the_data); free(t->Xu0); free(t->Xu); free(t->Ai); free(t->normals0); free(t->normals);
    free(t->Mi); free(t->Bi); free(t->Mib); free(t->Bib); free(t->Xub); free(t->normalsb); free(t->isCorner);
    free(t->Mid); free(t->Bid); free(t->Xud);
    free(t->nodes); free(t);
}

/* initializeWarping.F90:63-64 -> myCreateTree (kd_tree.F90:1365-1545).
 * pts: the initial surface (3*npts), uniquePts 3*nu (lexsorted), link
 * 1-based (npts), faceConn 0-based into pts. */
void *orc_create(const orc_options *opt, const double *pts, int npts, const double *uniquePts, int nu,
                 const int *link, const int *faceSizes, int nface, const int *faceConn, int nconn) {
    octx *t = calloc(1, sizeof(octx));
    t->opt = *opt;
    t->bucket = opt->bucketSize;
    t->nface = nface; t->nconn = nconn; t->nlink = npts; t->n = nu;
    t->facePtr = malloc(sizeof(int) * (nface + 1));
    t->facePtr[0] = 0;
    for (int i = 1; i <= nface; i++) t->facePtr[i] = faceSizes[i - 1] + t->facePtr[i - 1];
    int n = nu;
    t->the_data = malloc(sizeof(double) * 3 * n);
    memcpy(t->the_data, uniquePts, sizeof(double) * 3 * n);
    t->indexes = malloc(sizeof(int) * n);
    for (int j = 0; j < n; j++) t->indexes[j] = j + 1;
    build_range(t, 1, n);
    set_centers(t, 0);
    t->Xu0 = malloc(sizeof(double) * 3 * n);
    t->Xu = malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) t->Xu0[3 * i + k] = t->Xu[3 * i + k] = uniquePts[3 * (t->indexes[i] - 1) + k];
    int *inv = malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) inv[t->indexes[i] - 1] = i + 1;
    t->faceConn = malloc(sizeof(int) * nconn);
    for (int i = 0; i < nconn; i++) t->faceConn[i] = inv[link[faceConn[i]] - 1];
    /* nodeToElem as CSR, same order as the reference's fill loop */
    t->n2eStart = calloc(n + 1, sizeof(int));
    for (int f = 1; f <= nface; f++)
        for (int j = t->facePtr[f - 1]; j < t->facePtr[f]; j++) t->n2eStart[t->faceConn[j]]++;
    for (int i = 0; i < n; i++) t->n2eStart[i + 1] += t->n2eStart[i];
    int *fill = calloc(n, sizeof(int));
    t->n2e = malloc(sizeof(int) * (t->n2eStart[n] > 0 ? t->n2eStart[n] : 1));
    for (int f = 1; f <= nface; f++)
        for (int j = t->facePtr[f - 1]; j < t->facePtr[f]; j++) {
            int node = t->faceConn[j] - 1;
            t->n2e[t->n2eStart[node] + fill[node]++] = f;
        }
    free(fill);
    t->XuInd = calloc(n, sizeof(int));
    for (int i = 0; i < npts; i++)
        if (t->XuInd[inv[link[i] - 1] - 1] == 0) t->XuInd[inv[link[i] - 1] - 1] = i + 1;
    free(inv);
    double Xcen[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) Xcen[k] = Xcen[k] + t->Xu0[3 * i + k];
    for (int k = 0; k < 3; k++) Xcen[k] = Xcen[k] / n;
    t->Ldef0 = 0.0;
    for (int i = 0; i < n; i++) {
        double dx[3];
        for (int k = 0; k < 3; k++) dx[k] = t->the_data[3 * i + k] - Xcen[k];
        double d = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        if (d > t->Ldef0) t->Ldef0 = d;
    }
    t->Mi = calloc(9 * (size_t)n, sizeof(double));
    t->Bi = calloc(3 * (size_t)n, sizeof(double));
    t->Ai = calloc(n, sizeof(double));
    t->normals = calloc(3 * (size_t)n, sizeof(double));
    t->normals0 = calloc(3 * (size_t)n, sizeof(double));
    t->isCorner = calloc(n, sizeof(int));
    t->Mib = calloc(9 * (size_t)n, sizeof(double));
    t->Bib = calloc(3 * (size_t)n, sizeof(double));
    t->Xub = calloc(3 * (size_t)n, sizeof(double));
    t->normalsb = calloc(3 * (size_t)n, sizeof(double));
    t->Mid = calloc(9 * (size_t)n, sizeof(double));
    t->Bid = calloc(3 * (size_t)n, sizeof(double));
    t->Xud = calloc(3 * (size_t)n, sizeof(double));
    compute_nodal_properties(t, pts, 1);
    determine_corners(t);
    set_ai(t, 0);
    t->Ldef = t->Ldef0 * opt->LdefFact;
    t->alphaToBexp = pow(opt->alpha, opt->bExp);
    t->aExp = opt->aExp;
    t->bExp = opt->bExp;
    t->errTol = opt->errTol;
    for (int i = 0; i < NERR - 1; i++) t->orstar[i] = 1.0 / (RSTAR[i + 1] - RSTAR[i]);
    compute_errors(t);
    return t;
}

/* setSymmetryPlanes (initializeWarping.F90:511-531): n <= 6 planes given
 * by a point and a normal (used as given, as the reference does). */
int orc_set_symmetry_planes(void *h, const double *pts, const double *normals, int n) {
    octx *t = h;
    if (n < 0 || n > 6) return -1;
    t->nplanes = n;
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) {
            t->symmPts[j][k] = pts[3 * j + k];
            t->symmNormals[j][k] = normals[3 * j + k];
        }
    return 0;
}

/* denominator0 (initializeWarping.F90:240-247), summed over the mirror
 * passes into one running accumulator */
void orc_den0(void *h, const double *xv0, long nvol, double *den0) {
    octx *t = h;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (long j = 0; j < nvol; j++) {
        double d = 0.0, r[3];
        for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
            mirror_pt(t, xv0 + 3 * j, r, kk);
            wi_estimate(t, 0, r, &d);
        }
        den0[j] = d;
    }
}

static void warp_setup(octx *t, const double *xs) {
    compute_nodal_properties(t, xs, 0);
    t->Ldef = t->Ldef0 * t->opt.LdefFact;
    t->alphaToBexp = pow(t->opt.alpha, t->opt.bExp);
    t->errTol = t->opt.errTol;
    set_data(t, 0);
}

/* Wall seconds of the last orc_warp / orc_warp_deriv: [0] surface set-up
 * (nodal properties, setData, and for the reverse setData_b + nodal
 * reverse), [1] the loop over volume nodes. */
static double g_times[2];
void orc_last_times(double *out) { out[0] = g_times[0]; out[1] = g_times[1]; }
static double wall(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* warpMesh (warpMesh.F90:38-70): writes xv and the denominator. */
void orc_warp(void *h, const double *xs, const double *xv0, long nvol, const double *den0, double *xv, double *den) {
    octx *t = h;
    double t0 = wall();
    warp_setup(t, xs);
    double t1 = wall();
    int fast = t->opt.evalMode == 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (long j = 0; j < nvol; j++) {
        const double *r0 = xv0 + 3 * j;
        double numer[3] = {0, 0, 0}, dsum = 0.0;
        for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
            double r[3], num[3] = {0, 0, 0}, d = 0.0;
            mirror_pt(t, r0, r, kk);
            if (!fast) eval_exact(t, &t->nodes[0], r, num, &d);
            else eval_node(t, 0, r, num, &d, den0[j], NULL);
            mirror_num(t, num, kk);
            for (int k = 0; k < 3; k++) numer[k] = numer[k] + num[k];
            dsum = dsum + d;
        }
        double oden = 1.0 / dsum;
        for (int k = 0; k < 3; k++) xv[3 * j + k] = r0[k] + numer[k] * oden;
        den[j] = dsum;
    }
    g_times[0] = t1 - t0;
    g_times[1] = wall() - t1;
}

/* Load-balancing costs of initializeWarping.F90:258-265: per volume node,
 * the sum over mirror passes of dryRun / (unique surface nodes), with the
 * init-time errTol. */
void orc_costs(void *h, const double *xv0, long nvol, const double *den0, double *costs) {
    octx *t = h;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (long j = 0; j < nvol; j++) {
        double cost = 0.0;
        for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
            double r[3], c = 0.0;
            mirror_pt(t, xv0 + 3 * j, r, kk);
            dry_run(t, 0, r, den0[j], &c);
            cost = cost + c / t->n;
        }
        costs[j] = cost;
    }
}

/* Per-node work counts for the fast mode: V internal visits, A accepted
 * clusters, L leaf point pairs, E error checks. */
void orc_counts(void *h, const double *xs, const double *xv0, long nvol, const double *den0, double *out) {
    octx *t = h;
    warp_setup(t, xs);
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (long j = 0; j < nvol; j++) {
        ocount c = {0, 0, 0, 0};
        for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
            double r[3], num[3] = {0, 0, 0}, d = 0.0;
            mirror_pt(t, xv0 + 3 * j, r, kk);
            eval_node(t, 0, r, num, &d, den0[j], &c);
        }
        out[4 * j] = c.V; out[4 * j + 1] = c.A; out[4 * j + 2] = c.L; out[4 * j + 3] = c.E;
    }
}

/* warpDeriv (warpDeriv.F90:45-112): dxs (3*npts) = (dXv/dXs)^T dxv.  den is
 * the denominator of the last warp. */
void orc_warp_deriv(void *h, const double *xs, const double *xv0, long nvol, const double *den0, const double *den,
                    const double *dxv, double *dxs) {
    octx *t = h;
    double t0 = wall();
    warp_setup(t, xs);
    double t1 = wall();
    int n = t->n, nn = t->nnodes, fast = t->opt.evalMode == 1;
    int nth = g_threads;
    double **pM = malloc(sizeof(double *) * nth), **pB = malloc(sizeof(double *) * nth);
    double **qM = malloc(sizeof(double *) * nth), **qB = malloc(sizeof(double *) * nth);
    for (int k = 0; k < nth; k++) {
        pM[k] = calloc(9 * (size_t)n, sizeof(double)); pB[k] = calloc(3 * (size_t)n, sizeof(double));
        qM[k] = calloc(9 * (size_t)nn, sizeof(double)); qB[k] = calloc(3 * (size_t)nn, sizeof(double));
    }
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        long chunk = (nvol + nth - 1) / nth, j0 = tid * chunk, j1 = j0 + chunk < nvol ? j0 + chunk : nvol;
        for (long j = j0; j < j1; j++) {
            for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
                double r[3];
                mirror_pt(t, xv0 + 3 * j, r, kk);
                double oden = 1.0 / den[j];
                double numb[3] = {dxv[3 * j] * oden, dxv[3 * j + 1] * oden, dxv[3 * j + 2] * oden};
                mirror_num_b(t, numb, kk);
                if (!fast) eval_exact_b(t, &t->nodes[0], r, numb, pM[tid], pB[tid]);
                else eval_node_b(t, 0, r, numb, den0[j], pM[tid], pB[tid], qM[tid], qB[tid]);
            }
        }
    }
    memset(t->Mib, 0, sizeof(double) * 9 * n);
    memset(t->Bib, 0, sizeof(double) * 3 * n);
    for (int k = 0; k < nth; k++) {
        for (long i = 0; i < 9L * n; i++) t->Mib[i] += pM[k][i];
        for (long i = 0; i < 3L * n; i++) t->Bib[i] += pB[k][i];
    }
    for (int id = 0; id < nn; id++) {
        onode *np = &t->nodes[id];
        memset(np->Mib, 0, sizeof(np->Mib));
        memset(np->Bib, 0, sizeof(np->Bib));
        for (int k = 0; k < nth; k++) {
            for (int q = 0; q < 9; q++) np->Mib[q] += qM[k][9 * id + q];
            for (int q = 0; q < 3; q++) np->Bib[q] += qB[k][3 * id + q];
        }
    }
    for (int k = 0; k < nth; k++) { free(pM[k]); free(pB[k]); free(qM[k]); free(qB[k]); }
    free(pM); free(pB); free(qM); free(qB);
    double t2 = wall();
    set_data_b(t, 0);
    memset(dxs, 0, sizeof(double) * 3 * t->nlink);
    compute_nodal_properties_b(t, xs, dxs);
    g_times[0] = (t1 - t0) + (wall() - t2);
    g_times[1] = t2 - t1;
}

/* warpDerivFwd (warpDerivFwd.F90:62-99): xvd (3*nvol) = (dXv/dXs) xsd.
 * den is the denominator of the last warp. */
void orc_warp_deriv_fwd(void *h, const double *xs, const double *xsd, const double *xv0, long nvol,
                        const double *den0, const double *den, double *xvd) {
    octx *t = h;
    compute_nodal_properties_d(t, xs, xsd);
    t->Ldef = t->Ldef0 * t->opt.LdefFact;
    t->alphaToBexp = pow(t->opt.alpha, t->opt.bExp);
    t->errTol = t->opt.errTol;
    set_data_d(t, 0);
    int fast = t->opt.evalMode == 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (long j = 0; j < nvol; j++) {
        double numer[3] = {0, 0, 0};
        for (int kk = 1; kk <= (1 << t->nplanes); kk++) {
            double r[3], numd[3] = {0, 0, 0};
            mirror_pt(t, xv0 + 3 * j, r, kk);
            if (!fast) eval_exact_d(t, &t->nodes[0], r, numd);
            else eval_node_d(t, 0, r, numd, den0[j]);
            mirror_num_d(t, numd, kk);
            for (int k = 0; k < 3; k++) numer[k] = numer[k] + numd[k];
        }
        double oden = 1.0 / den[j];
        for (int k = 0; k < 3; k++) xvd[3 * j + k] = numer[k] * oden;
    }
}

/* ------------------------------------------------------------------ */
/* Introspection for the tests                                         */
/* ------------------------------------------------------------------ */
int orc_num_nodes(void *h) { return ((octx *)h)->nnodes; }
int orc_num_unique(void *h) { return ((octx *)h)->n; }
double orc_ldef0(void *h) { return ((octx *)h)->Ldef0; }

/* Nodal state after the last warp/init: Xu0 (3n), Ai (n), normals0 (3n),
 * Mi (9n), Bi (3n), isCorner (n), indexes (n, 1-based), XuInd (n, 1-based) */
void orc_get_surface(void *h, double *Xu0, double *Ai, double *normals0, double *Mi, double *Bi, int *isCorner,
                     int *indexes, int *XuInd) {
    octx *t = h;
    int n = t->n;
    if (Xu0) memcpy(Xu0, t->Xu0, sizeof(double) * 3 * n);
    if (Ai) memcpy(Ai, t->Ai, sizeof(double) * n);
    if (normals0) memcpy(normals0, t->normals0, sizeof(double) * 3 * n);
    if (Mi) memcpy(Mi, t->Mi, sizeof(double) * 9 * n);
    if (Bi) memcpy(Bi, t->Bi, sizeof(double) * 3 * n);
    if (isCorner) memcpy(isCorner, t->isCorner, sizeof(int) * n);
    if (indexes) memcpy(indexes, t->indexes, sizeof(int) * n);
    if (XuInd) memcpy(XuInd, t->XuInd, sizeof(int) * n);
}

/* Tree nodes in creation (DFS pre-) order: ints 6 per node (dnum, l, u, n,
 * left, right; children -1 for leaves), doubles 17 per node (X[3], oradius,
 * Ai, err[12]). */
void orc_get_tree(void *h, int *ints, double *dbl) {
    octx *t = h;
    for (int id = 0; id < t->nnodes; id++) {
        const onode *np = &t->nodes[id];
        int *ip = ints + 6 * id;
        ip[0] = np->dnum; ip[1] = np->l; ip[2] = np->u; ip[3] = np->n; ip[4] = np->left; ip[5] = np->right;
        double *dp = dbl + 17 * id;
        dp[0] = np->X[0]; dp[1] = np->X[1]; dp[2] = np->X[2]; dp[3] = np->oradius; dp[4] = np->Ai;
        for (int k = 0; k < NERR; k++) dp[5 + k] = np->err[k];
    }
}
