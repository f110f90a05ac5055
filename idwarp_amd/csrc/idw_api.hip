// Context management and C ABI of libidwarp_hip.so (declared in
// include/idwarp_hip.h).  One context holds one mesh instance: the volume
// nodes in traversal order, the surface tree and all per-call buffers, on one
// device and one stream.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "idw_internal.h"
#include "idw_kernels.h"

namespace {

thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(IDW_EHIP, "%s: %s", #x, hipGetErrorString(e_));  \
    } while (0)

// getSpherePts.F90:20-40 nominal vertices (phi = 2.1180339887498949,
// constants.F90:32, the reference's own constant).
void sphere_nominal(double *v) {
    const double o = 1.0, z = 0.0, p = 2.1180339887498949, ip = 1.0 / p;
    const double t[20][3] = {{-o, -o, -o}, {o, -o, -o}, {-o, o, -o}, {o, o, -o}, {-o, -o, o}, {o, -o, o}, {-o, o, o},
                             {o, o, o},    {z, -ip, -p}, {z, ip, -p}, {z, -ip, p}, {z, ip, p}, {-ip, -p, z}, {ip, -p, z},
                             {-ip, p, z},  {ip, p, z},  {-p, z, -ip}, {-p, z, ip}, {p, z, -ip}, {p, z, ip}};
    std::memcpy(v, t, sizeof(t));
}

// Index of the first non-finite value of a host array, or -1.
int64_t first_non_finite(const double *a, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(a[k])) return k;
    return -1;
}

bool same_numeric_options(const idw_options &a, const idw_options &b) {
    return a.LdefFact == b.LdefFact && a.alpha == b.alpha && a.aExp == b.aExp && a.bExp == b.bExp &&
           a.errTol == b.errTol && a.evalMode == b.evalMode && a.precision == b.precision &&
           a.nearWallFactor == b.nearWallFactor;
}

double plan_near_t2(const idw_ctx *c);
double plan_far_t2(const idw_ctx *c);

}  // namespace

struct idw_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    idw_options opt{};
    idw_options initOpt{};
    idw_options denOpt{};
    bool volSet = false, surfInit = false, denValid = false, haveDxs = false;
    idw::HostSurface hs;
    idw::DevSurface S;
    idw::DevVolume V;
    std::vector<idw::DevPlan> plans;  // one per symmetry-plane pass
    bool planValid = false;  // plans written by the last forward at denOpt
    double orderTol = -1.0, orderT2 = -1.0;  // key of the near-wall ordering of the volume nodes
    int nplanes = 0;  // symmetry planes; 2^nplanes passes (setSymmetryPlanes)
    double symmPts[idw::MAX_PLANES][3] = {}, symmNormals[idw::MAX_PLANES][3] = {};
    int64_t npts = 0;
    double *xv0w = nullptr, *xv = nullptr, *xs = nullptr, *dxs = nullptr, *dxv = nullptr;
    double *dxsIn = nullptr;  // host-API tangent seed (3 * npts)
    double *partial = nullptr;
    int nchunk = 0;
    double bboxLo[3] = {0, 0, 0}, bboxScale[3] = {0, 0, 0};  // quantisation of the volume box for the curve keys
    std::vector<void *> allocs;
    std::vector<size_t> allocBytes;  // the size of each of allocs
    int64_t devBytes = 0;            // their sum
    double *W = nullptr;   // exact mode, storeWeights: every pair's weight, [pass][surface point][sorted volume node]
    int64_t wCount = 0;    // weights it holds
    bool wValid = false;   // written by the last warp at the current options
    std::vector<int32_t> solverInd;
    hipEvent_t ev[8] = {};
    double ms[6] = {0, 0, 0, 0, 0, 0};
    bool fwdTimed = false, revTimed = false;
    bool capturing = false;  // work is being recorded into a HIP graph: no events, syncs or allocations
    bool batchPlan = false;  // a batch is running: keep the filled plans (the options keep reusePlan as set)
    // fill walks of the *_dev entry points report a plan that disagreed with
    // its counting pass through these pinned flags (one per mirror pass),
    // read at a later call once the event is reached
    unsigned int *mmHost = nullptr;
    hipEvent_t mmEvent = nullptr;
    bool mmPending = false;
    int64_t failAllocAfter = -1;  // test hook IDW_TEST_FAIL_ALLOC_AFTER: device allocations allowed before failing
    int64_t nAllocs = 0;
    // tile parameters of the volume kernels (idw_tune) and what they were
    // chosen for: evalMode, precision, errTol, near-wall distance, sizes, planes
    idw::Tuning tune;
    bool tuned = false;
    bool tuneFixed = false;  // set by idw_set_tuning: batches keep these parameters
    double tuneKey[7] = {-1, -1, -1, -1, -1, -1, -1};
    double tuneMs[idw::TUNE_SLOTS];
    double *detPart = nullptr;     // deterministic fast reverse: 12 sums per plan slot
    int64_t detPartCap = 0;        // slots it holds
    unsigned *detKeys = nullptr, *detKeys2 = nullptr;  // the plan transpose's sort: keys in and out,
    int32_t *detVals = nullptr;                        // slots in,
    unsigned char *detTmp = nullptr;                   // and the sort's temporary storage
    int64_t detSortCap = 0;
    size_t detTmpBytes = 0;
    int64_t *detCount = nullptr;      // chunks per record (nrec + 1, the last 0) before their scan
    int64_t detCountCap = 0;
    double *detChunkPart = nullptr;   // 12 partial sums per chunk
    int64_t detChunkCap = 0;
    hipEvent_t evOrder = nullptr;  // orders the context stream with a caller's stream (allocation-free entry point)
    double *scratch = nullptr;     // this device's copy of the allocation-free reverse's scratch array
    uint64_t gen = 0;        // bumped whenever device buffers, plans or options change (recorded graphs check it)
    std::string restartFile;        // denominator estimate + costs file (initializeWarping.F90:98-312)
    std::string restartStatus;      // "", "loaded", "saved"
    std::vector<double> costs;      // load-balancing costs of the last estimate (warp order)
};

namespace {

template <class T>
int dalloc(idw_ctx *c, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (c->failAllocAfter >= 0 && c->nAllocs >= c->failAllocAfter)
        return fail(IDW_ENOMEM, "hipMalloc(%zu bytes): failure injected by IDW_TEST_FAIL_ALLOC_AFTER=%lld",
                    count * sizeof(T), (long long)c->failAllocAfter);
    ++c->nAllocs;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        // reported here: clear it, or the next launch check would report it again
        (void)hipGetLastError();
        return fail(IDW_ENOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    c->allocs.push_back(q);
    c->allocBytes.push_back(count * sizeof(T));
    c->devBytes += (int64_t)(count * sizeof(T));
    *p = static_cast<T *>(q);
    ++c->gen;
    return IDW_OK;
}

template <class T>
void dfree(idw_ctx *c, T *&p) {
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)p);
    if (it != c->allocs.end()) {
        const size_t at = (size_t)(it - c->allocs.begin());
        c->devBytes -= (int64_t)c->allocBytes[at];
        c->allocBytes.erase(c->allocBytes.begin() + (std::ptrdiff_t)at);
        c->allocs.erase(it);
    }
    (void)hipFree(p);
    p = nullptr;
    ++c->gen;
}

// Timing events of a warp / reverse product (not while a graph is captured).
#define REC(c, i)                                                          \
    do {                                                                   \
        if (!(c)->capturing) HIPCHK(hipEventRecord((c)->ev[i], (c)->st)); \
    } while (0)

#define TRY(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_ != IDW_OK) return r_; \
    } while (0)

idw::WParams make_params(const idw_ctx *c, bool init) {
    const idw_options &o = init ? c->initOpt : c->opt;
    idw::WParams P{};
    P.Ldef = c->hs.ldef0 * o.LdefFact;
    P.a5 = std::pow(o.alpha, o.bExp);
    P.aExp = c->initOpt.aExp;
    P.bExp = c->initOpt.bExp;
    P.errTol = o.errTol;
    const double nw = o.nearWallFactor * P.Ldef;
    P.nearT2 = nw * nw;
    const double a = P.aExp, b = P.bExp;
    if (a == 3.0 && b == 5.0) {
        P.expMode = idw::EXP_35;
    } else if (a == std::floor(a) && b == std::floor(b) && a >= 0 && b >= 0 && a <= 64 && b <= 64) {
        P.expMode = idw::EXP_INT;
        P.ia = (int)a;
        P.ib = (int)b;
    } else {
        P.expMode = idw::EXP_REAL;
    }
    return P;
}

int ensure_device(idw_ctx *c) {
    HIPCHK(hipSetDevice(c->device));
    return IDW_OK;
}

void drop_plan(idw_ctx *c);

int nloop(const idw_ctx *c) { return 1 << c->nplanes; }

// Reflection of mirror pass kk (0-based; the reference's kk - 1).
idw::Mirror mirror_of(const idw_ctx *c, int kk) {
    idw::Mirror M;
    M.n = c->nplanes;
    M.mask = kk;
    for (int j = 0; j < c->nplanes; ++j)
        for (int k = 0; k < 3; ++k) {
            M.p[j][k] = c->symmPts[j][k];
            M.nrm[j][k] = c->symmNormals[j][k];
        }
    return M;
}

int pass_of(const idw_ctx *c, int kk) {
    if (nloop(c) == 1) return idw::PASS_ONLY;
    return kk == 0 ? idw::PASS_FIRST : (kk == nloop(c) - 1 ? idw::PASS_LAST : idw::PASS_MID);
}

// Accumulators of the mirror passes (only with symmetry planes), and in
// exact mode the packed seeded nodes of the reverse (positions, and per
// block of 256 nodes its offset, the total last).
int ensure_mirror_buffers(idw_ctx *c) {
    if (c->opt.evalMode == IDW_EVAL_EXACT && c->V.n && !c->V.cpos) {
        TRY(dalloc(c, &c->V.cpos, c->V.n));
        TRY(dalloc(c, &c->V.seedOff, (size_t)((c->V.n + 255) / 256) + 1));
    }
    if (nloop(c) == 1 || c->V.acc) return IDW_OK;
    TRY(dalloc(c, &c->V.acc, c->V.n));
    TRY(dalloc(c, &c->V.volm, c->V.n));
    return IDW_OK;
}

int restart_file(idw_ctx *c);

int compute_den0(idw_ctx *c) {
    if (!c->volSet || !c->surfInit || c->V.n == 0) return IDW_OK;
    for (int kk = 0; kk < nloop(c); ++kk) idw::launch_den0(c->st, c->S, c->V, make_params(c, true), mirror_of(c, kk));
    HIPCHK(hipGetLastError());
    if (!c->restartFile.empty()) TRY(restart_file(c));
    HIPCHK(hipStreamSynchronize(c->st));
    c->denValid = false;
    drop_plan(c);
    return IDW_OK;
}

// Restart file (initializeWarping.F90:98-312), one process: int32 node
// count, then the denominator estimate and the load-balancing costs, each
// nodes x float64 in the volume nodes' own order.  A file whose count matches
// and whose estimate agrees with the device's at every 100th node (relative
// 1e-10: the device's powers and sums round differently from the Fortran's)
// is loaded; otherwise the estimate is kept, the costs are computed and the
// file is written.
int restart_file(idw_ctx *c) {
    const int64_t n = c->V.n;
    std::vector<double4> vol(n);
    std::vector<int32_t> perm(n);
    HIPCHK(hipMemcpyAsync(vol.data(), c->V.vol, sizeof(double4) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(perm.data(), c->V.perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<double> den(n), fden(n), fcost(n);
    for (int64_t k = 0; k < n; ++k) den[perm[k]] = vol[k].w;
    bool good = false;
    if (FILE *f = std::fopen(c->restartFile.c_str(), "rb")) {
        int32_t cnt = -1;
        good = std::fread(&cnt, sizeof(cnt), 1, f) == 1 && cnt == n &&
               std::fread(fden.data(), sizeof(double), n, f) == (size_t)n &&
               std::fread(fcost.data(), sizeof(double), n, f) == (size_t)n;
        std::fclose(f);
        for (int64_t j = 0; good && j < n; j += 100)
            if (!(std::fabs(1.0 - den[j] / fden[j]) <= 1e-10)) good = false;
    }
    if (good) {
        for (int64_t k = 0; k < n; ++k) vol[k].w = fden[perm[k]];
        HIPCHK(hipMemcpyAsync(c->V.vol, vol.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->costs = std::move(fcost);
        c->restartStatus = "loaded";
        return IDW_OK;
    }
    double *d = nullptr;
    TRY(dalloc(c, &d, n));
    for (int kk = 0; kk < nloop(c); ++kk)
        idw::launch_dry_run(c->st, c->S, c->V, mirror_of(c, kk), c->initOpt.errTol, d);
    HIPCHK(hipGetLastError());
    std::vector<double> cost(n);
    HIPCHK(hipMemcpyAsync(cost.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c, d);
    c->costs.assign(n, 0.0);
    for (int64_t k = 0; k < n; ++k) c->costs[perm[k]] = cost[k];
    FILE *f = std::fopen(c->restartFile.c_str(), "wb");
    if (!f) return fail(IDW_EINVAL, "cannot write restart file %s", c->restartFile.c_str());
    const int32_t cnt = (int32_t)n;
    const bool ok = std::fwrite(&cnt, sizeof(cnt), 1, f) == 1 &&
                    std::fwrite(den.data(), sizeof(double), n, f) == (size_t)n &&
                    std::fwrite(c->costs.data(), sizeof(double), n, f) == (size_t)n;
    if (std::fclose(f) != 0 || !ok) return fail(IDW_EINVAL, "cannot write restart file %s", c->restartFile.c_str());
    c->restartStatus = "saved";
    return IDW_OK;
}

void drop_plan(idw_ctx *c) {
    for (idw::DevPlan &P : c->plans) {
        dfree(c, P.entries);
        dfree(c, P.detPos);
        dfree(c, P.recStart);
        dfree(c, P.chunkOff);
        dfree(c, P.chunkRec);
        P.chunkCap = P.nchunk = 0;
        dfree(c, P.waveOff);
        dfree(c, P.waveSplit);
        dfree(c, P.waveCount);
        dfree(c, P.groupSparse);
        dfree(c, P.mismatch);
    }
    c->plans.assign(nloop(c), idw::DevPlan());
    c->orderTol = c->orderT2 = -1.0;
    c->planValid = false;
    ++c->gen;
}

// Size the per-wave plan for the current errTol: one counting walk and an
// exclusive scan, redone only when errTol, the volume nodes or the surface
// change (like the denominator estimate, the counts depend on init-time data
// only).
double plan_near_t2(const idw_ctx *c) {
    if (c->opt.precision != IDW_PREC_MIXED) return 0.0;
    const double t = c->opt.nearWallFactor * c->hs.ldef0 * c->opt.LdefFact;
    return t * t;
}

// Squared distance beyond which mixed precision keeps a term in FP64 (0: no
// such test, FP64 precision).
double plan_far_t2(const idw_ctx *c) {
    if (c->opt.precision != IDW_PREC_MIXED) return 0.0;
    const double t = idw::FAR_FACTOR * c->hs.ldef0 * c->opt.LdefFact;
    return t * t;
}

// Mixed precision: move the volume nodes that need FP64 buckets behind the
// others (Hilbert order kept within both groups), so that the other waves
// run FP32 throughout.  Redone when errTol or the near-wall distance change.
int order_near_wall(idw_ctx *c) {
    const double nt2 = plan_near_t2(c);
    if (nt2 <= 0.0 || c->V.n == 0) return IDW_OK;
    if (c->orderTol == c->opt.errTol && c->orderT2 == nt2) return IDW_OK;
    const int64_t n = c->V.n;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    int32_t *ids = nullptr, *order = nullptr, *perm2 = nullptr;
    double4 *vol2 = nullptr;
    TRY(dalloc(c, &keys, n));
    TRY(dalloc(c, &keys2, n));
    TRY(dalloc(c, &ids, n));
    TRY(dalloc(c, &order, n));
    idw::launch_near_keys(c->st, c->S, c->V, c->opt.errTol, nt2, c->bboxLo, c->bboxScale, keys, ids);
    HIPCHK(hipGetLastError());
    HIPCHK(idw::sort_pairs(c->st, n, keys, ids, keys2, order));
    dfree(c, keys);
    dfree(c, keys2);
    dfree(c, ids);
    TRY(dalloc(c, &vol2, n));
    TRY(dalloc(c, &perm2, n));
    idw::launch_permute_vol(c->st, n, order, c->V.vol, c->V.perm, vol2, perm2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c, order);
    dfree(c, c->V.vol);
    dfree(c, c->V.perm);
    c->V.vol = vol2;
    c->V.perm = perm2;
    c->denValid = false;  // den is stored in traversal order
    c->orderTol = c->opt.errTol;
    c->orderT2 = nt2;
    for (idw::DevPlan &P : c->plans) P.sized = P.filled = false;
    return IDW_OK;
}

int size_one_plan(idw_ctx *c, int kk);

int size_plan(idw_ctx *c) {
    if ((int)c->plans.size() != nloop(c)) drop_plan(c);
    TRY(order_near_wall(c));
    for (int kk = 0; kk < nloop(c); ++kk) TRY(size_one_plan(c, kk));
    return IDW_OK;
}

int size_one_plan(idw_ctx *c, int kk) {
    idw::DevPlan &P = c->plans[kk];
    if (P.sized && P.errTol == c->opt.errTol) return IDW_OK;
    P.filled = false;
    const int64_t nwave = ((c->V.n + 255) / 256) * 4;
    if (!P.waveOff || P.nwave != nwave) {
        dfree(c, P.waveOff);
        dfree(c, P.waveSplit);
        dfree(c, P.waveCount);
        dfree(c, P.groupSparse);
        if (!P.mismatch) TRY(dalloc(c, &P.mismatch, 1));
        TRY(dalloc(c, &P.waveOff, nwave + 1));
        TRY(dalloc(c, &P.waveSplit, nwave));
        TRY(dalloc(c, &P.waveCount, nwave + 1));
        TRY(dalloc(c, &P.groupSparse, nwave));
        P.nwave = nwave;
    }
    HIPCHK(hipMemsetAsync(P.waveCount, 0, sizeof(int64_t) * (nwave + 1), c->st));
    if (c->V.n) idw::launch_plan_count(c->st, c->S, c->V, mirror_of(c, kk), c->opt.errTol, P.waveCount);
    HIPCHK(hipGetLastError());
    HIPCHK(idw::exclusive_scan(c->st, P.waveCount, P.waveOff, nwave + 1));
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, P.waveOff + nwave, sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (total > P.capacity) {
        dfree(c, P.entries);
        dfree(c, P.detPos);
        TRY(dalloc(c, &P.entries, (size_t)total));
        P.capacity = total;
    }
    P.transposed = false;
    P.total = total;
    P.errTol = c->opt.errTol;
    P.sized = true;
    return IDW_OK;
}

int check_plan(idw_ctx *c) {
    c->mmPending = false;  // superseded: the synchronous check below reads the device flags
    for (idw::DevPlan &P : c->plans) {
        if (!P.mismatch) continue;
        unsigned int mm = 0;
        HIPCHK(hipMemcpyAsync(&mm, P.mismatch, sizeof(mm), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (mm) {
            c->planValid = false;
            P.sized = false;
            return fail(IDW_EHIP, "fast-mode plan of %u waves disagreed with its counting pass", mm);
        }
    }
    return IDW_OK;
}

// The outcome of the fill walks of an earlier *_dev call, if they have
// completed: a plan that disagreed with its counting pass is dropped (the
// next warp sizes and fills it again) and reported by this call.
int poll_plan(idw_ctx *c) {
    if (!c->mmPending || c->capturing) return IDW_OK;
    const hipError_t q = hipEventQuery(c->mmEvent);
    if (q == hipErrorNotReady) return IDW_OK;
    c->mmPending = false;
    if (q != hipSuccess) return fail(IDW_EHIP, "hipEventQuery: %s", hipGetErrorString(q));
    for (size_t kk = 0; kk < c->plans.size() && kk < (size_t)(1 << idw::MAX_PLANES); ++kk)
        if (c->mmHost[kk]) {
            c->planValid = false;
            c->plans[kk].sized = false;
            return fail(IDW_EHIP, "the fast-mode plan written by an earlier warp disagreed with its counting pass "
                                  "in %u waves; that warp's results are invalid", c->mmHost[kk]);
        }
    return IDW_OK;
}

// Mixed precision: the near-wall distance below which pairs stay FP64 must
// keep the FP32 weights of every farther pair finite.
int check_mixed_range(const idw_ctx *c) {
    if (c->opt.precision != IDW_PREC_MIXED) return IDW_OK;
    const double f = c->opt.nearWallFactor;
    if (!(f >= idw::mixed_near_wall_floor(c->initOpt.aExp, c->initOpt.bExp, c->opt.alpha)))
        return fail(IDW_EINVAL, "nearWallFactor %g is below the mixed-precision floor %g for the tree's exponents", f,
                    idw::mixed_near_wall_floor(c->initOpt.aExp, c->initOpt.bExp, c->opt.alpha));
    if (!(f * c->hs.ldef0 * c->opt.LdefFact >= idw::MIN_NEAR_WALL_DISTANCE))
        return fail(IDW_EINVAL, "mixed precision needs nearWallFactor * Ldef >= %g (got %g)", idw::MIN_NEAR_WALL_DISTANCE,
                    f * c->hs.ldef0 * c->opt.LdefFact);
    return IDW_OK;
}

// Surface nodal properties and, in fast mode, the cluster data.
int surface_pass(idw_ctx *c, const double *xs, bool clusters) {
    idw::launch_face_props(c->st, c->S, xs);
    idw::launch_node_props(c->st, c->S, xs, 0, c->opt.useRotations);
    if (clusters) idw::launch_cluster_data(c->st, c->S);
    HIPCHK(hipGetLastError());
    return IDW_OK;
}

int run_warp(idw_ctx *c, const double *xs, double *xv) {
    if (!c->surfInit) return fail(IDW_ESTATE, "warp before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    const int fast = c->opt.evalMode == IDW_EVAL_FAST;
    TRY(poll_plan(c));
    TRY(check_mixed_range(c));
    idw::WParams P = make_params(c, false);
    if (fast) TRY(size_plan(c));
    TRY(ensure_mirror_buffers(c));
    REC(c, 0);
    TRY(surface_pass(c, xs, fast));
    REC(c, 1);
    // exact mode, storeWeights: room for every pair's weight of every pass
    // (the weights follow from the undeformed surface and volume and the
    // options, so any warp at the same options writes the same ones)
    const bool storeW = !fast && c->opt.storeWeights && c->V.n;
    c->wValid = false;
    if (storeW) {
        const int64_t need = (int64_t)nloop(c) * idw::wrow(c->V.n) * c->S.n;
        if (c->wCount != need) {
            if (c->capturing) return fail(IDW_ESTATE, "a recorded warp with stored weights needs an earlier one");
            dfree(c, c->W);
            c->wCount = 0;
            TRY(dalloc(c, &c->W, (size_t)need));
            c->wCount = need;
        }
    } else if (c->W && !c->capturing) {
        dfree(c, c->W);  // storeWeights is off: give the memory back
        c->wCount = 0;
    }
    std::vector<char> fill(nloop(c), 0);
    const double nt2 = fast ? plan_near_t2(c) : 0.0, ft2 = fast ? plan_far_t2(c) : 0.0;
    if (fast) {
        for (int kk = 0; kk < nloop(c); ++kk) {
            idw::DevPlan &pl = c->plans[kk];
            const bool current = pl.filled && pl.errTol == c->opt.errTol && pl.nearT2 == nt2 && pl.farT2 == ft2;
            if (!(c->opt.reusePlan || c->batchPlan) || !current) {
                if (c->capturing && !current)
                    return fail(IDW_ESTATE, "a captured warp needs the plan of a previous warp at the same options");
                fill[kk] = 1;
            }
        }
    }
    for (int kk = 0; kk < nloop(c); ++kk) {
        if (!fill[kk]) continue;
        idw::DevPlan &pl = c->plans[kk];
        idw::launch_zero_u32(c->st, pl.mismatch, 1);
        if (c->V.n) idw::launch_plan_fill(c->st, c->S, c->V, mirror_of(c, kk), c->opt.errTol, nt2, ft2, pl, c->tune);
        HIPCHK(hipGetLastError());
        if (!c->capturing && kk < (1 << idw::MAX_PLANES)) {
            HIPCHK(hipMemcpyAsync(c->mmHost + kk, pl.mismatch, sizeof(unsigned int), hipMemcpyDeviceToHost, c->st));
            c->mmPending = true;
        }
        pl.filled = true;
        pl.transposed = false;
        pl.nearT2 = nt2;
        pl.farT2 = ft2;
    }
    if (c->mmPending && !c->capturing) HIPCHK(hipEventRecord(c->mmEvent, c->st));
    REC(c, 2);
    for (int kk = 0; kk < nloop(c) && c->V.n; ++kk) {
        P.pass = pass_of(c, kk);
        idw::launch_forward(c->st, c->S, c->V, P, mirror_of(c, kk), fast, c->opt.precision, xv,
                            fast ? &c->plans[kk] : nullptr, c->tune, 0, -1,
                            storeW ? c->W + (int64_t)kk * idw::wrow(c->V.n) * c->S.n : nullptr);
    }
    HIPCHK(hipGetLastError());
    REC(c, 3);
    c->wValid = storeW;
    c->denOpt = c->opt;
    c->denValid = true;
    c->planValid = fast;
    c->fwdTimed = !c->capturing;
    return IDW_OK;
}

// Volume chunks of the exact reverse: enough block rows (chunks x 256-node
// surface blocks) to fill the device, each chunk at least 64 volume nodes.
int exact_rev_chunks(const idw_ctx *c, int shift) {
    const int64_t nbs = (c->S.n + 255) / 256;
    int nchunk = (int)std::min<int64_t>(128, std::max<int64_t>(1, (4096 + nbs - 1) / nbs));
    nchunk = shift > 0 ? std::min(256, nchunk << shift) : std::max(1, nchunk >> (-shift));
    while (nchunk > 1 && c->V.n / nchunk < 64) nchunk /= 2;
    return nchunk;
}

int exact_rev_chunks(const idw_ctx *c) { return exact_rev_chunks(c, c->tune.exactChunkShift); }

int ensure_partial(idw_ctx *c) {
    const int nchunk = exact_rev_chunks(c);
    if (c->partial && nchunk == c->nchunk) return IDW_OK;
    dfree(c, c->partial);
    TRY(dalloc(c, &c->partial, (size_t)nchunk * c->S.n * 12));
    c->nchunk = nchunk;
    return IDW_OK;
}

// Deterministic fast reverse: the plan's slots listed by term record (a
// stable sort keeps each record's slots in plan order) and the per-slot
// buffer, made when a filled plan is first reversed.
int ensure_transposed(idw_ctx *c, idw::DevPlan &P) {
    if (P.transposed) return IDW_OK;
    if (c->capturing) return fail(IDW_ESTATE, "a recorded deterministic reverse needs an earlier one of the same plan");
    if (P.total >= INT32_MAX) return fail(IDW_EINVAL, "deterministic reverse: %lld plan entries (fewer than 2^31)",
                                          (long long)P.total);
    const int nrec = c->S.nnodes + c->S.n;
    if (!P.detPos) TRY(dalloc(c, &P.detPos, (size_t)std::max<int64_t>(P.capacity, 1)));
    if (!P.recStart) TRY(dalloc(c, &P.recStart, (size_t)nrec + 1));
    // the sort's buffers are kept: a plan refilled at every warp is sorted
    // again at every reverse, without allocations after the first
    if (c->detSortCap < P.total) {
        dfree(c, c->detKeys);
        dfree(c, c->detKeys2);
        dfree(c, c->detVals);
        c->detSortCap = 0;
        TRY(dalloc(c, &c->detKeys, (size_t)P.total));
        TRY(dalloc(c, &c->detKeys2, (size_t)P.total));
        TRY(dalloc(c, &c->detVals, (size_t)P.total));
        c->detSortCap = P.total;
    }
    int bits = 1;  // keys run to nrec (slots without a record)
    while (bits < 32 && (1ll << bits) <= nrec) ++bits;
    if (P.total > 0) {
        size_t bytes = 0;
        HIPCHK(idw::sort_pairs32(c->st, nullptr, bytes, P.total, c->detKeys, c->detVals, c->detKeys2, P.detPos, bits));
        if (bytes > c->detTmpBytes) {
            dfree(c, c->detTmp);
            c->detTmpBytes = 0;
            TRY(dalloc(c, &c->detTmp, bytes));
            c->detTmpBytes = bytes;
        }
        idw::launch_plan_recs(c->st, P, nrec, c->detKeys, c->detVals);
        HIPCHK(hipGetLastError());
        bytes = c->detTmpBytes;
        HIPCHK(idw::sort_pairs32(c->st, c->detTmp, bytes, P.total, c->detKeys, c->detVals, c->detKeys2, P.detPos,
                                 bits));
    }
    idw::launch_rec_bounds(c->st, nrec, P.total, c->detKeys2, P.recStart);
    HIPCHK(hipGetLastError());
    // each record's range cut into chunks, and the record of every chunk
    if (!P.chunkOff) TRY(dalloc(c, &P.chunkOff, (size_t)nrec + 1));
    if (c->detCountCap < nrec + 1) {
        dfree(c, c->detCount);
        c->detCountCap = 0;
        TRY(dalloc(c, &c->detCount, (size_t)nrec + 1));
        c->detCountCap = nrec + 1;
    }
    HIPCHK(hipMemsetAsync(c->detCount + nrec, 0, sizeof(int64_t), c->st));
    idw::launch_rec_chunks(c->st, nrec, P.recStart, c->detCount);
    HIPCHK(hipGetLastError());
    HIPCHK(idw::exclusive_scan(c->st, c->detCount, P.chunkOff, nrec + 1));
    int64_t nch = 0;
    HIPCHK(hipMemcpy(&nch, P.chunkOff + nrec, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (nch > P.chunkCap) {
        dfree(c, P.chunkRec);
        P.chunkCap = 0;
        TRY(dalloc(c, &P.chunkRec, (size_t)nch));
        P.chunkCap = nch;
    }
    P.nchunk = nch;
    idw::launch_chunk_recs(c->st, nrec, P.chunkOff, P.chunkRec);
    HIPCHK(hipGetLastError());
    P.transposed = true;
    return IDW_OK;
}

int ensure_det_part(idw_ctx *c) {
    int64_t need = 1, nch = 1;
    for (const idw::DevPlan &P : c->plans) {
        need = std::max(need, P.total);
        nch = std::max(nch, P.nchunk);
    }
    if (c->capturing && !(c->detPart && c->detPartCap >= need && c->detChunkPart && c->detChunkCap >= nch))
        return fail(IDW_ESTATE, "a recorded deterministic reverse needs an earlier one of the same plan");
    if (!(c->detPart && c->detPartCap >= need)) {
        dfree(c, c->detPart);
        c->detPartCap = 0;
        TRY(dalloc(c, &c->detPart, (size_t)need * 12));
        c->detPartCap = need;
    }
    if (!(c->detChunkPart && c->detChunkCap >= nch)) {
        dfree(c, c->detChunkPart);
        c->detChunkCap = 0;
        TRY(dalloc(c, &c->detChunkPart, (size_t)nch * 12));
        c->detChunkCap = nch;
    }
    return IDW_OK;
}

bool det_ready(const idw_ctx *c) {
    int64_t need = 1, nch = 1;
    for (const idw::DevPlan &P : c->plans) {
        if (!P.transposed) return false;
        need = std::max(need, P.total);
        nch = std::max(nch, P.nchunk);
    }
    return c->detPart && c->detPartCap >= need && c->detChunkPart && c->detChunkCap >= nch;
}

// The kernels of a reverse product, given the exact mode's chunk partials
// (nchunk * n_unique * 12 doubles); enqueues work only.
int deriv_kernels(idw_ctx *c, const double *xs, const double *dxv, double *dxs, double *partial, int nchunk) {
    const int fast = c->opt.evalMode == IDW_EVAL_FAST;
    REC(c, 4);
    TRY(surface_pass(c, xs, false));
    idw::launch_reverse_prep(c->st, c->S, fast || c->V.n == 0);
    HIPCHK(hipGetLastError());
    REC(c, 5);
    const bool det = fast && c->opt.deterministic && c->V.n;
    if (det && !det_ready(c)) return fail(IDW_ESTATE, "the deterministic reverse's plan transpose is not built");
    // exact mode, storeWeights: the last warp's weights (at these options) instead of recomputed ones
    const bool useW = !fast && c->opt.storeWeights && c->wValid && c->V.n;
    for (int kk = 0; kk < nloop(c) && c->V.n; ++kk) {
        idw::launch_reverse(c->st, c->S, c->V, make_params(c, false), mirror_of(c, kk), fast, c->opt.precision,
                            partial, nchunk, dxv, fast ? &c->plans[kk] : nullptr, c->tune,
                            det ? c->detPart : nullptr,
                            useW ? c->W + (int64_t)kk * idw::wrow(c->V.n) * c->S.n : nullptr);
        // per term, the slots of this pass in plan order (stored by the first pass, added by later ones)
        if (det) idw::launch_det_reduce(c->st, c->S, c->plans[kk], c->detPart, c->detChunkPart, kk > 0);
    }
    HIPCHK(hipGetLastError());
    REC(c, 6);
    idw::launch_nodal_reverse(c->st, c->S, xs, fast, c->opt.useRotations, dxs, c->npts);
    HIPCHK(hipGetLastError());
    REC(c, 7);
    c->revTimed = !c->capturing;
    return IDW_OK;
}

bool deriv_needs_warp(const idw_ctx *c) {
    // the denominator (and the plan) depend on the options, not on the surface coordinates
    return !c->denValid || !same_numeric_options(c->denOpt, c->opt) ||
           (c->opt.evalMode == IDW_EVAL_FAST && !c->planValid) ||
           (c->opt.evalMode == IDW_EVAL_EXACT && c->opt.storeWeights && !c->wValid);
}

int run_deriv(idw_ctx *c, const double *xs, const double *dxv, double *dxs) {
    if (!c->surfInit) return fail(IDW_ESTATE, "warp_deriv before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    const int fast = c->opt.evalMode == IDW_EVAL_FAST;
    TRY(poll_plan(c));
    TRY(check_mixed_range(c));
    if (deriv_needs_warp(c)) TRY(run_warp(c, xs, c->xv));
    if (!fast) TRY(ensure_partial(c));
    TRY(ensure_mirror_buffers(c));
    if (fast && c->opt.deterministic && c->V.n) {
        for (idw::DevPlan &P : c->plans) TRY(ensure_transposed(c, P));
        TRY(ensure_det_part(c));
    }
    return deriv_kernels(c, xs, dxv, dxs, c->partial, c->nchunk);
}

// warpDerivFwd (warpDerivFwd.F90:62-99): dxv = (dXv/dXs) dxs.  The tangent
// point and cluster records replace the per-call records of the last warp
// (every warp and reverse product rebuilds them), and the forward kernels
// replay the fast mode's plan with them; the denominator is recomputed with
// the same weights in the same order as the forward's.
int run_deriv_fwd(idw_ctx *c, const double *xs, const double *dxs, double *dxv) {
    if (!c->surfInit) return fail(IDW_ESTATE, "warp_deriv_fwd before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    const int fast = c->opt.evalMode == IDW_EVAL_FAST;
    TRY(poll_plan(c));
    TRY(check_mixed_range(c));
    if (fast && (!c->planValid || !same_numeric_options(c->denOpt, c->opt))) TRY(run_warp(c, xs, c->xv));
    TRY(ensure_mirror_buffers(c));
    idw::WParams P = make_params(c, false);
    P.tangent = 1;
    REC(c, 0);
    idw::launch_props_tangent(c->st, c->S, xs, dxs, c->opt.useRotations);
    if (fast) idw::launch_cluster_data(c->st, c->S);
    HIPCHK(hipGetLastError());
    REC(c, 1);
    REC(c, 2);
    for (int kk = 0; kk < nloop(c) && c->V.n; ++kk) {
        P.pass = pass_of(c, kk);
        idw::launch_forward(c->st, c->S, c->V, P, mirror_of(c, kk), fast, c->opt.precision, dxv,
                            fast ? &c->plans[kk] : nullptr, c->tune);
    }
    HIPCHK(hipGetLastError());
    REC(c, 3);
    c->fwdTimed = !c->capturing;
    return IDW_OK;
}

// ---------------------------------------------------------------------------
// Autotuner of the volume kernels' tile parameters
// ---------------------------------------------------------------------------
// Each candidate of each parameter is timed on the context's own data at the
// current options (one launch to warm up, then the mean of two, HIP events on
// the context stream), with the other parameters at their current choice;
// the fastest is kept.  Parameters: the walk's gather buffer, the forward's
// chunk of staged terms, the mixed reverse's merge window (fast mode); the
// FP32 group and the reverse's volume chunks (exact mode).
void tune_key(const idw_ctx *c, double *k) {
    k[0] = c->opt.evalMode;
    k[1] = c->opt.precision;
    k[2] = c->opt.evalMode == IDW_EVAL_FAST ? c->opt.errTol : 0.0;
    k[3] = plan_near_t2(c);
    k[4] = (double)c->V.n;
    k[5] = (double)c->S.n;
    k[6] = c->nplanes;
}

template <class F>
int time_launches(idw_ctx *c, F &&launch, double *ms) {
    launch();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    for (int r = 0; r < 2; ++r) launch();
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    HIPCHK(hipEventSynchronize(c->ev[1]));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    *ms = t / 2.0;
    return IDW_OK;
}

int run_tune(idw_ctx *c) {
    if (!c->surfInit) return fail(IDW_ESTATE, "tuning before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    if (c->capturing) return fail(IDW_ESTATE, "tuning while recording a batch");
    for (double &t : c->tuneMs) t = -1.0;
    const int fast = c->opt.evalMode == IDW_EVAL_FAST, mixed = c->opt.precision == IDW_PREC_MIXED;
    // a warp at the current options sizes and fills the plans and gives the denominator
    TRY(run_warp(c, c->xs, c->xv));
    HIPCHK(hipStreamSynchronize(c->st));
    TRY(check_plan(c));
    if (c->V.n == 0) {
        c->tuned = true;
        tune_key(c, c->tuneKey);
        return IDW_OK;
    }
    idw::WParams P = make_params(c, false);
    auto pick = [&](int slot0, const int *cand, int ncand, int32_t &param, auto &&launch) -> int {
        int best = param;
        double bestMs = INFINITY;
        for (int k = 0; k < ncand; ++k) {
            param = cand[k];
            double ms = 0.0;
            TRY(time_launches(c, launch, &ms));
            c->tuneMs[slot0 + k] = ms;
            if (ms < bestMs) {
                bestMs = ms;
                best = cand[k];
            }
        }
        param = best;
        return IDW_OK;
    };
    auto forward = [&]() {
        for (int kk = 0; kk < nloop(c); ++kk) {
            P.pass = pass_of(c, kk);
            idw::launch_forward(c->st, c->S, c->V, P, mirror_of(c, kk), fast, c->opt.precision, c->xv,
                                fast ? &c->plans[kk] : nullptr, c->tune);
        }
    };
    if (fast) {
        const double nt2 = plan_near_t2(c), ft2 = plan_far_t2(c);
        static const int wbufs[3] = {64, 128, 256}, fchs[3] = {16, 32, 64}, mqs[3] = {8, 16, 32};
        TRY(pick(idw::TUNE_WBUF, wbufs, 3, c->tune.wbuf, [&]() {
            for (int kk = 0; kk < nloop(c); ++kk) {
                idw::launch_zero_u32(c->st, c->plans[kk].mismatch, 1);
                idw::launch_plan_fill(c->st, c->S, c->V, mirror_of(c, kk), c->opt.errTol, nt2, ft2, c->plans[kk],
                                      c->tune);
            }
        }));
        HIPCHK(hipStreamSynchronize(c->st));
        TRY(check_plan(c));
        TRY(pick(idw::TUNE_FCH, fchs, 3, c->tune.fch, forward));
        if (mixed) {
            // a seed of ones on every node (every term has work, as a real seed's)
            idw::launch_fill(c->st, c->dxv, 3 * c->V.n, 1.0);
            TRY(pick(idw::TUNE_MQ, mqs, 3, c->tune.mq, [&]() {
                idw::launch_reverse_prep(c->st, c->S, 1);
                for (int kk = 0; kk < nloop(c); ++kk)
                    idw::launch_reverse(c->st, c->S, c->V, P, mirror_of(c, kk), 1, c->opt.precision, nullptr, 0, c->dxv,
                                        &c->plans[kk], c->tune);
            }));
        }
    } else {
        static const int gs[3] = {16, 32, 64}, shifts[3] = {-1, 0, 1};
        if (mixed) TRY(pick(idw::TUNE_EXACT_G, gs, 3, c->tune.exactG, forward));
        // partial sums for the largest chunk count, freed afterwards
        double *partial = nullptr;
        TRY(dalloc(c, &partial, (size_t)exact_rev_chunks(c, 1) * c->S.n * 12));
        idw::launch_fill(c->st, c->dxv, 3 * c->V.n, 1.0);
        TRY(ensure_mirror_buffers(c));
        const int r = pick(idw::TUNE_EXACT_CHUNKS, shifts, 3, c->tune.exactChunkShift, [&]() {
            const int nchunk = exact_rev_chunks(c);
            for (int kk = 0; kk < nloop(c); ++kk)
                idw::launch_reverse(c->st, c->S, c->V, P, mirror_of(c, kk), 0, c->opt.precision, partial, nchunk,
                                    c->dxv, nullptr, c->tune);
        });
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(c, partial);
        TRY(r);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    // the last forward candidate wrote the warp of the current surface with
    // another parameter: warp again with the chosen ones
    TRY(run_warp(c, c->xs, c->xv));
    HIPCHK(hipStreamSynchronize(c->st));
    c->tuned = true;
    tune_key(c, c->tuneKey);
    ++c->gen;
    return check_plan(c);
}

bool tune_current(const idw_ctx *c) {
    if (!c->tuned) return false;
    double k[7];
    tune_key(c, k);
    for (int i = 0; i < 7; ++i)
        if (k[i] != c->tuneKey[i]) return false;
    return true;
}

void collect_timings(idw_ctx *c) {
    float t;
    if (c->fwdTimed) {
        for (int k = 0; k < 3; ++k)
            if (hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]) == hipSuccess) c->ms[k] = t;
    }
    if (c->revTimed) {
        for (int k = 0; k < 3; ++k)
            if (hipEventElapsedTime(&t, c->ev[4 + k], c->ev[5 + k]) == hipSuccess) c->ms[3 + k] = t;
    }
}

// Why the options are unusable, or nullptr.
const char *invalid_opts(const idw_options *o) {
    if (!o) return "null options";
    if (o->bucketSize < 1) return "bucketSize must be at least 1";
    if (!(o->evalMode == 0 || o->evalMode == 1)) return "evalMode must be exact (0) or fast (1)";
    if (!(o->precision == 0 || o->precision == 1)) return "precision must be fp64 (0) or mixed (1)";
    if (!(o->reusePlan == 0 || o->reusePlan == 1)) return "reusePlan must be 0 or 1";
    if (!(o->deterministic == 0 || o->deterministic == 1)) return "deterministic must be 0 or 1";
    if (!(o->storeWeights == 0 || o->storeWeights == 1)) return "storeWeights must be 0 or 1";
    if (o->storeWeights && o->evalMode != IDW_EVAL_EXACT) return "storeWeights needs evalMode exact";
    if (!std::isfinite(o->aExp) || !std::isfinite(o->bExp)) return "aExp and bExp must be finite";
    if (!(o->LdefFact > 0)) return "LdefFact must be positive";
    if (!(o->errTol >= 0)) return "errTol must be non-negative";
    if (!(o->nearWallFactor >= 0)) return "nearWallFactor must be non-negative";
    if (o->precision == IDW_PREC_MIXED &&
        !(o->nearWallFactor >= idw::mixed_near_wall_floor(o->aExp, o->bExp, o->alpha)))
        return "nearWallFactor is below the mixed-precision floor (FP32 weights of nearer pairs would overflow)";
    return nullptr;
}

}  // namespace

namespace idw {
hipError_t sort_pairs(hipStream_t st, int64_t n, uint64_t *keys, int32_t *ids, uint64_t *keys_out, int32_t *ids_out) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys_out, ids, ids_out, (int)n, 0, 63, st);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    e = hipMalloc(&tmp, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // returned to the caller, which reports it
        return e;
    }
    e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, keys_out, ids, ids_out, (int)n, 0, 63, st);
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    return e != hipSuccess ? e : e2;
}
}  // namespace idw

namespace idw {
hipError_t sort_pairs32(hipStream_t st, void *tmp, size_t &bytes, int64_t n, unsigned *keys, int32_t *vals,
                        unsigned *keys_out, int32_t *vals_out, int end_bit) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, keys_out, vals, vals_out, (int)n, 0, end_bit, st);
}

hipError_t exclusive_scan(hipStream_t st, const int64_t *in, int64_t *out, int64_t n) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    e = hipMalloc(&tmp, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // returned to the caller, which reports it
        return e;
    }
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st);
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    return e != hipSuccess ? e : e2;
}
}  // namespace idw

// ---------------------------------------------------------------------------
// Scratch of the allocation-free reverse product
// ---------------------------------------------------------------------------
// The exact mode's per-chunk partial sums of idw_warp_deriv_dev_noalloc live
// in this array, one copy per device, held until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): each device's address comes from
// hipGetSymbolAddress, taken once per device when its first context is
// created, and that address is what the kernels receive.
namespace {
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = int64_t(4) << 20;  // 32 MiB
constexpr int MAX_DEVICES = 64;
std::mutex g_scratchMu;
double *g_scratchPtr[MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
// With `device` current: this device's address of the scratch array.
int scratch_address(int device, double **out) {
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return fail(IDW_ENODEV, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(IDW_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratchPtr[device] = static_cast<double *>(p);
    }
    *out = g_scratchPtr[device];
    return IDW_OK;
}
}  // namespace

extern "C" {

void idw_default_options(idw_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->aExp = 3.0;
    o->bExp = 5.0;
    o->LdefFact = 1.0;
    o->alpha = 0.25;
    o->errTol = 5e-4;
    o->cornerAngle = 30.0;
    o->evalMode = IDW_EVAL_FAST;
    o->useRotations = 1;
    o->zeroCornerRotations = 1;
    o->bucketSize = 8;
    o->precision = IDW_PREC_FP64;
    o->reusePlan = 0;
    o->nearWallFactor = 1e-6;
}

const char *idw_last_error(void) { return t_err.c_str(); }

int idw_create(idw_ctx **out, int device, const idw_options *opt) {
    if (!out) return fail(IDW_EINVAL, "null output pointer");
    *out = nullptr;
    if (const char *why = invalid_opts(opt)) return fail(IDW_EINVAL, "invalid options: %s", why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(IDW_ENODEV, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(IDW_ENODEV, "device %d out of range (%d devices)", device, ndev);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(IDW_ENODEV, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    idw_ctx *c = new (std::nothrow) idw_ctx();
    if (!c) return fail(IDW_ENOMEM, "out of host memory");
    c->device = device;
    c->opt = *opt;
    for (double &t : c->tuneMs) t = -1.0;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    for (int k = 0; k < 8 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->mmEvent, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evOrder, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->mmHost, sizeof(unsigned int) << idw::MAX_PLANES);
    if (e == hipSuccess) std::memset(c->mmHost, 0, sizeof(unsigned int) << idw::MAX_PLANES);
    if (const char *fa = std::getenv("IDW_TEST_FAIL_ALLOC_AFTER")) c->failAllocAfter = std::atoll(fa);
    if (e != hipSuccess) {
        idw_destroy(c);
        return fail(IDW_EHIP, "stream/event creation: %s", hipGetErrorString(e));
    }
    if (const int r = scratch_address(device, &c->scratch); r != IDW_OK) {
        const std::string msg = t_err;
        idw_destroy(c);
        return fail(r, "%s", msg.c_str());
    }
    *out = c;
    return IDW_OK;
}

int idw_destroy(idw_ctx *c) {
    if (!c) return IDW_OK;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->allocs) (void)hipFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->mmEvent) (void)hipEventDestroy(c->mmEvent);
    if (c->evOrder) (void)hipEventDestroy(c->evOrder);
    if (c->mmHost) (void)hipHostFree(c->mmHost);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return IDW_OK;
}

int idw_set_options(idw_ctx *c, const idw_options *opt) {
    if (!c) return fail(IDW_EINVAL, "null context");
    if (const char *why = invalid_opts(opt)) return fail(IDW_EINVAL, "invalid options: %s", why);
    if (std::memcmp(&c->opt, opt, sizeof(idw_options)) != 0) ++c->gen;
    c->opt = *opt;
    return IDW_OK;
}

int idw_set_volume_nodes(idw_ctx *c, const double *xv0, int64_t ndof) {
    if (!c || (!xv0 && ndof)) return fail(IDW_EINVAL, "null argument");
    if (ndof < 0 || ndof % 3) return fail(IDW_EINVAL, "volume DOF count %lld is not a multiple of 3", (long long)ndof);
    const int64_t n = ndof / 3;
    if (n >= INT32_MAX / 2) return fail(IDW_EINVAL, "too many volume nodes for 32-bit indexing");
    if (const int64_t bad = first_non_finite(xv0, ndof); bad >= 0)
        return fail(IDW_EINVAL, "volume coordinate %lld is not finite", (long long)bad);
    TRY(ensure_device(c));
    dfree(c, c->xv0w); dfree(c, c->xv); dfree(c, c->dxv);
    dfree(c, c->V.vol); dfree(c, c->V.perm); dfree(c, c->V.den); dfree(c, c->V.numb);
    dfree(c, c->V.acc); dfree(c, c->V.volm); dfree(c, c->V.cpos); dfree(c, c->V.seedOff);
    dfree(c, c->partial);
    dfree(c, c->W);
    c->wCount = 0;
    c->volSet = false;
    c->denValid = false;
    drop_plan(c);
    c->V.n = n;
    TRY(dalloc(c, &c->xv0w, 3 * n));
    TRY(dalloc(c, &c->xv, 3 * n));
    TRY(dalloc(c, &c->dxv, 3 * n));
    TRY(dalloc(c, &c->V.vol, n));
    TRY(dalloc(c, &c->V.perm, n));
    TRY(dalloc(c, &c->V.den, n));
    TRY(dalloc(c, &c->V.numb, n));
    if (n) {
        HIPCHK(hipMemcpyAsync(c->xv0w, xv0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->xv, c->xv0w, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
        // Hilbert order of the volume nodes over their bounding box
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t j = 0; j < n; ++j)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], xv0[3 * j + k]);
                hi[k] = std::max(hi[k], xv0[3 * j + k]);
            }
        double span = 0.0;
        for (int k = 0; k < 3; ++k) span = std::max(span, hi[k] - lo[k]);
        double scale[3];
        for (int k = 0; k < 3; ++k) scale[k] = span > 0 ? 2097151.0 / span : 0.0;
        for (int k = 0; k < 3; ++k) {
            c->bboxLo[k] = lo[k];
            c->bboxScale[k] = scale[k];
        }
        uint64_t *keys = nullptr, *keys2 = nullptr;
        int32_t *ids = nullptr;
        TRY(dalloc(c, &keys, n));
        TRY(dalloc(c, &keys2, n));
        TRY(dalloc(c, &ids, n));
        idw::launch_curve_keys(c->st, n, c->xv0w, lo, scale, keys, ids);
        HIPCHK(hipGetLastError());
        HIPCHK(idw::sort_pairs(c->st, n, keys, ids, keys2, c->V.perm));
        dfree(c, keys);
        dfree(c, keys2);
        dfree(c, ids);
        idw::launch_build_vol(c->st, n, c->xv0w, c->V.perm, c->V.vol);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
    }
    c->volSet = true;
    return compute_den0(c);
}

int idw_set_symmetry_planes(idw_ctx *c, const double *pts, const double *normals, int64_t n) {
    if (!c || (n && (!pts || !normals))) return fail(IDW_EINVAL, "null argument");
    if (n < 0 || n > idw::MAX_PLANES) return fail(IDW_EINVAL, "%lld symmetry planes (at most %d)", (long long)n,
                                                  idw::MAX_PLANES);
    TRY(ensure_device(c));
    HIPCHK(hipStreamSynchronize(c->st));
    drop_plan(c);
    dfree(c, c->V.acc);
    dfree(c, c->V.volm);
    c->nplanes = (int)n;
    for (int j = 0; j < c->nplanes; ++j)
        for (int k = 0; k < 3; ++k) {
            c->symmPts[j][k] = pts[3 * j + k];
            c->symmNormals[j][k] = normals[3 * j + k];
        }
    drop_plan(c);
    c->denValid = false;
    // the denominator estimate sums over the mirror passes
    return compute_den0(c);
}

int idw_set_restart_file(idw_ctx *c, const char *path) {
    if (!c) return fail(IDW_EINVAL, "null context");
    c->restartFile = path ? path : "";
    c->restartStatus.clear();
    return IDW_OK;
}

int idw_get_restart_info(idw_ctx *c, int32_t *status, double *costs, int64_t n) {
    if (!c) return fail(IDW_EINVAL, "null context");
    if (status) *status = c->restartStatus == "loaded" ? 1 : (c->restartStatus == "saved" ? 2 : 0);
    if (costs) {
        if (n != (int64_t)c->costs.size()) return fail(IDW_EINVAL, "expected %lld costs", (long long)c->costs.size());
        std::copy(c->costs.begin(), c->costs.end(), costs);
    }
    return IDW_OK;
}

double idw_mixed_near_wall_floor(double aExp, double bExp, double alpha) {
    return idw::mixed_near_wall_floor(aExp, bExp, alpha);
}

int idw_point_reduce(const double *pts, int64_t n, double tol, double *unique_pts, int32_t *link, int64_t *n_unique) {
    if (n < 0 || (n && (!pts || !unique_pts || !link)) || !n_unique) return fail(IDW_EINVAL, "bad argument");
    if (n >= INT32_MAX) return fail(IDW_EINVAL, "too many points");
    *n_unique = idw::point_reduce(pts, n, tol, unique_pts, link);
    return IDW_OK;
}

int idw_initialize_warping(idw_ctx *c, const double *pts, int64_t n_pts, const double *unique_pts, int64_t n_unique,
                           const int32_t *link, const int32_t *face_sizes, int64_t n_faces, const int32_t *face_conn,
                           int64_t n_conn) {
    if (!c || !pts || !unique_pts || !link || (n_faces && (!face_sizes || !face_conn)))
        return fail(IDW_EINVAL, "null argument");
    if (n_pts <= 0 || n_unique <= 0 || n_unique > n_pts) return fail(IDW_EINVAL, "bad point counts");
    if (const int64_t bad = first_non_finite(pts, 3 * n_pts); bad >= 0)
        return fail(IDW_EINVAL, "surface coordinate %lld is not finite", (long long)bad);
    if (first_non_finite(unique_pts, 3 * n_unique) >= 0) return fail(IDW_EINVAL, "unique surface points are not finite");
    TRY(ensure_device(c));
    std::string msg;
    idw::HostSurface hs;
    if (!idw::build_surface(unique_pts, n_unique, link, n_pts, face_sizes, n_faces, face_conn, n_conn,
                            c->opt.bucketSize, hs, msg))
        return fail(IDW_EINVAL, "%s", msg.c_str());
    if (!idw::walk_offsets_fit((int64_t)hs.nodes.size()))
        return fail(IDW_EINVAL, "kd-tree of %lld nodes exceeds the walk's %lld-node limit (32-bit header offsets)",
                    (long long)hs.nodes.size(), (long long)idw::MAX_WALK_NODES);
    // free a previous surface
    idw::DevSurface &S = c->S;
    void *olds[] = {S.nodes, S.whdr, S.rec, S.rec32, S.internal, S.cut, S.topOrder, S.csum, S.ps, S.normals0, S.nprops, S.corner, S.xuInd,
                    S.facePtr, S.faceConn, S.n2eStart, S.n2e, S.slotStart, S.slots, S.fprops, S.adj,
                    S.ntmp, S.pb, S.fpropsd, c->xs, c->dxs, c->dxsIn, c->partial};
    for (void *p : olds) {
        void *q = p;
        dfree(c, q);
    }
    S = idw::DevSurface();
    c->xs = c->dxs = c->dxsIn = c->partial = nullptr;
    dfree(c, c->W);
    c->wCount = 0;
    c->wValid = false;
    c->surfInit = false;
    c->denValid = false;
    drop_plan(c);
    c->hs = std::move(hs);
    c->initOpt = c->opt;
    c->npts = n_pts;
    const idw::HostSurface &H = c->hs;
    S.n = (int)H.n;
    S.nnodes = (int)H.nodes.size();
    S.nInternal = H.nInternal;
    S.nface = (int)H.nface;
    TRY(dalloc(c, &S.nodes, S.nnodes));
    TRY(dalloc(c, &S.whdr, (size_t)S.nnodes + 1));
    TRY(dalloc(c, &S.rec, (size_t)S.nnodes + S.n));
    TRY(dalloc(c, &S.rec32, (size_t)S.nnodes + S.n));
    S.pd = S.rec + S.nnodes;
    S.pd32 = S.rec32 + S.nnodes;
    TRY(dalloc(c, &S.internal, S.nInternal));
    TRY(dalloc(c, &S.csum, (size_t)S.nnodes * idw::CSUM));
    TRY(dalloc(c, &S.ps, S.n));
    TRY(dalloc(c, &S.normals0, 3 * (size_t)S.n));
    TRY(dalloc(c, &S.nprops, S.n));
    TRY(dalloc(c, &S.corner, S.n));
    TRY(dalloc(c, &S.xuInd, S.n));
    TRY(dalloc(c, &S.facePtr, H.nface + 1));
    TRY(dalloc(c, &S.faceConn, H.nconn));
    TRY(dalloc(c, &S.n2eStart, S.n + 1));
    TRY(dalloc(c, &S.n2e, H.nconn));
    TRY(dalloc(c, &S.slotStart, S.n + 1));
    TRY(dalloc(c, &S.slots, H.nconn));
    TRY(dalloc(c, &S.fprops, H.nface));
    TRY(dalloc(c, &S.fpropsd, H.nface));
    TRY(dalloc(c, &S.adj, ((size_t)S.nnodes + S.n) * idw::ADJ_STRIDE));
    S.nadj = S.adj;
    S.padj = S.adj + (size_t)S.nnodes * idw::ADJ_STRIDE;
    TRY(dalloc(c, &S.ntmp, (size_t)S.n * 8));
    TRY(dalloc(c, &S.pb, 3 * (size_t)H.nconn));
    TRY(dalloc(c, &c->xs, 3 * (size_t)n_pts));
    TRY(dalloc(c, &c->dxs, 3 * (size_t)n_pts));
    TRY(dalloc(c, &c->dxsIn, 3 * (size_t)n_pts));
    std::vector<int32_t> internal;
    for (int k = 0; k < S.nnodes; ++k)
        if (!H.nodes[k].leaf) internal.push_back(k);
    // bottom-up order of the cluster data: the subtrees rooted at the cut
    // depth (and the buckets above it), each with its deepest depth, then
    // the internal nodes above the cut by decreasing depth
    int maxDepth = 0;
    for (const auto &nd : H.nodes) maxDepth = std::max(maxDepth, nd.depth);
    if (maxDepth >= idw::MAX_LEVELS) return fail(IDW_EINVAL, "kd-tree depth %d exceeds %d", maxDepth, idw::MAX_LEVELS - 1);
    const int cutDepth = std::min(maxDepth, idw::CUT_DEPTH);
    std::vector<int32_t> cut, topOrder;
    for (int k = 0; k < S.nnodes; ++k) {
        const auto &nd = H.nodes[k];
        if (nd.depth == cutDepth || (nd.leaf && nd.depth < cutDepth)) {
            int deepest = nd.depth;
            for (int j = k; j < nd.skip; ++j) deepest = std::max(deepest, H.nodes[j].depth);
            cut.push_back(k);
            cut.push_back(deepest);
        }
    }
    S.nCut = (int)cut.size() / 2;
    for (int d = cutDepth - 1; d >= 0 && S.nInternal; --d) {
        const int start = (int)topOrder.size();
        for (int k : internal)
            if (H.nodes[k].depth == d) topOrder.push_back(k);
        S.top.start[S.top.n] = start;
        S.top.count[S.top.n] = (int)topOrder.size() - start;
        ++S.top.n;
    }
    TRY(dalloc(c, &S.cut, cut.size()));
    if (!topOrder.empty()) TRY(dalloc(c, &S.topOrder, topOrder.size()));
    std::vector<double4> ps(S.n);
    for (int i = 0; i < S.n; ++i) ps[i] = make_double4(H.xu0[3 * i], H.xu0[3 * i + 1], H.xu0[3 * i + 2], 0.0);
    hipStream_t st = c->st;
    HIPCHK(hipMemcpyAsync(S.internal, internal.data(), sizeof(int32_t) * internal.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S.cut, cut.data(), sizeof(int32_t) * cut.size(), hipMemcpyHostToDevice, st));
    if (!topOrder.empty())
        HIPCHK(hipMemcpyAsync(S.topOrder, topOrder.data(), sizeof(int32_t) * topOrder.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S.ps, ps.data(), sizeof(double4) * S.n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S.xuInd, H.xuInd.data(), sizeof(int32_t) * S.n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S.facePtr, H.facePtr.data(), sizeof(int32_t) * (H.nface + 1), hipMemcpyHostToDevice, st));
    if (H.nconn) {
        HIPCHK(hipMemcpyAsync(S.faceConn, H.faceConn.data(), sizeof(int32_t) * H.nconn, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.n2e, H.n2e.data(), sizeof(int32_t) * H.nconn, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S.slots, H.slots.data(), sizeof(int32_t) * H.nconn, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(S.n2eStart, H.n2eStart.data(), sizeof(int32_t) * (S.n + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S.slotStart, H.slotStart.data(), sizeof(int32_t) * (S.n + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->xs, pts, sizeof(double) * 3 * n_pts, hipMemcpyHostToDevice, st));
    // initial nodal properties: normals0, A_i (initNodalProperties, kd_tree.F90:1547-1578)
    idw::launch_face_props(st, S, c->xs);
    idw::launch_node_props(st, S, c->xs, 1, 0);
    const double PI = 3.1415926535897931;
    const double thr = c->opt.zeroCornerRotations ? std::cos(c->opt.cornerAngle * PI / 180.0) : -1.0;
    idw::launch_corners(st, S, thr);
    HIPCHK(hipGetLastError());
    std::vector<double4> ps2(S.n);
    HIPCHK(hipMemcpyAsync(ps2.data(), S.ps, sizeof(double4) * S.n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> Ai(S.n);
    for (int i = 0; i < S.n; ++i) Ai[i] = ps2[i].w;
    idw::set_node_areas(c->hs, Ai.data());
    HIPCHK(hipMemcpyAsync(S.nodes, c->hs.nodes.data(), sizeof(idw::NodeRec) * S.nnodes, hipMemcpyHostToDevice, st));
    std::vector<idw::WalkHdr> whdr(S.nnodes + 1);  // value-initialised: the pad record is zero
    for (int k = 0; k < S.nnodes; ++k) {
        const idw::NodeRec &r = c->hs.nodes[k];
        idw::WalkHdr &w = whdr[k];
        for (int i = 0; i < 3; ++i) w.X[i] = r.X[i];
        w.r = r.leaf ? r.rad : r.orad;
        w.skip = r.skip;
        w.depth = r.depth;
        w.l = r.l;
        w.n = r.u - r.l + 1;
    }
    HIPCHK(hipMemcpyAsync(S.whdr, whdr.data(), sizeof(idw::WalkHdr) * whdr.size(), hipMemcpyHostToDevice, st));
    static double sphere[60];
    sphere_nominal(sphere);
    idw::launch_errors(st, S, make_params(c, true), sphere);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    c->surfInit = true;
    return compute_den0(c);
}

int idw_set_surface_coordinates(idw_ctx *c, const double *xs, int64_t ndof) {
    if (!c || !xs) return fail(IDW_EINVAL, "null argument");
    if (!c->surfInit) return fail(IDW_ESTATE, "surface not initialized");
    if (ndof != 3 * c->npts)
        return fail(IDW_EINVAL, "Incorrect length of coordinates: expected %lld, received %lld", (long long)(3 * c->npts),
                    (long long)ndof);
    if (const int64_t bad = first_non_finite(xs, ndof); bad >= 0)
        return fail(IDW_EINVAL, "surface coordinate %lld is not finite", (long long)bad);
    TRY(ensure_device(c));
    HIPCHK(hipMemcpyAsync(c->xs, xs, sizeof(double) * ndof, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return IDW_OK;
}

int idw_get_surface_coordinates(idw_ctx *c, double *xs, int64_t ndof) {
    if (!c || !xs) return fail(IDW_EINVAL, "null argument");
    if (!c->surfInit) return fail(IDW_ESTATE, "surface not initialized");
    if (ndof != 3 * c->npts) return fail(IDW_EINVAL, "expected %lld values", (long long)(3 * c->npts));
    TRY(ensure_device(c));
    HIPCHK(hipMemcpyAsync(xs, c->xs, sizeof(double) * ndof, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return IDW_OK;
}

int idw_warp_mesh(idw_ctx *c) {
    if (!c) return fail(IDW_EINVAL, "null context");
    TRY(ensure_device(c));
    TRY(run_warp(c, c->xs, c->xv));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_timings(c);
    return check_plan(c);
}

int idw_warp_mesh_async(idw_ctx *c) {
    if (!c) return fail(IDW_EINVAL, "null context");
    TRY(ensure_device(c));
    return run_warp(c, c->xs, c->xv);
}

int idw_wait(idw_ctx *c) {
    if (!c) return fail(IDW_EINVAL, "null context");
    TRY(ensure_device(c));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_timings(c);
    return check_plan(c);
}

int idw_get_volume_coordinates(idw_ctx *c, double *xv, int64_t ndof) {
    if (!c || (!xv && ndof)) return fail(IDW_EINVAL, "null argument");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    if (ndof != 3 * c->V.n) return fail(IDW_EINVAL, "expected %lld values", (long long)(3 * c->V.n));
    TRY(ensure_device(c));
    if (ndof) HIPCHK(hipMemcpyAsync(xv, c->xv, sizeof(double) * ndof, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return IDW_OK;
}

int idw_set_external_mesh_indices(idw_ctx *c, const int32_t *ind, int64_t n) {
    if (!c || (n && !ind) || n < 0) return fail(IDW_EINVAL, "bad argument");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    for (int64_t k = 0; k < n; ++k)
        if (ind[k] < 0 || ind[k] >= 3 * c->V.n) return fail(IDW_EINVAL, "index %lld out of range", (long long)k);
    c->solverInd.assign(ind, ind + n);
    return IDW_OK;
}

int idw_get_solver_dof(idw_ctx *c, int64_t *n) {
    if (!c || !n) return fail(IDW_EINVAL, "null argument");
    *n = (int64_t)c->solverInd.size();
    return IDW_OK;
}

// gridTransform.F90:4-51 (one process: common == warp ordering)
int idw_warp_to_solver_grid(idw_ctx *c, const double *warp, int64_t wdof, double *solver, int64_t sdof) {
    if (!c || (wdof && !warp) || (sdof && !solver)) return fail(IDW_EINVAL, "null argument");
    if (wdof != 3 * c->V.n || sdof != (int64_t)c->solverInd.size()) return fail(IDW_EINVAL, "size mismatch");
    for (int64_t k = 0; k < sdof; ++k) solver[k] = warp[c->solverInd[k]];
    return IDW_OK;
}

// gridTransform.F90:53-99: scatter-add of the solver vector (ADD_VALUES).
int idw_solver_to_warp_grid(idw_ctx *c, const double *solver, int64_t sdof, double *warp, int64_t wdof) {
    if (!c || (wdof && !warp) || (sdof && !solver)) return fail(IDW_EINVAL, "null argument");
    if (wdof != 3 * c->V.n || sdof != (int64_t)c->solverInd.size()) return fail(IDW_EINVAL, "size mismatch");
    std::fill(warp, warp + wdof, 0.0);
    for (int64_t k = 0; k < sdof; ++k) warp[c->solverInd[k]] += solver[k];
    return IDW_OK;
}

int idw_warp_deriv(idw_ctx *c, const double *dxv, int64_t ndof) {
    if (!c || (!dxv && ndof)) return fail(IDW_EINVAL, "null argument");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    if (ndof != 3 * c->V.n) return fail(IDW_EINVAL, "expected %lld values", (long long)(3 * c->V.n));
    TRY(ensure_device(c));
    if (ndof) HIPCHK(hipMemcpyAsync(c->dxv, dxv, sizeof(double) * ndof, hipMemcpyHostToDevice, c->st));
    TRY(run_deriv(c, c->xs, c->dxv, c->dxs));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_timings(c);
    TRY(check_plan(c));
    c->haveDxs = true;
    return IDW_OK;
}

int idw_warp_deriv_async(idw_ctx *c, const double *dxv, int64_t ndof) {
    if (!c || (!dxv && ndof)) return fail(IDW_EINVAL, "null argument");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    if (ndof != 3 * c->V.n) return fail(IDW_EINVAL, "expected %lld values", (long long)(3 * c->V.n));
    TRY(ensure_device(c));
    // from pageable memory the copy has read dxv when it returns
    if (ndof) HIPCHK(hipMemcpyAsync(c->dxv, dxv, sizeof(double) * ndof, hipMemcpyHostToDevice, c->st));
    TRY(run_deriv(c, c->xs, c->dxv, c->dxs));
    c->haveDxs = true;
    return IDW_OK;
}

int idw_get_dxs(idw_ctx *c, double *dxs, int64_t ndof) {
    if (!c || !dxs) return fail(IDW_EINVAL, "null argument");
    if (!c->haveDxs) return fail(IDW_ESTATE, "no warp_deriv result");
    if (ndof != 3 * c->npts) return fail(IDW_EINVAL, "expected %lld values", (long long)(3 * c->npts));
    TRY(ensure_device(c));
    HIPCHK(hipMemcpyAsync(dxs, c->dxs, sizeof(double) * ndof, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return IDW_OK;
}

int idw_warp_mesh_dev(idw_ctx *c, const double *d_xs, double *d_xv, void *stream) {
    if (!c || !d_xs || !d_xv) return fail(IDW_EINVAL, "null argument");
    TRY(ensure_device(c));
    hipStream_t user = static_cast<hipStream_t>(stream);
    // order the context stream after the caller's stream, and back
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e, user));
    HIPCHK(hipStreamWaitEvent(c->st, e, 0));
    int r = run_warp(c, d_xs, d_xv);
    if (r == IDW_OK) {
        HIPCHK(hipEventRecord(e, c->st));
        HIPCHK(hipStreamWaitEvent(user, e, 0));
    }
    (void)hipEventDestroy(e);
    return r;
}

int idw_warp_deriv_dev(idw_ctx *c, const double *d_xs, const double *d_dxv, double *d_dxs, void *stream) {
    if (!c || !d_xs || !d_dxv || !d_dxs) return fail(IDW_EINVAL, "null argument");
    TRY(ensure_device(c));
    hipStream_t user = static_cast<hipStream_t>(stream);
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e, user));
    HIPCHK(hipStreamWaitEvent(c->st, e, 0));
    int r = run_deriv(c, d_xs, d_dxv, d_dxs);
    if (r == IDW_OK) {
        HIPCHK(hipEventRecord(e, c->st));
        HIPCHK(hipStreamWaitEvent(user, e, 0));
    }
    (void)hipEventDestroy(e);
    return r;
}

int idw_warp_deriv_dev_noalloc(idw_ctx *c, const double *d_xs, const double *d_dxv, double *d_dxs, void *stream) {
    if (!c || !d_xs || !d_dxv || !d_dxs) return fail(IDW_EINVAL, "null argument");
    if (!c->surfInit) return fail(IDW_ESTATE, "warp_deriv before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    const int fast = c->opt.evalMode == IDW_EVAL_FAST;
    int nchunk = 0;
    if (!fast) {
        const int64_t per = 12 * (int64_t)c->S.n;  // doubles of one volume chunk's partial sums
        if (per > NOALLOC_SCRATCH_DOUBLES)
            return fail(IDW_EINVAL,
                        "the allocation-free reverse holds %lld bytes of scratch: exact mode over %d unique surface "
                        "nodes needs %lld (at most %lld nodes)",
                        (long long)(8 * NOALLOC_SCRATCH_DOUBLES), c->S.n, (long long)(8 * per),
                        (long long)(NOALLOC_SCRATCH_DOUBLES / 12));
        nchunk = (int)std::min<int64_t>(exact_rev_chunks(c), NOALLOC_SCRATCH_DOUBLES / per);
    }
    if (c->capturing) return fail(IDW_ESTATE, "the allocation-free reverse is not recorded into batches");
    TRY(poll_plan(c));
    TRY(check_mixed_range(c));
    if (deriv_needs_warp(c))
        return fail(IDW_ESTATE, "the allocation-free reverse needs a warp of this context at the current options first");
    if (nloop(c) > 1 && !c->V.acc) return fail(IDW_ESTATE, "the allocation-free reverse needs a warp with the planes first");
    if (!fast && c->V.n && !c->V.cpos) return fail(IDW_ESTATE, "the allocation-free reverse needs an exact-mode warp first");
    if (fast && c->opt.deterministic && c->V.n && !det_ready(c))
        return fail(IDW_ESTATE, "the allocation-free deterministic reverse needs an allocating one of this plan first");
    TRY(ensure_device(c));
    hipStream_t user = static_cast<hipStream_t>(stream);
    HIPCHK(hipEventRecord(c->evOrder, user));
    HIPCHK(hipStreamWaitEvent(c->st, c->evOrder, 0));
    const int r = deriv_kernels(c, d_xs, d_dxv, d_dxs, fast ? nullptr : c->scratch, nchunk);
    if (r == IDW_OK) {
        HIPCHK(hipEventRecord(c->evOrder, c->st));
        HIPCHK(hipStreamWaitEvent(user, c->evOrder, 0));
    }
    return r;
}

int idw_warp_deriv_fwd(idw_ctx *c, const double *dxs, int64_t ndof_s, double *dxv, int64_t ndof_v) {
    if (!c || (!dxs && ndof_s) || (!dxv && ndof_v)) return fail(IDW_EINVAL, "null argument");
    if (!c->surfInit) return fail(IDW_ESTATE, "warp_deriv_fwd before initialize_warping");
    if (!c->volSet) return fail(IDW_ESTATE, "no volume nodes set");
    if (ndof_s != 3 * c->npts) return fail(IDW_EINVAL, "dxs: expected %lld values", (long long)(3 * c->npts));
    if (ndof_v != 3 * c->V.n) return fail(IDW_EINVAL, "dxv: expected %lld values", (long long)(3 * c->V.n));
    TRY(ensure_device(c));
    HIPCHK(hipMemcpyAsync(c->dxsIn, dxs, sizeof(double) * ndof_s, hipMemcpyHostToDevice, c->st));
    TRY(run_deriv_fwd(c, c->xs, c->dxsIn, c->dxv));
    if (ndof_v) HIPCHK(hipMemcpyAsync(dxv, c->dxv, sizeof(double) * ndof_v, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_timings(c);
    TRY(check_plan(c));
    return IDW_OK;
}

int idw_warp_deriv_fwd_dev(idw_ctx *c, const double *d_xs, const double *d_dxs, double *d_dxv, void *stream) {
    if (!c || !d_xs || !d_dxs || !d_dxv) return fail(IDW_EINVAL, "null argument");
    TRY(ensure_device(c));
    hipStream_t user = static_cast<hipStream_t>(stream);
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e, user));
    HIPCHK(hipStreamWaitEvent(c->st, e, 0));
    int r = run_deriv_fwd(c, d_xs, d_dxs, d_dxv);
    if (r == IDW_OK) {
        HIPCHK(hipEventRecord(e, c->st));
        HIPCHK(hipStreamWaitEvent(user, e, 0));
    }
    (void)hipEventDestroy(e);
    return r;
}

int idw_get_sizes(idw_ctx *c, int64_t *n_vol, int64_t *n_pts, int64_t *n_unique, int64_t *n_tree_nodes) {
    if (!c) return fail(IDW_EINVAL, "null context");
    if (n_vol) *n_vol = c->V.n;
    if (n_pts) *n_pts = c->npts;
    if (n_unique) *n_unique = c->S.n;
    if (n_tree_nodes) *n_tree_nodes = c->S.nnodes;
    return IDW_OK;
}

int idw_get_ldef0(idw_ctx *c, double *ldef0) {
    if (!c || !ldef0) return fail(IDW_EINVAL, "null argument");
    if (!c->surfInit) return fail(IDW_ESTATE, "surface not initialized");
    *ldef0 = c->hs.ldef0;
    return IDW_OK;
}

int idw_get_denominator0(idw_ctx *c, double *den0, int64_t n) {
    if (!c || (!den0 && n)) return fail(IDW_EINVAL, "null argument");
    if (!c->volSet || !c->surfInit) return fail(IDW_ESTATE, "not initialized");
    if (n != c->V.n) return fail(IDW_EINVAL, "expected %lld values", (long long)c->V.n);
    TRY(ensure_device(c));
    std::vector<double4> vol(n);
    std::vector<int32_t> perm(n);
    if (n) {
        HIPCHK(hipMemcpyAsync(vol.data(), c->V.vol, sizeof(double4) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(perm.data(), c->V.perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    for (int64_t k = 0; k < n; ++k) den0[perm[k]] = vol[k].w;
    return IDW_OK;
}

int idw_get_timings(idw_ctx *c, double *ms6) {
    if (!c || !ms6) return fail(IDW_EINVAL, "null argument");
    TRY(ensure_device(c));
    HIPCHK(hipStreamSynchronize(c->st));
    collect_timings(c);
    for (int k = 0; k < 6; ++k) ms6[k] = c->ms[k];
    return check_plan(c);
}

int idw_tune(idw_ctx *c, int force) {
    if (!c) return fail(IDW_EINVAL, "null context");
    TRY(ensure_device(c));
    if (!force && tune_current(c)) return IDW_OK;
    c->tuneFixed = false;
    return run_tune(c);
}

int idw_set_tuning(idw_ctx *c, const int32_t *params5) {
    if (!c || !params5) return fail(IDW_EINVAL, "null argument");
    auto in = [](int32_t v, std::initializer_list<int32_t> l) { return std::find(l.begin(), l.end(), v) != l.end(); };
    if (!in(params5[0], {16, 32, 64}) || !in(params5[1], {8, 16, 32}) || !in(params5[2], {64, 128, 256}) ||
        !in(params5[3], {16, 32, 64}) || !in(params5[4], {-1, 0, 1}))
        return fail(IDW_EINVAL, "tuning parameters outside their candidates (16/32/64, 8/16/32, 64/128/256, 16/32/64, "
                                "-1/0/1)");
    c->tune.fch = params5[0];
    c->tune.mq = params5[1];
    c->tune.wbuf = params5[2];
    c->tune.exactG = params5[3];
    c->tune.exactChunkShift = params5[4];
    c->tuned = false;
    c->tuneFixed = true;
    ++c->gen;
    return IDW_OK;
}

int idw_get_tuning(idw_ctx *c, int32_t *params5, double *ms, int64_t nms) {
    if (!c || !params5) return fail(IDW_EINVAL, "null argument");
    if (ms && nms != idw::TUNE_SLOTS) return fail(IDW_EINVAL, "expected %d timing slots", idw::TUNE_SLOTS);
    params5[0] = c->tune.fch;
    params5[1] = c->tune.mq;
    params5[2] = c->tune.wbuf;
    params5[3] = c->tune.exactG;
    params5[4] = c->tune.exactChunkShift;
    if (ms)
        for (int k = 0; k < idw::TUNE_SLOTS; ++k) ms[k] = c->tuned ? c->tuneMs[k] : -1.0;
    return c->tuned ? (tune_current(c) ? 1 : 2) : 0;
}

int idw_fast_counts(idw_ctx *c, double *counts8) {
    if (!c || !counts8) return fail(IDW_EINVAL, "null argument");
    if (!c->volSet || !c->surfInit) return fail(IDW_ESTATE, "not initialized");
    TRY(ensure_device(c));
    unsigned long long *d = nullptr;
    TRY(dalloc(c, &d, 8));
    HIPCHK(hipMemsetAsync(d, 0, 8 * sizeof(unsigned long long), c->st));
    for (int kk = 0; kk < nloop(c) && c->V.n; ++kk)
        idw::launch_fast_count(c->st, c->S, c->V, mirror_of(c, kk), c->opt.errTol, d);
    unsigned long long h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c, d);
    for (int k = 0; k < 8; ++k) counts8[k] = (double)h[k];
    return IDW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many warps and reverse products over several meshes, recorded
// once into a HIP graph and replayed with one launch.
// ---------------------------------------------------------------------------
struct idw_batch {
    struct Op {
        idw_ctx *c;
        int kind;  // 0 warp, 1 warp_deriv
        const double *xs, *dxv;
        double *out;
    };
    int device = 0;
    std::vector<Op> ops;
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int layout = -1;  // 0: one graph branch per mesh, 1: one chain
    double ms[2] = {-1.0, -1.0};
    std::vector<hipEvent_t> events;                      // events recorded into the graph
    hipEvent_t in = nullptr, out = nullptr;              // ordering with the caller's stream
    std::vector<std::pair<idw_ctx *, uint64_t>> sig;     // context generations at recording time
    int64_t recordings = 0, launches = 0;                // warm-up + recording passes, graph launches
};

namespace {

int batch_op(idw_batch::Op &o) {
    return o.kind == 0 ? run_warp(o.c, o.xs, o.out) : run_deriv(o.c, o.xs, o.dxv, o.out);
}

std::vector<idw_ctx *> batch_contexts(const idw_batch *b) {
    std::vector<idw_ctx *> cs;
    for (const auto &o : b->ops)
        if (std::find(cs.begin(), cs.end(), o.c) == cs.end()) cs.push_back(o.c);
    return cs;
}

int batch_event(idw_batch *b, hipEvent_t *e) {
    HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    b->events.push_back(*e);
    return IDW_OK;
}

void batch_drop_graph(idw_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    b->events.clear();
}

bool batch_current(const idw_batch *b) {
    for (const auto &p : b->sig)
        if (p.first->gen != p.second) return false;
    return b->exec != nullptr;
}

// Record every op into a graph: layout 0 forks one branch per mesh (each on
// its context's stream) and joins them; layout 1 chains the meshes.
int batch_capture(idw_batch *b, int layout) {
    batch_drop_graph(b);
    const std::vector<idw_ctx *> cs = batch_contexts(b);
    HIPCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeGlobal));
    hipEvent_t fork;
    int r = batch_event(b, &fork);
    if (r == IDW_OK && hipEventRecord(fork, b->origin) != hipSuccess) r = fail(IDW_EHIP, "event record in capture");
    hipStream_t prev = b->origin;
    for (size_t k = 0; r == IDW_OK && k < cs.size(); ++k) {
        idw_ctx *c = cs[k];
        hipEvent_t dep = fork;
        if (layout == 1 && k > 0) {
            r = batch_event(b, &dep);
            if (r == IDW_OK && hipEventRecord(dep, prev) != hipSuccess) r = fail(IDW_EHIP, "event record in capture");
        }
        if (r == IDW_OK && hipStreamWaitEvent(c->st, dep, 0) != hipSuccess) r = fail(IDW_EHIP, "stream join in capture");
        c->capturing = true;
        for (auto &o : b->ops)
            if (r == IDW_OK && o.c == c) r = batch_op(o);
        c->capturing = false;
        prev = c->st;
    }
    // join every branch back into the origin stream
    for (size_t k = 0; r == IDW_OK && k < cs.size(); ++k) {
        hipEvent_t j;
        r = batch_event(b, &j);
        if (r == IDW_OK && hipEventRecord(j, cs[k]->st) != hipSuccess) r = fail(IDW_EHIP, "event record in capture");
        if (r == IDW_OK && hipStreamWaitEvent(b->origin, j, 0) != hipSuccess) r = fail(IDW_EHIP, "join in capture");
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->origin, &g);
    if (r != IDW_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (ee != hipSuccess) return fail(IDW_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->layout = layout;
    b->sig.clear();
    for (idw_ctx *c : cs) b->sig.push_back({c, c->gen});
    return IDW_OK;
}

int batch_time(idw_batch *b, float *ms) {
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    HIPCHK(hipEventRecord(t0, b->origin));
    for (int k = 0; k < 2; ++k) HIPCHK(hipGraphLaunch(b->exec, b->origin));
    HIPCHK(hipEventRecord(t1, b->origin));
    HIPCHK(hipEventSynchronize(t1));
    HIPCHK(hipEventElapsedTime(ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return IDW_OK;
}

}  // namespace

extern "C" {

int idw_batch_create(idw_batch **out, int device) {
    if (!out) return fail(IDW_EINVAL, "null argument");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return fail(IDW_ENODEV, "no HIP device %d", device);
    HIPCHK(hipSetDevice(device));
    idw_batch *b = new (std::nothrow) idw_batch();
    if (!b) return fail(IDW_ENOMEM, "out of host memory");
    b->device = device;
    if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->out, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(IDW_EHIP, "stream or event creation failed");
    }
    *out = b;
    return IDW_OK;
}

int idw_batch_destroy(idw_batch *b) {
    if (!b) return IDW_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->origin);
    batch_drop_graph(b);
    (void)hipEventDestroy(b->in);
    (void)hipEventDestroy(b->out);
    (void)hipStreamDestroy(b->origin);
    delete b;
    return IDW_OK;
}

static int batch_add(idw_batch *b, idw_ctx *c, int kind, const double *xs, const double *dxv, double *out) {
    if (!b || !c || !xs || !out || (kind == 1 && !dxv)) return fail(IDW_EINVAL, "null argument");
    if (c->device != b->device) return fail(IDW_EINVAL, "context on device %d, batch on device %d", c->device, b->device);
    batch_drop_graph(b);  // recorded again at the next run
    b->ops.push_back({c, kind, xs, dxv, out});
    return IDW_OK;
}

int idw_batch_add_warp(idw_batch *b, idw_ctx *c, const double *d_xs, double *d_xv) {
    return batch_add(b, c, 0, d_xs, nullptr, d_xv);
}

int idw_batch_add_warp_deriv(idw_batch *b, idw_ctx *c, const double *d_xs, const double *d_dxv, double *d_dxs) {
    return batch_add(b, c, 1, d_xs, d_dxv, d_dxs);
}

int idw_batch_run(idw_batch *b, void *hip_stream) {
    if (!b) return fail(IDW_EINVAL, "null batch");
    if (b->ops.empty()) return IDW_OK;
    HIPCHK(hipSetDevice(b->device));
    hipStream_t user = static_cast<hipStream_t>(hip_stream);
    if (!batch_current(b)) {
        // run every op once in order: sizes and fills the fast-mode plans and
        // allocates the per-context buffers, none of which may happen in a
        // graph.  The plans are then kept: within a batch every warp of a mesh
        // uses the same accept/recurse decisions (they depend on the mesh and
        // errTol only).
        batch_drop_graph(b);
        ++b->recordings;
        std::vector<idw_ctx *> cs = batch_contexts(b);
        HIPCHK(hipEventRecord(b->in, user));
        // the autotuner picks each mesh's tile parameters for its own size and
        // options before anything is recorded (it keeps an earlier choice made
        // for the same options and sizes); not for a deterministic context,
        // whose bits would then depend on a timing
        for (idw_ctx *c : cs)
            if (!c->tuneFixed && !c->opt.deterministic && !tune_current(c)) {
                HIPCHK(hipStreamWaitEvent(c->st, b->in, 0));
                TRY(run_tune(c));
            }
        HIPCHK(hipEventRecord(b->in, user));
        for (idw_ctx *c : cs) {
            HIPCHK(hipStreamWaitEvent(c->st, b->in, 0));
            c->batchPlan = true;
        }
        int r = IDW_OK;
        for (auto &o : b->ops)
            if (r == IDW_OK) r = batch_op(o);
        for (idw_ctx *c : cs) {
            if (r == IDW_OK && hipStreamSynchronize(c->st) != hipSuccess) r = fail(IDW_EHIP, "batch warm-up failed");
            if (r == IDW_OK) r = check_plan(c);
        }
        if (r != IDW_OK) {
            for (idw_ctx *c : cs) c->batchPlan = false;
            return r;
        }
        // choose the graph layout by timing both (two replays each)
        float t = 0.f;
        for (int layout = 0; layout < 2 && r == IDW_OK; ++layout) {
            r = batch_capture(b, layout);
            if (r == IDW_OK) r = batch_time(b, &t);
            b->ms[layout] = t / 2.0;
        }
        if (r == IDW_OK) r = batch_capture(b, b->ms[0] <= b->ms[1] ? 0 : 1);
        for (idw_ctx *c : cs) c->batchPlan = false;
        TRY(r);
    }
    HIPCHK(hipEventRecord(b->in, user));
    HIPCHK(hipStreamWaitEvent(b->origin, b->in, 0));
    HIPCHK(hipGraphLaunch(b->exec, b->origin));
    ++b->launches;
    HIPCHK(hipEventRecord(b->out, b->origin));
    HIPCHK(hipStreamWaitEvent(user, b->out, 0));
    return IDW_OK;
}

int idw_device_bytes(idw_ctx *c, int64_t *bytes) {
    if (!c || !bytes) return fail(IDW_EINVAL, "null argument");
    *bytes = c->devBytes;
    return IDW_OK;
}

int idw_batch_stats(idw_batch *b, int64_t *recordings, int64_t *launches) {
    if (!b) return fail(IDW_EINVAL, "null batch");
    if (recordings) *recordings = b->recordings;
    if (launches) *launches = b->launches;
    return IDW_OK;
}

int idw_batch_info(idw_batch *b, int *layout, double *ms2) {
    if (!b) return fail(IDW_EINVAL, "null batch");
    if (layout) *layout = b->layout;
    if (ms2) {
        ms2[0] = b->ms[0];
        ms2[1] = b->ms[1];
    }
    return IDW_OK;
}

}  // extern "C"
