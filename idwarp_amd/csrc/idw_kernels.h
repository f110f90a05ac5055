// Device-side structures and kernel launchers of the volume warp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/idwarp_hip.h"
#include "idw_internal.h"

namespace idw {

constexpr int ADJ_STRIDE = 16;  // doubles per adjoint record (12 used, 128 B aligned)

enum { EXP_35 = 0, EXP_INT = 1, EXP_REAL = 2 };

constexpr int MAX_PLANES = 6;

// Mixed precision keeps in FP64 every fast-mode term farther than
// FAR_FACTOR * Ldef from an active lane of the wave.
constexpr double FAR_FACTOR = 2.0;

// One pass of the symmetry-plane loop (warpMesh.F90:46-64): plane j (0-based)
// reflects the volume node when bit j of mask is set, i.e. the reference's
// getMirrorPt test mod(kk-1, 2^(j+1)) >= 2^j with mask = kk-1.
struct Mirror {
    int32_t n = 0, mask = 0;
    double p[MAX_PLANES][3] = {};
    double nrm[MAX_PLANES][3] = {};
};

// How a forward launch combines its sums with those of the other mirror
// passes: ONLY writes Xv; FIRST / MID store / add (numerator, denominator)
// into the accumulator; LAST adds the accumulator and writes Xv.
enum { PASS_ONLY = 0, PASS_FIRST = 1, PASS_MID = 2, PASS_LAST = 3 };

// Scalars of one warp (warpMesh.F90:39-41): Ldef = Ldef0 * LdefFact and
// alpha^bExp with the current options; the exponents are those frozen into
// the tree at initialisation.
struct WParams {
    double Ldef, a5, aExp, bExp, errTol, nearT2;
    int32_t expMode, ia, ib;
    int32_t tangent;  // 1: the point/cluster records hold tangents; write num/den only (warpDerivFwd)
    int32_t pass;     // PASS_* of this launch
};

constexpr int MAX_LEVELS = 64;  // tree depths the cluster-data pass handles
constexpr int CUT_DEPTH = 9;    // the cluster-data pass's split: subtrees below, one block above

// levels of the internal nodes above the cut, deepest first
struct TopLevels {
    int n = 0;
    int start[CUT_DEPTH] = {}, count[CUT_DEPTH] = {};
};
constexpr int CSUM = 16;        // per-node sums of the cluster data (15 used)

struct DevSurface {
    int n = 0, nnodes = 0, nInternal = 0, nface = 0;
    NodeRec *nodes = nullptr;
    WalkHdr *whdr = nullptr;  // nnodes + 1 walk headers
    int32_t *internal = nullptr;
    // the cluster data bottom up (csum: CSUM sums per node): below the cut
    // depth, nCut subtrees (cut: root, deepest depth of its subtree; a bucket
    // above the cut is a subtree of its own), then the internal nodes above
    // the cut one depth at a time from the deepest (topOrder, levels in top)
    int32_t *cut = nullptr, *topOrder = nullptr;
    int nCut = 0;
    TopLevels top;
    double *csum = nullptr;
    double4 *ps = nullptr;  // Xu0, A (tree order)
    // term records: nnodes cluster records (index = tree node), then n point
    // records (index nnodes + tree-order point); pd / pd32 point at the latter
    Rec64 *rec = nullptr, *pd = nullptr;
    Rec32 *rec32 = nullptr, *pd32 = nullptr;
    double *normals0 = nullptr;
    double4 *nprops = nullptr;  // sumArea, sumNormal of the last evaluation
    int32_t *corner = nullptr, *xuInd = nullptr;
    int32_t *facePtr = nullptr, *faceConn = nullptr, *n2eStart = nullptr, *n2e = nullptr;
    int32_t *slotStart = nullptr, *slots = nullptr;
    double4 *fprops = nullptr;  // dA = area/nPts, unit normal
    double4 *fpropsd = nullptr; // their tangents (warpDerivFwd)
    // adjoint records indexed like rec (ADJ_STRIDE doubles each): nadj = adj
    // (clusters), padj = adj + nnodes * ADJ_STRIDE (points)
    double *adj = nullptr, *padj = nullptr, *nadj = nullptr, *ntmp = nullptr, *pb = nullptr;
};

// One term of the forward walk: the record (accepted cluster, or one point
// of an evaluated bucket) and the lanes of the wave it serves.
struct alignas(16) PlanEntry {
    unsigned long long mask;
    int32_t rec;   // index into the term records (DevSurface::rec)
    int32_t node;  // tree node where the walk found the term (the cluster, or the point's bucket)
};

// Per-wave plan of the fast mode (valid for one volume ordering, tree and errTol).
struct DevPlan {
    PlanEntry *entries = nullptr;
    int64_t *waveOff = nullptr;  // nwave + 1 offsets
    // first FP64 entry of each wave: the walk writes the terms evaluated in
    // FP32 from the front of the wave's range and those kept in FP64 (a lane
    // near the wall, mixed precision) from the back
    int64_t *waveSplit = nullptr;
    int64_t *waveCount = nullptr;
    unsigned int *mismatch = nullptr;
    // per group of walk waves of the mixed reverse: 1 when its seeds were
    // sparse and the group reverse took it (written by each reverse)
    uint32_t *groupSparse = nullptr;
    // deterministic reverse: the plan's slots sorted by record (detPos) and
    // each record's first sorted position (recStart, nrec + 1)
    int32_t *detPos = nullptr;
    int64_t *recStart = nullptr;
    // each record's sorted range cut into chunks of DET_CHUNK slots: chunk
    // offsets per record (nrec + 1) and the record of every chunk
    int64_t *chunkOff = nullptr;
    int32_t *chunkRec = nullptr;
    int64_t nchunk = 0, chunkCap = 0;
    bool transposed = false;  // detPos / recStart / chunks describe the current entries
    int64_t nwave = 0, total = 0, capacity = 0;
    double errTol = -1.0, nearT2 = -1.0, farT2 = -1.0;  // key of the filled plan
    double orderTol = -1.0, orderT2 = -1.0;  // key of the near-wall ordering of the volume nodes
    bool sized = false, filled = false;
};

// Tile parameters of the volume kernels, chosen per context by the
// autotuner (idw_tune); the defaults are those of an untuned context.
struct Tuning {
    int32_t fch = 32;         // fast forward: terms staged per chunk (16, 32, 64)
    int32_t mq = 32;          // mixed fast reverse: merge window per walk wave (8, 16, 32)
    int32_t wbuf = 128;       // walk: plan entries gathered in LDS per kind (64, 128, 256)
    int32_t exactG = 64;      // exact mixed forward: surface points staged per chunk (16, 32, 64)
    int32_t exactChunkShift = 0;  // exact reverse: volume chunks = the default count * 2^shift (-1, 0, 1)
};

// Slots of the autotuner's timing table (ms per launch of each candidate):
// walk buffer 64/128/256, forward chunk 16/32/64, merge window 8/16/32, exact
// forward chunk 16/32/64, exact reverse chunks x1/2, x1, x2.
enum { TUNE_WBUF = 0, TUNE_FCH = 3, TUNE_MQ = 6, TUNE_EXACT_G = 9, TUNE_EXACT_CHUNKS = 12, TUNE_SLOTS = 15 };

struct DevVolume {
    int64_t n = 0;
    double4 *vol = nullptr;   // x, y, z, denominator0 in traversal (Hilbert) order
    int32_t *perm = nullptr;  // traversal order -> warp order
    double *den = nullptr;    // denominator of the last warp, traversal order
    double4 *numb = nullptr;  // reverse seed dXv/den, traversal order
    double4 *acc = nullptr;   // numerator and denominator summed over mirror passes (symmetry planes only)
    double4 *volm = nullptr;  // mirrored positions of one pass for the stored-weight exact reverse (symmetry planes only)
    // exact reverse: positions of the seeded nodes (mirrored for the pass) packed in traversal order, and per
    // block of 256 nodes the offset of its first (grid_for(n, 256) + 1 entries, the total last)
    double4 *cpos = nullptr;
    int32_t *seedOff = nullptr;
};

void launch_face_props(hipStream_t st, const DevSurface &S, const double *xs);
void launch_node_props(hipStream_t st, const DevSurface &S, const double *xs, int initial, int useRot);
// face and node property tangents for a surface tangent dxs; writes tangent
// point records into S.pd / S.pd32
void launch_props_tangent(hipStream_t st, const DevSurface &S, const double *xs, const double *dxs, int useRot);
void launch_corners(hipStream_t st, const DevSurface &S, double threshold);
void launch_cluster_data(hipStream_t st, const DevSurface &S);
void launch_errors(hipStream_t st, const DevSurface &S, const WParams &P, const double *sphere60);
void launch_den0(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M);
// load-balancing costs (dryRun) of one mirror pass, added into cost (traversal order)
void launch_dry_run(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                    double *cost);
void launch_curve_keys(hipStream_t st, int64_t n, const double *xv0, double lo[3], double scale[3], uint64_t *keys,
                   int32_t *ids);
void launch_build_vol(hipStream_t st, int64_t n, const double *xv0, const int32_t *perm, double4 *vol);
void launch_near_keys(hipStream_t st, const DevSurface &S, const DevVolume &V, double errTol, double nearT2,
                      const double lo[3], const double scale[3], uint64_t *keys, int32_t *ids);
void launch_permute_vol(hipStream_t st, int64_t n, const int32_t *order, const double4 *vin, const int32_t *pin,
                        double4 *vout, int32_t *pout);
void launch_fill(hipStream_t st, double *a, int64_t n, double v);
// zero n counters with a kernel (safe inside graph capture, unlike a memset node)
void launch_zero_u32(hipStream_t st, unsigned int *a, int64_t n);
// Row length of the stored weights: n_vol rounded up to 4.
__host__ __device__ inline int64_t wrow(int64_t nvol) { return (nvol + 3) & ~(int64_t)3; }
// fast mode: blocks [blk0, blk0 + nblk) of 256 volume nodes (nblk < 0: to the end);
// exact mode with wout: every pair's weight stored at wout[i wrow(n_vol) + k]
void launch_forward(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                    int fast, int prec, double *xv, const DevPlan *plan, const Tuning &T, int64_t blk0 = 0,
                    int64_t nblk = -1, double *wout = nullptr);
void launch_prep_numb(hipStream_t st, const DevVolume &V, const Mirror &M, const double *dxv);
void launch_plan_count(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                       int64_t *waveCount);
void launch_plan_fill(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                      double nearT2, double farT2, const DevPlan &plan, const Tuning &T, int64_t blk0 = 0,
                      int64_t nblk = -1);
// detPart (fast mode, deterministic): per-slot sums instead of atomics;
// wst (exact mode): the forward's stored weights instead of recomputed ones
void launch_reverse(hipStream_t st, const DevSurface &S, const DevVolume &V, const WParams &P, const Mirror &M,
                    int fast, int prec, double *partial, int nchunk, const double *dxv, const DevPlan *plan,
                    const Tuning &T, double *detPart = nullptr, const double *wst = nullptr);
// the deterministic reverse's transpose of a plan and its reduction into the adjoint records
void launch_plan_recs(hipStream_t st, const DevPlan &plan, int nrec, unsigned *keys, int32_t *vals);
void launch_rec_bounds(hipStream_t st, int nrec, int64_t total, const unsigned *keys, int64_t *recStart);
constexpr int DET_CHUNK = 512;  // slots of a record summed by one wave
void launch_rec_chunks(hipStream_t st, int nrec, const int64_t *recStart, int64_t *count);
void launch_chunk_recs(hipStream_t st, int nrec, const int64_t *chunkOff, int32_t *chunkRec);
void launch_det_reduce(hipStream_t st, const DevSurface &S, const DevPlan &plan, const double *part,
                       double *chunkPart, int add);
hipError_t exclusive_scan(hipStream_t st, const int64_t *in, int64_t *out, int64_t n);
void launch_reverse_prep(hipStream_t st, const DevSurface &S, int fast);
void launch_fast_count(hipStream_t st, const DevSurface &S, const DevVolume &V, const Mirror &M, double errTol,
                       unsigned long long *out);
void launch_nodal_reverse(hipStream_t st, const DevSurface &S, const double *xs, int fast, int useRot, double *dxs,
                          int64_t npts);
// Stable sort of 32-bit keys (bits [0, end_bit)) with int32 values into keys_out / vals_out; with
// tmp null, sets bytes to the temporary storage it needs and sorts nothing.
hipError_t sort_pairs32(hipStream_t st, void *tmp, size_t &bytes, int64_t n, unsigned *keys, int32_t *vals,
                        unsigned *keys_out, int32_t *vals_out, int end_bit);
// Sort keys/ids pairs (curve order); temp storage allocated inside.
hipError_t sort_pairs(hipStream_t st, int64_t n, uint64_t *keys, int32_t *ids, uint64_t *keys_out, int32_t *ids_out);

}  // namespace idw
