/*
 * idwarp_hip.h -- C ABI of the MI355X volume-warp library (libidwarp_hip.so).
 *
 * The entry points replace the f2py module `libidwarp` that IDWarp's
 * USMesh binds (src/f2py/libidwarp.pyf).  Each reference binding is cited
 * beside its replacement.  All state lives in an opaque context (one per
 * mesh instance) instead of Fortran module globals; this replaces the
 * private shared-library copies of idwarp/MExt.py:29-63.
 *
 * Conventions
 *  - Return 0 on success, a negative IDW_E* code on failure; the message of
 *    the last failure on the calling thread is in idw_last_error().  No entry
 *    point aborts the process (the reference's EChk calls MPI_Abort,
 *    src/utils/EChk.F90:12-25).
 *  - Coordinates are FP64, xyz-interleaved, flat (3*n values), as the
 *    reference's numpy arrays.  Host buffers are caller-owned and copied.
 *  - *_dev entry points take device pointers that are borrowed for the call
 *    and enqueue on the given hipStream_t (NULL = default stream); they
 *    return once the work is enqueued.
 *  - A context is not re-entrant; distinct contexts may run concurrently on
 *    distinct streams.
 */
#ifndef IDWARP_HIP_H
#define IDWARP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IDW_OK = 0,
    IDW_EINVAL = -1,  /* bad argument or size mismatch */
    IDW_ESTATE = -2,  /* call out of order (e.g. warp before surface definition) */
    IDW_EHIP = -3,    /* HIP runtime error */
    IDW_ENOMEM = -4,  /* device or host allocation failed */
    IDW_ENODEV = -5   /* no usable gfx950 device */
};

enum { IDW_EVAL_EXACT = 0, IDW_EVAL_FAST = 1 };       /* gridInput.f90:16-17 */
enum { IDW_PREC_FP64 = 0, IDW_PREC_MIXED = 1 };

/* Options of USMesh._getDefaultOptions (UnstructuredMesh.py:124-145) that
 * reach the numerics, plus the device-only ones. */
typedef struct idw_options {
    double aExp;            /* 3.0 */
    double bExp;            /* 5.0 */
    double LdefFact;        /* 1.0 */
    double alpha;           /* 0.25 */
    double errTol;          /* 5e-4 */
    double cornerAngle;     /* 30.0 degrees */
    int32_t evalMode;       /* IDW_EVAL_FAST */
    int32_t useRotations;   /* 1 */
    int32_t zeroCornerRotations; /* 1 */
    int32_t bucketSize;     /* 8 (kd_tree.F90:14 default 18 is overridden by USMesh) */
    int32_t precision;      /* IDW_PREC_FP64 or IDW_PREC_MIXED (device only) */
    int32_t reusePlan;      /* fast mode: 0 = walk the tree in every warp (as the reference does);
                               1 = keep the walk's plan while errTol, the grid and the surface
                               definition are unchanged (its decisions depend on nothing else) */
    double nearWallFactor;  /* mixed mode: buckets with a volume node closer than
                               nearWallFactor*Ldef stay FP64 (1e-6: keeps q^5 < 1e30 in FP32).
                               With precision mixed it must be at least the floor that keeps
                               q^aExp and alpha^bExp q^bExp below 1e32 for q = 1/nearWallFactor
                               (9.95e-8 for the defaults), and nearWallFactor*Ldef >= 1e-12;
                               smaller values are rejected with IDW_EINVAL */
    int32_t deterministic;  /* fast mode: 1 = the reverse product is bit-for-bit the same on
                               every run (the reference sums serially, kd_tree.F90:942-1051):
                               each plan entry's sums go to a slot of their own and are added
                               per term in plan order, with FP64 weights in either precision;
                               0 (default) = FP64 atomics, equal to rounding only.  The bits
                               hold for fixed tile parameters (the forward's chunking orders
                               its sums), so batches do not search them for such a context */
    int32_t storeWeights;   /* exact mode: 1 = the warp keeps every (volume, surface) pair's
                               weight, as used, for the reverse products that follow at the
                               same options (passes x n_unique x n_vol x 8 bytes of device
                               memory, n_vol rounded up to 4; the reference recomputes them,
                               kd_tree.F90:984-1013);
                               0 (default) = recompute.  Fast mode rejects 1 */
} idw_options;

typedef struct idw_ctx idw_ctx;

/* Fill *opt with the reference defaults. */
void idw_default_options(idw_options *opt);

/* Message of the last failure on this thread ("" if none). */
const char *idw_last_error(void);

/* Create a context on HIP device `device`.  Replaces USMesh.__init__'s
 * MExt load + initpetsc + _setMeshOptions (UnstructuredMesh.py:50-106). */
int idw_create(idw_ctx **out, int device, const idw_options *opt);
int idw_destroy(idw_ctx *ctx);

/* _setMeshOptions (UnstructuredMesh.py:1051-1069): gridInput values used by
 * the next warp.  As in the reference, the tree exponents, error tables and
 * denominator estimate keep their values from idw_initialize_warping. */
int idw_set_options(idw_ctx *ctx, const idw_options *opt);

/* createcommongrid (libidwarp.pyf:14-17, src/IO/createGrid.F90): the
 * undeformed volume nodes, 3*n values.  Must be
 * set before idw_initialize_warping. */
int idw_set_volume_nodes(idw_ctx *ctx, const double *xv0, int64_t ndof);

/* setsymmetryplanes (libidwarp.pyf:69-73, initializeWarping.F90:511-531):
 * n <= 6 planes, each a point and a normal (3*n values each, used as given);
 * every warp, reverse and tangent then sums over the 2^n mirror images of
 * each volume node (getMirrorPt / getMirrorNumerator, :533-659).  n = 0
 * removes them.  May be called before or after idw_initialize_warping (the
 * denominator estimate is recomputed). */
int idw_set_symmetry_planes(idw_ctx *ctx, const double *pts, const double *normals, int64_t n);

/* The restartfile argument of initializewarping (libidwarp.pyf:55-68,
 * initializeWarping.F90:98-312): the file used by the next
 * idw_initialize_warping (and by idw_set_symmetry_planes after it) for the
 * denominator estimate.  Same layout as the reference's single-process file:
 * int32 node count, then n_vol float64 estimates and n_vol float64
 * load-balancing costs (dryRun, kd_tree.F90:789-830), in volume-node order.
 * A file with the right count whose estimates agree with the device's at
 * every 100th node (relative 1e-10) is loaded; otherwise the file is
 * (re)written.  NULL or "" = no file. */
int idw_set_restart_file(idw_ctx *ctx, const char *path);

/* Outcome of the last restart-file use (0 none, 1 loaded, 2 written) and
 * the costs it holds (n = n_vol; costs may be NULL).  No counterpart in the
 * reference. */
int idw_get_restart_info(idw_ctx *ctx, int32_t *status, double *costs, int64_t n);

/* Smallest nearWallFactor that precision mixed accepts for these exponents
 * and alpha (see idw_options).  Context-free.  No counterpart in the
 * reference, which has no mixed precision. */
double idw_mixed_near_wall_floor(double aExp, double bExp, double alpha);

/* pointreduce (libidwarp.pyf:195-202, src/utils/pointReduce.F90:1-83):
 * merge points within tol.  pts and unique_pts are 3*n; link (n) is 1-based
 * into unique_pts.  Context-free. */
int idw_point_reduce(const double *pts, int64_t n, double tol, double *unique_pts, int32_t *link,
                     int64_t *n_unique);

/* initializewarping (libidwarp.pyf:55-68, src/warp/initializeWarping.F90):
 * pts 3*n_pts initial surface (local order), unique_pts 3*n_unique (sorted
 * by z, y, x), link 1-based (n_pts), face_sizes (n_faces), face_conn
 * 0-based into pts.  Builds the kd-tree, nodal properties, error tables and
 * the denominator estimate on the device. */
int idw_initialize_warping(idw_ctx *ctx, const double *pts, int64_t n_pts, const double *unique_pts,
                           int64_t n_unique, const int32_t *link, const int32_t *face_sizes, int64_t n_faces,
                           const int32_t *face_conn, int64_t n_conn);

/* setsurfacecoordinates / getsurfacecoordinates (libidwarp.pyf:78-87). */
int idw_set_surface_coordinates(idw_ctx *ctx, const double *xs, int64_t ndof);
int idw_get_surface_coordinates(idw_ctx *ctx, double *xs, int64_t ndof);

/* warpmesh (libidwarp.pyf:75-76, src/warp/warpMesh.F90): warp the volume
 * for the current surface coordinates; result stays on the device. */
int idw_warp_mesh(idw_ctx *ctx);

/* Non-blocking forms of idw_warp_mesh and idw_warp_deriv, for warping
 * several contexts at once (each context has its own stream): they enqueue
 * the work and return; idw_wait waits for the context's stream and reports
 * what the blocking forms report (a plan that disagreed with its counting
 * pass).  Later calls on the context are ordered after the work, so
 * idw_get_volume_coordinates and idw_get_dxs read its results.  No
 * counterpart in the reference, whose MultiUSMesh warps its instances one
 * after another (MultiUnstructuredMesh.py:677-755). */
int idw_warp_mesh_async(idw_ctx *ctx);
int idw_warp_deriv_async(idw_ctx *ctx, const double *dxv, int64_t ndof);
int idw_wait(idw_ctx *ctx);

/* getvolumecoordinates (libidwarp.pyf:45-48): copy the warped grid out. */
int idw_get_volume_coordinates(idw_ctx *ctx, double *xv, int64_t ndof);

/* setexternalmeshindices (libidwarp.pyf:40-43, setExternalMeshIndices.F90):
 * 0-based per-DOF indices into the common grid, n = solver DOF count. */
int idw_set_external_mesh_indices(idw_ctx *ctx, const int32_t *ind, int64_t n);
int idw_get_solver_dof(idw_ctx *ctx, int64_t *n);

/* warp_to_solver_grid / solver_to_warp_grid (libidwarp.pyf:88-100,
 * gridTransform.F90:4-99).  The reverse scatter sums duplicated indices. */
int idw_warp_to_solver_grid(idw_ctx *ctx, const double *warp_grid, int64_t wdof, double *solver_grid, int64_t sdof);
int idw_solver_to_warp_grid(idw_ctx *ctx, const double *solver_grid, int64_t sdof, double *warp_grid, int64_t wdof);

/* warpderiv (libidwarp.pyf:102-105, src/warp/warpDeriv.F90): dXs =
 * (dXv/dXs)^T dxv for the current surface; dxv is 3*n_vol in warp order.
 * Uses the denominator of the last warp at the same surface and options,
 * recomputing it if there was none. */
int idw_warp_deriv(idw_ctx *ctx, const double *dxv, int64_t ndof);

/* getdxs (libidwarp.pyf:114-117): 3*n_pts result of the last warp_deriv. */
int idw_get_dxs(idw_ctx *ctx, double *dxs, int64_t ndof);

/* warpderivfwd (libidwarp.pyf:107-112, src/warp/warpDerivFwd.F90:62-99):
 * dxv (3*n_vol, warp order) = (dXv/dXs) dxs for the current surface and
 * options; dxs is 3*n_pts.  The fast mode replays the accept/recurse
 * decisions of the last warp (a warp runs first if there is none for the
 * current options). */
int idw_warp_deriv_fwd(idw_ctx *ctx, const double *dxs, int64_t ndof_s, double *dxv, int64_t ndof_v);

/* Device-resident forms of warpMesh and warpDeriv for frameworks that keep
 * coordinates on the GPU (torch autograd wrapper).  d_xs: 3*n_pts surface,
 * d_xv: 3*n_vol output, d_dxv: 3*n_vol seed, d_dxs: 3*n_pts output.
 * They do not wait for the device: a fast-mode plan whose fill walk
 * disagreed with its counting pass (the host forms report it at once) is
 * reported with IDW_EHIP by the first later call on the context that finds
 * the warp completed, or by idw_get_timings, which waits for it.  The plan
 * is then dropped and rebuilt by the next warp.  Coordinates are not checked
 * for NaN here (the host forms reject non-finite input with IDW_EINVAL). */
int idw_warp_mesh_dev(idw_ctx *ctx, const double *d_xs, double *d_xv, void *hip_stream);
int idw_warp_deriv_dev(idw_ctx *ctx, const double *d_xs, const double *d_dxv, double *d_dxs, void *hip_stream);
/* Allocation-free form of idw_warp_deriv_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code being captured into HIP graphs by the caller, serving loops,
 * harnesses that forbid allocation).  It calls no hipMalloc, hipFree, event
 * or stream creation, nor any allocating library routine.  Requires a warp
 * of this context at the current options first (IDW_ESTATE otherwise); in
 * fast mode it replays that warp's plan.  Exact mode keeps its per-chunk
 * partial sums in a 32 MiB array of the library's own (one per device, held
 * until the process ends), which bounds the surface to 349,525 unique nodes
 * (IDW_EINVAL beyond).  Every call on a device shares that device's array:
 * calls that overlap on one device, from any context or thread, race and
 * must be ordered by the caller.  Same arguments and result as
 * idw_warp_deriv_dev (exact mode may split the volume into fewer chunks, so
 * sums can differ in the last bits). */
int idw_warp_deriv_dev_noalloc(idw_ctx *ctx, const double *d_xs, const double *d_dxv, double *d_dxs, void *hip_stream);

/* Device-resident warpDerivFwd: d_dxs 3*n_pts tangent, d_dxv 3*n_vol output. */
int idw_warp_deriv_fwd_dev(idw_ctx *ctx, const double *d_xs, const double *d_dxs, double *d_dxv, void *hip_stream);

/* Sizes and derived scalars. */
int idw_get_sizes(idw_ctx *ctx, int64_t *n_vol, int64_t *n_pts, int64_t *n_unique, int64_t *n_tree_nodes);
int idw_get_ldef0(idw_ctx *ctx, double *ldef0);

/* Denominator estimate of the fast mode (initializeWarping.F90:240-247,
 * getWiEstimate, kd_tree.F90:751-787): one value per volume node, warp
 * order; n = n_vol. */
int idw_get_denominator0(idw_ctx *ctx, double *den0, int64_t n);

/* Per-phase timing of the last warp / warp_deriv (host or _dev form),
 * measured with HIP events on the context's stream, in milliseconds:
 * [0] nodal properties + cluster data, [1] tree walk recording the plan
 * (fast mode; 0 when reused or exact), [2] forward volume kernel,
 * [3] reverse set-up (nodal properties, zero fill), [4] reverse volume
 * kernel(s), [5] nodal reverse.  Synchronises the context's stream.  No
 * counterpart in the reference (which has no timers). */
int idw_get_timings(idw_ctx *ctx, double *ms6);

/* Work of the fast mode for the current options.  Per volume node, summed
 * over all nodes: [0] internal-node visits, [1] accepted clusters, [2] leaf
 * point pairs, [3] error-table checks (the terms of dryRun's cost model,
 * kd_tree.F90:789-830).  Per 64-node wave of the device walk, summed over
 * waves: [4] internal-node visits, [5] visits with an accepting lane,
 * [6] bucket visits, [7] bucket point evaluations.  [0..3] give the
 * algorithmic flops; [0..3] / (64 [4..7]) the SIMD lane utilisation.
 * No counterpart in the reference. */
int idw_fast_counts(idw_ctx *ctx, double *counts8);

/* Autotuner (no counterpart in the reference): times the candidates of the
 * volume kernels' tile parameters on this context's mesh at its current
 * options and keeps the fastest of each.  Fast mode: the walk's LDS gather
 * buffer (64, 128, 256 plan entries), the forward's chunk of staged terms
 * (16, 32, 64) and, in mixed precision, the reverse's merge window (8, 16,
 * 32 entries per walk wave); exact mode: the mixed forward's FP32 group (4,
 * 8, 16 surface points) and the reverse's volume chunks (half, the default,
 * double).  Results do not depend on the choice beyond rounding (exact
 * mode's chunking; the order of FP32 partial sums).  Warps the current
 * surface.  Keeps an earlier choice made for the same evaluation mode,
 * precision, errTol, sizes and planes unless force is set.  idw_batch_run
 * calls it for every context it records. */
int idw_tune(idw_ctx *ctx, int force);
/* The chosen parameters {forward chunk, merge window, walk buffer, exact
 * forward chunk (surface points staged per chunk by the mixed exact
 * forward: 16, 32, 64), exact chunk shift (-1, 0, 1)} and, if ms is not NULL, the 15
 * timings of the last search in ms per launch (walk buffer 64/128/256,
 * forward chunk 16/32/64, merge window 8/16/32, exact chunk 16/32/64, exact
 * chunks x1/2, x1, x2; -1 where not applicable).  Returns 0 if never tuned,
 * 1 if tuned for the current options and sizes, 2 if tuned for others. */
int idw_get_tuning(idw_ctx *ctx, int32_t *params5, double *ms, int64_t nms);
/* Set the parameters (same order and candidates as idw_get_tuning) without
 * a search; the context counts as untuned afterwards, and batches keep these
 * parameters instead of searching (until idw_tune is called). */
int idw_set_tuning(idw_ctx *ctx, const int32_t *params5);
/* The deterministic reverse also needs memory: per plan entry 96 bytes of
 * sums, 4 of transpose and 12 for its sort (with the sort's temporary
 * storage), and 8 bytes per term record; the first reverse after a plan
 * fill sorts the plan's entries. */

/* Batches (no counterpart in the reference, which warps one surface per
 * call): many warps and reverse products over one or more contexts on one
 * device, recorded at the first idw_batch_run into a HIP graph and replayed
 * by every later run with a single launch.  Ops take device pointers (3*n_pts
 * surface, 3*n_vol volume) that must stay valid while the batch exists.
 * The first run executes every op once in order (it sizes and fills the
 * fast-mode plans, which are then kept for the batch: within a batch each
 * mesh's warps share their accept/recurse decisions), then records the ops
 * twice, as one graph branch per context and as one chain, times both and
 * keeps the faster.  A context that changes (options, buffers, surface
 * definition) after recording is recorded again at the next run.  Destroy a
 * batch before the contexts it uses. */
typedef struct idw_batch idw_batch;
int idw_batch_create(idw_batch **out, int device);
int idw_batch_destroy(idw_batch *batch);
int idw_batch_add_warp(idw_batch *batch, idw_ctx *ctx, const double *d_xs, double *d_xv);
int idw_batch_add_warp_deriv(idw_batch *batch, idw_ctx *ctx, const double *d_xs, const double *d_dxv,
                             double *d_dxs);
/* Enqueue the whole batch after the work on hip_stream (NULL = default
 * stream); later work on hip_stream waits for it. */
int idw_batch_run(idw_batch *batch, void *hip_stream);
/* The chosen layout (0: a branch per context, 1: a chain) and the measured
 * ms of one replay of each layout. */
int idw_batch_info(idw_batch *batch, int *layout, double *ms2);
/* How often the batch was (re)recorded (warm-up run, layouts timed, graph
 * recorded) and how many graph launches its runs made. */
int idw_batch_stats(idw_batch *batch, int64_t *recordings, int64_t *launches);

/* Device memory the context holds, in bytes: its surface, volume, plan and
 * stored-weight buffers (not the allocation-free reverse's scratch array,
 * which every context on a device shares).  No counterpart in the reference,
 * whose arrays live in host memory. */
int idw_device_bytes(idw_ctx *ctx, int64_t *bytes);

#ifdef __cplusplus
}
#endif

#endif
