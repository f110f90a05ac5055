// Internal layout of the volume-warp library: host-built surface/tree data
// and the device records the kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace idw {

constexpr int NERR = 12;

// kd-tree node in depth-first pre-order (parent, left subtree, right subtree)
// as created by build_tree (kd_tree.F90:140-182).  One record per node; the
// traversal kernels read it with wave-uniform (scalar) loads.
// The first 64 bytes (NodeHdr) are what the walk needs at every visit and
// are fetched with one scalar load.
struct alignas(16) NodeRec {
    double X[3];      // centroid of the node's points (setCenters, kd_tree.F90:558-593)
    double orad;      // 1 / max distance from X to a point of the node
    int32_t l, u;     // points [l, u] (0-based, inclusive) in tree order
    int32_t skip;     // index of the next node after this subtree (== node count at the end)
    int32_t depth;    // root = 0
    int32_t leaf;     // 1 for a bucket
    int32_t n;        // u - l + 1
    double rad;       // 1 / orad
    double err[NERR]; // RMS denominator error at rstar radii (computeErrors, :832-880)
    double Ai;        // summed area weights (setAi, kd_tree.F90:536-556)
};
static_assert(sizeof(NodeRec) == 176, "NodeRec layout");

// What the fast walk reads at every visit, one 64-byte cache line per node,
// in a copy of its own (with one zeroed record past the last node, so that
// the walk can prefetch both successors of any node without clamping their
// indices).  A node is a bucket exactly when skip == its index + 1.
struct alignas(64) WalkHdr {
    double X[3];
    double r;              // orad for an internal node, rad for a bucket
    int32_t skip, depth;
    int32_t l, n;          // the bucket's points
    int32_t pad[4];
};
static_assert(sizeof(WalkHdr) == 64, "WalkHdr layout");

// The walk addresses headers by 32-bit byte offsets: the tree's nodes plus
// the pad record must fit below 4 GiB of headers (67,108,863 nodes).
constexpr int64_t MAX_WALK_NODES = (int64_t)(0x100000000ll / (int64_t)sizeof(WalkHdr)) - 1;
inline bool walk_offsets_fit(int64_t nnodes) { return nnodes >= 1 && nnodes <= MAX_WALK_NODES; }

// Mixed precision evaluates weights A (q^aExp + alpha^bExp q^bExp), q = Ldef/d,
// in FP32 for every pair farther than nearWallFactor * Ldef, so q <=
// 1 / nearWallFactor there.  The smallest nearWallFactor that keeps both
// powers below 1e32 (six decades under the FP32 maximum, for the area factor
// and the sums of up to 128 weighted terms).
double mixed_near_wall_floor(double aExp, double bExp, double alpha);
// The FP32 fast-mode weights form 1/d^3 before scaling by Ldef^3: the
// near-wall distance itself must keep 1/d^3 below 1e36.
constexpr double MIN_NEAR_WALL_DISTANCE = 1e-12;

struct alignas(16) NodeHdr {
    double X[3];
    double orad;
    int32_t l, u, skip, depth, leaf, n;
    double rad;
};
static_assert(sizeof(NodeHdr) == 64, "NodeHdr layout");

// Per-call record of one term of the IDW sum.  Surface points and tree
// clusters share the layout, so the fast-mode kernels treat both alike: the
// term's position x, weight factor A, and K, d such that the displacement it
// contributes at r is K (r - x) + d.
//  * surface point i (record nnodes + i, tree order): x = Xu0_i, A = A_i,
//    K = M_i - I and d = Xu_i - Xu0_i, so that M_i r + b_i - r equals
//    K (r - Xu0_i) + d (b_i = Xu_i - M_i Xu0_i, kd_tree.F90:1648);
//  * internal tree node (record = node index): x = centroid X, A = Ai,
//    K = mean(M_i) - I and d = K X + mean(b_i) (setData,
//    kd_tree.F90:595-617), so that mean(M) r + mean(b) - r equals K (r - X) + d.
struct alignas(16) Rec64 {
    double x[3];
    double A;
    double K[9];  // row-major
    double d[3];
};
static_assert(sizeof(Rec64) == 128, "Rec64 layout");
struct alignas(16) Rec32 {
    double x[3];  // positions stay FP64: r - x is formed before rounding
    float A;
    float K[9];
    float d[3];
};
static_assert(sizeof(Rec32) == 80, "Rec32 layout");

// Host-side result of the surface definition (myCreateTree, kd_tree.F90:1365-1545).
struct HostSurface {
    int64_t npts = 0;       // surface points incl. duplicates (Xs size / 3)
    int64_t n = 0;          // unique nodes
    int64_t nface = 0, nconn = 0;
    std::vector<int32_t> facePtr;    // nface + 1
    std::vector<int32_t> faceConn;   // 0-based tree index
    std::vector<int32_t> n2eStart, n2e;      // node -> faces (0-based), duplicates kept, face order
    std::vector<int32_t> slotStart, slots;   // node -> positions in faceConn (for the reverse gather)
    std::vector<int32_t> xuInd;      // tree node -> first point index (0-based)
    std::vector<double> xu0;         // 3*n tree order
    std::vector<NodeRec> nodes;      // DFS pre-order
    int32_t maxDepth = 0, nInternal = 0;
    double ldef0 = 0.0;
};

// Build the tree, permutation, connectivity maps and Ldef0 from the reduced
// surface.  uniquePts: 3*nu (already lexsorted), link 1-based (npts),
// faceConn 0-based into the npts points.  Returns false with msg on error.
bool build_surface(const double *uniquePts, int64_t nu, const int32_t *link, int64_t npts, const int32_t *faceSizes,
                   int64_t nface, const int32_t *faceConn, int64_t nconn, int bucket, HostSurface &out, std::string &msg);

// setAi (kd_tree.F90:536-556) on the host tree once nodal areas are known.
void set_node_areas(HostSurface &s, const double *Ai);

// pointReduce (src/utils/pointReduce.F90): returns number of unique points.
int64_t point_reduce(const double *pts, int64_t n, double tol, double *unique, int32_t *link);

}  // namespace idw
