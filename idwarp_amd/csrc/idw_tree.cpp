// Host-side surface set-up: duplicate-point merge, kd-tree build, connectivity
// maps and Ldef0.  This is init-time work (not on the timed warp path); the
// tree must be the reference's exactly, because the fast mode's accept/recurse
// decisions depend on node membership, centroids and radii.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "idw_internal.h"

namespace idw {

// pointReduce.F90:46-77: visit points in order; an unassigned point opens a
// new unique point and every point within tol (inclusive, kd_tree2.f90:1537)
// is linked to it, overwriting earlier links exactly as the reference does.
int64_t point_reduce(const double *pts, int64_t n, double tol, double *unique, int32_t *link) {
    if (n <= 0) return 0;
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return pts[3 * a] < pts[3 * b]; });
    std::vector<int64_t> rank(n);
    for (int64_t q = 0; q < n; ++q) rank[order[q]] = q;
    const double tol2 = tol * tol;
    std::fill(link, link + n, 0);
    std::vector<int64_t> hits;
    int64_t nu = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (link[i] != 0) continue;
        hits.clear();
        const double *p = pts + 3 * i;
        auto probe = [&](int64_t q) -> bool {
            const int64_t j = order[q];
            const double dx = pts[3 * j] - p[0];
            if (dx * dx > tol2) return false;  // sorted by x: nothing further this way
            const double dy = pts[3 * j + 1] - p[1], dz = pts[3 * j + 2] - p[2];
            if (dx * dx + dy * dy + dz * dz <= tol2) hits.push_back(j);
            return true;
        };
        for (int64_t q = rank[i]; q < n && probe(q); ++q) {
        }
        for (int64_t q = rank[i] - 1; q >= 0 && probe(q); --q) {
        }
        std::memcpy(unique + 3 * nu, p, 3 * sizeof(double));
        ++nu;
        for (int64_t j : hits) link[j] = static_cast<int32_t>(nu);
    }
    return nu;
}

double mixed_near_wall_floor(double aExp, double bExp, double alpha) {
    // q^e <= c  <=>  q <= c^(1/e); the floor is 1 / the smallest such q
    const double cap = 1e32;
    double qmax = INFINITY;
    if (aExp > 0.0) qmax = std::min(qmax, std::pow(cap, 1.0 / aExp));
    const double ab = std::pow(std::fabs(alpha), bExp);
    if (bExp > 0.0 && ab > 0.0) qmax = std::min(qmax, std::pow(cap / ab, 1.0 / bExp));
    return std::isfinite(qmax) ? 1.0 / qmax : 0.0;
}

namespace {

struct Builder {
    const double *data;         // unique points, 3*n, original (lexsorted) order
    std::vector<int32_t> idx;   // 0-based permutation (reference's `indexes` - 1)
    int bucket;

    double coord(int c, int32_t p) const { return data[3 * p + c]; }

    // most_spread_coordinate / spread_in_coordinate (kd_tree.F90:223-299):
    // spread over at most the first 201 entries of the range; ties keep the
    // lowest axis.
    int widest_axis(int l, int u) const {
        const int ul = std::min(u, l + 200);
        int best = 0;
        double bestSpread = -1.0;
        for (int c = 0; c < 3; ++c) {
            double lo = coord(c, idx[l]), hi = lo;
            for (int i = l + 1; i <= ul; ++i) {
                const double v = coord(c, idx[i]);
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            if (hi - lo > bestSpread) {
                bestSpread = hi - lo;
                best = c;
            }
        }
        return best;
    }

    // select_on_coordinate (kd_tree.F90:190-222): first-element pivot
    // partition, narrowing to the side holding position k.  The resulting
    // order inside the range is part of the tree's definition (it feeds the
    // 201-entry spread of the children), so the exchange sequence is kept.
    void select(int c, int k, int l, int u) {
        while (l < u) {
            const int32_t pivot = idx[l];
            const double pv = coord(c, pivot);
            int m = l;
            for (int i = l + 1; i <= u; ++i)
                if (coord(c, idx[i]) < pv) std::swap(idx[++m], idx[i]);
            std::swap(idx[l], idx[m]);
            if (m <= k) l = m + 1;
            if (m >= k) u = m - 1;
        }
    }
};

}  // namespace

bool build_surface(const double *uniquePts, int64_t nu, const int32_t *link, int64_t npts, const int32_t *faceSizes,
                   int64_t nface, const int32_t *faceConn, int64_t nconn, int bucket, HostSurface &s, std::string &msg) {
    if (nu <= 0) {
        msg = "surface has no unique points";
        return false;
    }
    if (nu > INT32_MAX / 2 || nconn > INT32_MAX) {
        msg = "surface too large for 32-bit indices";
        return false;
    }
    s = HostSurface();
    s.npts = npts;
    s.n = nu;
    s.nface = nface;
    s.nconn = nconn;
    s.facePtr.assign(nface + 1, 0);
    for (int64_t f = 0; f < nface; ++f) {
        if (faceSizes[f] < 3 || faceSizes[f] > 20) {
            msg = "face sizes must be in [3, 20]";
            return false;
        }
        s.facePtr[f + 1] = s.facePtr[f] + faceSizes[f];
    }
    if (s.facePtr[nface] != nconn) {
        msg = "sum(faceSizes) != len(faceConn)";
        return false;
    }
    for (int64_t i = 0; i < npts; ++i)
        if (link[i] < 1 || link[i] > nu) {
            msg = "link out of range";
            return false;
        }
    for (int64_t i = 0; i < nconn; ++i)
        if (faceConn[i] < 0 || faceConn[i] >= npts) {
            msg = "faceConn out of range";
            return false;
        }

    // --- tree (build_tree_for_range, kd_tree.F90:140-182), DFS pre-order ---
    Builder b{uniquePts, std::vector<int32_t>(nu), bucket};
    std::iota(b.idx.begin(), b.idx.end(), 0);
    struct Item { int l, u, depth; };
    std::vector<Item> stack{{0, (int)nu - 1, 0}};
    std::vector<int32_t> rightChild;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        NodeRec r{};
        r.l = it.l;
        r.u = it.u;
        r.n = it.u - it.l + 1;
        r.depth = it.depth;
        s.maxDepth = std::max(s.maxDepth, it.depth);
        const int id = (int)s.nodes.size();
        if (it.u - it.l <= bucket) {
            r.leaf = 1;
            s.nodes.push_back(r);
            rightChild.push_back(-1);
        } else {
            const int c = b.widest_axis(it.l, it.u);
            const int m = (it.l + it.u + 2) / 2 - 1;  // (l+u)/2 in the reference's 1-based indices
            b.select(c, m, it.l, it.u);
            r.leaf = 0;
            s.nodes.push_back(r);
            rightChild.push_back(0);
            ++s.nInternal;
            stack.push_back({m + 1, it.u, it.depth + 1});  // right, visited after the left subtree
            stack.push_back({it.l, m, it.depth + 1});
            (void)id;
        }
    }
    // skip pointers: subtree sizes, children of node k are k+1 and k+1+size(k+1)
    const int nn = (int)s.nodes.size();
    std::vector<int32_t> size(nn, 1);
    for (int k = nn - 1; k >= 0; --k)
        if (!s.nodes[k].leaf) {
            const int left = k + 1, right = left + size[left];
            size[k] = 1 + size[left] + size[right];
        }
    for (int k = 0; k < nn; ++k) s.nodes[k].skip = k + size[k];

    // setCenters (kd_tree.F90:558-593): plain mean and 1 / max radius
    for (auto &r : s.nodes) {
        double X[3] = {0.0, 0.0, 0.0};
        for (int i = r.l; i <= r.u; ++i)
            for (int c = 0; c < 3; ++c) X[c] = X[c] + uniquePts[3 * b.idx[i] + c];
        for (int c = 0; c < 3; ++c) r.X[c] = X[c] / r.n;
        double radius = 0.0;
        for (int i = r.l; i <= r.u; ++i) {
            const double *p = uniquePts + 3 * b.idx[i];
            const double dx = p[0] - r.X[0], dy = p[1] - r.X[1], dz = p[2] - r.X[2];
            radius = std::max(radius, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        // an internal node holds distinct points, so its radius is far above
        // the range where the walk's square root would need rescaling
        if (!r.leaf && !(radius > 1e-100)) {
            msg = "degenerate kd-tree node: points " + std::to_string(r.l) + ".." + std::to_string(r.u) +
                  " lie within 1e-100 of their centroid";
            return false;
        }
        r.orad = 1.0 / radius;
        r.rad = 1.0 / r.orad;
    }

    // Xu0 in tree order and the inverse permutation
    s.xu0.resize(3 * nu);
    std::vector<int32_t> inv(nu);
    for (int64_t i = 0; i < nu; ++i) {
        for (int c = 0; c < 3; ++c) s.xu0[3 * i + c] = uniquePts[3 * b.idx[i] + c];
        inv[b.idx[i]] = (int32_t)i;
    }
    // faceConn -> tree index (kd_tree.F90:1452-1455)
    s.faceConn.resize(nconn);
    for (int64_t i = 0; i < nconn; ++i) s.faceConn[i] = inv[link[faceConn[i]] - 1];
    // nodeToElem (kd_tree.F90:1460-1483): faces in increasing order, a face
    // listed once per occurrence of the node in it; slots keep the position.
    s.n2eStart.assign(nu + 1, 0);
    for (int64_t i = 0; i < nconn; ++i) ++s.n2eStart[s.faceConn[i] + 1];
    for (int64_t i = 0; i < nu; ++i) s.n2eStart[i + 1] += s.n2eStart[i];
    s.slotStart = s.n2eStart;
    s.n2e.resize(nconn);
    s.slots.resize(nconn);
    std::vector<int32_t> fill(nu, 0);
    for (int64_t f = 0; f < nface; ++f)
        for (int32_t j = s.facePtr[f]; j < s.facePtr[f + 1]; ++j) {
            const int32_t node = s.faceConn[j];
            const int32_t at = s.n2eStart[node] + fill[node]++;
            s.n2e[at] = (int32_t)f;
            s.slots[at] = j;
        }
    // XuInd (kd_tree.F90:1488-1498): first point linked to each node
    s.xuInd.assign(nu, -1);
    for (int64_t i = 0; i < npts; ++i) {
        const int32_t t = inv[link[i] - 1];
        if (s.xuInd[t] < 0) s.xuInd[t] = (int32_t)i;
    }
    // Ldef0 (kd_tree.F90:1501-1512): centroid over tree order, distances over
    // the unique points in their given order.
    double Xc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < nu; ++i)
        for (int c = 0; c < 3; ++c) Xc[c] = Xc[c] + s.xu0[3 * i + c];
    for (int c = 0; c < 3; ++c) Xc[c] = Xc[c] / nu;
    s.ldef0 = 0.0;
    for (int64_t i = 0; i < nu; ++i) {
        const double dx = uniquePts[3 * i] - Xc[0], dy = uniquePts[3 * i + 1] - Xc[1], dz = uniquePts[3 * i + 2] - Xc[2];
        s.ldef0 = std::max(s.ldef0, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    return true;
}

// setAi (kd_tree.F90:536-556): leaf sums in point order, internal nodes as
// left + right (post-order, i.e. reverse pre-order).
void set_node_areas(HostSurface &s, const double *Ai) {
    const int nn = (int)s.nodes.size();
    std::vector<int32_t> size(nn, 1);
    for (int k = nn - 1; k >= 0; --k) {
        NodeRec &r = s.nodes[k];
        if (r.leaf) {
            double a = 0.0;
            for (int i = r.l; i <= r.u; ++i) a = a + Ai[i];
            r.Ai = a;
        } else {
            const int left = k + 1, right = left + size[left];
            size[k] = 1 + size[left] + size[right];
            r.Ai = (s.nodes[left].Ai + s.nodes[right].Ai);
        }
    }
}

}  // namespace idw
