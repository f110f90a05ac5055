// Host-side checks of libidwarp_hip's C++ (built with AddressSanitizer and
// UndefinedBehaviorSanitizer by idwarp_amd.build.build_host_checks; run on a
// machine without a GPU by tests/test_host_checks.py).
//
//   host_checks tree IN OUT   build the kd-tree and surface maps of the
//                             surface in IN with idw::build_surface and
//                             idw::set_node_areas, write them to OUT
//   host_checks reduce IN OUT idw::point_reduce of the points in IN
//   host_checks api           every C-ABI entry point on its no-device and
//                             bad-argument paths
//   host_checks limits        the walk's header-offset limit and the
//                             mixed-precision near-wall floor
//
// Binary files are little-endian: int64 counts followed by the arrays.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../idwarp_amd/csrc/idw_internal.h"
#include "../../include/idwarp_hip.h"

namespace {

struct Reader {
    FILE *f;
    int64_t i64() {
        int64_t v = 0;
        if (std::fread(&v, sizeof(v), 1, f) != 1) throw std::string("short read");
        return v;
    }
    template <class T>
    std::vector<T> arr(int64_t n) {
        std::vector<T> v(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != (size_t)n) throw std::string("short read");
        return v;
    }
};

struct Writer {
    FILE *f;
    void i64(int64_t v) { std::fwrite(&v, sizeof(v), 1, f); }
    template <class T>
    void arr(const std::vector<T> &v) {
        i64((int64_t)v.size());
        if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    }
};

int fails = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL %s (%s)\n", what, idw_last_error());
        ++fails;
    }
}

// IN: nu, unique (3 nu, lexsorted), npts, link (npts, 1-based), nface,
// faceSizes, nconn, faceConn (0-based into the npts points), bucket, Ai (nu,
// tree order; may be empty).  OUT: per node l, u, skip, depth, leaf (int32
// each), X (3), orad, Ai (double each); xu0; xuInd; faceConn (tree index);
// n2eStart; n2e; ldef0.
int tree(const char *in, const char *out) {
    FILE *fi = std::fopen(in, "rb");
    if (!fi) return 2;
    Reader r{fi};
    const int64_t nu = r.i64();
    auto uniq = r.arr<double>(3 * nu);
    const int64_t npts = r.i64();
    auto link = r.arr<int32_t>(npts);
    const int64_t nface = r.i64();
    auto fs = r.arr<int32_t>(nface);
    const int64_t nconn = r.i64();
    auto conn = r.arr<int32_t>(nconn);
    const int bucket = (int)r.i64();
    const int64_t na = r.i64();
    auto Ai = r.arr<double>(na);
    std::fclose(fi);
    idw::HostSurface s;
    std::string msg;
    if (!idw::build_surface(uniq.data(), nu, link.data(), npts, fs.data(), nface, conn.data(), nconn, bucket, s, msg)) {
        std::printf("build_surface: %s\n", msg.c_str());
        return 3;
    }
    if (na == nu) idw::set_node_areas(s, Ai.data());
    std::vector<int32_t> ints;
    std::vector<double> dbl;
    for (const idw::NodeRec &n : s.nodes) {
        for (int32_t v : {n.l, n.u, n.skip, n.depth, n.leaf}) ints.push_back(v);
        for (double v : {n.X[0], n.X[1], n.X[2], n.orad, n.Ai}) dbl.push_back(v);
    }
    FILE *fo = std::fopen(out, "wb");
    if (!fo) return 2;
    Writer w{fo};
    w.arr(ints);
    w.arr(dbl);
    w.arr(s.xu0);
    w.arr(s.xuInd);
    w.arr(s.faceConn);
    w.arr(s.n2eStart);
    w.arr(s.n2e);
    w.arr(std::vector<double>{s.ldef0});
    std::fclose(fo);
    return 0;
}

// IN: n, pts (3 n), tol.  OUT: unique (3 nu), link (n).
int reduce(const char *in, const char *out) {
    FILE *fi = std::fopen(in, "rb");
    if (!fi) return 2;
    Reader r{fi};
    const int64_t n = r.i64();
    auto pts = r.arr<double>(3 * n);
    auto tol = r.arr<double>(1);
    std::fclose(fi);
    std::vector<double> uniq(3 * n);
    std::vector<int32_t> link(n);
    const int64_t nu = idw::point_reduce(pts.data(), n, tol[0], uniq.data(), link.data());
    uniq.resize(3 * nu);
    FILE *fo = std::fopen(out, "wb");
    if (!fo) return 2;
    Writer w{fo};
    w.arr(uniq);
    w.arr(link);
    std::fclose(fo);
    return 0;
}

int api() {
    idw_options o;
    idw_default_options(&o);
    idw_ctx *c = nullptr;
    expect(idw_create(&c, 0, &o) == IDW_ENODEV && c == nullptr, "create without a device");
    expect(std::strlen(idw_last_error()) > 0, "message of the last failure");
    expect(idw_create(nullptr, 0, &o) == IDW_EINVAL, "create into null");
    expect(idw_create(&c, 0, nullptr) == IDW_EINVAL, "create with null options");
    idw_options bad = o;
    bad.bucketSize = 0;
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "bucketSize 0");
    bad = o;
    bad.precision = IDW_PREC_MIXED;
    bad.nearWallFactor = 0.0;
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "mixed precision, nearWallFactor 0");
    bad.nearWallFactor = 1e-6;
    expect(idw_create(&c, 0, &bad) == IDW_ENODEV, "mixed precision, nearWallFactor 1e-6 (valid options)");
    bad = o;
    bad.deterministic = 2;
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "deterministic 2");
    bad.deterministic = 1;
    expect(idw_create(&c, 0, &bad) == IDW_ENODEV, "deterministic 1 (valid options)");
    bad = o;
    bad.storeWeights = 1;  // the default evalMode is fast
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "storeWeights in fast mode");
    bad.evalMode = IDW_EVAL_EXACT;
    expect(idw_create(&c, 0, &bad) == IDW_ENODEV, "storeWeights in exact mode (valid options)");
    bad.storeWeights = 2;
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "storeWeights 2");
    bad = o;
    bad.errTol = NAN;
    expect(idw_create(&c, 0, &bad) == IDW_EINVAL, "NaN errTol");
    expect(idw_destroy(nullptr) == IDW_OK, "destroy null");
    // every context entry point rejects a null context instead of dereferencing it
    double d[6] = {0, 0, 0, 0, 0, 0};
    int32_t i[2] = {0, 0};
    int64_t n64 = 0;
    int32_t st = 0;
    expect(idw_set_options(nullptr, &o) == IDW_EINVAL, "set_options");
    expect(idw_set_volume_nodes(nullptr, d, 3) == IDW_EINVAL, "set_volume_nodes");
    expect(idw_set_symmetry_planes(nullptr, d, d, 1) == IDW_EINVAL, "set_symmetry_planes");
    expect(idw_set_restart_file(nullptr, "x") == IDW_EINVAL, "set_restart_file");
    expect(idw_get_restart_info(nullptr, &st, nullptr, 0) == IDW_EINVAL, "get_restart_info");
    expect(idw_initialize_warping(nullptr, d, 1, d, 1, i, i, 0, i, 0) == IDW_EINVAL, "initialize_warping");
    expect(idw_set_surface_coordinates(nullptr, d, 3) == IDW_EINVAL, "set_surface_coordinates");
    expect(idw_get_surface_coordinates(nullptr, d, 3) == IDW_EINVAL, "get_surface_coordinates");
    expect(idw_warp_mesh(nullptr) == IDW_EINVAL, "warp_mesh");
    expect(idw_warp_mesh_async(nullptr) == IDW_EINVAL, "warp_mesh_async");
    expect(idw_warp_deriv_async(nullptr, d, 3) == IDW_EINVAL, "warp_deriv_async");
    expect(idw_wait(nullptr) == IDW_EINVAL, "wait");
    expect(idw_warp_deriv_dev_noalloc(nullptr, d, d, d, nullptr) == IDW_EINVAL, "warp_deriv_dev_noalloc");
    expect(idw_batch_stats(nullptr, &n64, &n64) == IDW_EINVAL, "batch_stats");
    expect(idw_device_bytes(nullptr, &n64) == IDW_EINVAL, "device_bytes");
    expect(idw_get_volume_coordinates(nullptr, d, 3) == IDW_EINVAL, "get_volume_coordinates");
    expect(idw_set_external_mesh_indices(nullptr, i, 1) == IDW_EINVAL, "set_external_mesh_indices");
    expect(idw_get_solver_dof(nullptr, &n64) == IDW_EINVAL, "get_solver_dof");
    expect(idw_warp_to_solver_grid(nullptr, d, 3, d, 3) == IDW_EINVAL, "warp_to_solver_grid");
    expect(idw_solver_to_warp_grid(nullptr, d, 3, d, 3) == IDW_EINVAL, "solver_to_warp_grid");
    expect(idw_warp_deriv(nullptr, d, 3) == IDW_EINVAL, "warp_deriv");
    expect(idw_get_dxs(nullptr, d, 3) == IDW_EINVAL, "get_dxs");
    expect(idw_warp_deriv_fwd(nullptr, d, 3, d, 3) == IDW_EINVAL, "warp_deriv_fwd");
    expect(idw_warp_mesh_dev(nullptr, d, d, nullptr) == IDW_EINVAL, "warp_mesh_dev");
    expect(idw_warp_deriv_dev(nullptr, d, d, d, nullptr) == IDW_EINVAL, "warp_deriv_dev");
    expect(idw_warp_deriv_fwd_dev(nullptr, d, d, d, nullptr) == IDW_EINVAL, "warp_deriv_fwd_dev");
    expect(idw_get_sizes(nullptr, &n64, &n64, &n64, &n64) == IDW_EINVAL, "get_sizes");
    expect(idw_get_ldef0(nullptr, d) == IDW_EINVAL, "get_ldef0");
    expect(idw_get_denominator0(nullptr, d, 1) == IDW_EINVAL, "get_denominator0");
    expect(idw_get_timings(nullptr, d) == IDW_EINVAL, "get_timings");
    expect(idw_fast_counts(nullptr, d) == IDW_EINVAL, "fast_counts");
    idw_batch *b = nullptr;
    expect(idw_batch_create(&b, 0) == IDW_ENODEV && b == nullptr, "batch_create without a device");
    expect(idw_batch_create(nullptr, 0) == IDW_EINVAL, "batch_create into null");
    expect(idw_batch_destroy(nullptr) == IDW_OK, "batch_destroy null");
    expect(idw_batch_add_warp(nullptr, nullptr, d, d) == IDW_EINVAL, "batch_add_warp");
    expect(idw_batch_add_warp_deriv(nullptr, nullptr, d, d, d) == IDW_EINVAL, "batch_add_warp_deriv");
    expect(idw_batch_run(nullptr, nullptr) == IDW_EINVAL, "batch_run");
    int layout = 0;
    expect(idw_batch_info(nullptr, &layout, d) == IDW_EINVAL, "batch_info");
    // context-free routines
    double pts[12] = {0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 1e-13}, uq[12];
    int32_t link[4];
    expect(idw_point_reduce(pts, 4, 1e-12, uq, link, &n64) == IDW_OK && n64 == 2, "point_reduce");
    expect(link[0] == 1 && link[1] == 2 && link[2] == 1 && link[3] == 2, "point_reduce links");
    expect(idw_point_reduce(pts, 0, 1e-12, uq, link, &n64) == IDW_OK && n64 == 0, "point_reduce of nothing");
    expect(idw_point_reduce(nullptr, 4, 1e-12, uq, link, &n64) == IDW_EINVAL, "point_reduce of null");
    return fails ? 1 : 0;
}

int limits() {
    expect(idw::walk_offsets_fit(1), "one node");
    expect(idw::walk_offsets_fit(67108863), "67,108,863 nodes: header offsets up to 2^32 - 64");
    expect(!idw::walk_offsets_fit(67108864), "67,108,864 nodes: the pad header would sit at 2^32");
    expect(!idw::walk_offsets_fit(0), "no nodes");
    expect(idw::MAX_WALK_NODES == 67108863, "the limit's value");
    const double f = idw::mixed_near_wall_floor(3.0, 5.0, 0.25);
    // q^3 <= 1e32 and 0.25^5 q^5 <= 1e32 with q = 1/f: the second binds
    const double q = 1.0 / f;
    expect(std::fabs(std::pow(0.25, 5.0) * std::pow(q, 5.0) / 1e32 - 1.0) < 1e-12, "floor of the default exponents");
    expect(f > 9.9e-8 && f < 1e-7 && f < 1e-6, "default nearWallFactor above the floor");
    expect(idw_mixed_near_wall_floor(3.0, 5.0, 0.25) == f, "exported floor");
    expect(idw::mixed_near_wall_floor(0.0, 0.0, 0.25) == 0.0, "no powers, no floor");
    expect(std::fabs(idw::mixed_near_wall_floor(10.0, 5.0, 0.25) - std::pow(1e32, -0.1)) < 1e-15,
           "aExp 10 binds");
    return fails ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        int rc = 2;
        if (mode == "tree" && argc == 4) rc = tree(argv[2], argv[3]);
        else if (mode == "reduce" && argc == 4) rc = reduce(argv[2], argv[3]);
        else if (mode == "api") rc = api();
        else if (mode == "limits") rc = limits();
        else std::printf("usage: host_checks tree|reduce IN OUT | api | limits\n");
        if (rc == 0) std::printf("ok %s\n", mode.c_str());
        return rc;
    } catch (const std::string &e) {
        std::printf("error: %s\n", e.c_str());
        return 2;
    }
}
