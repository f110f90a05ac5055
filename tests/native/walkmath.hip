// Test-only library: the fast-mode walk's device arithmetic (idw_walk_math.h)
// evaluated on arrays, so that tests can compare it bit for bit with sqrt()
// and with getError (kd_tree.F90:882-906) restated on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "idw_walk_math.h"

namespace {

__constant__ double2 t_rseg[idw::NERR - 1] = {
    {2.0, 1.0 / (3.0 - 2.0)},       {3.0, 1.0 / (4.0 - 3.0)},       {4.0, 1.0 / (5.0 - 4.0)},
    {5.0, 1.0 / (7.5 - 5.0)},       {7.5, 1.0 / (10.0 - 7.5)},      {10.0, 1.0 / (20.0 - 10.0)},
    {20.0, 1.0 / (40.0 - 20.0)},    {40.0, 1.0 / (80.0 - 40.0)},    {80.0, 1.0 / (160.0 - 80.0)},
    {160.0, 1.0 / (320.0 - 160.0)}, {320.0, 1.0 / (640.0 - 320.0)}};

__global__ void k_sqrt(const double *x, double *fast, double *lib, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fast[i] = idw::sqrt_normal(x[i]);
    lib[i] = sqrt(x[i]);
}

// err: n error tables of NERR+1 words (the table and the word after it);
// d[i] >= 2 evaluated against table i
__global__ void k_error(const double *err, const double *d, double *out, int64_t n) {
    __shared__ idw::SegCell cells[idw::ERR_CELLS];
    if (threadIdx.x < idw::ERR_CELLS) cells[threadIdx.x] = idw::seg_cell(threadIdx.x, t_rseg, 640.0);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // a NodeRec whose error table and Ai are the test's words
    idw::NodeRec r = {};
    for (int k = 0; k < idw::NERR; ++k) r.err[k] = err[i * (idw::NERR + 1) + k];
    r.Ai = err[i * (idw::NERR + 1) + idw::NERR];
    out[i] = idw::get_error_cells(&r, d[i], cells);
}

template <typename F>
int run(int64_t n, F launch) {
    if (launch() != hipSuccess) return 1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

}  // namespace

extern "C" int walkmath_sqrt(const double *hx, double *hfast, double *hlib, int64_t n) {
    double *d = nullptr;
    if (n <= 0) return 0;
    if (hipMalloc(&d, 3 * n * sizeof(double)) != hipSuccess) return 3;
    int rc = hipMemcpy(d, hx, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess ? 0 : 4;
    if (!rc)
        rc = run(n, [&] {
            k_sqrt<<<(unsigned)((n + 255) / 256), 256>>>(d, d + n, d + 2 * n, n);
            return hipGetLastError();
        });
    if (!rc && hipMemcpy(hfast, d + n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    if (!rc && hipMemcpy(hlib, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    hipFree(d);
    return rc;
}

extern "C" int walkmath_error(const double *herr, const double *hd, double *hout, int64_t n) {
    double *d = nullptr;
    if (n <= 0) return 0;
    const int64_t w = (idw::NERR + 1) * n;
    if (hipMalloc(&d, (w + 2 * n) * sizeof(double)) != hipSuccess) return 3;
    int rc = hipMemcpy(d, herr, w * sizeof(double), hipMemcpyHostToDevice) == hipSuccess ? 0 : 4;
    if (!rc && hipMemcpy(d + w, hd, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = 4;
    if (!rc)
        rc = run(n, [&] {
            k_error<<<(unsigned)((n + 255) / 256), 256>>>(d, d + w, d + w + n, n);
            return hipGetLastError();
        });
    if (!rc && hipMemcpy(hout, d + w + n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    hipFree(d);
    return rc;
}
