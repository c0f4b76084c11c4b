// Deviation storage for the FP32-storage variant (no reference counterpart):
// f(x, y, vx, vy) = feq(vx, vy) + g(x, y, vx, vy) with g stored in float32 and
// feq a separable equilibrium, feq(vx, vy) = fx(vx) fy(vy) (the float64
// product, rounded once), from float64 profiles fx[nvx], fy[nvy]: the
// rank-one part of the (x, y) mean of the initial f, fx = its sums over vy
// and fy = its sums over vx divided by its total (for a mean that is a
// product a(vx) b(vy), as for every initial condition of the reference,
// fx fy is that mean).  feq is constant in x and y, so the drifts leave it
// unchanged and run on g alone; the kick K is linear and maps a product
// fx(vx) fy(vy) to (Kvx fx)(vx) (Kvy fy)(vy) (each velocity kick acts on its
// own profile), so the kick of f = feq + g is K(g) + (Kvx fx)(Kvy fy) - fx fy:
// the float32 kick of g plus a float64 outer-product correction per (x, y)
// plane (kick_rows_kernel with DEV).  Storage and arithmetic rounding then
// scale with |g| (the perturbation) instead of |f|.
//
//   split: M[m] = (sum_p f[p][m]) / P (compensated), fx[i] = sum_j M[i][j],
//          fy[j] = (sum_i M[i][j]) / sum_ij M[i][j], g = float(f - fx[i] fy[j]),
//          rho_shift[0] = ((sum_ij fx[i] fy[j]) * dvx) * dvy - 1
//   join:  f = fx[i] fy[j] + g (float64 and/or rounded to float32)
// with p = x * ny + y < P = nx * ny and m = i * nvy + j < M = nvx * nvy; the
// profiles are stored together, feq[0 .. nvx) = fx, feq[nvx .. nvx + nvy) = fy.
#include "vp_common.hpp"

namespace vp {

constexpr int kDevBlocks = 64;  // plane blocks of the column sums

// partial[b][m] = compensated sum of f[p][m] over the b-th block of planes.
__global__ __launch_bounds__(256) void dev_colsum_kernel(const double* __restrict__ f, int64_t P, int64_t M,
                                                         double2* __restrict__ partial) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (m >= M) return;
    const int64_t b = blockIdx.y;
    const int64_t p0 = b * P / kDevBlocks, p1 = (b + 1) * P / kDevBlocks;
    KSum k{0.0, 0.0};
    for (int64_t p = p0; p < p1; ++p) ksum_add(k, f[p * M + m]);
    partial[b * M + m] = make_double2(k.s, k.c);
}

// mean[m] = (merged block partials) / P.
__global__ __launch_bounds__(256) void dev_mean_kernel(const double2* __restrict__ partial, int64_t P, int64_t M,
                                                       double* __restrict__ mean) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (m >= M) return;
    KSum k{0.0, 0.0};
    for (int b = 0; b < kDevBlocks; ++b) {
        const double2 v = partial[b * M + m];
        ksum_merge(k, KSum{v.x, v.y});
    }
    mean[m] = (k.s + k.c) / static_cast<double>(P);
}

// One workgroup: the profiles from the mean, then rho_shift.
__global__ __launch_bounds__(1024) void dev_profiles_kernel(const double* __restrict__ mean, int nvx, int nvy,
                                                            double dvx, double dvy, double* __restrict__ feq,
                                                            double* __restrict__ rho_shift) {
    __shared__ double scratch[2 * 1024 / kWave];
    double* fx = feq;
    double* fy = feq + nvx;
    // the total of the mean
    KSum k{0.0, 0.0};
    for (int m = threadIdx.x; m < nvx * nvy; m += 1024) ksum_add(k, mean[m]);
    const double total = block_ksum(k, scratch);  // every thread gets it
    for (int i = threadIdx.x; i < nvx; i += 1024) {
        KSum r{0.0, 0.0};
        for (int j = 0; j < nvy; ++j) ksum_add(r, mean[static_cast<int64_t>(i) * nvy + j]);
        fx[i] = r.s + r.c;
    }
    for (int j = threadIdx.x; j < nvy; j += 1024) {
        KSum c{0.0, 0.0};
        for (int i = 0; i < nvx; ++i) ksum_add(c, mean[static_cast<int64_t>(i) * nvy + j]);
        fy[j] = (c.s + c.c) / total;
    }
    __syncthreads();  // fx, fy written
    // rho shift from the products as the kernels form them
    KSum t{0.0, 0.0};
    for (int m = threadIdx.x; m < nvx * nvy; m += 1024) ksum_add(t, fx[m / nvy] * fy[m % nvy]);
    const double sum = block_ksum(t, scratch);
    if (threadIdx.x == 0) rho_shift[0] = ((sum * dvx) * dvy) - 1.0;
}

__global__ __launch_bounds__(256) void dev_split_kernel(const double* __restrict__ f, const double* __restrict__ feq,
                                                        int64_t n, int nvx, int nvy, float* __restrict__ g) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256;
    const int64_t M = static_cast<int64_t>(nvx) * nvy;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += step) {
        const int64_t m = i % M;
        g[i] = static_cast<float>(f[i] - feq[m / nvy] * feq[nvx + m % nvy]);
    }
}

__global__ __launch_bounds__(256) void dev_join_kernel(const float* __restrict__ g, const double* __restrict__ feq,
                                                       int64_t n, int nvx, int nvy, float* __restrict__ out32,
                                                       double* __restrict__ out64) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256;
    const int64_t M = static_cast<int64_t>(nvx) * nvy;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += step) {
        const int64_t m = i % M;
        const double h = feq[m / nvy] * feq[nvx + m % nvy] + static_cast<double>(g[i]);
        if (out64) out64[i] = h;
        if (out32) out32[i] = static_cast<float>(h);
    }
}

static int elementwise_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return static_cast<int>(b < 65536 ? b : 65536);
}

}  // namespace vp

static bool dev_shape_ok(const int64_t shape[4]) {
    return shape && shape[0] >= 1 && shape[1] >= 1 && shape[2] >= 1 && shape[3] >= 1;
}

extern "C" size_t vp_dev_split_workspace(const int64_t shape[4]) {
    if (!dev_shape_ok(shape)) return 0;
    const size_t M = static_cast<size_t>(shape[2]) * shape[3];
    return static_cast<size_t>(vp::kDevBlocks) * M * sizeof(double2) + M * sizeof(double);
}

extern "C" int vp_dev_split_f32(const double* f, const int64_t shape[4], double dvx, double dvy, double* feq,
                                double* rho_shift, float* g, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_RETURN_IF(!f || !feq || !rho_shift || !g || !ws, VP_ERR_ARG);
    VP_RETURN_IF(!dev_shape_ok(shape), VP_ERR_SHAPE);
    VP_RETURN_IF(shape[2] > (1 << 15) || shape[3] > (1 << 15), VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < vp_dev_split_workspace(shape), VP_ERR_WORKSPACE);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t P = shape[0] * shape[1], M = shape[2] * shape[3];
    const int nvx = static_cast<int>(shape[2]), nvy = static_cast<int>(shape[3]);
    double2* partial = static_cast<double2*>(ws);
    double* mean = reinterpret_cast<double*>(partial + static_cast<int64_t>(vp::kDevBlocks) * M);
    const unsigned mb = static_cast<unsigned>((M + 255) / 256);
    hipLaunchKernelGGL(vp::dev_colsum_kernel, dim3(mb, vp::kDevBlocks), dim3(256), 0, s, f, P, M, partial);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(vp::dev_mean_kernel, dim3(mb), dim3(256), 0, s, partial, P, M, mean);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(vp::dev_profiles_kernel, dim3(1), dim3(1024), 0, s, mean, nvx, nvy, dvx, dvy, feq,
                       rho_shift);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(vp::dev_split_kernel, dim3(vp::elementwise_grid(P * M)), dim3(256), 0, s, f, feq, P * M, nvx,
                       nvy, g);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_dev_join_f32(const float* g, const double* feq, const int64_t shape[4], float* f32_out,
                               double* f64_out, vp_stream_t stream) {
    VP_RETURN_IF(!g || !feq || (!f32_out && !f64_out), VP_ERR_ARG);
    VP_RETURN_IF(!dev_shape_ok(shape), VP_ERR_SHAPE);
    const int64_t P = shape[0] * shape[1], M = shape[2] * shape[3];
    hipLaunchKernelGGL(vp::dev_join_kernel, dim3(vp::elementwise_grid(P * M)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g, feq, P * M, static_cast<int>(shape[2]),
                       static_cast<int>(shape[3]), f32_out, f64_out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}
