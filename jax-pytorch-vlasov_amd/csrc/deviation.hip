// Deviation storage for the FP32-storage variant (no reference counterpart):
// f(x, y, vx, vy) = feq(vx, vy) + g(x, y, vx, vy) with feq the (x, y) mean of
// the initial f rounded to float32 and g stored in float32.  feq is defined
// as those float32 values (every use widens them exactly to float64), so its
// storage loses nothing; g absorbs the difference.  feq is constant in
// x and y, so the drifts leave it unchanged and run on g alone; the kick
// evaluates feq + g in float64 (kick_rows_kernel with DEV).  Storage rounding
// then scales with |g| (the perturbation) instead of |f|.
//
//   split: feq[m] = float((sum_p f[p][m]) / P) (compensated sum),
//          g = float(f - feq), rho_shift[0] = ((sum_m feq[m]) * dvx) * dvy - 1
//   join:  f = feq + g (float64 and/or rounded to float32)
// with p = x * ny + y < P = nx * ny and m = vx * nvy + vy < M = nvx * nvy.
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

// One workgroup: feq[m] from the block partials, then rho_shift.
__global__ __launch_bounds__(1024) void dev_feq_kernel(const double2* __restrict__ partial, int64_t P, int64_t M,
                                                       double dvx, double dvy, float* __restrict__ feq,
                                                       double* __restrict__ rho_shift) {
    __shared__ double red_s[1024];
    __shared__ double red_c[1024];
    KSum tot{0.0, 0.0};
    for (int64_t m = threadIdx.x; m < M; m += 1024) {
        KSum k{0.0, 0.0};
        for (int b = 0; b < kDevBlocks; ++b) {
            const double2 v = partial[b * M + m];
            ksum_merge(k, KSum{v.x, v.y});
        }
        const float e = static_cast<float>((k.s + k.c) / static_cast<double>(P));
        feq[m] = e;
        ksum_add(tot, static_cast<double>(e));
    }
    red_s[threadIdx.x] = tot.s;
    red_c[threadIdx.x] = tot.c;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) {
            KSum a{red_s[threadIdx.x], red_c[threadIdx.x]};
            ksum_merge(a, KSum{red_s[threadIdx.x + h], red_c[threadIdx.x + h]});
            red_s[threadIdx.x] = a.s;
            red_c[threadIdx.x] = a.c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) rho_shift[0] = (((red_s[0] + red_c[0]) * dvx) * dvy) - 1.0;
}

__global__ __launch_bounds__(256) void dev_split_kernel(const double* __restrict__ f, const float* __restrict__ feq,
                                                        int64_t n, int64_t M, float* __restrict__ g) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += step)
        g[i] = static_cast<float>(f[i] - static_cast<double>(feq[i % M]));
}

__global__ __launch_bounds__(256) void dev_join_kernel(const float* __restrict__ g, const float* __restrict__ feq,
                                                       int64_t n, int64_t M, float* __restrict__ out32,
                                                       double* __restrict__ out64) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += step) {
        const double h = static_cast<double>(feq[i % M]) + static_cast<double>(g[i]);
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
    return static_cast<size_t>(vp::kDevBlocks) * shape[2] * shape[3] * sizeof(double2);
}

extern "C" int vp_dev_split_f32(const double* f, const int64_t shape[4], double dvx, double dvy, float* feq,
                                double* rho_shift, float* g, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_RETURN_IF(!f || !feq || !rho_shift || !g || !ws, VP_ERR_ARG);
    VP_RETURN_IF(!dev_shape_ok(shape), VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < vp_dev_split_workspace(shape), VP_ERR_WORKSPACE);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t P = shape[0] * shape[1], M = shape[2] * shape[3];
    double2* partial = static_cast<double2*>(ws);
    hipLaunchKernelGGL(vp::dev_colsum_kernel, dim3(static_cast<unsigned>((M + 255) / 256), vp::kDevBlocks),
                       dim3(256), 0, s, f, P, M, partial);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(vp::dev_feq_kernel, dim3(1), dim3(1024), 0, s, partial, P, M, dvx, dvy, feq, rho_shift);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(vp::dev_split_kernel, dim3(vp::elementwise_grid(P * M)), dim3(256), 0, s, f, feq, P * M, M,
                       g);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_dev_join_f32(const float* g, const float* feq, const int64_t shape[4], float* f32_out,
                               double* f64_out, vp_stream_t stream) {
    VP_RETURN_IF(!g || !feq || (!f32_out && !f64_out), VP_ERR_ARG);
    VP_RETURN_IF(!dev_shape_ok(shape), VP_ERR_SHAPE);
    const int64_t P = shape[0] * shape[1], M = shape[2] * shape[3];
    hipLaunchKernelGGL(vp::dev_join_kernel, dim3(vp::elementwise_grid(P * M)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g, feq, P * M, M, f32_out, f64_out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}
