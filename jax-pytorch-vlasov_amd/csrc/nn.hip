// Kernels of the Fourier-operator surrogate (training and inference):
// a batched, strided GEMM on the matrix cores (BF16 operands, FP32
// accumulation: v_mfma_f32_16x16x16_bf16) and the elementwise pieces
// (GELU, bias, mean-squared error, Adam).  Every tensor is float32 in
// memory; GEMM operands are rounded to bfloat16 (round to nearest even)
// when staged into LDS, so weights and activations enter the MFMA as BF16.
//
// No reference counterpart: the surrogate is an addition of this library.
#include "vp_common.hpp"

namespace vp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16 | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<unsigned short>(u >> 16);
}

constexpr int kTile = 64;  // output tile (kTile x kTile) per workgroup
constexpr int kKStep = 16;

// C[b](i, j) = alpha * sum_k A[b](i, k) B[b](k, j) + beta * C[b](i, j)
// X[b](r, s) = X[b * sXb + r * sXr + s * sXs] (any strides: transposes are
// stride swaps).  256 threads = 4 waves; wave w owns the 32 x 32 quadrant
// (w / 2, w % 2) of the 64 x 64 tile as 2 x 2 MFMA blocks of 16 x 16.
// Operand layout of v_mfma_f32_16x16x16_bf16 (lane l, i = l % 16,
// q = l / 16): A row i, k = 4q..4q+3; B column i, k = 4q..4q+3; result
// column i, rows 4q..4q+3.
// Split K: grid.z = batch * splitk; split s covers k in [s * kc, (s+1) * kc)
// and, when splitk > 1, adds alpha * (its partial product) to C atomically
// (C was scaled by beta beforehand).  Staging reads 16 consecutive k (or 64
// consecutive rows) per instruction group, whichever stride is 1.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, int splitk, int kc, float alpha,
                                                        const float* __restrict__ A, int64_t sAb, int64_t sAi,
                                                        int64_t sAk, const float* __restrict__ B, int64_t sBb,
                                                        int64_t sBk, int64_t sBj, float beta, float* __restrict__ C,
                                                        int64_t sCb, int64_t sCi, int64_t sCj) {
    __shared__ unsigned short As[kTile][kKStep + 4];  // [i][k], padded rows
    __shared__ unsigned short Bs[kTile][kKStep + 4];  // [j][k]
    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const int b = blockIdx.z / splitk;
    const int split = blockIdx.z - b * splitk;
    const int kbeg = split * kc;
    const int kend = (kbeg + kc < K) ? kbeg + kc : K;
    const int i0 = blockIdx.y * kTile;
    const int j0 = blockIdx.x * kTile;
    const float* Ab = A + b * sAb;
    const float* Bb = B + b * sBb;
    const bool a_k_fast = sAk == 1 || sAi != 1;  // which index runs along consecutive threads
    const bool b_k_fast = sBk == 1 || sBj != 1;
    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int wi = (wave >> 1) * 32;
    const int wj = (wave & 1) * 32;
    const int li = lane & 15;
    const int lq = lane >> 4;
    for (int k0 = kbeg; k0 < kend; k0 += kKStep) {
        // stage 64 x 16 of A and 16 x 64 of B (4 elements per thread each), zero outside
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 256 * e;  // 0 .. 1023
            {
                const int r = a_k_fast ? idx >> 4 : idx & 63;
                const int k = a_k_fast ? idx & 15 : idx >> 6;
                const int gi = i0 + r, gk = k0 + k;
                As[r][k] = to_bf16((gi < M && gk < kend) ? Ab[gi * sAi + gk * sAk] : 0.f);
            }
            {
                const int r = b_k_fast ? idx >> 4 : idx & 63;
                const int k = b_k_fast ? idx & 15 : idx >> 6;
                const int gj = j0 + r, gk = k0 + k;
                Bs[r][k] = to_bf16((gj < N && gk < kend) ? Bb[gk * sBk + gj * sBj] : 0.f);
            }
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            bf16x4 a;
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = static_cast<short>(As[wi + 16 * x + li][4 * lq + u]);
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                bf16x4 bv;
#pragma unroll
                for (int u = 0; u < 4; ++u) bv[u] = static_cast<short>(Bs[wj + 16 * y + li][4 * lq + u]);
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[x][y], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* Cb = C + b * sCb;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gi = i0 + wi + 16 * x + 4 * lq + u;
                const int gj = j0 + wj + 16 * y + li;
                if (gi < M && gj < N) {
                    float* p = Cb + gi * sCi + gj * sCj;
                    const float v = alpha * acc[x][y][u];
                    if (splitk > 1)
                        atomicAdd(p, v);
                    else
                        *p = (beta == 0.f) ? v : v + beta * *p;
                }
            }
}

// C[b](i, j) = beta * C[b](i, j) (beta == 0: written as 0 without reading C)
__global__ void scale_kernel(float* __restrict__ C, int M, int N, int batch, int64_t sCb, int64_t sCi, int64_t sCj,
                             float beta) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t tot = static_cast<int64_t>(M) * N * batch;
    if (idx >= tot) return;
    const int j = static_cast<int>(idx % N);
    const int64_t rest = idx / N;
    const int i = static_cast<int>(rest % M);
    const int b = static_cast<int>(rest / M);
    float* p = C + b * sCb + i * sCi + j * sCj;
    *p = beta == 0.f ? 0.f : beta * *p;
}

__global__ void gelu_fwd_kernel(const float* __restrict__ x, float* __restrict__ y, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const float v = x[i];
        y[i] = 0.5f * v * (1.f + erff(v * 0.70710678118654752f));
    }
}

// dx = dy * (Phi(x) + x phi(x))
__global__ void gelu_bwd_kernel(const float* __restrict__ x, const float* __restrict__ dy, float* __restrict__ dx,
                                int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const float v = x[i];
        const float cdf = 0.5f * (1.f + erff(v * 0.70710678118654752f));
        const float pdf = 0.39894228040143268f * expf(-0.5f * v * v);
        dx[i] = dy[i] * (cdf + v * pdf);
    }
}

// y[r][c] += bias[c]
__global__ void bias_add_kernel(float* __restrict__ y, const float* __restrict__ bias, int64_t rows, int cols) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < rows * cols) y[i] += bias[i % cols];
}

// db[c] += sum over this workgroup's 512 rows of dy[r][c] (db zeroed first)
__global__ __launch_bounds__(256) void bias_grad_kernel(const float* __restrict__ dy, float* __restrict__ db,
                                                        int64_t rows, int cols) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rg = threadIdx.x >> 6;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 512;
    const int64_t r1 = r0 + 512 < rows ? r0 + 512 : rows;
    float s = 0.f;
    if (c < cols)
        for (int64_t r = r0 + rg; r < r1; r += 4) s += dy[r * cols + c];
    red[rg][threadIdx.x & 63] = s;
    __syncthreads();
    if (rg == 0 && c < cols)
        atomicAdd(&db[c], red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ void zero_kernel(float* __restrict__ p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// grad = 2 (pred - target) / n ; partial[block] = sum of squared errors of the block
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                  float* __restrict__ grad, float* __restrict__ partial, int64_t n) {
    __shared__ float red[256];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    float e2 = 0.f;
    if (i < n) {
        const float d = pred[i] - target[i];
        if (grad) grad[i] = 2.f * d / static_cast<float>(n);
        e2 = d * d;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Weighted squared error: partial[block] = sum of w d^2 over the block,
// grad = 2 w d (d = pred - target).  With w = 1/n it is mse_kernel.
__global__ __launch_bounds__(256) void wmse_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                   const float* __restrict__ w, float* __restrict__ grad,
                                                   float* __restrict__ partial, int64_t n) {
    __shared__ float red[256];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    float e2 = 0.f;
    if (i < n) {
        const float d = pred[i] - target[i];
        const float wd = w[i] * d;
        if (grad) grad[i] = 2.f * wd;
        e2 = wd * d;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (bias-corrected), float32 master weights.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Encoding of the surrogate's initial condition f0 [n][nx][nv] (float64):
// one workgroup per run reduces its f0 on the run's own grid (x_i = i L/nx,
// v_j = -Vmax + j dv, dv = 2 Vmax/(nv - 1); the mode's phase depends on i
// only, so L is not needed) to
//   out[r][0..2] = (1/nx) sum_i rho(x_i) (1, cos(2 pi i/nx), sin(2 pi i/nx)),
//                  rho(x_i) = sum_j f(x_i, v_j) dv   (density and its first
//                  spatial Fourier mode, times 1/nx; the phase has e^{-i})
//   out[r][3..5] = sum_j fbar(v_j) v_j^p dv, p = 0, 2, 4,
//                  fbar = (1/nx) sum_i f(x_i, .)      (velocity moments)
// Thread j owns velocity columns j, j + 256, ...; compensated sums.
__global__ __launch_bounds__(256) void f0_moments_kernel(const double* __restrict__ f0, int nx, int nv,
                                                         const double* __restrict__ vmax, double* __restrict__ out) {
    __shared__ double scratch[2 * 256 / kWave];
    const int r = blockIdx.x;
    const double* f = f0 + static_cast<int64_t>(r) * nx * nv;
    const double V = vmax[r];
    const double dv = 2.0 * V / static_cast<double>(nv - 1);
    KSum a0{0, 0}, ac{0, 0}, as{0, 0}, m0{0, 0}, m2{0, 0}, m4{0, 0};
    for (int j = threadIdx.x; j < nv; j += blockDim.x) {
        KSum c0{0, 0}, c1{0, 0}, s1{0, 0};
        for (int i = 0; i < nx; ++i) {
            const double v = f[static_cast<int64_t>(i) * nv + j];
            double sn, cs;
            sincospi(2.0 * static_cast<double>(i) / static_cast<double>(nx), &sn, &cs);
            ksum_add(c0, v);
            ksum_add(c1, v * cs);
            ksum_add(s1, -v * sn);
        }
        const double col = c0.s + c0.c;  // sum over x of f(., v_j)
        ksum_add(a0, col * dv);
        ksum_add(ac, (c1.s + c1.c) * dv);
        ksum_add(as, (s1.s + s1.c) * dv);
        const double vj = -V + static_cast<double>(j) * dv;
        const double fb = col / static_cast<double>(nx);
        ksum_add(m0, fb * dv);
        ksum_add(m2, fb * vj * vj * dv);
        ksum_add(m4, fb * (vj * vj) * (vj * vj) * dv);
    }
    const double inv = 1.0 / static_cast<double>(nx);
    const double r0 = block_ksum(a0, scratch) * inv;
    const double rc = block_ksum(ac, scratch) * inv;
    const double rs = block_ksum(as, scratch) * inv;
    const double q0 = block_ksum(m0, scratch);
    const double q2 = block_ksum(m2, scratch);
    const double q4 = block_ksum(m4, scratch);
    if (threadIdx.x == 0) {
        double* o = out + 6 * static_cast<int64_t>(r);
        o[0] = r0;
        o[1] = rc;
        o[2] = rs;
        o[3] = q0;
        o[4] = q2;
        o[5] = q4;
    }
}

static dim3 grid1d(int64_t n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

}  // namespace vp

extern "C" int vp_nn_gemm(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
                          int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
                          int64_t sCb, int64_t sCi, int64_t sCj, vp_stream_t stream) {
    VP_RETURN_IF(!A || !B || !C, VP_ERR_ARG);
    VP_RETURN_IF(batch < 1 || M < 1 || N < 1 || K < 1 || batch > 65535, VP_ERR_SHAPE);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int tm = (M + vp::kTile - 1) / vp::kTile, tn = (N + vp::kTile - 1) / vp::kTile;
    VP_RETURN_IF(tm > 65535, VP_ERR_SHAPE);
    // split K until the launch has ~512 workgroups, keeping >= 256 k per split
    const int64_t tiles = static_cast<int64_t>(tm) * tn * batch;
    int splitk = 1;
    while (tiles * splitk * 2 <= 512 && K / (splitk * 2) >= 256 && batch * splitk * 2 <= 65535) splitk *= 2;
    const int kc = ((K + splitk - 1) / splitk + vp::kKStep - 1) / vp::kKStep * vp::kKStep;
    if (splitk > 1) {
        const int64_t tot = static_cast<int64_t>(M) * N * batch;
        hipLaunchKernelGGL(vp::scale_kernel, vp::grid1d(tot), dim3(256), 0, s, C, M, N, batch, sCb, sCi, sCj, beta);
        VP_CHECK_LAUNCH();
    }
    const dim3 grid(tn, tm, batch * splitk);
    hipLaunchKernelGGL(vp::gemm_bf16_kernel, grid, dim3(256), 0, s, M, N, K, splitk, kc, alpha, A, sAb, sAi, sAk, B,
                       sBb, sBk, sBj, beta, C, sCb, sCi, sCj);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_f0_moments(const double* f0, int n, int nx, int nv, const double* vmax, double* out,
                                vp_stream_t stream) {
    VP_RETURN_IF(!f0 || !vmax || !out, VP_ERR_ARG);
    VP_RETURN_IF(n < 1 || nx < 1 || nv < 2 || n > 0x7fffffff / 2, VP_ERR_SHAPE);
    hipLaunchKernelGGL(vp::f0_moments_kernel, dim3(n), dim3(256), 0, static_cast<hipStream_t>(stream), f0, nx, nv,
                       vmax, out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_gelu(const float* x, float* y, int64_t n, vp_stream_t stream) {
    VP_RETURN_IF(!x || !y || n < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(vp::gelu_fwd_kernel, vp::grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), x, y, n);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_gelu_bwd(const float* x, const float* dy, float* dx, int64_t n, vp_stream_t stream) {
    VP_RETURN_IF(!x || !dy || !dx || n < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(vp::gelu_bwd_kernel, vp::grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), x, dy, dx,
                       n);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_bias_add(float* y, const float* bias, int64_t rows, int cols, vp_stream_t stream) {
    VP_RETURN_IF(!y || !bias || rows < 1 || cols < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(vp::bias_add_kernel, vp::grid1d(rows * cols), dim3(256), 0, static_cast<hipStream_t>(stream),
                       y, bias, rows, cols);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_bias_grad(const float* dy, float* db, int64_t rows, int cols, vp_stream_t stream) {
    VP_RETURN_IF(!dy || !db || rows < 1 || cols < 1, VP_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(vp::zero_kernel, vp::grid1d(cols), dim3(256), 0, s, db, static_cast<int64_t>(cols));
    VP_CHECK_LAUNCH();
    const dim3 grid((cols + 63) / 64, static_cast<unsigned>((rows + 511) / 512));
    hipLaunchKernelGGL(vp::bias_grad_kernel, grid, dim3(256), 0, s, dy, db, rows, cols);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// partial must hold ceil(n / 256) floats
extern "C" int vp_nn_mse(const float* pred, const float* target, float* grad, float* partial, int64_t n,
                         vp_stream_t stream) {
    VP_RETURN_IF(!pred || !target || !partial || n < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(vp::mse_kernel, vp::grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), pred, target,
                       grad, partial, n);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_wmse(const float* pred, const float* target, const float* w, float* grad, float* partial,
                          int64_t n, vp_stream_t stream) {
    VP_RETURN_IF(!pred || !target || !w || !partial || n < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(vp::wmse_kernel, vp::grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), pred, target,
                       w, grad, partial, n);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

extern "C" int vp_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                          float eps, int step, vp_stream_t stream) {
    VP_RETURN_IF(!w || !g || !m || !v || n < 1 || step < 1, VP_ERR_ARG);
    const float c1 = 1.f - powf(b1, static_cast<float>(step));
    const float c2 = 1.f - powf(b2, static_cast<float>(step));
    hipLaunchKernelGGL(vp::adam_kernel, vp::grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v, n,
                       lr, b1, b2, eps, c1, c2);
    VP_CHECK_LAUNCH();
    return VP_OK;
}
