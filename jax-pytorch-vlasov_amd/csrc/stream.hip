// STREAM-style HBM calibration kernels (copy / read / write of float64
// arrays with 16-byte lanes), used to report the measured bandwidth ceiling
// beside the nominal one.  The reference ships the same kind of benchmark
// (simulations/stream/pytorch/src/stream.py:92-105).
#include "vp_common.hpp"

namespace vp {

typedef double v2d __attribute__((ext_vector_type(2)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_copy_kernel(const v2d* __restrict__ a, v2d* __restrict__ b,
                                                          int64_t n2) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256 * U;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * U + threadIdx.x; base < n2; base += step) {
        v2d v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + u * 256;
            if (i < n2) v[u] = NT ? __builtin_nontemporal_load(a + i) : a[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + u * 256;
            if (i < n2) {
                if (NT)
                    __builtin_nontemporal_store(v[u], b + i);
                else
                    b[i] = v[u];
            }
        }
    }
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_read_kernel(const v2d* __restrict__ a, double* __restrict__ out,
                                                          int64_t n2) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256 * U;
    double acc = 0.0;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * U + threadIdx.x; base < n2; base += step) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + u * 256;
            if (i < n2) {
                const v2d v = NT ? __builtin_nontemporal_load(a + i) : a[i];
                acc += v.x + v.y;
            }
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_write_kernel(v2d* __restrict__ b, int64_t n2, double val) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256 * U;
    const v2d v = {val, val};
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * U + threadIdx.x; base < n2; base += step) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + u * 256;
            if (i < n2) {
                if (NT)
                    __builtin_nontemporal_store(v, b + i);
                else
                    b[i] = v;
            }
        }
    }
}

template <int U, bool NT>
static void launch_stream(int kind, const double* a, double* b, int64_t n2, int grid, hipStream_t s) {
    if (kind == 0)
        hipLaunchKernelGGL((stream_copy_kernel<U, NT>), dim3(grid), dim3(256), 0, s,
                           reinterpret_cast<const v2d*>(a), reinterpret_cast<v2d*>(b), n2);
    else if (kind == 1)
        hipLaunchKernelGGL((stream_read_kernel<U, NT>), dim3(grid), dim3(256), 0, s,
                           reinterpret_cast<const v2d*>(a), b, n2);
    else
        hipLaunchKernelGGL((stream_write_kernel<U, NT>), dim3(grid), dim3(256), 0, s,
                           reinterpret_cast<v2d*>(b), n2, 1.0);
}

}  // namespace vp

// kind: 0 copy a -> b (n doubles each), 1 read a (b receives `grid` partial
// sums), 2 write b.  unroll in {1, 2, 4, 8}; grid = number of 256-thread
// workgroups (grid-stride loop).  n must be even; a, b 16-byte aligned.
extern "C" int vp_stream_f64(int kind, const double* a, double* b, int64_t n, int unroll, int nontemporal, int grid,
                             vp_stream_t stream) {
    VP_RETURN_IF(kind < 0 || kind > 2 || grid < 1 || n < 2 || (n & 1), VP_ERR_ARG);
    VP_RETURN_IF((kind != 2 && !a) || !b, VP_ERR_ARG);
    VP_RETURN_IF((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15, VP_ERR_ARG);
    const int64_t n2 = n / 2;
    hipStream_t s = static_cast<hipStream_t>(stream);
#define VP_S(U)                                                            \
    if (nontemporal)                                                       \
        vp::launch_stream<U, true>(kind, a, b, n2, grid, s);               \
    else                                                                   \
        vp::launch_stream<U, false>(kind, a, b, n2, grid, s);
    switch (unroll) {
        case 1: VP_S(1) break;
        case 2: VP_S(2) break;
        case 4: VP_S(4) break;
        case 8: VP_S(8) break;
        default: return VP_ERR_ARG;
    }
#undef VP_S
    VP_CHECK_LAUNCH();
    return VP_OK;
}
