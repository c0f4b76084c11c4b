// Directional sweep with the Kokkos implementation's semantics, for
// cross-checking this library against Kokkos runs (vlasov_poisson.hpp of
// simulations/vlasov2d_2v/kokkos, K2 below).  It differs from the PyTorch
// form the rest of the library follows in three places:
//   * the departure point is a physical coordinate, grid[i] - vel * dt
//     (K2:463-464, K2:592-593), not an index;
//   * periodic lines wrap it with fmod by the box length L after
//     subtracting xmin (K2:293-316); Dirichlet lines return 0 when it lies
//     outside [xmin, xmin + L] and read taps outside [0, n) as 0
//     (K2:318-339);
//   * the weights are the degree-5 Lagrange basis at nodes -2..3 in Horner
//     form (lagrange_interp_5_uniform, K2:244-291), summed left to right.
// One thread per output element, taps read from global memory: a
// cross-checking mode, not a hot path.
#include "vp_common.hpp"

namespace vp {

// Horner-form basis of the 6-point uniform Lagrange interpolant at x in
// [0, 1) (node offsets -2..3), the products in K2's order.
__device__ __forceinline__ double kokkos_interp(const double y[6], double x) {
    const double b0 = x * (1.0 / 20.0 + x * (-1.0 / 24.0 + x * (-1.0 / 24.0 + x * (1.0 / 24.0 + x * (-1.0 / 120.0)))));
    const double b1 = x * (-1.0 / 2.0 + x * (2.0 / 3.0 + x * (-1.0 / 24.0 + x * (-1.0 / 6.0 + x * (1.0 / 24.0)))));
    const double b2 = 1.0 + x * (-1.0 / 3.0 + x * (-5.0 / 4.0 + x * (5.0 / 12.0 + x * (1.0 / 4.0 + x * (-1.0 / 12.0)))));
    const double b3 = x * (1.0 + x * (2.0 / 3.0 + x * (-7.0 / 12.0 + x * (-1.0 / 6.0 + x * (1.0 / 12.0)))));
    const double b4 = x * (-1.0 / 4.0 + x * (-1.0 / 24.0 + x * (7.0 / 24.0 + x * (1.0 / 24.0 + x * (-1.0 / 24.0)))));
    const double x2 = x * x;
    const double b5 = x * (1.0 / 30.0 + x2 * (-1.0 / 24.0 + x2 * (1.0 / 120.0)));
    return b0 * y[0] + b1 * y[1] + b2 * y[2] + b3 * y[3] + b4 * y[4] + b5 * y[5];
}

struct KokkosSweep {
    int64_t shape[4];
    int64_t vs[4];
    int axis;
    int64_t stride;  // elements between neighbours along the swept axis
    double dt, L, xmin, inv_dx;
};

template <bool PERIODIC>
__global__ __launch_bounds__(256) void kokkos_sweep_kernel(const double* __restrict__ fin, double* __restrict__ fout,
                                                           const double* __restrict__ grid,
                                                           const double* __restrict__ vel, KokkosSweep K,
                                                           int64_t total) {
    const int n = static_cast<int>(K.shape[K.axis]);
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t rem = e, voff = 0;
        int i = 0;
        for (int d = 3; d >= 0; --d) {
            const int64_t id = rem % K.shape[d];
            rem /= K.shape[d];
            voff += id * K.vs[d];
            if (d == K.axis) i = static_cast<int>(id);
        }
        const double* line = fin + (e - static_cast<int64_t>(i) * K.stride);
        const double dp = grid[i] - vel[voff] * K.dt;
        double y[6];
        double out;
        if (PERIODIC) {
            const double m = fmod(dp - K.xmin, K.L);
            const int il = static_cast<int>(floor(m * K.inv_dx));
            const double x = m * K.inv_dx - static_cast<double>(il);
            int first = il - 2;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                first += first < 0 ? n : 0;  // K2:305-310: re-tested for every tap
                int idx = first + k;
                idx -= idx >= n ? n : 0;
                // K2 reads outside the line here when the departure point lies
                // more than L - 2 dx below xmin (or n < 6, or inv_dx is not
                // n / L); this wraps it instead (identity for every index
                // K2 keeps in [0, n))
                idx %= n;
                idx += idx < 0 ? n : 0;
                y[k] = line[static_cast<int64_t>(idx) * K.stride];
            }
            out = kokkos_interp(y, x);
        } else {
            const double sft = dp - K.xmin;
            if (!(sft >= 0.0 && sft <= K.L)) {
                out = 0.0;
            } else {
                const int il = static_cast<int>(floor(sft * K.inv_dx));
                const double x = sft * K.inv_dx - static_cast<double>(il);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int idx = il + k - 2;
                    y[k] = (idx >= 0 && idx < n) ? line[static_cast<int64_t>(idx) * K.stride] : 0.0;
                }
                out = kokkos_interp(y, x);
            }
        }
        fout[e] = out;
    }
}

}  // namespace vp

extern "C" int vp_advect_kokkos_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                                    const double* grid, const double* vel, const int64_t vel_strides[4],
                                    double dt, double L, double xmin, double inv_dx, int periodic,
                                    vp_stream_t stream) {
    VP_RETURN_IF(!f_in || !f_out || !shape || !grid || !vel || !vel_strides, VP_ERR_ARG);
    VP_RETURN_IF(axis < 0 || axis > 3, VP_ERR_ARG);
    VP_RETURN_IF(f_in == f_out, VP_ERR_ARG);
    vp::KokkosSweep K;
    int64_t total = 1;
    for (int d = 0; d < 4; ++d) {
        VP_RETURN_IF(shape[d] < 1, VP_ERR_SHAPE);
        K.shape[d] = shape[d];
        K.vs[d] = vel_strides[d];
        total *= shape[d];
    }
    VP_RETURN_IF(shape[axis] > 0x7fffffffLL, VP_ERR_SHAPE);
    VP_RETURN_IF(!(L > 0.0) || !(inv_dx > 0.0), VP_ERR_ARG);
    K.axis = axis;
    K.stride = 1;
    for (int d = axis + 1; d < 4; ++d) K.stride *= shape[d];
    K.dt = dt;
    K.L = L;
    K.xmin = xmin;
    K.inv_dx = inv_dx;
    const int64_t want = (total + 255) / 256;
    const int blocks = static_cast<int>(want < 65536 ? want : 65536);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (periodic)
        hipLaunchKernelGGL(vp::kokkos_sweep_kernel<true>, dim3(blocks), dim3(256), 0, s, f_in, f_out, grid, vel, K,
                           total);
    else
        hipLaunchKernelGGL(vp::kokkos_sweep_kernel<false>, dim3(blocks), dim3(256), 0, s, f_in, f_out, grid, vel, K,
                           total);
    VP_CHECK_LAUNCH();
    return VP_OK;
}
