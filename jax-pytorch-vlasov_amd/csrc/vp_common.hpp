// Shared device helpers: error plumbing, the 6-point Lagrange weights in
// the reference's power form, departure-point arithmetic and compensated
// (Neumaier) float64 accumulation.
//
// The whole library is compiled with -ffp-contract=off: every a*b+c below
// rounds twice, exactly like PyTorch-CPU eager evaluates the reference's
// expressions, so a sweep reproduces the reference element for element.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vlasov_amd.h"

#define VP_CHECK_LAUNCH()                                  \
    do {                                                   \
        hipError_t e_ = hipGetLastError();                 \
        if (e_ != hipSuccess) return static_cast<int>(e_); \
    } while (0)

#define VP_RETURN_IF(cond, code) \
    do {                         \
        if (cond) return (code); \
    } while (0)

#define VP_TRY_HIP(expr)                                       \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return static_cast<int>(e_);     \
    } while (0)

#define VP_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != VP_OK) return rc_; \
    } while (0)

namespace vp {

constexpr int kWave = 64;

// Launch-shape knob value (vp_tune_set; 0 = automatic), defined in fast2d.hip.
int knob(int k);

// ---------------------------------------------------------------------------
// Lagrange weights on nodes -2..3 (lagrange5_interp_sum, vlasov1D1V.py:163-189;
// lagrange5_weights, vlasov1D1V.py:191-210).  Same divisions by literal
// constants and the same left-to-right association as the reference, so the
// float64 results are bit-identical to PyTorch-CPU eager.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void lagrange5_weights(T a, T w[6]) {
    const T a2 = a * a;
    const T a3 = a2 * a;
    const T a4 = a3 * a;
    const T a5 = a4 * a;
    w[0] = (a / T(20)) - (a2 / T(24)) - (a3 / T(24)) + (a4 / T(24)) - (a5 / T(120));
    w[1] = -(a / T(2)) + (T(2) * a2 / T(3)) - (a3 / T(24)) - (a4 / T(6)) + (a5 / T(24));
    w[2] = T(1) - (a / T(3)) - (T(5) * a2 / T(4)) + (T(5) * a3 / T(12)) + (a4 / T(4)) - (a5 / T(12));
    w[3] = a + (T(2) * a2 / T(3)) - (T(7) * a3 / T(12)) - (a4 / T(6)) + (a5 / T(12));
    w[4] = -(a / T(4)) - (a2 / T(24)) + (T(7) * a3 / T(24)) + (a4 / T(24)) - (a5 / T(24));
    w[5] = (a / T(30)) - (a3 / T(24)) + (a5 / T(120));
}

// Departure point of grid index n for velocity vel:
// dep = n - (vel*dt)/dx ; il = floor(dep) ; alpha = dep - il
// (advect_1d_vectorized, vlasov1D1V.py:229-238).
struct Departure {
    int64_t il;
    double alpha;
};

__device__ __forceinline__ Departure departure(int n, double vel, double dt, double dx) {
    const double disp = (vel * dt) / dx;
    const double dep = static_cast<double>(n) - disp;
    const double fl = floor(dep);
    Departure d;
    d.alpha = dep - fl;
    d.il = static_cast<int64_t>(fl);
    return d;
}

// j mod n in [0, n).  Taps lie within one period of the line almost always:
// a compare and an add there, the 64-bit remainder (tens of instructions)
// only beyond.
__device__ __forceinline__ int wrap_index(int64_t j, int n) {
    const int64_t nn = n;
    if (j >= -nn && j < 2 * nn) return static_cast<int>(j < 0 ? j + nn : (j >= nn ? j - nn : j));
    int64_t r = j % n;
    return static_cast<int>(r < 0 ? r + n : r);
}

// 6-tap sum w0*f0 + ... + w5*f5, left to right (vlasov1D1V.py:182-189).
template <typename T>
__device__ __forceinline__ T tap_sum(const T w[6], const T v[6]) {
    return w[0] * v[0] + w[1] * v[1] + w[2] * v[2] + w[3] * v[3] + w[4] * v[4] + w[5] * v[5];
}

// Float64 weights in the storage type S of a sweep.  double: as computed.
// float (FP32 storage): each weight rounded to nearest, then the one of
// smallest magnitude replaced by 1 minus the (exact, float64) sum of the
// other five, rounded to float, so the six float weights sum to 1 within
// half an ulp of the smallest weight instead of within a few ulp of the
// largest.  Without that the rounding bias of the weights (~1e-7) scales f
// by a velocity-dependent factor in every sweep, a systematic drift that
// the field energy integrates over hundreds of steps.
template <typename S>
__device__ __forceinline__ void round_weights(const double w[6], S out[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = static_cast<S>(w[k]);
}

template <>
__device__ __forceinline__ void round_weights<float>(const double w[6], float out[6]) {
    int m = 0;
    double am = fabs(w[0]);
#pragma unroll
    for (int k = 1; k < 6; ++k) {
        const bool smaller = fabs(w[k]) < am;
        am = smaller ? fabs(w[k]) : am;
        m = smaller ? k : m;
    }
    double rest = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        out[k] = static_cast<float>(w[k]);
        rest += (k == m) ? 0.0 : static_cast<double>(out[k]);
    }
    const float bal = static_cast<float>(1.0 - rest);
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = (k == m) ? bal : out[k];
}

// ---------------------------------------------------------------------------
// Neumaier compensated summation.
// ---------------------------------------------------------------------------
struct KSum {
    double s;
    double c;
};

// Neumaier's compensated sum: c collects the exact rounding error of every
// s + x.  The error is computed by Knuth's branch-free TwoSum, which yields
// the same exact value as Neumaier's |s| >= |x| branches (both are exact),
// without the comparison and selects.
__device__ __forceinline__ void ksum_add(KSum& k, double x) {
    const double t = k.s + x;
    const double z = t - k.s;
    k.c += (k.s - (t - z)) + (x - z);
    k.s = t;
}

__device__ __forceinline__ void ksum_merge(KSum& a, const KSum& b) {
    ksum_add(a, b.s);
    a.c += b.c;
}

__device__ __forceinline__ KSum wave_ksum(KSum k) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
        KSum o;
        o.s = __shfl_xor(k.s, off, kWave);
        o.c = __shfl_xor(k.c, off, kWave);
        ksum_merge(k, o);
    }
    return k;
}

// A float64 moved across lanes by DPP (both 32-bit halves), a register move
// with no LDS round trip; lanes the row mask leaves out keep `old`.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double old, double x) {
    const long long xb = __double_as_longlong(x), ob = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(ob), static_cast<int>(xb), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(ob >> 32), static_cast<int>(xb >> 32), CTRL, ROW_MASK,
                                               0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Sum over each aligned group of `width` (1, 2, 4, ..., 64; wave-uniform)
// consecutive lanes by DPP moves, in registers: pairs (quad permutation),
// pairs of pairs, 8 lanes (half-row mirror), 16 lanes (row mirror); every
// lane of a group of up to 16 ends with the group's sum.  Groups of 32 add
// row 0 into row 1 and row 2 into row 3 (row_bcast:15), groups of 64 then
// rows 0-1 into rows 2-3 (row_bcast:31): there the group's last lane holds
// the sum.  Merges are compensated when COMP.
template <bool COMP>
__device__ __forceinline__ KSum group_ksum(KSum k, int width) {
    auto step = [&](double os, double oc) {
        if (COMP) {
            ksum_merge(k, KSum{os, oc});
        } else {
            k.s += os;
        }
    };
    if (width >= 2) step(dpp_f64<0xB1>(0.0, k.s), COMP ? dpp_f64<0xB1>(0.0, k.c) : 0.0);    // quad_perm [1, 0, 3, 2]
    if (width >= 4) step(dpp_f64<0x4E>(0.0, k.s), COMP ? dpp_f64<0x4E>(0.0, k.c) : 0.0);    // quad_perm [2, 3, 0, 1]
    if (width >= 8) step(dpp_f64<0x141>(0.0, k.s), COMP ? dpp_f64<0x141>(0.0, k.c) : 0.0);  // row_half_mirror
    if (width >= 16) step(dpp_f64<0x140>(0.0, k.s), COMP ? dpp_f64<0x140>(0.0, k.c) : 0.0);  // row_mirror
    if (width >= 32)
        step(dpp_f64<0x142, 0xA>(0.0, k.s), COMP ? dpp_f64<0x142, 0xA>(0.0, k.c) : 0.0);  // row_bcast:15 into rows 1, 3
    if (width >= 64)
        step(dpp_f64<0x143, 0xC>(0.0, k.s), COMP ? dpp_f64<0x143, 0xC>(0.0, k.c) : 0.0);  // row_bcast:31 into rows 2, 3
    return k;
}

// Row sums of the R rows every lane holds (acc[R], R = 1, 2 or 4) over aligned
// groups of W = 32 or 64 lanes, halving the rows a lane holds while reducing:
// at the exchanges with the lane one and two apart (quad permutations) the
// lane keeps half of its rows (the lower half where the exchange's lane bit
// is clear) and merges the partner's copy of that half, so each step merges
// at most R / 2 rows instead of R.  Row 2 b0 + b1 (R = 4; row b0 for R = 2)
// ends in the lanes whose low bits are b0, b1; those lanes then merge over
// row rotations by 4 and 8 (which keep the lane index mod 4) and across 16
// (and 32) lanes by cross-row swaps.  The result: lane j < R of each group holds the sum of row
// rows_ksum_row<R>(j).  Compensated merges when COMP.
template <int R>
__device__ __forceinline__ int rows_ksum_row(int j) {
    return R == 4 ? 2 * (j & 1) + ((j >> 1) & 1) : (R == 2 ? (j & 1) : 0);
}

template <bool COMP>
__device__ __forceinline__ void ksum_merge_c(KSum& a, const KSum& b) {
    if (COMP) {
        ksum_merge(a, b);
    } else {
        a.s += b.s;
    }
}

template <int CTRL, bool COMP>
__device__ __forceinline__ KSum dpp_ksum(const KSum& k) {
    return KSum{dpp_f64<CTRL>(0.0, k.s), COMP ? dpp_f64<CTRL>(0.0, k.c) : 0.0};
}

// One halving exchange: keep rows [0, M/2) (upper == false) or [M/2, M) and
// merge the partner's copy of them; the kept rows move to r[0 .. M/2).
template <int CTRL, bool COMP, int M>
__device__ __forceinline__ void rows_split(KSum* r, bool upper) {
    constexpr int H = M / 2;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const KSum keep = upper ? r[H + i] : r[i];
        const KSum send = upper ? r[i] : r[H + i];
        KSum k = keep;
        ksum_merge_c<COMP>(k, dpp_ksum<CTRL, COMP>(send));
        r[i] = k;
    }
}

// The value of lane ^ 16 (ROWS16) or lane ^ 32, by the gfx950 cross-row
// swaps (VALU, no LDS round trip).  With x as both operands,
// v_permlane16_swap returns every 16-lane row filled with the even rows' values
// (first result) and with the odd rows' (second); v_permlane32_swap likewise
// with the lower and upper halves of the wave.  A lane takes the other kind.
template <bool ROWS16>
__device__ __forceinline__ double lane_swap_f64(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = static_cast<unsigned>(b), hi = static_cast<unsigned>(b >> 32);
    const bool upper = ROWS16 ? (lane & 16) != 0 : (lane & 32) != 0;
    unsigned plo, phi;
    if constexpr (ROWS16) {
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        plo = upper ? l[0] : l[1];
        phi = upper ? h[0] : h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        plo = upper ? l[0] : l[1];
        phi = upper ? h[0] : h[1];
    }
    return __longlong_as_double((static_cast<long long>(phi) << 32) | plo);
}

template <bool COMP, int R, int W>
__device__ __forceinline__ KSum rows_ksum(KSum (&acc)[R], int lane) {
    static_assert(R == 1 || R == 2 || R == 4, "1, 2 or 4 rows per lane");
    static_assert(W == 32 || W == 64, "groups of 32 or 64 lanes");
    KSum r[R];
#pragma unroll
    for (int i = 0; i < R; ++i) r[i] = acc[i];
    if constexpr (R >= 2) rows_split<0xB1, COMP, R>(r, (lane & 1) != 0);  // partner: lane ^ 1
    if constexpr (R == 4) rows_split<0x4E, COMP, 2>(r, (lane & 2) != 0);  // partner: lane ^ 2
    KSum k = r[0];
    if constexpr (R == 1) ksum_merge_c<COMP>(k, dpp_ksum<0xB1, COMP>(k));  // lane ^ 1
    if constexpr (R <= 2) ksum_merge_c<COMP>(k, dpp_ksum<0x4E, COMP>(k));  // lane ^ 2
    ksum_merge_c<COMP>(k, dpp_ksum<0x124, COMP>(k));                      // row_ror:4
    ksum_merge_c<COMP>(k, dpp_ksum<0x128, COMP>(k));                      // row_ror:8
    ksum_merge_c<COMP>(k, KSum{lane_swap_f64<true>(k.s, lane), COMP ? lane_swap_f64<true>(k.c, lane) : 0.0});
    if constexpr (W == 64)
        ksum_merge_c<COMP>(k, KSum{lane_swap_f64<false>(k.s, lane), COMP ? lane_swap_f64<false>(k.c, lane) : 0.0});
    return k;
}

// Block-wide compensated sum; every thread gets the result.  `scratch` must
// hold 2 * (blockDim.x / 64) doubles.
__device__ __forceinline__ double block_ksum(KSum k, double* scratch) {
    k = wave_ksum(k);
    const int lane = threadIdx.x % kWave;
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) {
        scratch[2 * wid] = k.s;
        scratch[2 * wid + 1] = k.c;
    }
    __syncthreads();
    KSum t{0.0, 0.0};
    for (int w = 0; w < nw; ++w) {
        KSum o{scratch[2 * w], scratch[2 * w + 1]};
        ksum_merge(t, o);
    }
    __syncthreads();
    return t.s + t.c;
}

}  // namespace vp
