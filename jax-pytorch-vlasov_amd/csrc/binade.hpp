// Lagrange weights of the departure form from per-line binade tables.
//
// A sweep whose velocity is constant along a line has one displacement
// disp = (vel dt)/dx for the whole line; element r departs from
// t = RN(r - disp).  Within a binade of t the rounding grid is uniform and r
// lies on it, so frac(t), and with it the six weights, depends only on the
// binade and disp: one weight set per binade, computed from any r whose t
// falls in it, equals the element-wise computation bit for bit, without
// its divisions.  Shared by the generic sweeps and the ensemble.
#pragma once
#include "vp_common.hpp"

namespace vp {

template <typename T>
__device__ __forceinline__ void weights_as(double alpha, T w[6]) {
    double wd[6];
    lagrange5_weights<double>(alpha, wd);
    round_weights<T>(wd, w);
}

// Binade weight tables of the departure form.  For a line of displacement
// disp = (vel dt)/dx and an integer index r, t = r - disp (rounded) gives
// il = floor(t) and alpha = t - il.  When t lies in a binade [2^b, 2^(b+1))
// with b >= 0, r is an even multiple of that binade's spacing u, so the
// rounding of r - disp is r plus -disp rounded to the grid uZ (ties to an
// even multiple, unaffected by r): alpha = frac(that) depends on (b, disp)
// only.  One weight set per binade (from any r whose t falls there) is then
// bit-identical to computing the weights element by element, without the
// divisions.  |t| < 1 keeps the element-wise weights.
constexpr int kBinades = 16;  // t in [1, 2^16)

template <typename T>
__device__ __forceinline__ void binade_weights(double disp, int b, T w[6]) {
    const double lo = ldexp(1.0, b);
    const double hi = 2.0 * lo;
    const double c0 = ceil(lo + disp);
    for (int d = -1; d <= 2; ++d) {
        const double t = (c0 + d) - disp;  // the kernel's (double)r - disp for r = c0 + d
        if (t >= lo && t < hi) {
            weights_as<T>(t - floor(t), w);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = T(0);  // no index of this line reaches binade b
}

// The same for negative departure points, t in (-2^(b+1), -2^b]: the
// binade's grid is again uniform with r on it, so alpha depends on (b, disp)
// only.
template <typename T>
__device__ __forceinline__ void binade_weights_neg(double disp, int b, T w[6]) {
    const double lo = ldexp(1.0, b);
    const double hi = 2.0 * lo;
    const double c0 = floor(disp - lo);
    for (int d = 1; d >= -2; --d) {
        const double t = (c0 + d) - disp;
        if (-t >= lo && -t < hi) {
            weights_as<T>(t - floor(t), w);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = T(0);
}

// Signed binade code of a departure point: e - 1 for t >= 1, -e for t <= -1
// (|t| = m 2^e, m in [0.5, 1)); codes rise with t.  |t| < 1 or >= 2^kBinades
// has no code (kNoCode).
constexpr int kNoCode = 1 << 20;
__device__ __forceinline__ int binade_code(double t) {
    int e;
    frexp(t, &e);
    if (t >= 1.0 && e <= kBinades) return e - 1;
    if (t <= -1.0 && e <= kBinades) return -e;
    return kNoCode;
}

template <typename T>
__device__ __forceinline__ void code_weights(double disp, int q, T w[6]) {
    if (q >= 0)
        binade_weights<T>(disp, q, w);
    else
        binade_weights_neg<T>(disp, -q - 1, w);
}

// Departure form: weights of index r from a segment's table of binade codes
// [q0, q0 + nq) (tab[(code - q0) * ts6 + k], ts6 = 6), else computed.
template <typename T>
__device__ __forceinline__ void departure_weights(double dep, double fl, const T* tab, int q0, int nq, T w[6]) {
    const int slot = binade_code(dep) - q0;
    if (slot >= 0 && slot < nq) {
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = tab[slot * 6 + k];
    } else {
        weights_as<T>(dep - fl, w);
    }
}

// departure_weights for a thread walking along one line: the weights of the
// last table code stay in registers and are read again only when the code
// changes (departure points without a code compute theirs).
template <typename T>
struct LineWeights {
    T w[6];
    int cur = kNoCode;
    __device__ __forceinline__ const T* get(double dep, double fl, const T* tab, int q0, int nq) {
        const int code = binade_code(dep);
        const int slot = code - q0;
        if (code == kNoCode || slot < 0 || slot >= nq) {
            weights_as<T>(dep - fl, w);
            cur = kNoCode;
        } else if (code != cur) {
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = tab[slot * 6 + k];
            cur = code;
        }
        return w;
    }
};

// Binade codes a segment of departure points [dlo, dhi] reaches, clipped to
// the coded range: [q0, q0 + nq).
__device__ __forceinline__ void segment_codes(double dlo, double dhi, int& q0, int& nq) {
    int a = binade_code(dlo), b = binade_code(dhi);
    if (a == kNoCode) a = dlo < -1.0 ? -kBinades : 0;
    if (b == kNoCode) b = dhi > 1.0 ? kBinades - 1 : -1;
    q0 = a;
    nq = b - a + 1;
}

}  // namespace vp
