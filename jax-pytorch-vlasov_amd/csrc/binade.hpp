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

// x / d for a constant integer d in [3, 120] without the division sequence:
// q = RN(x y) with y = RN(1/d), r = x - q d (exact by fma), RN(q + r y).
// Equal to RN(x / d) for x = 0 and for every normal x whose quotient is
// normal: q is within 2 ulp u of x/d, so r is a small multiple of u/2 and
// exact, and q + r y differs from x/d by less than 2^-52 u, while x/d (x a
// multiple of 2u, d >= 3) lies at least u/(2d) >= u/240 from every midpoint
// of the double grid, so both round to the same double.  (Checked on 4e9
// normal numerators for each divisor below.)
__device__ __forceinline__ double div_const(double x, double d, double y) {
    const double q = x * y;
    const double r = fma(-q, d, x);
    return fma(r, y, q);
}

// lagrange5_weights<double> with div_const, for alpha = 0 or alpha in
// [2^-52, 1) (the fraction of a departure point with |t| >= 1): every
// numerator (alpha^k times 1, 2, 5 or 7) is then 0 or normal (>= 2^-260)
// and so is every quotient, so the weights are bit-identical to the
// divisions of the reference's expression (vlasov1D1V.py:191-210).
template <typename T>
__device__ __forceinline__ void weights_binade(double a, T w[6]) {
    constexpr double i3 = 1.0 / 3.0, i6 = 1.0 / 6.0, i12 = 1.0 / 12.0, i20 = 1.0 / 20.0, i24 = 1.0 / 24.0,
                     i30 = 1.0 / 30.0, i120 = 1.0 / 120.0;
    const double a2 = a * a;
    const double a3 = a2 * a;
    const double a4 = a3 * a;
    const double a5 = a4 * a;
    const double a3_24 = div_const(a3, 24.0, i24), a4_24 = div_const(a4, 24.0, i24);
    const double a5_24 = div_const(a5, 24.0, i24), a5_12 = div_const(a5, 12.0, i12);
    const double a5_120 = div_const(a5, 120.0, i120), a2_24 = div_const(a2, 24.0, i24);
    const double a4_6 = div_const(a4, 6.0, i6), tw_a2_3 = div_const(2.0 * a2, 3.0, i3);
    double wd[6];
    wd[0] = div_const(a, 20.0, i20) - a2_24 - a3_24 + a4_24 - a5_120;
    wd[1] = -(a / 2.0) + tw_a2_3 - a3_24 - a4_6 + a5_24;
    wd[2] = 1.0 - div_const(a, 3.0, i3) - (5.0 * a2 / 4.0) + div_const(5.0 * a3, 12.0, i12) + (a4 / 4.0) - a5_12;
    wd[3] = a + tw_a2_3 - div_const(7.0 * a3, 12.0, i12) - a4_6 + a5_12;
    wd[4] = -(a / 4.0) - a2_24 + div_const(7.0 * a3, 24.0, i24) + a4_24 - a5_24;
    wd[5] = div_const(a, 30.0, i30) - a3_24 + a5_120;
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
    double ts = 0.0;
    bool found = false;
#pragma unroll
    for (int d = -1; d <= 2; ++d) {
        const double t = (c0 + d) - disp;  // the kernel's (double)r - disp for r = c0 + d
        const bool in = !found && t >= lo && t < hi;
        ts = in ? t : ts;
        found = found || in;
    }
    if (found) {
        weights_binade<T>(ts - floor(ts), w);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = T(0);  // no index of this line reaches binade b
    }
}

// The same for negative departure points, t in (-2^(b+1), -2^b]: the
// binade's grid is again uniform with r on it, so alpha depends on (b, disp)
// only.
template <typename T>
__device__ __forceinline__ void binade_weights_neg(double disp, int b, T w[6]) {
    const double lo = ldexp(1.0, b);
    const double hi = 2.0 * lo;
    const double c0 = floor(disp - lo);
    double ts = 0.0;
    bool found = false;
#pragma unroll
    for (int d = 1; d >= -2; --d) {
        const double t = (c0 + d) - disp;
        const bool in = !found && -t >= lo && -t < hi;
        ts = in ? t : ts;
        found = found || in;
    }
    if (found) {
        weights_binade<T>(ts - floor(ts), w);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = T(0);
    }
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
