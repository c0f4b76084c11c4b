// Semi-Lagrangian directional sweeps (generic forms).
//
// advect_1d (vlasov2D2V.py:356-413) and advect_1d_vectorized
// (vlasov1D1V.py:212-256) materialise the swept axis last, build int64
// [B, N, 6] index tensors and gather.  Here the array is viewed in place as
// [outer][n][inner] (inner = stride of the swept axis) and every line is
// staged once through LDS:
//   * strided axis (inner > 1): a tile of n rows x C contiguous columns is
//     loaded with coalesced row segments, each column is one line;
//   * contiguous axis (inner == 1): a tile of L whole lines.
// Each element then reads its 6 taps from LDS.  One HBM read and one HBM
// write per element.
#include "binade.hpp"

namespace vp {

// Velocity offset of element (outer o, line index r, inner i): the outer
// and inner flat indices are split back into the array dims they cover.
struct VelMap {
    int64_t od[3];
    int64_t os[3];
    int no;
    int64_t id[3];
    int64_t is[3];
    int ni;
    int64_t s_line;
};

__device__ __forceinline__ int64_t split_offset(int64_t flat, const int64_t* dims,
                                                const int64_t* strides, int nd) {
    int64_t off = 0;
    for (int d = nd - 1; d >= 0; --d) {
        const int64_t q = flat / dims[d];
        off += (flat - q * dims[d]) * strides[d];
        flat = q;
    }
    return off;
}

// Fill tab[line * ls + b * 6 + k] for `lines` lines (velocities
// vel_of(line)) with the whole workgroup.  Lanes that read the tables of
// different lines at once (strided sweeps) use an odd line stride ls, so
// that they fall in different LDS banks.
constexpr int kTabStride = kBinades * 6 + 1;
template <typename T, typename VelOf>
__device__ __forceinline__ void fill_binade_tables(T* tab, int lines, int ls, VelOf vel_of, double dt, double dx) {
    for (int e = threadIdx.x; e < lines * kBinades; e += blockDim.x) {
        const int line = e / kBinades, b = e % kBinades;
        const double disp = (vel_of(line) * dt) / dx;
        T w[6];
        binade_weights<T>(disp, b, w);
#pragma unroll
        for (int k = 0; k < 6; ++k) tab[line * ls + b * 6 + k] = w[k];
    }
}

template <typename T, bool PERIODIC>
__device__ __forceinline__ T interp_w(const T* tile, int64_t il, const T w[6], int n, int row_stride, int col) {
    T v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int64_t j = il + (k - 2);
        if (PERIODIC) {
            v[k] = tile[wrap_index(j, n) * row_stride + col];
        } else {
            const int jc = static_cast<int>(j < 0 ? 0 : (j > n - 1 ? n - 1 : j));
            const T valid = (j >= 0 && j < n) ? T(1) : T(0);
            v[k] = tile[jc * row_stride + col] * valid;
        }
    }
    return tap_sum<T>(w, v);
}

// Departure-form value of index r from a line's binade table (tab[b][6]).
template <typename T, bool PERIODIC>
__device__ __forceinline__ T departure_value_tab(const T* tile, int r, double disp, const T* tab, int n,
                                                 int row_stride, int col) {
    const double dep = static_cast<double>(r) - disp;
    const double fl = floor(dep);
    T w[6];
    if (dep >= 1.0 && dep < static_cast<double>(1 << kBinades)) {
        int e;
        frexp(dep, &e);  // dep = m 2^e, m in [0.5, 1): binade e - 1
        const T* tw = tab + (e - 1) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = tw[k];
    } else {
        weights_as<T>(dep - fl, w);
    }
    return interp_w<T, PERIODIC>(tile, static_cast<int64_t>(fl), w, n, row_stride, col);
}

template <typename T, bool PERIODIC>
__device__ __forceinline__ T interp_from_lds(const T* tile, int64_t il, double alpha, int n,
                                             int row_stride, int col) {
    T w[6];
    weights_as<T>(alpha, w);
    T v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int64_t j = il + (k - 2);
        if (PERIODIC) {
            v[k] = tile[wrap_index(j, n) * row_stride + col];
        } else {
            const int jc = static_cast<int>(j < 0 ? 0 : (j > n - 1 ? n - 1 : j));
            const T valid = (j >= 0 && j < n) ? T(1) : T(0);
            v[k] = tile[jc * row_stride + col] * valid;
        }
    }
    return tap_sum<T>(w, v);
}

// Sweep forms.
//   kDeparture: the reference sweep (advect_1d_vectorized, vlasov1D1V.py:212-256):
//     dep = i - (v dt)/dx, il = floor(dep), alpha = dep - il, taps il - 2 .. il + 3.
//   kReflected: the roll/shift forms of solvers 1 and 2 (advect_1d_matrix_diag,
//     vlasov1D1V.py:258-337; advect_v_static, vlasov1D1V.py:479-501):
//     d = (v dt)/dx, s = floor(d), alpha = d - s, tap k (offset o = k - 2) reads
//     i - s - o with weight w_k(alpha); `window` >= 0 keeps only taps with
//     |s + o| <= window + 2 (the static window of solver 2), < 0 keeps all.
//   kStatic: kReflected with the shift s and the six weights given per line
//     (advect_x_static, vlasov1D1V.py:408-448): periodic only.
enum SweepForm { kDeparture = 0, kReflected = 1, kStatic = 2 };

struct SweepParams {
    const double* vel;     // kDeparture, kReflected
    const int64_t* shift;  // kStatic
    const double* W;       // kStatic: W[k * wk + offset]
    int64_t wk;
    double dt, dx;
    int window;
};

template <typename T, bool PERIODIC, int FORM>
__device__ __forceinline__ T sweep_value(const T* tile, int r, int64_t poff, const SweepParams& P, int n,
                                         int row_stride, int col) {
    if (FORM == kDeparture) {
        const Departure d = departure(r, P.vel[poff], P.dt, P.dx);
        return interp_from_lds<T, PERIODIC>(tile, d.il, d.alpha, n, row_stride, col);
    }
    int64_t sh;
    T w[6];
    if (FORM == kReflected) {
        const double disp = (P.vel[poff] * P.dt) / P.dx;
        const double fl = floor(disp);
        weights_as<T>(disp - fl, w);
        sh = static_cast<int64_t>(fl);
    } else {
        sh = P.shift[poff];
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = static_cast<T>(P.W[k * P.wk + poff]);
    }
    T v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int64_t j = r - sh - (k - 2);
        if (PERIODIC) {
            v[k] = tile[wrap_index(j, n) * row_stride + col];
        } else {
            const int jc = static_cast<int>(j < 0 ? 0 : (j > n - 1 ? n - 1 : j));
            const int64_t m = sh + (k - 2);
            const bool in_window = P.window < 0 || (m >= -(P.window + 2) && m <= P.window + 2);
            const T valid = (j >= 0 && j < n && in_window) ? T(1) : T(0);
            v[k] = tile[jc * row_stride + col] * valid;
        }
    }
    return tap_sum<T>(w, v);
}

// Per-line coefficients.  kDeparture: the displacement disp = (v dt)/dx
// (weights per element, from binade tables).  Shift forms: the shift sh and
// the six weights, fixed for the line, and the taps the static window keeps
// (Dirichlet lines).  key(): the departure point of index r is r - key()
// for the first tap's row (both forms), which bounds a segment's taps.
template <typename T, bool PERIODIC, int FORM>
struct LineSweep {
    double disp;
    int64_t sh;
    T w[6];
    T keep[6];
    __device__ __forceinline__ void init(int64_t poff, const SweepParams& P) {
        disp = 0.0;
        sh = 0;
        if (FORM == kDeparture) {
            disp = (P.vel[poff] * P.dt) / P.dx;
            return;
        }
        if (FORM == kReflected) {
            const double d = (P.vel[poff] * P.dt) / P.dx;
            const double fl = floor(d);
            weights_as<T>(d - fl, w);
            sh = static_cast<int64_t>(fl);
        } else {
            sh = P.shift[poff];
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = static_cast<T>(P.W[k * P.wk + poff]);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int64_t m = sh + (k - 2);
            keep[k] = (P.window < 0 || (m >= -(P.window + 2) && m <= P.window + 2)) ? T(1) : T(0);
        }
    }
    // taps of index r read rows floor(r - key()) - 2 .. floor(r - key()) + 3
    __device__ __forceinline__ double key() const {
        return FORM == kDeparture ? disp : static_cast<double>(sh) + 1.0;
    }
    // Value of index r with the taps read from memory at src[j * stride]
    // (the arithmetic of sweep_value / interp_from_lds).
    __device__ __forceinline__ T value_global(const T* src, int64_t stride, int r, int n) const {
        T wl[6];
        int64_t jt;  // row of tap 0
        int64_t dir;
        if (FORM == kDeparture) {
            const double dep = static_cast<double>(r) - disp;
            const double fl = floor(dep);
            weights_as<T>(dep - fl, wl);
            jt = static_cast<int64_t>(fl) - 2;
            dir = 1;
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) wl[k] = w[k];
            jt = r - sh + 2;
            dir = -1;
        }
        T v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int64_t j = jt + dir * k;
            if (PERIODIC) {
                v[k] = src[static_cast<int64_t>(wrap_index(j, n)) * stride];
            } else {
                const int64_t jc = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
                const bool ok = j >= 0 && j < n && (FORM == kDeparture || keep[k] != T(0));
                v[k] = src[jc * stride] * (ok ? T(1) : T(0));
            }
        }
        return tap_sum<T>(wl, v);
    }
    // Shift forms: value of index r from staged rows, jb = staged index of
    // row r - sh (taps at jb + 2 .. jb - 3).
    __device__ __forceinline__ T value_staged_shift(const T* tile, int jb, int stride) const {
        T v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const T x = tile[(jb + 2 - k) * stride];
            v[k] = PERIODIC ? x : x * keep[k];
        }
        return tap_sum<T>(w, v);
    }
};

constexpr int kTilePrefetch = 8;  // whole-line tiles: elements per thread loaded before the table fill

// Tile = n rows (the swept axis) x C columns (contiguous inner elements).
// TAB (sweep parameters constant along each line): for the departure form,
// binade weight tables of the C lines after the tile in LDS; for the shift
// forms, each thread's line coefficients (shift, weights) computed once.
template <typename T, bool PERIODIC, int FORM, bool TAB>
__global__ __launch_bounds__(256) void advect_strided_kernel(
    const T* __restrict__ fin, T* __restrict__ fout, int n, int64_t inner, int C,
    int64_t n_chunks, SweepParams P, VelMap vm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* tile = reinterpret_cast<T*>(smem_raw);
    const int64_t blk = blockIdx.x;
    const int64_t o = blk / n_chunks;
    const int64_t i0 = (blk - o * n_chunks) * C;
    const int c = threadIdx.x % C;
    const int r0 = threadIdx.x / C;
    const int rstep = blockDim.x / C;
    const bool cval = (i0 + c) < inner;
    const int64_t base = o * static_cast<int64_t>(n) * inner + i0 + c;
    const int64_t ooff = split_offset(o, vm.od, vm.os, vm.no);
    T* tab = tile + static_cast<int64_t>(n) * C;
    // this thread's first kTilePrefetch rows load while the tables are computed
    T pre[kTilePrefetch];
#pragma unroll
    for (int k = 0; k < kTilePrefetch; ++k) {
        const int r = r0 + k * rstep;
        pre[k] = (cval && r < n) ? fin[base + r * inner] : T(0);
    }
    if (TAB && FORM == kDeparture) {
        const int64_t lim = inner - i0;
        fill_binade_tables<T>(tab, C, kTabStride, [&](int line) {
            return line < lim ? P.vel[ooff + split_offset(i0 + line, vm.id, vm.is, vm.ni)] : 0.0;
        }, P.dt, P.dx);
    }
    if (cval) {
#pragma unroll
        for (int k = 0; k < kTilePrefetch; ++k) {
            const int r = r0 + k * rstep;
            if (r < n) tile[r * C + c] = pre[k];
        }
        for (int r = r0 + kTilePrefetch * rstep; r < n; r += rstep) tile[r * C + c] = fin[base + r * inner];
    }
    __syncthreads();
    if (!cval) return;
    const int64_t voff = ooff + split_offset(i0 + c, vm.id, vm.is, vm.ni);
    if (TAB && FORM == kDeparture) {
        const double disp = (P.vel[voff] * P.dt) / P.dx;
        const T* mytab = tab + c * kTabStride;
        for (int r = r0; r < n; r += rstep)
            fout[base + r * inner] = departure_value_tab<T, PERIODIC>(tile, r, disp, mytab, n, C, c);
        return;
    }
    if (TAB) {
        LineSweep<T, PERIODIC, FORM> ls;
        ls.init(voff, P);
        for (int r = r0; r < n; r += rstep) fout[base + r * inner] = ls.value_global(tile + c, C, r, n);
        return;
    }
    for (int r = r0; r < n; r += rstep)
        fout[base + r * inner] = sweep_value<T, PERIODIC, FORM>(tile, r, voff + r * vm.s_line, P, n, C, c);
}

// ---------------------------------------------------------------------------
// Long lines (n > 256) whose sweep parameters are constant along each line
// (velocity stride 0 along the swept axis): the line is cut into segments,
// and a workgroup stages only the elements its segment's taps reach.
// ---------------------------------------------------------------------------

// Strided lines: a workgroup takes KC = 256 B of contiguous columns and a
// segment of R rows and stages the rows its taps reach (the segment widened
// by the columns' spread of key() and the stencil; wrapped for periodic
// lines, zero outside [0, n) otherwise), and for the departure form each
// column's weights for up to kSegBinades binade codes from the segment's
// lowest departure point up (an odd column stride puts the columns' tables
// in different LDS banks).  A segment whose rows would not fit reads its
// taps from global memory.
constexpr int kSegBinades = 4;  // with 112-row segments: 38 KB of LDS, four workgroups per CU
constexpr int kSegMaxRows = 128;  // rows per segment at most (its staged rows: kSegMaxRows + 12)
constexpr int kSegPrefetch = 8;  // contiguous segments: staged elements per thread loaded before the table fill

template <typename T, bool PERIODIC, int FORM>
__global__ __launch_bounds__(256) void advect_seg_kernel(const T* __restrict__ fin, T* __restrict__ fout, int n,
                                                         int64_t inner, int64_t n_chunks, int n_segs, int R,
                                                         int rows_cap, SweepParams P, VelMap vm) {
    constexpr int KC = 256 / static_cast<int>(sizeof(T));
    constexpr int NG = 256 / KC;
    constexpr int TS = FORM == kDeparture ? kSegBinades * 6 + 1 : 0;
    // staged rows per thread loaded at once: all of them for 8-byte T (18),
    // 18 of up to 35 for 4-byte T (more would cut the waves per SIMD)
    constexpr int PF = (kSegMaxRows + 12 + NG - 1) / NG < 18 ? (kSegMaxRows + 12 + NG - 1) / NG : 18;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* tab = reinterpret_cast<T*>(smem_raw);  // [KC][TS]
    T* tile = tab + KC * TS;                    // [rows_cap][KC]
    __shared__ double keys[KC];
    __shared__ int span[4];  // first staged row, staged rows (0: read global), first code, codes
    const int64_t blk = blockIdx.x;
    const int64_t chunk = blk % n_chunks;
    const int64_t rest = blk / n_chunks;
    const int seg = static_cast<int>(rest % n_segs);
    const int64_t o = rest / n_segs;
    const int c = threadIdx.x % KC;
    const int g = threadIdx.x / KC;
    const int64_t i0 = chunk * KC;
    const int64_t lim = inner - i0;
    const int nc = lim < KC ? static_cast<int>(lim) : KC;
    const bool cval = c < nc;
    const int r0 = seg * R;
    const int r1 = (r0 + R < n) ? r0 + R : n;
    const int64_t ooff = split_offset(o, vm.od, vm.os, vm.no);
    const int64_t base = o * static_cast<int64_t>(n) * inner + i0 + c;
    LineSweep<T, PERIODIC, FORM> ls;
    ls.disp = 0.0;
    ls.sh = 0;
    if (cval) ls.init(ooff + split_offset(i0 + c, vm.id, vm.is, vm.ni), P);  // columns past the array read nothing
    if (g == 0) keys[c] = cval ? ls.key() : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        double kmin = keys[0], kmax = keys[0];
        for (int k = 1; k < nc; ++k) {
            kmin = fmin(kmin, keys[k]);
            kmax = fmax(kmax, keys[k]);
        }
        // departure points of rows [r0, r1) lie in [RN(r0 - kmax), RN(r1 - 1 - kmin)]
        // (rounding is monotone); one row of margin each side
        const double dlo = static_cast<double>(r0) - kmax;
        const double lo = floor(dlo) - 3.0;
        const double hi = floor(static_cast<double>(r1 - 1) - kmin) + 4.0;
        const bool fits = hi - lo + 1.0 <= static_cast<double>(rows_cap) && lo > -1073741824.0 && hi < 1073741824.0;
        span[0] = fits ? static_cast<int>(lo) : 0;
        span[1] = fits ? static_cast<int>(hi - lo) + 1 : 0;
        int q0, nq;
        segment_codes(dlo, static_cast<double>(r1 - 1) - kmin, q0, nq);
        span[2] = q0;
        span[3] = nq < kSegBinades ? nq : kSegBinades;
    }
    __syncthreads();
    const int lo = span[0];
    const int cnt = span[1];
    if (cnt == 0) {
        if (!cval) return;
        const T* src = fin + base;
        for (int r = r0 + g; r < r1; r += NG) fout[base + static_cast<int64_t>(r) * inner] = ls.value_global(src, inner, r, n);
        return;
    }
    const int q0 = span[2];
    const int nq = span[3];
    auto staged = [&](int rr) -> T {
        const int row = lo + rr;
        T v = T(0);
        if (cval) {
            if (PERIODIC)
                v = fin[base + static_cast<int64_t>(wrap_index(row, n)) * inner];
            else if (row >= 0 && row < n)
                v = fin[base + static_cast<int64_t>(row) * inner];
        }
        return v;
    };
    // the first PF staged rows of this thread are loading at once, while the
    // weight tables are computed (a load per row in turn leaves one load in flight)
    T pre[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int rr = g + k * NG;
        pre[k] = rr < cnt ? staged(rr) : T(0);
    }
    if (FORM == kDeparture) {
        // code-major over the threads: the KC * nq entries in use occupy the
        // first threads, so whole waves skip the divisions (slots >= nq are
        // never read)
        for (int e = threadIdx.x; e < KC * nq; e += blockDim.x) {
            const int q = e / KC, line = e % KC;
            T w[6];
            code_weights<T>(keys[line], q0 + q, w);
#pragma unroll
            for (int k = 0; k < 6; ++k) tab[line * TS + q * 6 + k] = w[k];
        }
    }
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int rr = g + k * NG;
        if (rr < cnt) tile[rr * KC + c] = pre[k];
    }
    for (int rr = g + PF * NG; rr < cnt; rr += NG) tile[rr * KC + c] = staged(rr);
    __syncthreads();
    if (!cval) return;
    if (FORM == kDeparture) {
        const T* mytab = tab + c * TS;
        LineWeights<T> lw;
        for (int r = r0 + g; r < r1; r += NG) {
            const double dep = static_cast<double>(r) - ls.disp;
            const double fl = floor(dep);
            const T* w = lw.get(dep, fl, mytab, q0, nq);
            int j0 = static_cast<int>(fl) - 2 - lo;  // staged row of tap 0, in [1, cnt - 7] for finite disp
            j0 = j0 < 0 ? 0 : (j0 > cnt - 6 ? cnt - 6 : j0);
            T v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = tile[(j0 + k) * KC + c];
            fout[base + static_cast<int64_t>(r) * inner] = tap_sum<T>(w, v);
        }
    } else {
        const int sh = static_cast<int>(ls.sh);  // |sh| < 2^30 when the segment fits
        for (int r = r0 + g; r < r1; r += NG) {
            int jb = r - sh - lo;  // staged row of r - sh, in [4, cnt - 3]
            jb = jb < 3 ? 3 : (jb > cnt - 3 ? cnt - 3 : jb);
            fout[base + static_cast<int64_t>(r) * inner] = ls.value_staged_shift(tile + c, jb, KC);
        }
    }
}

// Contiguous lines: a workgroup takes a segment of S indices of one line.
// The whole line shares one key(), so the segment's taps span at most S + 9
// consecutive elements (wrapped, or zero outside [0, n)), staged in LDS with,
// for the departure form, the line's weights for every binade code the
// segment reaches (at most 2 kBinades).  An out-of-range key() reads the
// taps from global memory.
template <typename T, bool PERIODIC, int FORM>
__global__ __launch_bounds__(256) void advect_cseg_kernel(const T* __restrict__ fin, T* __restrict__ fout, int n,
                                                          int S, int n_segs, SweepParams P, VelMap vm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* tile = reinterpret_cast<T*>(smem_raw);  // [S + 16]
    __shared__ T tab[2 * kBinades * 6];
    const int64_t line = blockIdx.x / n_segs;
    const int seg = static_cast<int>(blockIdx.x - line * n_segs);
    const int r0 = seg * S;
    const int r1 = (r0 + S < n) ? r0 + S : n;
    const T* src = fin + line * n;
    T* dst = fout + line * n;
    LineSweep<T, PERIODIC, FORM> ls;
    ls.init(split_offset(line, vm.od, vm.os, vm.no), P);
    const double key = ls.key();
    const double dlo = static_cast<double>(r0) - key;
    const double dhi = static_cast<double>(r1 - 1) - key;
    const double lo_d = floor(dlo) - 3.0;
    const double hi_d = floor(dhi) + 4.0;
    const bool fits = hi_d - lo_d + 1.0 <= static_cast<double>(S + 16) && lo_d > -1073741824.0 && hi_d < 1073741824.0;
    if (!fits) {
        for (int r = r0 + static_cast<int>(threadIdx.x); r < r1; r += blockDim.x) dst[r] = ls.value_global(src, 1, r, n);
        return;
    }
    const int lo = static_cast<int>(lo_d);
    const int cnt = static_cast<int>(hi_d - lo_d) + 1;
    auto staged = [&](int idx) -> T {
        const int row = lo + idx;
        T v = T(0);
        if (PERIODIC)
            v = src[wrap_index(row, n)];
        else if (row >= 0 && row < n)
            v = src[row];
        return v;
    };
    T pre[kSegPrefetch];  // loading while the weight table is computed
#pragma unroll
    for (int k = 0; k < kSegPrefetch; ++k) {
        const int idx = static_cast<int>(threadIdx.x) + k * static_cast<int>(blockDim.x);
        pre[k] = idx < cnt ? staged(idx) : T(0);
    }
    int q0 = 0, nq = 0;
    if (FORM == kDeparture) {
        segment_codes(dlo, dhi, q0, nq);  // nq <= 2 kBinades
        if (static_cast<int>(threadIdx.x) < nq) {
            T w[6];
            code_weights<T>(ls.disp, q0 + static_cast<int>(threadIdx.x), w);
#pragma unroll
            for (int k = 0; k < 6; ++k) tab[threadIdx.x * 6 + k] = w[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kSegPrefetch; ++k) {
        const int idx = static_cast<int>(threadIdx.x) + k * static_cast<int>(blockDim.x);
        if (idx < cnt) tile[idx] = pre[k];
    }
    for (int idx = static_cast<int>(threadIdx.x) + kSegPrefetch * static_cast<int>(blockDim.x); idx < cnt;
         idx += blockDim.x)
        tile[idx] = staged(idx);
    __syncthreads();
    if (FORM == kDeparture) {
        LineWeights<T> lw;
        for (int r = r0 + static_cast<int>(threadIdx.x); r < r1; r += blockDim.x) {
            const double dep = static_cast<double>(r) - ls.disp;
            const double fl = floor(dep);
            const T* w = lw.get(dep, fl, tab, q0, nq);
            int j0 = static_cast<int>(fl) - 2 - lo;  // staged index of tap 0, in [1, cnt - 7]
            j0 = j0 < 0 ? 0 : (j0 > cnt - 6 ? cnt - 6 : j0);
            T v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = tile[j0 + k];
            dst[r] = tap_sum<T>(w, v);
        }
    } else {
        const int sh = static_cast<int>(ls.sh);
        for (int r = r0 + static_cast<int>(threadIdx.x); r < r1; r += blockDim.x) {
            int jb = r - sh - lo;
            jb = jb < 3 ? 3 : (jb > cnt - 3 ? cnt - 3 : jb);
            dst[r] = ls.value_staged_shift(tile, jb, 1);
        }
    }
}

// Tile = L whole contiguous lines of length n; wave w handles lines w, w+4, ...
// TAB: binade weight tables of the L lines after the tile in LDS.
template <typename T, bool PERIODIC, int FORM, bool TAB>
__global__ __launch_bounds__(256) void advect_contig_kernel(
    const T* __restrict__ fin, T* __restrict__ fout, int64_t n_lines, int n, int L, SweepParams P, VelMap vm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* tile = reinterpret_cast<T*>(smem_raw);
    const int64_t l0 = static_cast<int64_t>(blockIdx.x) * L;
    const int64_t nl = (n_lines - l0) < L ? (n_lines - l0) : L;
    const int64_t cnt = nl * n;
    const T* src = fin + l0 * n;
    T* tab = tile + static_cast<int64_t>(L) * n;
    // this thread's first kTilePrefetch elements load while the tables are computed
    T pre[kTilePrefetch];
#pragma unroll
    for (int k = 0; k < kTilePrefetch; ++k) {
        const int64_t idx = threadIdx.x + static_cast<int64_t>(k) * blockDim.x;
        pre[k] = idx < cnt ? src[idx] : T(0);
    }
    if (TAB && FORM == kDeparture)
        fill_binade_tables<T>(tab, static_cast<int>(nl), kBinades * 6, [&](int line) {
            return P.vel[split_offset(l0 + line, vm.od, vm.os, vm.no)];
        }, P.dt, P.dx);
#pragma unroll
    for (int k = 0; k < kTilePrefetch; ++k) {
        const int64_t idx = threadIdx.x + static_cast<int64_t>(k) * blockDim.x;
        if (idx < cnt) tile[idx] = pre[k];
    }
    for (int64_t idx = threadIdx.x + static_cast<int64_t>(kTilePrefetch) * blockDim.x; idx < cnt; idx += blockDim.x)
        tile[idx] = src[idx];
    __syncthreads();
    const int lane = threadIdx.x % kWave;
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    for (int l = wid; l < nl; l += nw) {
        const int64_t voff = split_offset(l0 + l, vm.od, vm.os, vm.no);
        const T* line = tile + static_cast<int64_t>(l) * n;
        T* dst = fout + (l0 + l) * n;
        if (TAB && FORM == kDeparture) {
            const double disp = (P.vel[voff] * P.dt) / P.dx;
            const T* ltab = tab + l * (kBinades * 6);
            for (int j = lane; j < n; j += kWave) dst[j] = departure_value_tab<T, PERIODIC>(line, j, disp, ltab, n, 1, 0);
            continue;
        }
        if (TAB) {
            LineSweep<T, PERIODIC, FORM> ls;
            ls.init(voff, P);
            for (int j = lane; j < n; j += kWave) dst[j] = ls.value_global(line, 1, j, n);
            continue;
        }
        for (int j = lane; j < n; j += kWave)
            dst[j] = sweep_value<T, PERIODIC, FORM>(line, j, voff + j * vm.s_line, P, n, 1, 0);
    }
}

template <typename T, int FORM>
static int launch_sweep(const T* fin, T* fout, const int64_t shape[4], int axis, const SweepParams& P,
                        const int64_t vs[4], int periodic, hipStream_t s) {
    VP_RETURN_IF(!fin || !fout || !shape || !vs, VP_ERR_ARG);
    VP_RETURN_IF(axis < 0 || axis > 3, VP_ERR_ARG);
    for (int d = 0; d < 4; ++d) VP_RETURN_IF(shape[d] < 1, VP_ERR_SHAPE);
    const int64_t n64 = shape[axis];
    VP_RETURN_IF(n64 > 8192, VP_ERR_UNSUPPORTED);
    const int n = static_cast<int>(n64);
    VelMap vm{};
    int64_t outer = 1, inner = 1;
    for (int d = 0; d < axis; ++d) {
        vm.od[vm.no] = shape[d];
        vm.os[vm.no] = vs[d];
        ++vm.no;
        outer *= shape[d];
    }
    for (int d = axis + 1; d < 4; ++d) {
        vm.id[vm.ni] = shape[d];
        vm.is[vm.ni] = vs[d];
        ++vm.ni;
        inner *= shape[d];
    }
    vm.s_line = vs[axis];
    const int block = 256;
    // sweep parameters constant along each line: binade tables (departure form)
    // or per-line shift and weights (shift forms)
    const bool tab = vm.s_line == 0 && (FORM != kDeparture || n < (1 << kBinades));
    const size_t tab_line = tab && FORM == kDeparture ? static_cast<size_t>(kTabStride) * sizeof(T) : 0;
    const int64_t kMaxLds = 96 * 1024;
    const int64_t cus = 256;  // spread small problems over at least ~2 workgroups per CU
    // segmented kernels: long lines with the sweep parameters constant along each line
    const bool seg_ok = vm.s_line == 0 && n > 256;
    if (inner > 1 && seg_ok && inner >= 256 / static_cast<int64_t>(sizeof(T))) {
        constexpr int KC = 256 / static_cast<int>(sizeof(T));
        const int64_t chunks = (inner + KC - 1) / KC;
        // rows per segment: 112 (its tile and tables fit four workgroups per CU in LDS) when that
        // gives at least two workgroups per CU, else 64 or 32
        int R = knob(VP_KNOB_SEG_ROWS);
        if (R < 16 || R > kSegMaxRows) {  // 128 rows: 42 KB of LDS, under the default 64 KB
            R = 112;
            if (outer * chunks * ((n + R - 1) / R) < 2 * cus) {
                R = 64;
                while (R > 32 && outer * chunks * ((n + R - 1) / R) < 2 * cus) R /= 2;
            }
        }
        const int rows_cap = R + 12;  // the segment, a displacement spread of up to 4 rows and the stencil
        const int n_segs = (n + R - 1) / R;
        const int64_t blocks = outer * chunks * n_segs;
        VP_RETURN_IF(blocks > 0x7fffffffLL, VP_ERR_SHAPE);
        const size_t ts = FORM == kDeparture ? kSegBinades * 6 + 1 : 0;
        const size_t lds = (static_cast<size_t>(KC) * ts + static_cast<size_t>(rows_cap) * KC) * sizeof(T);
        if (periodic)
            hipLaunchKernelGGL((advect_seg_kernel<T, true, FORM>), dim3(blocks), dim3(block), lds, s, fin, fout, n,
                               inner, chunks, n_segs, R, rows_cap, P, vm);
        else
            hipLaunchKernelGGL((advect_seg_kernel<T, false, FORM>), dim3(blocks), dim3(block), lds, s, fin, fout, n,
                               inner, chunks, n_segs, R, rows_cap, P, vm);
    } else if (inner > 1) {
        int C = 64;
        while (C > 1 && C / 2 >= inner) C /= 2;
        while (C > 1 && static_cast<int64_t>(n) * C * sizeof(T) + C * tab_line > kMaxLds) C /= 2;
        while (C > 8 && outer * ((inner + C - 1) / C) < 2 * cus) C /= 2;
        VP_RETURN_IF(static_cast<int64_t>(n) * C * sizeof(T) + C * tab_line > kMaxLds, VP_ERR_UNSUPPORTED);
        const int64_t chunks = (inner + C - 1) / C;
        const int64_t blocks = outer * chunks;
        VP_RETURN_IF(blocks > 0x7fffffffLL, VP_ERR_SHAPE);
        const size_t lds = static_cast<size_t>(n) * C * sizeof(T) + C * tab_line;
#define VP_S(PER, TB)                                                                                               \
    {                                                                                                               \
        static bool attr = false;                                                                                   \
        if (!attr) {                                                                                                \
            VP_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&advect_strided_kernel<T, PER, FORM, TB>), \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds))); \
            attr = true;                                                                                            \
        }                                                                                                           \
        hipLaunchKernelGGL((advect_strided_kernel<T, PER, FORM, TB>), dim3(blocks), dim3(block), lds, s, fin, fout, \
                           n, inner, C, chunks, P, vm);                                                             \
    }
        if (periodic) {
            if (tab) VP_S(true, true) else VP_S(true, false)
        } else {
            if (tab) VP_S(false, true) else VP_S(false, false)
        }
#undef VP_S
    } else if (inner == 1 && seg_ok) {
        int S = knob(VP_KNOB_SEG_CONTIG);
        if (S < 256 || S > 4096 || (S & (S - 1))) S = n >= 2048 ? 2048 : 1024;
        while (S > 256 && S / 2 >= n) S /= 2;
        const int n_segs = (n + S - 1) / S;
        const int64_t blocks = outer * n_segs;
        VP_RETURN_IF(blocks > 0x7fffffffLL, VP_ERR_SHAPE);
        const size_t lds = static_cast<size_t>(S + 16) * sizeof(T);
        if (periodic)
            hipLaunchKernelGGL((advect_cseg_kernel<T, true, FORM>), dim3(blocks), dim3(block), lds, s, fin, fout, n,
                               S, n_segs, P, vm);
        else
            hipLaunchKernelGGL((advect_cseg_kernel<T, false, FORM>), dim3(blocks), dim3(block), lds, s, fin, fout, n,
                               S, n_segs, P, vm);
    } else {
        int L = n >= 4096 ? 1 : static_cast<int>(4096 / n);
        if (L > outer) L = static_cast<int>(outer);
        while (L > 1 && static_cast<int64_t>(L) * (n * sizeof(T) + tab_line) > kMaxLds) L /= 2;
        while (L > 1 && (outer + L - 1) / L < 2 * cus) L /= 2;
        VP_RETURN_IF(static_cast<int64_t>(L) * (n * sizeof(T) + tab_line) > kMaxLds, VP_ERR_UNSUPPORTED);
        const int64_t blocks = (outer + L - 1) / L;
        VP_RETURN_IF(blocks > 0x7fffffffLL, VP_ERR_SHAPE);
        const size_t lds = static_cast<size_t>(L) * (n * sizeof(T) + tab_line);
#define VP_C(PER, TB)                                                                                               \
    {                                                                                                               \
        static bool attr = false;                                                                                   \
        if (!attr) {                                                                                                \
            VP_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&advect_contig_kernel<T, PER, FORM, TB>),  \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds))); \
            attr = true;                                                                                            \
        }                                                                                                           \
        hipLaunchKernelGGL((advect_contig_kernel<T, PER, FORM, TB>), dim3(blocks), dim3(block), lds, s, fin, fout,  \
                           outer, n, L, P, vm);                                                                     \
    }
        if (periodic) {
            if (tab) VP_C(true, true) else VP_C(true, false)
        } else {
            if (tab) VP_C(false, true) else VP_C(false, false)
        }
#undef VP_C
    }
    VP_CHECK_LAUNCH();
    return VP_OK;
}

template <typename T>
static int launch_advect(const T* fin, T* fout, const int64_t shape[4], int axis, const double* vel,
                         const int64_t vs[4], double dt, double dx, int periodic, hipStream_t s) {
    VP_RETURN_IF(!vel, VP_ERR_ARG);
    SweepParams P{vel, nullptr, nullptr, 0, dt, dx, -1};
    return launch_sweep<T, kDeparture>(fin, fout, shape, axis, P, vs, periodic, s);
}

}  // namespace vp

extern "C" int vp_advect_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                             const double* vel, const int64_t vel_strides[4], double dt, double dx,
                             int periodic, vp_stream_t stream) {
    return vp::launch_advect<double>(f_in, f_out, shape, axis, vel, vel_strides, dt, dx, periodic,
                                     static_cast<hipStream_t>(stream));
}

extern "C" int vp_advect_f32(const float* f_in, float* f_out, const int64_t shape[4], int axis,
                             const double* vel, const int64_t vel_strides[4], double dt, double dx,
                             int periodic, vp_stream_t stream) {
    return vp::launch_advect<float>(f_in, f_out, shape, axis, vel, vel_strides, dt, dx, periodic,
                                    static_cast<hipStream_t>(stream));
}

extern "C" int vp_advect_reflected_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                                       const double* vel, const int64_t vel_strides[4], double dt, double dx,
                                       int periodic, int window, vp_stream_t stream) {
    VP_RETURN_IF(!vel, VP_ERR_ARG);
    VP_RETURN_IF(periodic && window >= 0, VP_ERR_ARG);
    vp::SweepParams P{vel, nullptr, nullptr, 0, dt, dx, window};
    return vp::launch_sweep<double, vp::kReflected>(f_in, f_out, shape, axis, P, vel_strides, periodic,
                                                    static_cast<hipStream_t>(stream));
}

extern "C" int vp_advect_static_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                                    const int64_t* shift, const double* W, int64_t w_tap_stride,
                                    const int64_t param_strides[4], vp_stream_t stream) {
    VP_RETURN_IF(!shift || !W, VP_ERR_ARG);
    vp::SweepParams P{nullptr, shift, W, w_tap_stride, 0.0, 1.0, -1};
    return vp::launch_sweep<double, vp::kStatic>(f_in, f_out, shape, axis, P, param_strides, 1,
                                                 static_cast<hipStream_t>(stream));
}
