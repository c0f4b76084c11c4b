// Ensemble of independent 1D1V Vlasov-Poisson instances of different sizes
// (f_i[nx_i][nv_i] in one pool), stepped together: every launch covers all
// instances, three per step (X drift pair + rho, field solve, V kick).
//
// Each instance follows the reference step (vlasov_step, vlasov1D1V.py:574-622):
// X(dt/2) periodic along x with velocity v_j, rho = sum_v f dv - 1, E from
// the 1D Poisson solve, V(dt) along v with E(x) (zero outside [0, nv)),
// X(dt/2).  The sweeps evaluate the reference arithmetic element by element
// (departure, power-form weights, left-to-right six-tap sum; the weights of
// a line come from its binade tables, bit-identical to the element-wise
// computation), so they are bit-identical to the single-instance path; only
// the charge-density sum order differs (compensated here).  The closing
// X(dt/2) of a step and the opening X(dt/2) of the next run as one launch
// (two passes through LDS).
//
// Work lists: a set-up kernel turns the instance descriptors into
// descriptors of equal-sized work items, X drift: a block of 16..64 columns
// of one instance with all its nx rows (128..512-byte row segments, at most
// 4096 elements), with the binade codes its departure points reach; kick: a
// block of whole rows.  The sweep kernels are persistent, two workgroups of
// 512 threads per CU (80 KB of LDS each, so that one computes while the
// other waits at a barrier): each loads the next item's tile (and its
// velocities or fields) into registers while it computes the current one
// from LDS.  Items are handed out dynamically (a counter per sweep: each
// workgroup takes its first item by its index, then the one after next from
// the counter while it works), so that workgroups that drew small items
// take more.  Each counter is
// cleared by the kernel launched before its sweep (set-up, field solve,
// kick), never by a memset, so a captured graph replays correctly.
//
// Inside an item a thread sweeps a run of up to 8 consecutive elements of
// one line (X: rows of a column; kick: an 8-aligned chunk of a row).  Along
// a line the displacement is constant, so floor(t) rises by one per element:
// floor(RN(x)) is floor(x) or floor(x) + 1 (when x rounds up to an
// integer), so floor(t_i) - floor(t_0) - i is -1, 0 or 1.  Every thread
// first reads the window of run + 7 values its stencils can reach into
// registers; after a barrier it computes its run from them and writes the
// results in place.  Each element reads its six weights from the column's
// LDS table slot of its binade code (three 16-byte reads): keeping one set
// per run and switching only where the code changes made waves diverge.
#include "binade.hpp"
#include "field_dev.hpp"

namespace vp {

constexpr int kEnsThreads = 256;  // field-solve kernel
constexpr int kEnsMaxNx = 256;
constexpr int kEnsMaxNv = 512;
constexpr int kEnsNT = 512;       // threads of the persistent sweep kernels
constexpr int kEnsPer = 8;        // tile elements per thread (tiles of at most 4096 elements); run length
constexpr int kEnsTileMax = kEnsNT * kEnsPer;
constexpr int kEnsWin = kEnsPer + 7;  // a run's stencils, one more value on each side
constexpr int kEnsLdsDoubles = 10176;  // 79.5 KB of dynamic LDS per sweep workgroup (two per CU)
constexpr int kEnsMaxCols = 64;

// X drift LDS: tile [nx][cols + 1] (nx * cols <= 4096), column
// displacements [64], column tables.
constexpr int kEnsXTile = kEnsTileMax + kEnsMaxNx;
constexpr int kEnsXDisp = kEnsXTile;
constexpr int kEnsXTab = kEnsXDisp + kEnsMaxCols;
constexpr int kEnsXTabDoubles = kEnsLdsDoubles - kEnsXTab;
static_assert(kEnsXTabDoubles >= kEnsTileMax, "X element-wise output buffer");

// Kick LDS: rows of kEnsKPad zeros, nv values, kEnsKPad zeros, flat index P
// stored at P + P / 8 (one slack double every 8, so that the windows of
// consecutive threads, 8 apart, fall in different banks); row tables; per
// row displacement, special row, first code, code count.
constexpr int kEnsKPad = 8;
constexpr int kEnsKRows = 32;  // rows per kick item at most
constexpr int kEnsKTile = ((kEnsTileMax + kEnsKRows * 2 * kEnsKPad + 32) * 9) / 8;
constexpr int kEnsKCodes = 14;                       // binade codes per row in the tables
constexpr int kEnsKT = (kEnsKCodes + 2) * 6 + 2;     // codes + 2 special rows; 2 mod 4 (16-byte rows)
constexpr int kEnsKTab = kEnsKTile;
constexpr int kEnsKAux = kEnsLdsDoubles - 4 * kEnsKRows;  // disp[32], then int rs[32], q0[32], nq[32]
static_assert(kEnsKTab + kEnsKRows * kEnsKT <= kEnsKAux, "kick tables");
static_assert(kEnsKTab + kEnsTileMax <= kEnsKAux, "kick element-wise output buffer");

__device__ __forceinline__ int kidx(int p) { return p + (p >> 3); }

// Stride of the per-column X tables of ns slots: 6 ns rounded up to 2 mod 4
// doubles (16-byte aligned slots; the 16-byte reads of lanes on consecutive
// columns spread over 16 bank groups).
__host__ __device__ inline int ens_xtab_stride(int ns) {
    const int b = 6 * ns;
    return b + (b % 4 == 0 ? 2 : 4);
}

// X-drift work item.  nq < 0: the tables do not cover it (element-wise).
struct XItem {
    int64_t f;     // pool offset of (row 0, first column)
    int64_t v;     // tables offset of the first column's velocity
    int nx, nv;    // rows; row stride
    int ncol;      // columns inside the grid (the rest of the block reads and moves zeros)
    int shift;     // log2(block width)
    int q0, nq;    // binade codes the item's departure points reach
    double dx;
};
static_assert(sizeof(XItem) == 48, "XItem layout");

// Kick work item: rows [r0, r0 + rows) of one instance.
struct KItem {
    int64_t f;    // pool offset of the first row
    int e;        // field offset: instance * kEnsMaxNx + r0
    int nv, rows;
    int cpr;      // 8-element chunks per row
    double dv;
};
static_assert(sizeof(KItem) == 32, "KItem layout");

// Column-block width of an X-drift item: the largest power of two in
// [16, 64] with width * nx <= tile (tile >= 4096: at least 16, 128-byte
// row segments, for nx <= 256).
__host__ __device__ inline int ens_xcols(int nx, int tile) {
    int c = 16;
    while (c < kEnsMaxCols && 2 * c * nx <= tile) c *= 2;
    return c;
}

// Row-block height of a kick item: the largest power of two in [1, 32] with
// rows * nv <= tile and one thread per 8-element chunk of its rows.
__host__ __device__ inline int ens_krows(int nv, int tile) {
    const int cpr = (nv + kEnsPer - 1) / kEnsPer;
    int r = 1;
    while (r < kEnsKRows && 2 * r * nv <= tile && 2 * r * cpr <= kEnsNT) r *= 2;
    return r;
}

__host__ __device__ inline int ilog2(int x) {
    int l = 0;
    while ((1 << (l + 1)) <= x) ++l;
    return l;
}

// The only rows whose departure point t = RN(r - disp) can lie in (-1, 1),
// where binades do not apply: r - 1 < disp < r + 1, i.e. r = rs or rs + 1,
// rs = floor(disp).  Their weights sit in two slots after a line's codes.
__device__ __forceinline__ int ens_special_row(double disp) {
    return fabs(disp) < 1e9 ? static_cast<int>(floor(disp)) : -0x40000000;
}

// Weights of a special row (element-wise arithmetic).
__device__ __forceinline__ void ens_special_weights(double disp, int r, double* out) {
    const double dep = static_cast<double>(r) - disp;
    double w[6];
    weights_as<double>(dep - floor(dep), w);
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = w[k];
}

__device__ __forceinline__ void ens_code_weights(double disp, int code, double* out) {
    double w[6];
    code_weights<double>(disp, code, w);
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = w[k];
}

__device__ __forceinline__ void load_w6(const double* p, double (&w)[6]) {  // p 16-byte aligned
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double2 x = q[k];
        w[2 * k] = x.x;
        w[2 * k + 1] = x.y;
    }
}

// Work lists of both sweeps: offx[i] / offk[i] = first item of instance i
// (offx[ninst] = total) and the item descriptors.  One workgroup; instances
// in chunks of kEnsNT with a block-wide exclusive scan of the item counts;
// each thread then writes its instance's items (an X item's codes from the
// displacements of its columns, which the run keeps).
__global__ __launch_bounds__(kEnsNT) void ens_setup_kernel(const vp_ens_inst* __restrict__ inst, int ninst,
                                                           const double* __restrict__ tables, double half,
                                                           int xtile, int ktile, int* __restrict__ offx,
                                                           int* __restrict__ offk, XItem* __restrict__ itx,
                                                           KItem* __restrict__ itk, int* __restrict__ counters) {
    __shared__ int sx[kEnsNT], sk[kEnsNT];
    __shared__ int base[2];
    const int t = threadIdx.x;
    if (t == 0) base[0] = base[1] = 0;
    if (t < 2) counters[t] = 0;
    for (int b0 = 0; b0 < ninst; b0 += kEnsNT) {
        const int i = b0 + t;
        int cx = 0, ck = 0, cols = 32, rows = 1;
        vp_ens_inst I{};
        if (i < ninst) {
            I = inst[i];
            cols = ens_xcols(I.nx, xtile);
            rows = ens_krows(I.nv, ktile);
            cx = (I.nv + cols - 1) / cols;
            ck = (I.nx + rows - 1) / rows;
        }
        __syncthreads();  // base[] of the previous chunk is final
        sx[t] = cx;
        sk[t] = ck;
        __syncthreads();
        for (int off = 1; off < kEnsNT; off <<= 1) {  // inclusive scan (Hillis-Steele)
            const int ax = t >= off ? sx[t - off] : 0;
            const int ak = t >= off ? sk[t - off] : 0;
            __syncthreads();
            sx[t] += ax;
            sk[t] += ak;
            __syncthreads();
        }
        const int ox = base[0] + sx[t] - cx;
        const int ok = base[1] + sk[t] - ck;
        if (i < ninst) {
            offx[i] = ox;
            offk[i] = ok;
            const double* v = tables + I.v_off;
            const int shift = ilog2(cols);
            int cap = 0;  // most codes whose tables fit
            while (cols * ens_xtab_stride(cap + 3) <= kEnsXTabDoubles) ++cap;
            for (int q = 0; q < cx; ++q) {
                const int c0 = q * cols;
                const int ncol = I.nv - c0 < cols ? I.nv - c0 : cols;
                // columns past nv move zeros with displacement 0
                double dmin = ncol < cols ? 0.0 : HUGE_VAL, dmax = ncol < cols ? 0.0 : -HUGE_VAL;
                bool finite = true;
                for (int c = 0; c < ncol; ++c) {
                    const double d = (v[c0 + c] * half) / I.dx;  // departure(), vlasov1D1V.py:229-238
                    finite = finite && fabs(d) < 1e9;
                    dmin = fmin(dmin, d);
                    dmax = fmax(dmax, d);
                }
                // departure points t = RN(r - disp), r in [0, nx): within
                // [-dmax, (nx - 1) - dmin] (rounding is monotonic)
                int q0 = 0, nq = -1;
                if (finite) segment_codes(-dmax, static_cast<double>(I.nx - 1) - dmin, q0, nq);
                // covered: the codes fit the LDS, and one wrap brings every
                // first tap into [0, nx) (|disp| <= nx - 2)
                const bool covered = finite && fmax(fabs(dmin), fabs(dmax)) <= static_cast<double>(I.nx - 2) &&
                                     nq <= cap;
                XItem it;
                it.f = I.f_off + c0;
                it.v = I.v_off + c0;
                it.nx = I.nx;
                it.nv = I.nv;
                it.ncol = ncol;
                it.shift = shift;
                it.q0 = q0;
                it.nq = covered ? nq : -1;
                it.dx = I.dx;
                itx[ox + q] = it;
            }
            for (int q = 0; q < ck; ++q) {
                const int r0 = q * rows;
                KItem it;
                it.f = I.f_off + static_cast<int64_t>(r0) * I.nv;
                it.e = i * kEnsMaxNx + r0;
                it.nv = I.nv;
                it.rows = I.nx - r0 < rows ? I.nx - r0 : rows;
                it.cpr = (I.nv + kEnsPer - 1) / kEnsPer;
                it.dv = I.dv;
                itk[ok + q] = it;
            }
        }
        __syncthreads();  // every thread has read base[]
        if (t == kEnsNT - 1) {
            base[0] += sx[t];
            base[1] += sk[t];
        }
    }
    __syncthreads();
    if (t == 0) {
        offx[ninst] = base[0];
        offk[ninst] = base[1];
    }
}

// ---------------------------------------------------------------------------
// X drift
// ---------------------------------------------------------------------------
// Loads of an X tile (element e = row * cols + column) into registers
// (columns past the grid read 0); returns the velocity of this thread's
// column (0 past the grid).
__device__ __forceinline__ double x_load(const double* pool, const double* tables, const XItem& d,
                                         double (&pre)[kEnsPer]) {
    // element e = t + k * kEnsNT: column t mod cols for every k (kEnsNT is a
    // multiple of cols), row (t >> shift) + k * (kEnsNT >> shift)
    const int t = static_cast<int>(threadIdx.x);
    const int shift = d.shift, cols = 1 << shift, n = d.nx << shift;
    const int c = t & (cols - 1);
    const bool in_grid = c < d.ncol;
    const double* p = pool + d.f + static_cast<int64_t>(t >> shift) * d.nv + c;
    const int64_t step = static_cast<int64_t>(kEnsNT >> shift) * d.nv;
#pragma unroll
    for (int k = 0; k < kEnsPer; ++k) {
        pre[k] = (t + k * kEnsNT < n && in_grid) ? __builtin_nontemporal_load(p) : 0.0;
        p += step;
    }
    return in_grid ? tables[d.v + c] : 0.0;
}

// Window of the run of rows [r0, r0 + nr) (nr <= 8) of column c of the LDS
// tile (nx rows, row stride S): win[M] = row floor(r0 - disp) - 3 + M
// (periodic), M < nr + 7.  Compile-time indices throughout (here and in
// the run), so that the window lives in registers.
template <int M>
__device__ __forceinline__ void ens_xwin_at(double (&win)[kEnsWin], const double* tile, int S, int c, int nx, int nr,
                                            int j) {
    if constexpr (M < kEnsWin) {
        win[M] = M < nr + 7 ? tile[j * S + c] : 0.0;
        j += 1;
        j -= j >= nx ? nx : 0;
        ens_xwin_at<M + 1>(win, tile, S, c, nx, nr, j);
    }
}

// S: the tile's row stride, a compile-time constant, so that a window that
// does not wrap (the usual case) reads at immediate offsets from one address.
template <int S>
__device__ __forceinline__ void ens_xwin(const double* tile, int c, int r0, int nr, double disp, int nx,
                                         double (&win)[kEnsWin]) {
    int j = static_cast<int>(floor(static_cast<double>(r0) - disp)) - 3;  // in [-nx, 2nx): |disp| <= nx - 2
    j += j < 0 ? nx : 0;
    j -= j >= nx ? nx : 0;
    if (j + kEnsWin <= nx) {
        const double* q = tile + j * S + c;
#pragma unroll
        for (int M = 0; M < kEnsWin; ++M) win[M] = M < nr + 7 ? q[M * S] : 0.0;
    } else {
        ens_xwin_at<0>(win, tile, S, c, nx, nr, j);
    }
}

// Signed binade code of t (binade_code) from the exponent bits: for
// 1 <= |t| < 2^kBinades, b = floor(log2 |t|) = biased exponent - 1023 and
// the code is b (t > 0) or -b - 1 (t < 0); otherwise (|t| < 1, zero,
// subnormal, too large, inf, nan) kNoCode.
__device__ __forceinline__ int ens_code(double t) {
    const int hi = __double2hiint(t);
    const int b = ((hi >> 20) & 0x7ff) - 1023;
    const bool ok = static_cast<unsigned>(b) < static_cast<unsigned>(kBinades);
    return ok ? (hi >= 0 ? b : -b - 1) : kNoCode;
}

// Element I of a run: line index j0 + I, departure point dep, weights of its
// slot (the line's codes from q0, then the special rows at nq, nq + 1 for
// rows rs, rs + 1), read for every element; taps from the window at
// win[I + 1 + k], or (SEL, where rounding moved floor(t) by d = -1 or 1)
// one further or nearer.  Returns the six-tap sum.
template <int I, bool SEL>
__device__ __forceinline__ double ens_elem(const double (&win)[kEnsWin], int j0, int nr, double disp, double fl0,
                                           int rs, const double* tab, int q0, int nq) {
    const int j = j0 + I;
    const double dep = static_cast<double>(j) - disp;
    const int code = ens_code(dep);
    int slot = code != kNoCode ? code - q0 : nq + (j - rs);
    slot = I < nr ? slot : 0;
    double w[6];
    load_w6(tab + slot * 6, w);
    // (indices spelled out: constant from the start, the window stays in registers)
    double v[6] = {win[I + 1], win[I + 2], win[I + 3], win[I + 4], win[I + 5], win[I + 6]};
    if constexpr (SEL) {
        const double d = floor(dep) - (fl0 + static_cast<double>(I));
        if (d != 0.0) {
            const bool up = d > 0.0;
            v[0] = up ? win[I + 2] : win[I];
            v[1] = up ? win[I + 3] : win[I + 1];
            v[2] = up ? win[I + 4] : win[I + 2];
            v[3] = up ? win[I + 5] : win[I + 3];
            v[4] = up ? win[I + 6] : win[I + 4];
            v[5] = up ? win[I + 7] : win[I + 5];
        }
    }
    return tap_sum<double>(w, v);
}

// Whether rounding moved floor(t) of an element of the run off fl0 + i.
__device__ __forceinline__ bool ens_run_moved(int j0, int nr, double disp, double fl0) {
    bool moved = false;
#pragma unroll
    for (int i = 0; i < kEnsPer; ++i)
        moved = moved || (i < nr && floor(static_cast<double>(j0 + i) - disp) != fl0 + static_cast<double>(i));
    return moved;
}

// One X half-step of the run of rows [r0, r0 + nr) of column c from its
// window, written to the tile (row stride S); tables of the column at tab
// (codes q0 .. q0 + nq - 1, then the two special rows; stride 6).
template <int S, int I, bool SEL>
__device__ __forceinline__ void ens_xrun_at(double* col, int r0, int nr, double disp, double fl0, int rs,
                                            const double* tab, int q0, int nq, const double (&win)[kEnsWin]) {
    if constexpr (I < kEnsPer) {
        const double res = ens_elem<I, SEL>(win, r0, nr, disp, fl0, rs, tab, q0, nq);
        if (I < nr) col[I * S] = res;
        ens_xrun_at<S, I + 1, SEL>(col, r0, nr, disp, fl0, rs, tab, q0, nq, win);
    }
}

template <int S>
__device__ __forceinline__ void ens_xrun(double* tile, int c, int r0, int nr, double disp, int rs, const double* tab,
                                         int q0, int nq, const double (&win)[kEnsWin]) {
    const double fl0 = floor(static_cast<double>(r0) - disp);
    double* col = tile + r0 * S + c;  // row r0 of the thread's column: results at immediate offsets
    if (!ens_run_moved(r0, nr, disp, fl0)) {
        ens_xrun_at<S, 0, false>(col, r0, nr, disp, fl0, rs, tab, q0, nq, win);
    } else {
        asm volatile("");  // a branch of its own (rare)
        ens_xrun_at<S, 0, true>(col, r0, nr, disp, fl0, rs, tab, q0, nq, win);
    }
}

// The tile-pass part of an X item for a compile-time column count 1 << SH
// (row stride S = cols + 1).
template <int SH>
__device__ __forceinline__ void ens_xpass_tab(double* tile, int c, int r0, int nr, double disp, int rs,
                                              const double* tab, int q0, int nq, int nx) {
    constexpr int S = (1 << SH) + 1;
    double win[kEnsWin];
    ens_xwin<S>(tile, c, r0, nr, disp, nx, win);
    __syncthreads();  // every window has been read
    ens_xrun<S>(tile, c, r0, nr, disp, rs, tab, q0, nq, win);
}

// Element-wise X half-step of a whole tile (items the tables do not cover):
// out[r * cols + c] for every element, from the tile.
__device__ void ens_xpass_general(const double* tile, int S, int nx, int cols, int shift, const double* cdisp,
                                  double* out) {
    for (int e = threadIdx.x; e < (nx << shift); e += kEnsNT) {
        const int r = e >> shift, cc = e & (cols - 1);
        const double dep = static_cast<double>(r) - cdisp[cc];
        const double fl = floor(dep);
        double w[6], vv[6];
        weights_as<double>(dep - fl, w);
        int jr = wrap_index(static_cast<int64_t>(fl) - 2, nx);
        for (int k = 0; k < 6; ++k) {
            vv[k] = tile[jr * S + cc];
            jr += 1;
            jr -= (jr >= nx) ? nx : 0;
        }
        out[e] = tap_sum<double>(w, vv);
    }
}

// X drift (PASSES half-drifts) of every item of the list; with RHO the row
// sums of the result over the item's columns go to partial[g][x] as
// (sum, compensation).
// (two workgroups per CU: at most 128 VGPRs, four waves per SIMD)
template <int PASSES, bool RHO>
__global__ __launch_bounds__(kEnsNT) __attribute__((amdgpu_waves_per_eu(4))) void ens_xdrift_kernel(double* __restrict__ pool,
                                                            const double* __restrict__ tables, double half,
                                                            const int* __restrict__ offx, int ninst,
                                                            const XItem* __restrict__ items,
                                                            double2* __restrict__ partial, int* __restrict__ counter) {
    extern __shared__ __align__(16) double lds[];
    double* tile = lds;
    double* cdisp = lds + kEnsXDisp;
    double* tab = lds + kEnsXTab;  // column tables, or the output of the element-wise path
    __shared__ int s_next;
    const int total = offx[ninst];
    int g = blockIdx.x;
    if (g >= total) return;  // uniform over the workgroup
    const int t = threadIdx.x;
    if (t == 0) s_next = gridDim.x + atomicAdd(counter, 1);
    double pre[kEnsPer];
    XItem nd = items[g];
    double vpre = x_load(pool, tables, nd, pre);
    __syncthreads();
    int gn = s_next;  // the next item
    XItem nn = nd;
    if (gn < total) nn = items[gn];
    while (g < total) {
        if (t == 0) s_next = gridDim.x + atomicAdd(counter, 1);  // the one after
        const XItem cur = nd;
        const int nx = cur.nx, shift = cur.shift, cols = 1 << shift, S = cols + 1;
        const int n = nx << shift;
        const int c = t & (cols - 1);
        {
            // element e = t + k * kEnsNT: column c, row (t >> shift) + k * (kEnsNT >> shift)
            double* p = tile + (t >> shift) * S + c;
            const int step = (kEnsNT >> shift) * S;
#pragma unroll
            for (int k = 0; k < kEnsPer; ++k) {
                if (t + k * kEnsNT < n) *p = pre[k];
                p += step;
            }
        }
        const double disp = (vpre * half) / cur.dx;  // departure(), vlasov1D1V.py:229-238
        if (t < cols) cdisp[t] = disp;
        __syncthreads();
        const int q0 = cur.q0, nq = cur.nq;
        const bool fast = nq >= 0;          // uniform
        const int TS = ens_xtab_stride(nq + 2);
        if (fast) {
            // binade codes from the low threads, the special rows from the
            // high ones (different waves: their arithmetic differs)
            for (int e = t; e < cols * nq; e += kEnsNT) {
                const int cc = e / nq, s = e - cc * nq;
                ens_code_weights(cdisp[cc], q0 + s, tab + cc * TS + s * 6);
            }
            for (int u = kEnsNT - 1 - t; u < 2 * cols; u += kEnsNT) {
                const int cc = u >> 1, sp = u & 1;
                const double d = cdisp[cc];
                ens_special_weights(d, ens_special_row(d) + sp, tab + cc * TS + (nq + sp) * 6);
            }
        }
        __syncthreads();
        // the next item's tile loads while this one is computed
        const int gnn = s_next;
        if (gn < total) {
            nd = nn;
            vpre = x_load(pool, tables, nd, pre);
            if (gnn < total) nn = items[gnn];
        }
        const int NG = kEnsNT >> shift;
        const int R = (nx + NG - 1) / NG;  // rows per thread (<= kEnsPer)
        const int r0 = (t >> shift) * R;
        const int nr = nx - r0 < R ? nx - r0 : R;
        const int rs = ens_special_row(disp);
        for (int pass = 0; pass < PASSES; ++pass) {
            if (fast) {
                // recompute the run's set-up in every pass (hoisted out of
                // the loop it would hold registers across it)
                double dp = disp;
                int r0p = r0;
                asm volatile("" : "+v"(dp), "+v"(r0p));
                const double* ct = tab + c * TS;
                if (shift == 4)
                    ens_xpass_tab<4>(tile, c, r0p, nr, dp, rs, ct, q0, nq, nx);
                else if (shift == 5)
                    ens_xpass_tab<5>(tile, c, r0p, nr, dp, rs, ct, q0, nq, nx);
                else
                    ens_xpass_tab<6>(tile, c, r0p, nr, dp, rs, ct, q0, nq, nx);
            } else {
                ens_xpass_general(tile, S, nx, cols, shift, cdisp, tab);
                __syncthreads();
                for (int e = t; e < n; e += kEnsNT) tile[(e >> shift) * S + (e & (cols - 1))] = tab[e];
            }
            __syncthreads();
        }
        if (c < cur.ncol) {
            const double* p = tile + (t >> shift) * S + c;
            const int step = (kEnsNT >> shift) * S;
            double* q = pool + cur.f + static_cast<int64_t>(t >> shift) * cur.nv + c;
            const int64_t qstep = static_cast<int64_t>(kEnsNT >> shift) * cur.nv;
#pragma unroll
            for (int k = 0; k < kEnsPer; ++k) {
                if (t + k * kEnsNT < n) __builtin_nontemporal_store(*p, q);
                p += step;
                q += qstep;
            }
        }
        if (RHO) {
            // row sums over the item's columns: tpr consecutive threads per row,
            // cols / tpr columns each, merged by DPP moves (columns past nv
            // hold exact zeros: their taps read zeros); the group's last lane
            // holds the row's sum
            int tpr = 1;
            while (tpr < 64 && tpr < cols && 2 * tpr * nx <= kEnsNT) tpr *= 2;
            const int row = t / tpr, part = t - row * tpr;
            const int cpt = cols / tpr;
            KSum k{0.0, 0.0};
            if (row < nx)
                for (int u = 0; u < cpt; ++u) ksum_add(k, tile[row * S + part * cpt + u]);
            k = group_ksum<true>(k, tpr);
            if (row < nx && part == tpr - 1) partial[static_cast<int64_t>(g) * kEnsMaxNx + row] = make_double2(k.s, k.c);
        }
        __syncthreads();  // the tile, tables and s_next are free for the next item
        g = gn;
        gn = gnn;
    }
}

// ---------------------------------------------------------------------------
// Field solve
// ---------------------------------------------------------------------------
// One instance per workgroup: rho from the X items' partial sums (merged in
// item order, eight loads at a time), E by poisson1d_block, E to
// Eout[inst * kEnsMaxNx + x], field energy ((1/2 sum E^2) dx) to
// energy[inst] (if energy is not null).
__global__ __launch_bounds__(kEnsThreads) void ens_poisson_kernel(const vp_ens_inst* __restrict__ inst,
                                                                  const double* __restrict__ tables,
                                                                  const double2* __restrict__ partial,
                                                                  const int* __restrict__ offx, int physics_mode,
                                                                  double* __restrict__ Eout,
                                                                  double* __restrict__ energy,
                                                                  int* __restrict__ kick_counter) {
    __shared__ double rho[kEnsMaxNx];
    __shared__ double El[kEnsMaxNx];
    __shared__ double2 tw[kEnsMaxNx];
    __shared__ double2 Ek[kEnsMaxNx / 2 + 1];
    __shared__ double2 Pk[kEnsMaxNx / 2 + 1];
    __shared__ double scratch[2 * kEnsThreads / kWave];
    if (blockIdx.x == 0 && threadIdx.x == 0) *kick_counter = 0;  // for the kick that follows
    const vp_ens_inst I = inst[blockIdx.x];
    const int nx = I.nx;
    double* E = Eout + static_cast<int64_t>(blockIdx.x) * kEnsMaxNx;
    if (!physics_mode) {
        for (int x = threadIdx.x; x < nx; x += blockDim.x) E[x] = 0.0;
        if (energy && threadIdx.x == 0) energy[blockIdx.x] = 0.0;
        return;
    }
    const int i0 = offx[blockIdx.x], i1 = offx[blockIdx.x + 1];
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        KSum k{0.0, 0.0};
        const double2* pp = partial + static_cast<int64_t>(i0) * kEnsMaxNx + x;
        int it = 0;
        for (; it + 8 <= i1 - i0; it += 8) {
            double2 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = pp[(it + u) * kEnsMaxNx];
#pragma unroll
            for (int u = 0; u < 8; ++u) ksum_merge(k, KSum{p[u].x, p[u].y});
        }
        for (; it < i1 - i0; ++it) ksum_merge(k, KSum{pp[it * kEnsMaxNx].x, pp[it * kEnsMaxNx].y});
        // vlasov1D1V.py:559-561: n = sum(f, 1) * dv; rho = n - 1
        rho[x] = (((k.s + k.c) * I.dv) * 1.0) - 1.0;
    }
    __syncthreads();
    const int nk = nx / 2 + 1;
    poisson1d_block(rho, nx, tables + I.k_off, tables + I.k_off + nk, El, nullptr, tw, Ek, Pk);
    __syncthreads();
    KSum k{0.0, 0.0};
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        const double e = El[x];
        E[x] = e;
        ksum_add(k, e * e);
    }
    const double total = block_ksum(k, scratch);
    if (energy && threadIdx.x == 0) energy[blockIdx.x] = ((0.5 * total) * I.dx) * 1.0;
}

// ---------------------------------------------------------------------------
// V kick
// ---------------------------------------------------------------------------
// Loads of a kick block (elements e = row * nv + j, contiguous in the pool)
// into registers; returns the field of row t (t < rows).
__device__ __forceinline__ double k_load(const double* pool, const double* Ein, const KItem& d,
                                         double (&pre)[kEnsPer]) {
    const int n = d.rows * d.nv;
#pragma unroll
    for (int k = 0; k < kEnsPer; ++k) {
        const int e = static_cast<int>(threadIdx.x) + k * kEnsNT;
        pre[k] = e < n ? __builtin_nontemporal_load(pool + d.f + e) : 0.0;
    }
    return static_cast<int>(threadIdx.x) < d.rows ? Ein[d.e + threadIdx.x] : 0.0;
}

// Window of the run of elements [j0, j0 + nr) (nr <= 8) of a row whose
// padded LDS line starts at flat index pb (element j at pb + kEnsKPad + j):
// win[M] = element floor(j0 - disp) - 3 + M, M < nr + 7.
template <int M>
__device__ __forceinline__ void ens_kwin_at(double (&win)[kEnsWin], const double* tile, int p0, int nr) {
    if constexpr (M < kEnsWin) {
        win[M] = M < nr + 7 ? tile[kidx(p0 + M)] : 0.0;
        ens_kwin_at<M + 1>(win, tile, p0, nr);
    }
}

__device__ __forceinline__ void ens_kwin(const double* tile, int pb, int j0, int nr, double disp,
                                         double (&win)[kEnsWin]) {
    // >= pb: |disp| < 5
    ens_kwin_at<0>(win, tile, pb + kEnsKPad - 3 + static_cast<int>(floor(static_cast<double>(j0) - disp)), nr);
}

// One kick of the run of elements [j0, j0 + nr) of a row from its window,
// written to the tile (element j at flat index pb + kEnsKPad + j); row
// tables at tab (codes q0 .. q0 + kEnsKCodes - 1, special rows at
// kEnsKCodes, kEnsKCodes + 1).
template <int I, bool SEL>
__device__ __forceinline__ void ens_krun_at(double* tile, int pb, int j0, int nr, double disp, double fl0, int rs,
                                            const double* tab, int q0, const double (&win)[kEnsWin]) {
    if constexpr (I < kEnsPer) {
        const double res = ens_elem<I, SEL>(win, j0, nr, disp, fl0, rs, tab, q0, kEnsKCodes);
        if (I < nr) tile[kidx(pb + kEnsKPad + j0 + I)] = res;
        ens_krun_at<I + 1, SEL>(tile, pb, j0, nr, disp, fl0, rs, tab, q0, win);
    }
}

__device__ __forceinline__ void ens_krun(double* tile, int pb, int j0, int nr, double disp, int rs,
                                         const double* tab, int q0, const double (&win)[kEnsWin]) {
    const double fl0 = floor(static_cast<double>(j0) - disp);
    if (!ens_run_moved(j0, nr, disp, fl0)) {
        ens_krun_at<0, false>(tile, pb, j0, nr, disp, fl0, rs, tab, q0, win);
    } else {
        asm volatile("");  // a branch of its own (rare)
        ens_krun_at<0, true>(tile, pb, j0, nr, disp, fl0, rs, tab, q0, win);
    }
}

// V kick (dt, velocity E(x)) of every item of the list (blocks of whole
// rows), taps outside [0, nv) read as zero (clamp + mask,
// vlasov1D1V.py:248-253; in LDS every row has kEnsKPad zeros on each side,
// which give the same sums: a masked tap adds a zero product).  Fast path
// (every row of the item has |displacement| < 5 cells, so every tap of a
// row lies within its pads, and its binade codes fit the table): one thread
// per 8-element chunk of a row.  Otherwise element-wise for the whole item.
__global__ __launch_bounds__(kEnsNT) __attribute__((amdgpu_waves_per_eu(4))) void ens_kick_kernel(double* __restrict__ pool,
                                                          const double* __restrict__ Ein, double dt,
                                                          const int* __restrict__ offk, int ninst,
                                                          const KItem* __restrict__ items, int* __restrict__ counter,
                                                          int* __restrict__ x_counter) {
    extern __shared__ __align__(16) double lds[];
    double* tile = lds;
    double* tab = lds + kEnsKTab;  // row tables, or the element-wise path's output
    double* kdisp = lds + kEnsKAux;
    int* krs = reinterpret_cast<int*>(kdisp + kEnsKRows);
    int* kq0 = krs + kEnsKRows;
    int* knq = kq0 + kEnsKRows;
    __shared__ int slow_item;
    __shared__ int s_next;
    if (blockIdx.x == 0 && threadIdx.x == 0) *x_counter = 0;  // for the X drift that follows
    const int total = offk[ninst];
    int g = blockIdx.x;
    if (g >= total) return;  // uniform over the workgroup
    const int t = threadIdx.x;
    if (t == 0) s_next = gridDim.x + atomicAdd(counter, 1);
    double pre[kEnsPer];
    KItem nd = items[g];
    double epre = k_load(pool, Ein, nd, pre);
    __syncthreads();
    int gn = s_next;  // the next item
    KItem nn = nd;
    if (gn < total) nn = items[gn];
    while (g < total) {
        if (t == 0) s_next = gridDim.x + atomicAdd(counter, 1);  // the one after
        const KItem cur = nd;
        const int nv = cur.nv, rows = cur.rows, cpr = cur.cpr;
        const int KS = nv + 2 * kEnsKPad;
        const int n = rows * nv;
        // element e = t + k * kEnsNT of the block: row rr, column jj
        const int qn = kEnsNT / nv, rn = kEnsNT - qn * nv;
        const int rr0 = t / nv, jj0 = t - rr0 * nv;
        {
            int rr = rr0, jj = jj0;
#pragma unroll
            for (int k = 0; k < kEnsPer; ++k) {
                if (t + k * kEnsNT < n) tile[kidx(rr * KS + kEnsKPad + jj)] = pre[k];
                rr += qn;
                jj += rn;
                if (jj >= nv) {
                    jj -= nv;
                    ++rr;
                }
            }
        }
        if (t < rows * 2 * kEnsKPad) {  // the pads
            const int rr = t / (2 * kEnsKPad), q = t - rr * 2 * kEnsKPad;
            tile[kidx(rr * KS + (q < kEnsKPad ? q : nv + q))] = 0.0;
        }
        if (t < 64) {  // one wave (rows <= 32): displacements, codes, and whether the tables cover the item
            bool bad = false;
            if (t < rows) {
                const double disp = (epre * dt) / cur.dv;  // departure(), vlasov1D1V.py:229-238
                int q0, nq;
                segment_codes(-disp, static_cast<double>(nv - 1) - disp, q0, nq);
                kdisp[t] = disp;
                krs[t] = ens_special_row(disp);
                kq0[t] = q0;
                knq[t] = nq;
                bad = !(fabs(disp) < 5.0) || nq > kEnsKCodes;
            }
            const bool any_bad = __any(bad);
            if (t == 0) slow_item = any_bad ? 1 : 0;
        }
        __syncthreads();
        const bool fast = slow_item == 0;  // uniform
        if (fast) {
            for (int e = t; e < rows * kEnsKCodes; e += kEnsNT) {
                const int l = e / kEnsKCodes, s = e - l * kEnsKCodes;
                if (s < knq[l]) ens_code_weights(kdisp[l], kq0[l] + s, tab + l * kEnsKT + s * 6);
            }
            for (int u = kEnsNT - 1 - t; u < rows * 2; u += kEnsNT) {
                const int l = u >> 1, sp = u & 1;
                ens_special_weights(kdisp[l], krs[l] + sp, tab + l * kEnsKT + (kEnsKCodes + sp) * 6);
            }
        }
        __syncthreads();
        const int gnn = s_next;
        if (gn < total) {
            nd = nn;
            epre = k_load(pool, Ein, nd, pre);
            if (gnn < total) nn = items[gnn];
        }
        if (fast) {
            const int rr = t / cpr, ch = t - rr * cpr;
            const int j0 = ch * kEnsPer;
            const bool act = rr < rows;
            const int nr = act ? (nv - j0 < kEnsPer ? nv - j0 : kEnsPer) : 0;
            const int rv = act ? rr : 0;
            const double disp = kdisp[rv];
            double win[kEnsWin];
            ens_kwin(tile, rv * KS, j0, nr, disp, win);
            __syncthreads();  // every window has been read
            ens_krun(tile, rv * KS, j0, nr, disp, krs[rv], tab + rv * kEnsKT, kq0[rv], win);
        } else {
            // element-wise, into the table region, then into the tile
            for (int e = t; e < n; e += kEnsNT) {
                const int rr = e / nv, jj = e - rr * nv;
                const double dep = static_cast<double>(jj) - kdisp[rr];
                const double fl = floor(dep);
                double w[6], vv[6];
                weights_as<double>(dep - fl, w);
                const int64_t il = static_cast<int64_t>(fl);
                for (int k = 0; k < 6; ++k) {
                    const int64_t jt = il + (k - 2);
                    const int jc = static_cast<int>(jt < 0 ? 0 : (jt > nv - 1 ? nv - 1 : jt));
                    const double m = (jt >= 0 && jt < nv) ? 1.0 : 0.0;
                    vv[k] = tile[kidx(rr * KS + kEnsKPad + jc)] * m;
                }
                tab[e] = tap_sum<double>(w, vv);
            }
            __syncthreads();
            for (int e = t; e < n; e += kEnsNT) {
                const int rr = e / nv;
                tile[kidx(rr * KS + kEnsKPad + (e - rr * nv))] = tab[e];
            }
        }
        __syncthreads();
        {
            int rr = rr0, jj = jj0;
#pragma unroll
            for (int k = 0; k < kEnsPer; ++k) {
                const int e = t + k * kEnsNT;
                if (e < n) __builtin_nontemporal_store(tile[kidx(rr * KS + kEnsKPad + jj)], pool + cur.f + e);
                rr += qn;
                jj += rn;
                if (jj >= nv) {
                    jj -= nv;
                    ++rr;
                }
            }
        }
        __syncthreads();  // the tile, tables and s_next are free for the next item
        g = gn;
        gn = gnn;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int ens_xtile() {
    const int c = knob(VP_KNOB_ENS_XTILE);
    return c == 2048 ? 2048 : 4096;
}

static int ens_ktile() {
    const int r = knob(VP_KNOB_ENS_KTILE);
    return r == 2048 ? 2048 : 4096;
}

// persistent workgroups per CU (the LDS lets two run at once)
static int ens_wgs_per_cu() {
    const int w = knob(VP_KNOB_ENS_WGS);
    return w == 1 ? 1 : 2;
}

static int ens_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cached[dev] = v;
    }
    return cached[dev];
}

static size_t a256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Upper bounds of the item counts (the smallest tiles give the most items).
static int64_t ens_max_xitems(int ninst, int max_nv) {
    return static_cast<int64_t>(ninst) * ((max_nv + 15) / 16);
}
static int64_t ens_max_kitems(int ninst, int max_nx, int max_nv) {
    int most = 0;
    for (int nv = 6; nv <= max_nv; ++nv) {  // the most row blocks any nv <= max_nv gives
        const int rows = ens_krows(nv, 2048);
        const int c = (max_nx + rows - 1) / rows;
        most = c > most ? c : most;
    }
    return static_cast<int64_t>(ninst) * most;
}

struct EnsWs {
    double2* partial;
    double* E;
    int* counters;  // X drift, kick
    int* offx;
    int* offk;
    XItem* itx;
    KItem* itk;
    size_t bytes;
};

static EnsWs ens_layout(void* ws, int ninst, int max_nx, int max_nv) {
    EnsWs L{};
    unsigned char* p = static_cast<unsigned char*>(ws);
    size_t off = 0;
    auto take = [&](size_t b) {
        unsigned char* q = p ? p + off : nullptr;
        off += a256(b);
        return q;
    };
    const int64_t nxi = ens_max_xitems(ninst, max_nv), nki = ens_max_kitems(ninst, max_nx, max_nv);
    L.partial = reinterpret_cast<double2*>(take(static_cast<size_t>(nxi) * kEnsMaxNx * sizeof(double2)));
    L.E = reinterpret_cast<double*>(take(static_cast<size_t>(ninst) * kEnsMaxNx * sizeof(double)));
    L.counters = reinterpret_cast<int*>(take(2 * sizeof(int)));
    L.offx = reinterpret_cast<int*>(take(static_cast<size_t>(ninst + 1) * sizeof(int)));
    L.offk = reinterpret_cast<int*>(take(static_cast<size_t>(ninst + 1) * sizeof(int)));
    L.itx = reinterpret_cast<XItem*>(take(static_cast<size_t>(nxi) * sizeof(XItem)));
    L.itk = reinterpret_cast<KItem*>(take(static_cast<size_t>(nki) * sizeof(KItem)));
    L.bytes = off;
    return L;
}

static int ens_set_lds() {
    static bool done = false;
    if (done) return VP_OK;
    const int bytes = kEnsLdsDoubles * static_cast<int>(sizeof(double));
    const void* fns[] = {reinterpret_cast<const void*>(&ens_xdrift_kernel<1, true>),
                         reinterpret_cast<const void*>(&ens_xdrift_kernel<1, false>),
                         reinterpret_cast<const void*>(&ens_xdrift_kernel<2, true>),
                         reinterpret_cast<const void*>(&ens_xdrift_kernel<2, false>),
                         reinterpret_cast<const void*>(&ens_kick_kernel)};
    for (const void* f : fns) VP_TRY_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done = true;
    return VP_OK;
}

}  // namespace vp

extern "C" size_t vp_ens_workspace(int ninst, int max_nx, int max_nv) {
    if (ninst < 1 || max_nx < 1 || max_nx > vp::kEnsMaxNx || max_nv < 1 || max_nv > vp::kEnsMaxNv) return 0;
    return vp::ens_layout(nullptr, ninst, max_nx, max_nv).bytes;  // valid for every tile size
}

extern "C" int vp_ens_run_f64(double* pool, const vp_ens_inst* inst, int ninst, int max_nx, int max_nv,
                              const double* tables, int nsteps, double dt, int physics_mode, double* energy,
                              void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_RETURN_IF(!pool || !inst || !tables || !ws || nsteps < 0, VP_ERR_ARG);
    VP_RETURN_IF(ninst < 1 || ninst > 65535, VP_ERR_SHAPE);
    VP_RETURN_IF(max_nx < 6 || max_nx > vp::kEnsMaxNx || max_nv < 6 || max_nv > vp::kEnsMaxNv, VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < vp_ens_workspace(ninst, max_nx, max_nv), VP_ERR_WORKSPACE);
    if (nsteps == 0) return VP_OK;
    VP_TRY(vp::ens_set_lds());
    hipStream_t s = static_cast<hipStream_t>(stream);
    const vp::EnsWs L = vp::ens_layout(ws, ninst, max_nx, max_nv);
    const int xtile = vp::ens_xtile(), ktile = vp::ens_ktile();
    const double half = dt / 2.0;
    hipLaunchKernelGGL(vp::ens_setup_kernel, dim3(1), dim3(vp::kEnsNT), 0, s, inst, ninst, tables, half, xtile,
                       ktile, L.offx, L.offk, L.itx, L.itk, L.counters);
    VP_CHECK_LAUNCH();
    const int wgs = vp::ens_cu_count() * vp::ens_wgs_per_cu();
    const size_t lds = static_cast<size_t>(vp::kEnsLdsDoubles) * sizeof(double);
    auto xdrift = [&](int passes, bool rho) -> int {
#define VP_X(P, R)                                                                                               \
    hipLaunchKernelGGL((vp::ens_xdrift_kernel<P, R>), dim3(wgs), dim3(vp::kEnsNT), lds, s, pool, tables, half, \
                       L.offx, ninst, L.itx, L.partial, L.counters)
        if (passes == 2) {
            if (rho) VP_X(2, true); else VP_X(2, false);
        } else {
            if (rho) VP_X(1, true); else VP_X(1, false);
        }
#undef VP_X
        VP_CHECK_LAUNCH();
        return VP_OK;
    };
    VP_TRY(xdrift(1, physics_mode != 0));
    for (int st = 0; st < nsteps; ++st) {
        hipLaunchKernelGGL(vp::ens_poisson_kernel, dim3(ninst), dim3(vp::kEnsThreads), 0, s, inst, tables, L.partial,
                           L.offx, physics_mode, L.E, energy ? energy + static_cast<int64_t>(st) * ninst : nullptr,
                           L.counters + 1);
        VP_CHECK_LAUNCH();
        hipLaunchKernelGGL(vp::ens_kick_kernel, dim3(wgs), dim3(vp::kEnsNT), lds, s, pool, L.E, dt, L.offk, ninst,
                           L.itk, L.counters + 1, L.counters);
        VP_CHECK_LAUNCH();
        if (st + 1 < nsteps)
            VP_TRY(xdrift(2, physics_mode != 0));
        else
            VP_TRY(xdrift(1, false));
    }
    return VP_OK;
}
