// Fused, in-place 2D2V sweeps for f[nx][ny][nvx][nvy] (float64).
//
// The generic sweeps (advect.hip) evaluate the departure point and the six
// Lagrange weights per element.  Here the weights are hoisted:
//
//  * drift_table_kernel: the spatial drifts have time-constant velocities
//    (vx for X, vy for Y), so the weights of every (row, velocity) pair are
//    computed once per call with the exact reference arithmetic
//    (vlasov2D2V.py:330-353) into a small table.
//  * drift_kernel: a persistent workgroup owns one column chunk (32 vy) of
//    a family of lines, keeps the weights of its (row, column) pairs in
//    registers, and streams tile after tile (rows x 32 columns) through LDS
//    with a one-tile register prefetch.  X drift (rows = x): the workgroup
//    fixes (vx, vy chunk) and walks y, and can apply the drift twice in LDS
//    (the X(dt/2) that ends one step and the X(dt/2) that starts the next).
//    Y drift (rows = y): the workgroup fixes (x, vy chunk), walks vx and
//    accumulates the charge density of its rows with Neumaier compensation.
//  * kick_kernel: for one (x, y) the whole (vx, vy) plane is a closed unit
//    for both kicks (vlasov2D2V.py:976-986), so one workgroup loads the plane
//    into LDS once, applies the Vx kick then the Vy kick, and writes it back.
//
// Tiles and planes are disjoint between workgroups and each is read whole
// before it is written, so all three kernels may run in place.
#include <type_traits>

#include "fast2d.hpp"

namespace vp {

constexpr int kCol = 32;       // columns (contiguous vy) per drift tile
constexpr int kDriftThreads = 1024;
constexpr int kGroups = kDriftThreads / kCol;  // row groups per drift tile

// Weights and wrapped left index for every (row r, velocity index j):
// W[(r*6 + k)*nvel + j], IL[r*nvel + j] = floor(r - (vel[j]*dt)/dx) mod n.
__global__ void drift_table_kernel(const double* __restrict__ vel, int nvel, int n, double dt, double dx,
                                   double* __restrict__ W, int* __restrict__ IL) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nvel * n) return;
    const int r = idx / nvel;
    const int j = idx - r * nvel;
    const Departure d = departure(r, vel[j], dt, dx);
    double w[6];
    lagrange5_weights<double>(d.alpha, w);
#pragma unroll
    for (int k = 0; k < 6; ++k) W[(static_cast<int64_t>(r) * 6 + k) * nvel + j] = w[k];
    IL[static_cast<int64_t>(r) * nvel + j] = wrap_index(d.il, n);
}

template <int R>
__device__ __forceinline__ double periodic_tap_sum(const double* tile, int il, const double (&w)[6], int n, int c) {
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int j = il + k - 2;
        j += (j < 0) ? n : 0;
        j -= (j >= n) ? n : 0;
        v[k] = tile[j * kCol + c];
    }
    return tap_sum<double>(w, v);
}

// Persistent line sweep; see the file comment.  Layout of one workgroup:
//   a  = fixed outer index (X: vx, Y: x),  q = vy chunk,  s = split of the loop
//   element (row r, column c, loop l) = fin[a*a_stride + q*32 + c + r*row_stride + l*loop_stride]
// LANE_W: the weight column is the vy index (Y drift) instead of a (X drift).
template <int R, int PASSES, bool RHO, bool LANE_W>
__global__ __launch_bounds__(kDriftThreads) void drift_kernel(
    const double* fin, double* fout, int nq, int S, int64_t a_stride, int64_t row_stride,
    int64_t loop_stride, int loop_total, const double* __restrict__ W, const int* __restrict__ IL, int nvel,
    double2* __restrict__ partial) {
    constexpr int n = R * kGroups;
    __shared__ double tile[n * kCol];
    const int t = threadIdx.x;
    const int c = t & (kCol - 1);
    const int g = t / kCol;
    int b = blockIdx.x;
    const int s = b % S;
    b /= S;
    const int q = b % nq;
    const int a = b / nq;
    const int64_t base = static_cast<int64_t>(a) * a_stride + q * kCol + c;
    const int j = LANE_W ? q * kCol + c : a;

    double w[R][6];
    int il[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = g + kGroups * i;
#pragma unroll
        for (int k = 0; k < 6; ++k) w[i][k] = W[(static_cast<int64_t>(r) * 6 + k) * nvel + j];
        il[i] = IL[static_cast<int64_t>(r) * nvel + j];
    }
    // Consume the weight loads here: otherwise the wait-count analysis of the
    // loop body also waits for the in-flight tile prefetch before each FMA.
#pragma unroll
    for (int i = 0; i < R; ++i) {
#pragma unroll
        for (int k = 0; k < 6; ++k) asm volatile("" ::"v"(w[i][k]));
        asm volatile("" ::"v"(il[i]));
    }
    const int l0 = static_cast<int>(static_cast<int64_t>(s) * loop_total / S);
    const int l1 = static_cast<int>(static_cast<int64_t>(s + 1) * loop_total / S);
    // Software pipeline: tile l is in LDS while tile l+1 is in flight into
    // `cur`.  The wait for `cur` comes after the stores of tile l, on the
    // first iteration as on every later one, so the wait-count analysis waits
    // for the loads only (stores may still be draining).
    double cur[R];
    KSum acc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i] = KSum{0.0, 0.0};
    if (l0 < l1) {
#pragma unroll
        for (int i = 0; i < R; ++i) cur[i] = fin[base + (g + kGroups * i) * row_stride + l0 * loop_stride];
#pragma unroll
        for (int i = 0; i < R; ++i) tile[(g + kGroups * i) * kCol + c] = cur[i];
        __syncthreads();
        const int ln = (l0 + 1 < l1) ? l0 + 1 : l0;
#pragma unroll
        for (int i = 0; i < R; ++i) cur[i] = fin[base + (g + kGroups * i) * row_stride + ln * loop_stride];
    }
    for (int l = l0; l < l1; ++l) {
        if (PASSES == 2) {
            double out[R];
#pragma unroll
            for (int i = 0; i < R; ++i) out[i] = periodic_tap_sum<R>(tile, il[i], w[i], n, c);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < R; ++i) tile[(g + kGroups * i) * kCol + c] = out[i];
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const double o = periodic_tap_sum<R>(tile, il[i], w[i], n, c);
            fout[base + (g + kGroups * i) * row_stride + l * loop_stride] = o;
            if (RHO) ksum_add(acc[i], o);
        }
        __syncthreads();  // every read of tile l done
#pragma unroll
        for (int i = 0; i < R; ++i) tile[(g + kGroups * i) * kCol + c] = cur[i];
        __syncthreads();
        const int ln = (l + 2 < l1) ? l + 2 : l1 - 1;  // the last iterations re-read a tile they own
#pragma unroll
        for (int i = 0; i < R; ++i) cur[i] = fin[base + (g + kGroups * i) * row_stride + ln * loop_stride];
    }
    if (RHO) {
        const int nparts = nq * S;
        // part-major layout [part][x * ny + y]: the Poisson kernel reads it coalesced
        const int64_t npoints = static_cast<int64_t>(gridDim.x / nparts) * n;
        const KSum k = rows_ksum<true, R, kCol>(acc, c);
        if (c < R) {
            const int r = g + kGroups * rows_ksum_row<R>(c);
            partial[static_cast<int64_t>(q * S + s) * npoints + static_cast<int64_t>(a) * n + r] =
                make_double2(k.s, k.c);
        }
    }
}

// Wait until at most N vector memory operations of this wave are outstanding.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Direct-to-LDS load of 16 bytes per lane: lane i's bytes land at LDS byte
// address lds + 16 * i.  Written as inline asm so that the compiler does not
// treat the pending LDS write as a hazard for every later LDS read (it would
// otherwise wait for all such loads, including the prefetches, before each
// read); the caller orders it with wait_vmcnt and a barrier.
template <bool NT>
__device__ __forceinline__ void load_lds_x4(const void* gp, uint32_t lds) {
    // M0 holds the wave's LDS base address; binding it as an operand lets the
    // compiler set it (and know it is live) instead of the asm clobbering it.
    if (NT)
        asm volatile("global_load_lds_dwordx4 %1, off nt" ::"{m0}"(lds), "v"(gp) : "memory");
    else
        asm volatile("global_load_lds_dwordx4 %1, off" ::"{m0}"(lds), "v"(gp) : "memory");
}

// Store with the non-temporal (streaming) hint when NT: f is far larger than
// every cache level, so nothing written by a sweep is re-read from cache.
template <bool NT, typename S>
__device__ __forceinline__ void store_f(S* p, S v) {
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <typename T>
__device__ __forceinline__ uint32_t lds_addr(T* p) {
    typedef __attribute__((address_space(3))) T* LdsPtr;
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>((LdsPtr)(p)));
}

// Tile geometry of the LDS-ring drift for storage type S: a tile row is 256
// contiguous bytes (32 float64 or 64 float32 columns); 1024 threads.
template <typename S>
struct RingGeom {
    static constexpr int col = 256 / static_cast<int>(sizeof(S));
    static constexpr int groups = kDriftThreads / col;
};

// Element of the LDS ring: one value of the storage type S (VW = 1), or a
// pair of neighbouring float32 columns (VW = 2) that share their weights and
// tap rows, moved as 8-byte loads and stores.
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <typename S, int VW>
struct LaneElem {
    using type = S;
};
template <>
struct LaneElem<float, 2> {
    using type = f32x2;
};

template <typename S>
__device__ __forceinline__ S tap_sum_e(const S (&w)[6], const S (&v)[6]) {
    return tap_sum<S>(w, v);
}
// Each column of the pair: the scalar float32 sum, term for term.
__device__ __forceinline__ f32x2 tap_sum_e(const float (&w)[6], const f32x2 (&v)[6]) {
    float a[6], b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a[k] = v[k].x;
        b[k] = v[k].y;
    }
    f32x2 o;
    o.x = tap_sum<float>(w, a);
    o.y = tap_sum<float>(w, b);
    return o;
}

// Six taps of one row from a tile of n (power of two) rows of 256 bytes.
// A = byte offset of (row il - 2 mod n, this thread's column); tap k is at
// (A + 256 k) mod 256 n: two integer operations per tap from one register per
// row, instead of six precomputed offsets (which the compiler would otherwise
// hoist out of the tile loop: 6 registers per row).
// The same six taps when none of them wraps around the tile (tap-0 row at
// most n - 6): one base address, the taps at immediate offsets.
template <typename E, typename S>
__device__ __forceinline__ E ring_tap_sum_direct(const E* tile, uint32_t A, const S (&w)[6]) {
    const char* p = reinterpret_cast<const char*>(tile) + A;
    E v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = *reinterpret_cast<const E*>(p + 256 * k);
    return tap_sum_e(w, v);
}

template <typename E, typename S, int N>
__device__ __forceinline__ E ring_tap_sum(const E* tile, uint32_t A, const S (&w)[6]) {
    const char* base = reinterpret_cast<const char*>(tile);
    E v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = *reinterpret_cast<const E*>(base + ((A + 256u * k) & (256u * N - 1)));
    return tap_sum_e(w, v);
}

// A pair of float32 columns with weights of their own (w0, w1), taps from
// the same rows (their departure rows agree): the two scalar sums.
__device__ __forceinline__ f32x2 tap_sum_e2(const float (&w0)[6], const float (&w1)[6], const f32x2 (&v)[6]) {
    float a[6], b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a[k] = v[k].x;
        b[k] = v[k].y;
    }
    f32x2 o;
    o.x = tap_sum<float>(w0, a);
    o.y = tap_sum<float>(w1, b);
    return o;
}

template <bool DIRECT, int N>
__device__ __forceinline__ f32x2 ring_pair_sum(const f32x2* tile, uint32_t A, const float (&w0)[6],
                                               const float (&w1)[6]) {
    const char* base = reinterpret_cast<const char*>(tile);
    f32x2 v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        v[k] = *reinterpret_cast<const f32x2*>(base + (DIRECT ? A + 256u * k : (A + 256u * k) & (256u * N - 1)));
    return tap_sum_e2(w0, w1, v);
}

// One float32 column's six taps at byte offset A (tap-0 row) of the ring
// tile, wrapping: for a pair whose two columns depart from different rows.
template <int N>
__device__ __forceinline__ float ring_col_sum(const f32x2* tile, uint32_t A, const float (&w)[6]) {
    const char* base = reinterpret_cast<const char*>(tile);
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = *reinterpret_cast<const float*>(base + ((A + 256u * k) & (256u * N - 1)));
    return tap_sum<float>(w, v);
}

// Same sweep as drift_kernel, but the tiles stream into an LDS ring of T
// slots with direct-to-LDS loads (global_load_lds_dwordx4: each lane moves 16
// bytes, a wave fills 1 KB = 4 tile rows), so T - 1 tiles are in flight per
// workgroup without holding registers; the registers keep the weights.
// S = storage (and interpolation) type of f: double, or float for the
// FP32-storage variant (weights from the float64 table rounded to float, the
// charge density summed in float64 from the stored float values).
// Rows: n = R * 1024 / col in {64, 128} (P = n / 64 load instructions per wave
// per tile).
// Per iteration i (tile l = l0 + i): wait until this wave's loads of tile l
// have landed, barrier, start the load of tile l + T - 1 into the slot freed
// by tile l - 1, compute and store tile l (R stores).  The loads and stores
// share the in-order vmcnt counter, so after the load of tile l this wave
// has issued exactly (T - 2) * P + min(i, T - 1) * R further vector memory
// operations; waiting for vmcnt <= that number waits for tile l and nothing
// younger.  For T > 2 the tail iterations re-load the last tile (an L2 hit)
// so the count stays the same to the end; for T = 2 the only load a wait
// depends on is the one issued in the iteration before, and the last
// iteration issues none.
// VW = 2 (float32 storage): each thread handles a pair of neighbouring
// columns as one 8-byte element, with the float64 drift's thread layout;
// results are those of VW = 1.  X drift: the pair shares its weights (they
// depend on the row's velocity).  Y drift (LANE_W): each column of the pair
// has its own weights and departure row; where the two departure rows differ
// for some lane of a wave (columns either side of a change of floor(-disp)),
// that wave reads the row's taps column by column.
// Units: the loop range of outer index a splits into nsplit parts; unit
// u = a * nsplit + s covers tiles [s * loop_total / nsplit, (s + 1) * ...).
// With one workgroup per unit (gridDim.x = na * nq * nsplit) workgroup b
// takes s = b % nsplit, q = (b / nsplit) % nq, a = b / (nsplit * nq).  With
// fewer (persistent) workgroups, G per column chunk (gridDim.x = nq * G; the
// Y drift only: its weights depend on the column, not on a), workgroup b
// takes chunk q = b % nq and walks the units [g U / G, (g + 1) U / G),
// g = b / nq, U = na * nsplit, as one stream of tiles: the ring prefetches
// across unit boundaries, the weights stay in registers, and ROT rotates the
// start of each unit's loop range by (13 u) mod its length, so that the
// workgroups running at one time read different loop indices (all starting
// at 0, they would read the same offsets within their rows at once).
// RHO: at the end of a unit the row sums of its tiles over the chunk's
// columns go to partial[(q * nsplit + s) * na * n + a * n + row].
template <typename S, int R, int PASSES, bool RHO, bool LANE_W, int T, bool NT, int VW = 1>
__global__ __launch_bounds__(kDriftThreads) void drift_dma_kernel(
    const S* fin_s, S* fout_s, int nq, int nsplit, int na, int rot, int64_t a_stride, int64_t row_stride,
    int64_t loop_stride, int loop_total, const double* __restrict__ W, const int* __restrict__ IL, int nvel,
    double2* __restrict__ partial) {
    using E = typename LaneElem<S, VW>::type;
    constexpr bool PAIRW = VW == 2 && LANE_W;  // pair with a weight set per column
    static_assert(VW == 1 || !RHO || PAIRW, "column pairs: rho with the Y drift only");
    static_assert(!(VW == 2 && LANE_W) || sizeof(S) == 4, "Y-drift column pairs: float32");
    static_assert(!PAIRW || PASSES == 1, "Y drift: one pass");
    const E* fin = reinterpret_cast<const E*>(fin_s);
    E* fout = reinterpret_cast<E*>(fout_s);
    a_stride /= VW;  // strides in elements of E
    row_stride /= VW;
    loop_stride /= VW;
    constexpr int KC = RingGeom<E>::col;
    constexpr int KG = RingGeom<E>::groups;
    constexpr int n = R * KG;
    constexpr int P = n / 64;
    constexpr int PER_LANE = 16 / static_cast<int>(sizeof(E));
    static_assert(P * 64 == n && P >= 1, "rows must be a multiple of 64");
    static_assert(T >= 2 && T <= 4, "ring of 2..4 tiles");  // the waits below cover T <= 4
    __shared__ __align__(16) E ring[T][n * KC];
    __shared__ __align__(16) E mid[PASSES == 2 ? n * KC : 1];  // first-pass result of the X(dt/2) pair
    const int t = threadIdx.x;
    const int c = t & (KC - 1);
    const int g = t / KC;
    const int wave = t >> 6;
    const int lane = t & 63;
    const int U = na * nsplit;
    int q, u_begin, u_end;
    if (static_cast<int>(gridDim.x) == U * nq) {  // one workgroup per unit
        const int b = blockIdx.x;
        q = (b / nsplit) % nq;
        u_begin = (b / (nsplit * nq)) * nsplit + b % nsplit;
        u_end = u_begin + 1;
    } else {
        const int G = static_cast<int>(gridDim.x) / nq, b = static_cast<int>(blockIdx.x);
        int gw;
        if ((rot & 2) && G % 8 == 0) {
            // consecutive workgroups (dealt round-robin to the 8 XCDs) in
            // groups of 8 per chunk: every XCD works on every chunk
            q = (b / 8) % nq;
            gw = (b / (8 * nq)) * 8 + b % 8;
        } else {
            q = b % nq;
            gw = b / nq;
        }
        u_begin = static_cast<int>(static_cast<int64_t>(gw) * U / G);
        u_end = static_cast<int>(static_cast<int64_t>(gw + 1) * U / G);
    }
    if (u_begin >= u_end) return;  // uniform over the workgroup
    const int64_t base = q * KC + c;  // + a * a_stride of the unit
    const int j = LANE_W ? (PAIRW ? 2 * (q * KC + c) : q * KC + c) : u_begin / nsplit;

    // Walk over the units' tiles (wave-uniform): tile k of unit u is loop
    // index l0 + (r0 + k) mod len.
    struct Walk {
        int u, l0, len, r0, k;
    };
    auto start = [&](Walk& w, int u) {
        const int s = u % nsplit;
        w.u = u;
        w.l0 = static_cast<int>(static_cast<int64_t>(s) * loop_total / nsplit);
        w.len = static_cast<int>(static_cast<int64_t>(s + 1) * loop_total / nsplit) - w.l0;
        w.r0 = (rot & 1) && w.len > 0 ? (13 * u) % w.len : 0;
        w.k = 0;
    };
    auto tile_of = [](const Walk& w) {
        const int x = w.r0 + w.k;
        return w.l0 + (x >= w.len ? x - w.len : x);
    };
    int ntiles = 0;
    for (int u = u_begin; u < u_end; ++u) {
        Walk w;
        start(w, u);
        ntiles += w.len;
    }
    if (ntiles == 0) return;  // uniform over the workgroup

    // Load instruction k of this wave covers tile rows 4m .. 4m+3, m = wave + 16k;
    // lane -> row 4m + lane/16, the 16 bytes at column PER_LANE*(lane%16).
    const E* src_lane[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int row = 4 * (wave + 16 * k) + (lane >> 4);
        src_lane[k] = fin + q * KC + PER_LANE * (lane & 15) + row * row_stride;
    }
    uint32_t dst_wave[T];
#pragma unroll
    for (int u = 0; u < T; ++u) dst_wave[u] = __builtin_amdgcn_readfirstlane(lds_addr(&ring[u][4 * wave * KC]));
    // Loads the tile under the issue walk into `slot` and moves the walk on;
    // at the end of the last unit it stays on the last tile.
    Walk iw;
    start(iw, u_begin);
    auto issue = [&](int slot) {
        const int64_t off = static_cast<int64_t>(iw.u / nsplit) * a_stride + static_cast<int64_t>(tile_of(iw)) * loop_stride;
#pragma unroll
        for (int k = 0; k < P; ++k) load_lds_x4<NT>(src_lane[k] + off, dst_wave[slot] + k * (16 * 1024));
        if (iw.k + 1 < iw.len)
            ++iw.k;
        else if (iw.u + 1 < u_end)
            start(iw, iw.u + 1);
    };
#pragma unroll
    for (int u = 0; u < T - 1; ++u) issue(u);

    // The weights load while the first tiles are in flight (their loads were
    // issued first: a workgroup lives for a few tiles, and the table reads
    // would otherwise delay its first tile by an L2 round trip).
    S w[R][6];  // weights (of the first column when PAIRW, else shared by a pair)
    uint32_t A[R];
    float w1[PAIRW ? R : 1][6];   // PAIRW: the second column's weights,
    uint32_t A1[PAIRW ? R : 1];   // the byte offset of its tap-0 row,
    bool split[PAIRW ? R : 1];    // and whether that row differs from the first column's
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = g + KG * i;
        double wd[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) wd[k] = W[(static_cast<int64_t>(r) * 6 + k) * nvel + j];
        round_weights<S>(wd, w[i]);
        const int il = IL[static_cast<int64_t>(r) * nvel + j];
        A[i] = static_cast<uint32_t>(((il - 2) & (n - 1)) * 256 + c * static_cast<int>(sizeof(E)));
        if constexpr (PAIRW) {
#pragma unroll
            for (int k = 0; k < 6; ++k) wd[k] = W[(static_cast<int64_t>(r) * 6 + k) * nvel + j + 1];
            round_weights<float>(wd, w1[i]);
            const int il1 = IL[static_cast<int64_t>(r) * nvel + j + 1];
            A1[i] = static_cast<uint32_t>(((il1 - 2) & (n - 1)) * 256 + c * static_cast<int>(sizeof(E)) + 4);
            split[i] = il1 != il;
        }
    }
    // Consume the weight loads here: otherwise the wait-count analysis of the
    // loop body also waits for the in-flight stores before each FMA.
#pragma unroll
    for (int i = 0; i < R; ++i) {
#pragma unroll
        for (int k = 0; k < 6; ++k) asm volatile("" ::"v"(w[i][k]));
        asm volatile("" ::"v"(A[i]));
        if constexpr (PAIRW) {
#pragma unroll
            for (int k = 0; k < 6; ++k) asm volatile("" ::"v"(w1[i][k]));
            asm volatile("" ::"v"(A1[i]));
        }
    }
    bool wsplit[PAIRW ? R : 1];  // some lane of the wave reads the row column by column
#pragma unroll
    for (int i = 0; i < (PAIRW ? R : 1); ++i) wsplit[i] = PAIRW && __any(split[i]);
    // float32 storage: rows whose taps wrap for no lane of the wave (all but
    // a few next to the line ends) read them at immediate offsets, dropping
    // two integer operations per tap; float32 interpolation costs as many
    // VALU cycles per element as float64 (no packed math), so the float
    // drifts are the ones this saves time in (2 %; float64: no change).
    constexpr bool DIRECT = sizeof(S) == 4;
    bool direct[R];
#pragma unroll
    for (int i = 0; i < R; ++i) direct[i] = DIRECT && __all((A[i] >> 8) <= static_cast<uint32_t>(n - 6));

    // Charge density of this thread's rows: compensated for float64 storage;
    // for float32 storage the float64 sum of at most 128 float32 values per
    // launch is already exact to ~1e-14, far below the storage rounding.
    constexpr bool COMP = sizeof(S) == 8;
    KSum acc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i] = KSum{0.0, 0.0};
    // End of a unit: the row sums of its tiles over the chunk's KC columns
    // (the lanes of one row group), merged in registers, halving the rows a
    // lane holds at each of the first exchanges (rows_ksum): this merge runs
    // after the unit's last tile with no memory work to hide it, at the end of
    // every short-lived workgroup (a butterfly of LDS lane shuffles over all R
    // rows: 818-824 us per 128^4 Y drift + rho; by DPP over all R rows:
    // 802-806), and is written by lanes 0 .. R-1 of the group.  Its stores join the vmcnt queue behind the
    // tile stores; the waits below then wait for a little more than they
    // need to, never less.
    auto flush = [&](int u) {
        const int64_t at = static_cast<int64_t>(q * nsplit + u % nsplit) * na * n + static_cast<int64_t>(u / nsplit) * n;
        if constexpr (R <= 4) {
            const KSum k = rows_ksum<COMP, R, KC>(acc, c);
            if (c < R) partial[at + g + KG * rows_ksum_row<R>(c)] = make_double2(k.s, COMP ? k.c : 0.0);
        } else {
#pragma unroll
            for (int ii = 0; ii < R; ++ii) {
                const KSum k = group_ksum<COMP>(acc[ii], KC);
                if (c == KC - 1) partial[at + g + KG * ii] = make_double2(k.s, COMP ? k.c : 0.0);
            }
        }
#pragma unroll
        for (int ii = 0; ii < R; ++ii) acc[ii] = KSum{0.0, 0.0};
    };
    Walk ow;  // the tile computed
    start(ow, u_begin);
    for (int i = 0; i < ntiles; ++i) {
        const int slot = i % T;
        if (i == 0) wait_vmcnt<(T - 2) * P>();
        else if (T > 2 && i == 1) wait_vmcnt<(T - 2) * P + R>();
        else if (T > 3 && i == 2) wait_vmcnt<(T - 2) * P + 2 * R>();
        else wait_vmcnt<(T - 2) * P + (T - 1) * R>();
        __syncthreads();
        if (T == 2) {
            // with two slots nothing waits for a load issued in the last
            // iteration, so the tail issues none
            if (i + 1 < ntiles) issue((i + 1) % T);
        } else {
            issue((i + T - 1) % T);  // past the end: the last tile again
        }
        const int l = tile_of(ow);
        const int64_t obase = base + static_cast<int64_t>(ow.u / nsplit) * a_stride;
        const E* tile = ring[slot];
#pragma unroll
        for (int ii = 0; ii < R; ++ii) asm volatile("" : "+v"(A[ii]));  // keep the tap offsets in the loop
        if (PASSES == 2) {
            // first pass into `mid` (free since the barrier above), second from it
#pragma unroll
            for (int ii = 0; ii < R; ++ii)
                mid[(g + KG * ii) * KC + c] = direct[ii] ? ring_tap_sum_direct<E, S>(tile, A[ii], w[ii])
                                                         : ring_tap_sum<E, S, n>(tile, A[ii], w[ii]);
            __syncthreads();
            tile = mid;
        }
#pragma unroll
        for (int ii = 0; ii < R; ++ii) {
            E o;
            if constexpr (PAIRW) {
                if (wsplit[ii]) {
                    o.x = ring_col_sum<n>(tile, A[ii], w[ii]);
                    o.y = ring_col_sum<n>(tile, A1[ii], w1[ii]);
                } else if (direct[ii]) {
                    o = ring_pair_sum<true, n>(tile, A[ii], w[ii], w1[ii]);
                } else {
                    o = ring_pair_sum<false, n>(tile, A[ii], w[ii], w1[ii]);
                }
            } else {
                o = direct[ii] ? ring_tap_sum_direct<E, S>(tile, A[ii], w[ii]) : ring_tap_sum<E, S, n>(tile, A[ii], w[ii]);
            }
            store_f<NT>(&fout[obase + (g + KG * ii) * row_stride + l * loop_stride], o);
            if constexpr (RHO) {
                if constexpr (PAIRW) {
                    acc[ii].s += static_cast<double>(o.x);
                    acc[ii].s += static_cast<double>(o.y);
                } else if (COMP) {
                    ksum_add(acc[ii], static_cast<double>(o));
                } else {
                    acc[ii].s += static_cast<double>(o);
                }
            }
        }
        if (ow.k + 1 < ow.len) {
            ++ow.k;
        } else {
            if constexpr (RHO) flush(ow.u);
            if (ow.u + 1 < u_end) start(ow, ow.u + 1);
        }
    }
    wait_vmcnt<0>();  // no direct-to-LDS load may still be in flight when the workgroup ends
}

// Dirichlet tap (clamp + zero mask, vlasov2D2V.py:340-345).
__device__ __forceinline__ double dirichlet_tap(const double* p, int j, int n, int stride) {
    const int jc = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
    const double valid = (j >= 0 && j < n) ? 1.0 : 0.0;
    return p[jc * stride] * valid;
}

// Vx kick then Vy kick on whole (vx, vy) planes; plane p = x*ny + y uses
// Ex[p], Ey[p].  Thread t owns column c = t % NVY and the RR consecutive rows
// starting at (t / NVY) * RR, so for NVY >= 64 one wave works on a single
// row (the row's tap indices are wave-uniform) and row offsets fit immediates.
// PREFETCH: load the next plane into registers while the Vy kick runs.
template <int NVX, int NVY, bool PREFETCH>
__global__ __launch_bounds__(1024) void kick_kernel(const double* fin, double* fout, int nplanes,
                                                    const double* __restrict__ Ex, const double* __restrict__ Ey,
                                                    double dt, double dvx, double dvy) {
    constexpr int T = 1024;
    constexpr int G = T / NVY;
    constexpr int RR = NVX / G;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* P = reinterpret_cast<double*>(smem_raw);        // [NVX][NVY]
    double* wx = P + NVX * NVY;                              // [NVX][6]
    double* wy = wx + NVX * 6;                               // [NVY][6]
    int* ilx = reinterpret_cast<int*>(wy + NVY * 6);         // [NVX]
    int* ily = ilx + NVX;                                    // [NVY]
    const int t = threadIdx.x;
    const int c = t % NVY;
    const int g = t / NVY;
    constexpr int64_t plane = static_cast<int64_t>(NVX) * NVY;
    double* Pt = P + g * RR * NVY + c;  // this thread's column slice of the plane

    double cur[RR];
    int p = blockIdx.x;
    if (PREFETCH && p < nplanes) {
#pragma unroll
        for (int i = 0; i < RR; ++i) cur[i] = (fin + p * plane + g * RR * NVY + c)[i * NVY];
    }
    for (; p < nplanes; p += gridDim.x) {
        __syncthreads();  // previous plane fully consumed
        if (PREFETCH) {
#pragma unroll
            for (int i = 0; i < RR; ++i) Pt[i * NVY] = cur[i];
        } else {
            const double* src = fin + p * plane + g * RR * NVY + c;
#pragma unroll
            for (int i = 0; i < RR; ++i) Pt[i * NVY] = src[i * NVY];
        }
        if (t < NVX) {
            const Departure d = departure(t, Ex[p], dt, dvx);
            double w[6];
            lagrange5_weights<double>(d.alpha, w);
#pragma unroll
            for (int k = 0; k < 6; ++k) wx[t * 6 + k] = w[k];
            ilx[t] = static_cast<int>(d.il < -8 ? -8 : (d.il > NVX + 8 ? NVX + 8 : d.il));
        } else if (t < NVX + NVY) {
            const int cc = t - NVX;
            const Departure d = departure(cc, Ey[p], dt, dvy);
            double w[6];
            lagrange5_weights<double>(d.alpha, w);
#pragma unroll
            for (int k = 0; k < 6; ++k) wy[cc * 6 + k] = w[k];
            ily[cc] = static_cast<int>(d.il < -8 ? -8 : (d.il > NVY + 8 ? NVY + 8 : d.il));
        }
        __syncthreads();
        double out[RR];
        // Vx kick: line along vx (stride NVY); row r is uniform over the wave.
#pragma unroll
        for (int i = 0; i < RR; ++i) {
            const int r = g * RR + i;
            // a wave covers one row when NVY >= 64: the index is then scalar
            const int l = NVY >= kWave ? __builtin_amdgcn_readfirstlane(ilx[r]) : ilx[r];
            double w[6], v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                w[k] = wx[r * 6 + k];
                const int j = l + k - 2;  // wave-uniform: scalar clamp and mask
                const int jc = j < 0 ? 0 : (j > NVX - 1 ? NVX - 1 : j);
                const double m = (j >= 0 && j < NVX) ? 1.0 : 0.0;
                v[k] = P[jc * NVY + c] * m;
            }
            out[i] = tap_sum<double>(w, v);
            __builtin_amdgcn_sched_barrier(0);  // keep rows apart: bounds the live registers
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RR; ++i) Pt[i * NVY] = out[i];
        __syncthreads();
        const int pn = p + gridDim.x;
        if (PREFETCH && pn < nplanes) {
#pragma unroll
            for (int i = 0; i < RR; ++i) cur[i] = (fin + pn * plane + g * RR * NVY + c)[i * NVY];
        }
        // Vy kick: contiguous line, weights and clamped taps of column c.
        double w[6], m[6];
        int jy[6];
        const int l = ily[c];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            w[k] = wy[c * 6 + k];
            const int j = l + k - 2;
            jy[k] = j < 0 ? 0 : (j > NVY - 1 ? NVY - 1 : j);
            m[k] = (j >= 0 && j < NVY) ? 1.0 : 0.0;
        }
        double* dst = fout + p * plane + g * RR * NVY + c;
        const double* Prow = P + g * RR * NVY;
#pragma unroll
        for (int i = 0; i < RR; ++i) {
            double v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = Prow[i * NVY + jy[k]] * m[k];
            dst[i * NVY] = tap_sum<double>(w, v);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Row-streaming Vx+Vy kick (out of place: fin and fout must not overlap).
// One lane per vy column; a (vx, vy) plane's rows are streamed in order.
// The Vx kick of row r reads input rows il(r)-2 .. il(r)+3 of the lane's
// column; when il(r) - r is the same for every row ("regular" plane, the
// normal case) those six values live in a register window refilled by one
// load per row, issued PF rows ahead; otherwise each tap is read from fin.
// The Vx result of row r goes through a 2-row LDS buffer to the Vy kick,
// which is stored right away.  Small LDS footprint -> many planes per CU.
template <typename S, int NVX, int NVY, int PF, bool NT>
__global__ __launch_bounds__(128) void kick_stream_kernel(const S* __restrict__ fin, S* __restrict__ fout,
                                                          int nplanes, const double* __restrict__ Ex,
                                                          const double* __restrict__ Ey, double dt, double dvx,
                                                          double dvy) {
    constexpr int PPW = 128 / NVY;  // planes per workgroup
    static_assert(NVX % PF == 0, "NVX must be a multiple of PF");
    __shared__ S wx[PPW][NVX * 6];
    __shared__ int ilx[PPW][NVX];
    __shared__ S rowbuf[2][PPW * NVY];
    __shared__ int irregular[1];
    const int t = threadIdx.x;
    const int pg = t / NVY;
    const int c = t % NVY;
    const int p = blockIdx.x * PPW + pg;  // the host launches exactly nplanes / PPW workgroups

    if (t == 0) irregular[0] = 0;
    const double ex = Ex[p];
    for (int r = c; r < NVX; r += NVY) {
        const Departure d = departure(r, ex, dt, dvx);
        double w[6];
        lagrange5_weights<double>(d.alpha, w);
        S ws[6];
        round_weights<S>(w, ws);
#pragma unroll
        for (int k = 0; k < 6; ++k) wx[pg][r * 6 + k] = ws[k];
        ilx[pg][r] = static_cast<int>(d.il < -8 ? -8 : (d.il > NVX + 8 ? NVX + 8 : d.il));
    }
    S wy[6], my[6];
    int jy[6];
    {
        const Departure d = departure(c, Ey[p], dt, dvy);
        double wyd[6];
        lagrange5_weights<double>(d.alpha, wyd);
        round_weights<S>(wyd, wy);
        const int l = static_cast<int>(d.il < -8 ? -8 : (d.il > NVY + 8 ? NVY + 8 : d.il));
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int j = l + k - 2;
            jy[k] = j < 0 ? 0 : (j > NVY - 1 ? NVY - 1 : j);
            my[k] = (j >= 0 && j < NVY) ? S(1) : S(0);
        }
    }
    __syncthreads();
    const int s0 = ilx[pg][0];
    for (int r = c; r < NVX; r += NVY)
        if (ilx[pg][r] - r != s0) irregular[0] = 1;  // one flag for the whole workgroup
    __syncthreads();
    const bool regular = irregular[0] == 0;  // uniform: every plane of the group takes one path

    const S* src = fin + static_cast<int64_t>(p) * NVX * NVY + c;
    S* dst = fout + static_cast<int64_t>(p) * NVX * NVY + c;
    S* rb0 = &rowbuf[0][pg * NVY];
    S* rb1 = &rowbuf[1][pg * NVY];
    const S* wrow = wx[pg];
    auto clampr = [](int j) { return j < 0 ? 0 : (j > NVX - 1 ? NVX - 1 : j); };
    auto maskr = [](int j) { return (j >= 0 && j < NVX) ? S(1) : S(0); };
    // Vy kick of row r from the Vx result ox (through the 2-row LDS buffer).
    auto vy_row = [&](int r, S ox) {
        S* rb = (r & 1) ? rb1 : rb0;
        rb[c] = ox;
        __syncthreads();
        S v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = rb[jy[k]] * my[k];
        store_f<NT>(&dst[r * NVY], tap_sum<S>(wy, v));
    };
    auto ld = [](const S* p) { return NT ? __builtin_nontemporal_load(p) : *p; };

    if (regular) {
        S win[5], q[PF];
#pragma unroll
        for (int k = 0; k < 5; ++k) win[k] = ld(&src[clampr(s0 - 2 + k) * NVY]) * maskr(s0 - 2 + k);
#pragma unroll
        for (int u = 0; u < PF; ++u) q[u] = ld(&src[clampr(s0 + 3 + u) * NVY]);
        for (int r0 = 0; r0 < NVX; r0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int r = r0 + u;
                const S* w = wrow + r * 6;
                const int jn = r + s0 + 3;  // row held by q[u]
                const S v5 = q[u] * maskr(jn);
                const S ox =
                    w[0] * win[0] + w[1] * win[1] + w[2] * win[2] + w[3] * win[3] + w[4] * win[4] + w[5] * v5;
                win[0] = win[1];
                win[1] = win[2];
                win[2] = win[3];
                win[3] = win[4];
                win[4] = v5;
                q[u] = ld(&src[clampr(jn + PF) * NVY]);
                vy_row(r, ox);
            }
        }
    } else {
        for (int r = 0; r < NVX; ++r) {
            const S* w = wrow + r * 6;
            const int l = ilx[pg][r];
            S v[6], wl[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                wl[k] = w[k];
                v[k] = ld(&src[clampr(l + k - 2) * NVY]) * maskr(l + k - 2);
            }
            vy_row(r, tap_sum<S>(wl, v));
        }
    }
}

// Row-split Vx+Vy kick of a (vx, vy) plane with 128 vy columns (out of
// place).  Same arithmetic as kick_stream_kernel, other work split: WPP
// waves per plane, wave w owns the RPW = NVX / WPP consecutive rows
// [w RPW, (w + 1) RPW), and lane l owns the column pair (2l, 2l+1), so every
// row moves as one 2-element vector load and store per lane (16 bytes for
// double).  Short row streams keep HBM busier than one stream per 128-row
// plane (measured: 8 waves of 16 rows beat 1 wave of 128 rows by 16 %).
// Even waves walk their rows downwards, odd waves upwards, so two
// neighbours reach their shared boundary (and the up-to-5 input rows both
// read there) at the same time and the second read finds the rows on chip.
// The Vx window slides with the walk; the six taps are always summed from
// the lowest row up, as in the reference.  The Vy taps of the just-computed
// row come from a zero-padded per-wave LDS row: taps outside [0, NVY) read 0
// there instead of being clamped and multiplied by a 0/1 mask (the same
// values: the mask product is +-0 and adding it changes no nonzero sum).
// When every lane's two columns have adjacent tap windows (the normal case)
// the pair shares one 7-value read.  A wave's LDS row is written and read by
// that wave only: the write is waited for (lgkmcnt) before the reads, no
// workgroup barrier.
// DEV (deviation storage, S = float): fin/fout hold g = f - feq with the
// separable feq(vx, vy) = fx(vx) fy(vy) (feq = fx | fy, float64 profiles).
// The kick is linear and maps fx fy to ax ay, ax = Kvx fx and ay = Kvy fy (the
// plane's Vx kick acts on the vx profile, its Vy kick on the vy profile), so
// K(feq + g) - feq = K(g) + (ax ay - fx fy): the float32 kick of g as in the
// plain kernel, plus a float64 correction that needs per plane the 1D kicks
// of the two profiles (ax and fx of every row in LDS, ay and fy of the
// lane's columns in registers) and nothing from memory per element.
// Rows of a plane whose departure offset may differ from the middle row's
// for the windowed kick walks (their taps are loaded directly).
constexpr int kKickMaxDev = 4;

template <typename S, int NVX, int WPP, int PF, bool NT, bool DEV, int NVY = 128>
__global__ __launch_bounds__(64 * WPP) void kick_rows_kernel(const S* __restrict__ fin, S* __restrict__ fout,
                                                             const double* __restrict__ feq,
                                                             const double* __restrict__ Ex,
                                                             const double* __restrict__ Ey, double dt, double dvx,
                                                             double dvy) {
    // NVY = 128: lane l owns the column pair (2l, 2l + 1), a row is one 16-byte (float64) load per
    // lane; NVY = 64: lane l owns column l.
    static_assert(NVY == 128 || NVY == 64, "64 or 128 vy columns");
    constexpr int CPL = NVY / 64;  // columns per lane
    constexpr int PAD = 16;  // >= 10 on the left, >= 12 on the right (taps of l in [-8, NVY + 8])
    constexpr int RPW = NVX / WPP;
    static_assert(NVX % WPP == 0 && RPW % PF == 0, "rows per wave must be a multiple of PF");
    static_assert(!DEV || sizeof(S) == 4, "deviation storage is float32");
    static_assert(!DEV || CPL == 2, "deviation storage: 128 vy columns");
    typedef S C;  // arithmetic type
    typedef S V2s_ __attribute__((ext_vector_type(2)));
    typedef C V2_ __attribute__((ext_vector_type(2)));
    typedef typename std::conditional<CPL == 2, V2s_, S>::type V2s;  // a lane's part of a row, stored
    typedef typename std::conditional<CPL == 2, V2_, C>::type V2;    // and in arithmetic
    __shared__ C wx[NVX * 6];
    __shared__ int ilx[NVX];
    __shared__ int ndev[1];
    __shared__ C rowbuf[WPP][2][NVY + 2 * PAD];
    __shared__ double axs[DEV ? NVX : 1], fxs[DEV ? NVX : 1];  // DEV: Kvx fx and fx of every row
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int64_t p = blockIdx.x;  // the host launches exactly one workgroup per plane

    if (t == 0) ndev[0] = 0;
    const double ex = Ex[p];
    for (int r = t; r < NVX; r += 64 * WPP) {
        const Departure d = departure(r, ex, dt, dvx);
        double w[6];
        lagrange5_weights<double>(d.alpha, w);
        C wc[6];
        round_weights<C>(w, wc);
#pragma unroll
        for (int k = 0; k < 6; ++k) wx[r * 6 + k] = wc[k];
        const int l = static_cast<int>(d.il < -8 ? -8 : (d.il > NVX + 8 ? NVX + 8 : d.il));
        ilx[r] = l;
        if constexpr (DEV) {
            // Vx kick of the vx profile (taps outside the plane read zero)
            double v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int j = l + k - 2;
                v[k] = (j >= 0 && j < NVX) ? feq[j] : 0.0;
            }
            axs[r] = tap_sum<double>(w, v);
            fxs[r] = feq[r];
        }
    }
    // zero pads of the wave's two row buffers: 2 x 2 x PAD = 64 entries, one per lane
    {
        const int b = lane >> 5, i = lane & 31;
        rowbuf[wv][b][i < PAD ? i : NVY + i] = C(0);
    }
    C wy[CPL][6];
    int ty[CPL];  // LDS index of tap 0 of each column
    double ay[DEV ? CPL : 1], fy[DEV ? CPL : 1];  // DEV: Kvy fy and fy of the lane's columns
    {
        const double ey = Ey[p];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const Departure d = departure(CPL * lane + j, ey, dt, dvy);
            double w[6];
            lagrange5_weights<double>(d.alpha, w);
            round_weights<C>(w, wy[j]);
            const int l = static_cast<int>(d.il < -8 ? -8 : (d.il > NVY + 8 ? NVY + 8 : d.il));
            ty[j] = PAD + l - 2;
            if constexpr (DEV) {
                const double* fyp = feq + NVX;
                double v[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int i = l + k - 2;
                    v[k] = (i >= 0 && i < NVY) ? fyp[i] : 0.0;
                }
                ay[j] = tap_sum<double>(w, v);
                fy[j] = fyp[CPL * lane + j];
            }
        }
    }
    bool pair = false;
    if constexpr (CPL == 2) pair = __ballot(ty[1] != ty[0] + 1) == 0;  // wave-uniform
    __syncthreads();
    // floor(RN(r - disp)) - r takes at most two values, floor(-disp) and one
    // more (where r - disp rounds up to an integer: with |disp| below half an
    // ulp of r, every row but the first ones).  The windowed walks below take
    // the middle row's offset s0 and load the taps of the few other rows
    // directly; planes where more rows differ take the per-row path.
    // Only the float64 and plain float32 kicks of 128 x 128 planes compile
    // the walk that skips such rows: in the others (deviation storage, 64 rows
    // or columns) its registers cost a wave per SIMD, so any such row sends
    // the plane to the per-row path there.
    constexpr bool WALK_CHECK = !DEV && NVX == 128 && NVY == 128;
    const int s0 = ilx[NVX / 2] - NVX / 2;
    for (int r = t; r < NVX; r += 64 * WPP)
        if (ilx[r] - r != s0) atomicAdd(&ndev[0], 1);
    __syncthreads();
    const int nd = ndev[0];
    const bool regular = WALK_CHECK ? nd <= kKickMaxDev : nd == 0;  // uniform over the workgroup

    const S* src = fin + p * NVX * NVY + CPL * lane;
    S* dst = fout + p * NVX * NVY + CPL * lane;
    // A row as loaded and as used (the same values).
    struct Raw {
        V2s g;
    };
    auto ld = [&](int row) -> Raw {
        Raw r;
        const V2s* a = reinterpret_cast<const V2s*>(src + row * NVY);
        r.g = NT ? __builtin_nontemporal_load(a) : *a;
        return r;
    };
    auto val = [](const Raw& r) -> V2 { return r.g; };
    auto clampr = [](int j) { return j < 0 ? 0 : (j > NVX - 1 ? NVX - 1 : j); };
    auto maskr = [](int j) { return (j >= 0 && j < NVX) ? C(1) : C(0); };
    // Vy kick of row r from the Vx results ox of the lane's two columns
    // (DEV: plus the correction of row r).
    auto vy_row = [&](int r, C ox0, C ox1) {
        C* rb = rowbuf[wv][r & 1];
        if constexpr (CPL == 1) {
            rb[PAD + lane] = ox0;
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            C a[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) a[k] = rb[ty[0] + k];
            const S res = tap_sum<C>(wy[0], a);
            S* out = dst + r * NVY;
            if (NT)
                __builtin_nontemporal_store(res, out);
            else
                *out = res;
            (void)ox1;
        } else {
            V2_ o;
            o.x = ox0;
            o.y = ox1;
            *reinterpret_cast<V2_*>(&rb[PAD + 2 * lane]) = o;
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            C a[6], b[6];
            if (pair) {
                C v[7];
#pragma unroll
                for (int i = 0; i < 7; ++i) v[i] = rb[ty[0] + i];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    a[k] = v[k];
                    b[k] = v[k + 1];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    a[k] = rb[ty[0] + k];
                    b[k] = rb[ty[1] + k];
                }
            }
            V2s res;
            if constexpr (DEV) {
                const double axr = axs[r], fxr = fxs[r];
                res.x = static_cast<float>(static_cast<double>(tap_sum<C>(wy[0], a)) + (axr * ay[0] - fxr * fy[0]));
                res.y = static_cast<float>(static_cast<double>(tap_sum<C>(wy[1], b)) + (axr * ay[1] - fxr * fy[1]));
            } else {
                res.x = tap_sum<C>(wy[0], a);
                res.y = tap_sum<C>(wy[1], b);
            }
            V2s* out = reinterpret_cast<V2s*>(dst + r * NVY);
            if (NT)
                __builtin_nontemporal_store(res, out);
            else
                *out = res;
        }
    };
    auto vx_pair = [&](int r, const V2 v[6], C& ox0, C& ox1) {
        C w[6], a[6], b[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            w[k] = wx[r * 6 + k];
            if constexpr (CPL == 2) {
                a[k] = v[k].x;
                b[k] = v[k].y;
            } else {
                a[k] = v[k];
            }
        }
        ox0 = tap_sum<C>(w, a);
        if constexpr (CPL == 2)
            ox1 = tap_sum<C>(w, b);
        else
            ox1 = C(0);
    };

    const int r_lo = wv * RPW;
    const bool down = WPP > 1 && (wv & 1) == 0;
    // CHECK: the plane has rows whose offset is not s0; the walks skip them
    // (they still feed the window) and the per-row path below computes them.
    // Without such rows the walks are compiled without the test.
    auto walk = [&](auto check) {
    constexpr bool CHECK = decltype(check)::value;
    if (!down) {
        // rows r_lo .. r_lo + RPW - 1 upwards; window = taps 0..4 of row r
        V2 win[5];
        Raw q[PF];
#pragma unroll
        for (int k = 0; k < 5; ++k) win[k] = val(ld(clampr(r_lo + s0 - 2 + k))) * maskr(r_lo + s0 - 2 + k);
#pragma unroll
        for (int u = 0; u < PF; ++u) q[u] = ld(clampr(r_lo + s0 + 3 + u));
        for (int i0 = 0; i0 < RPW; i0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int r = r_lo + i0 + u;
                const int jn = r + s0 + 3;  // row held by q[u]: tap 5 of row r
                const V2 v5 = val(q[u]) * maskr(jn);
                const V2 v[6] = {win[0], win[1], win[2], win[3], win[4], v5};
                const bool own = !CHECK || ilx[r] - r == s0;  // uniform; the other rows after the walk
                C ox0, ox1;
                if (own) vx_pair(r, v, ox0, ox1);
                win[0] = win[1];
                win[1] = win[2];
                win[2] = win[3];
                win[3] = win[4];
                win[4] = v5;
                if (i0 + u + PF < RPW) q[u] = ld(clampr(jn + PF));
                if (own) vy_row(r, ox0, ox1);
            }
        }
    } else {
        // rows r_lo + RPW - 1 .. r_lo downwards; window = taps 1..5 of row r
        const int r_hi = r_lo + RPW - 1;
        V2 win[5];
        Raw q[PF];
#pragma unroll
        for (int k = 0; k < 5; ++k) win[k] = val(ld(clampr(r_hi + s0 - 1 + k))) * maskr(r_hi + s0 - 1 + k);
#pragma unroll
        for (int u = 0; u < PF; ++u) q[u] = ld(clampr(r_hi + s0 - 2 - u));
        for (int i0 = 0; i0 < RPW; i0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int r = r_hi - i0 - u;
                const int jn = r + s0 - 2;  // row held by q[u]: tap 0 of row r
                const V2 v0 = val(q[u]) * maskr(jn);
                const V2 v[6] = {v0, win[0], win[1], win[2], win[3], win[4]};
                const bool own = !CHECK || ilx[r] - r == s0;  // uniform; the other rows after the walk
                C ox0, ox1;
                if (own) vx_pair(r, v, ox0, ox1);
                win[4] = win[3];
                win[3] = win[2];
                win[2] = win[1];
                win[1] = win[0];
                win[0] = v0;
                if (i0 + u + PF < RPW) q[u] = ld(clampr(jn - PF));
                if (own) vy_row(r, ox0, ox1);
            }
        }
    }
    };
    if (regular) {
        if constexpr (WALK_CHECK) {
            if (nd == 0)
                walk(std::integral_constant<bool, false>{});
            else
                walk(std::integral_constant<bool, true>{});
        } else {
            walk(std::integral_constant<bool, false>{});
        }
    }
    // The per-row path: every row of an irregular plane, or after a windowed
    // walk the few rows whose offset is not s0 (their taps from memory).
    if (!regular || nd > 0) {  // uniform
        for (int i = 0; i < RPW; ++i) {
            const int r = r_lo + i;
            const int l = ilx[r];
            if (regular && l - r == s0) continue;  // uniform
            V2 v[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = val(ld(clampr(l + k - 2))) * maskr(l + k - 2);
            C ox0, ox1;
            vx_pair(r, v, ox0, ox1);
            vy_row(r, ox0, ox1);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Launch-shape knobs (vp_tune_set); 0 = automatic choice.
static int g_knobs[VP_KNOB_COUNT] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

int knob(int k) { return (k >= 0 && k < VP_KNOB_COUNT) ? g_knobs[k] : 0; }

// Sweeps that use the streaming hint: 1 = X drift, 2 = Y drift, 4 = kick.
static int nt_mask() {
    const int k = knob(VP_KNOB_NONTEMPORAL);
    return k == 0 ? 7 : (k & 7);
}

static int cu_count() {
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

bool fast2d_kick_supported(const int64_t shape[4]) {
    auto ok = [](int64_t n) { return n == 32 || n == 64 || n == 128; };
    return ok(shape[2]) && ok(shape[3]);
}

bool fast2d_supported(const int64_t shape[4]) {
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    auto ok = [](int64_t n) { return n == 32 || n == 64 || n == 128; };
    return ok(nx) && ok(ny) && ok(nvx) && ok(nvy);
}

// FP32 storage: the drifts need 64 or 128 rows (LDS-ring kernel only) and
// whole 64-column (256-byte) vy chunks.
bool fast2d_supported_f32(const int64_t shape[4]) {
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    auto big = [](int64_t n) { return n == 64 || n == 128; };
    return big(nx) && big(ny) && big(nvy) && (nvx == 32 || nvx == 64 || nvx == 128);
}

static inline size_t a256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Launch plan of the Y drift: nsplit = vx ranges per x (the units; it fixes
// the rho partial layout), groups = workgroups per vy chunk, rot = walk
// options of persistent workgroups (bit 0: rotated unit starts, bit 1:
// chunks spread over the XCDs).  Measured at 128^4 (alternating runs):
// * float32 storage: persistent workgroups, one per CU (the register file
//   holds one), two units each (2 vx ranges per x), rotated starts: Y drift
//   + rho 0.49-0.51 -> 0.42 ms, Y drift 0.43-0.44 -> 0.43 ms, step 1.85 ->
//   1.71 ms; one unit each was slower (0.53 ms Y drift), so were 4 ranges,
//   the XCD spread and unrotated starts (workgroups that start together then
//   read the same offsets of their rows at once).
// * float64: one workgroup per unit, 12 ranges per x (10-11 tiles each; 1, 8,
//   16, 32 slower): persistent workgroups took 0.89 ms per Y drift + rho
//   against 0.83-0.84 (any range count, rotation or spread).
// * 64^4: one workgroup per unit, 12 ranges for both storage types (2-8,
//   16, 24, 32 slower; float32 Y drifts 9 % faster than at 5).
// Knob VP_KNOB_Y_UNITS: 1 = one workgroup per unit, 2 = persistent, 3 =
// persistent and spread, 4 = persistent, spread, unrotated (0 = as above).
struct YPlan {
    int nsplit;
    int groups;
    int rot;
};

static YPlan y_plan(int64_t nx, int64_t ny, int64_t nvx, int64_t nvy, int elem_bytes) {
    const int nq = static_cast<int>(nvy / (256 / elem_bytes));
    const int mode = knob(VP_KNOB_Y_UNITS);
    const bool persistent = ny == 128 && nq >= 1 && (mode >= 2 || (mode == 0 && elem_bytes == 4));
    int S = knob(VP_KNOB_Y_SPLIT);
    if (S <= 0) {
        if (persistent) {
            const int gq = cu_count() / nq;  // two units per workgroup
            S = 2 * gq > nx ? static_cast<int>(2 * gq / nx) : 1;
        } else if (nvx == 64) {
            S = 12;
        } else {
            S = elem_bytes == 8 ? static_cast<int>((3 * nvx + 16) / 32) : static_cast<int>((5 * nvx + 32) / 64);
        }
        if (S < 1) S = 1;
    }
    if (S > nvx) S = static_cast<int>(nvx);
    const int units = static_cast<int>(nx) * S;
    YPlan p{S, units, 0};
    if (persistent) {
        const int gq = cu_count() / (nq > 0 ? nq : 1);
        p.groups = gq >= 1 && gq < units ? gq : units;
        p.rot = mode == 3 ? 3 : (mode == 4 ? 2 : 1);
    }
    return p;
}

Fast2DWs fast2d_layout(const int64_t shape[4], void* ws, int elem_bytes, bool with_ftmp) {
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    const int col = 256 / elem_bytes;
    Fast2DWs L{};
    unsigned char* p = static_cast<unsigned char*>(ws);
    size_t off = 0;
    auto take = [&](size_t b) {
        unsigned char* q = p ? p + off : nullptr;
        off += a256(b);
        return q;
    };
    L.Wx = reinterpret_cast<double*>(take(static_cast<size_t>(nx) * 6 * nvx * sizeof(double)));
    L.ILx = reinterpret_cast<int*>(take(static_cast<size_t>(nx) * nvx * sizeof(int)));
    L.Wy = reinterpret_cast<double*>(take(static_cast<size_t>(ny) * 6 * nvy * sizeof(double)));
    L.ILy = reinterpret_cast<int*>(take(static_cast<size_t>(ny) * nvy * sizeof(int)));
    L.nparts = static_cast<int>((nvy / col) * y_plan(nx, ny, nvx, nvy, elem_bytes).nsplit);
    L.partial = reinterpret_cast<double2*>(take(static_cast<size_t>(nx) * ny * L.nparts * sizeof(double2)));
    L.pws_bytes = poisson2d_ws(nx, ny);
    L.pws = take(L.pws_bytes);
    L.ftmp = with_ftmp ? take(static_cast<size_t>(nx) * ny * nvx * nvy * elem_bytes) : nullptr;
    L.bytes = off;
    return L;
}

int fast2d_tables(const int64_t shape[4], const Fast2DWs& L, const double* vx, const double* vy, double half_dt,
                  double dx, double dy, hipStream_t s) {
    const int nx = static_cast<int>(shape[0]), ny = static_cast<int>(shape[1]);
    const int nvx = static_cast<int>(shape[2]), nvy = static_cast<int>(shape[3]);
    hipLaunchKernelGGL(drift_table_kernel, dim3((nx * nvx + 255) / 256), dim3(256), 0, s, vx, nvx, nx, half_dt, dx,
                       L.Wx, L.ILx);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(drift_table_kernel, dim3((ny * nvy + 255) / 256), dim3(256), 0, s, vy, nvy, ny, half_dt, dy,
                       L.Wy, L.ILy);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// Drift launch: the LDS-ring kernel (T = 2 or 3 tiles, default 2) when it
// covers the line length (64 or 128 rows), unless the DRIFT_DMA knob is 1 or
// fin is not 16-byte aligned (float64 only); otherwise the register-prefetch
// kernel (float64 only).  R = rows per thread.
template <typename S, int R, int PASSES, bool RHO, bool LANE_W, int VW = 1>
static int launch_drift(const S* fin, S* fout, int nq, int nsplit, int na, int groups, int rot, int64_t a_stride,
                        int64_t row_stride, int64_t loop_stride, int lt, const double* W, const int* IL, int nvel,
                        double2* partial, hipStream_t s) {
    const int64_t unit_blocks = static_cast<int64_t>(na) * nq * nsplit;  // one workgroup per unit
    if (!LANE_W || groups > na * nsplit) groups = na * nsplit;  // X drift: weights depend on a
    const int64_t blocks = static_cast<int64_t>(nq) * groups;
    constexpr bool f64 = sizeof(S) == 8;
    constexpr int rows = R * RingGeom<typename LaneElem<S, VW>::type>::groups;
    const int mode = knob(VP_KNOB_DRIFT_DMA);
    const bool nt = (nt_mask() & (LANE_W ? 2 : 1)) != 0;
    const bool aligned = (reinterpret_cast<uintptr_t>(fin) & 15) == 0;
    if constexpr (rows == 64 || rows == 128) {
        if ((mode != 1 || !f64) && aligned) {
#define VP_D(T, NT)                                                                                        \
    hipLaunchKernelGGL((drift_dma_kernel<S, R, PASSES, RHO, LANE_W, T, NT, VW>), dim3(blocks), dim3(kDriftThreads), \
                       0, s, fin, fout, nq, nsplit, na, groups < na * nsplit ? rot : 0, a_stride, row_stride,     \
                       loop_stride, lt, W, IL, nvel, partial)
            if (mode == 3) {
                if (nt) VP_D(3, true); else VP_D(3, false);
            } else {
                if (nt) VP_D(2, true); else VP_D(2, false);
            }
#undef VP_D
            return VP_OK;
        }
    }
    if constexpr (f64) {
        hipLaunchKernelGGL((drift_kernel<R, PASSES, RHO, LANE_W>), dim3(unit_blocks), dim3(kDriftThreads), 0, s, fin, fout,
                           nq, nsplit, a_stride, row_stride, loop_stride, lt, W, IL, nvel, partial);
        return VP_OK;
    }
    return VP_ERR_UNSUPPORTED;
}

// X drift (passes = 1 or 2) from fin to fout (may alias).
template <typename S>
static int xdrift(const S* fin, S* fout, const int64_t shape[4], const Fast2DWs& L, int passes, hipStream_t s) {
    constexpr int KC = RingGeom<S>::col, KG = RingGeom<S>::groups;
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    if (nvy % KC != 0) return VP_ERR_UNSUPPORTED;
    const int nq = static_cast<int>(nvy / KC);
    int S_ = knob(VP_KNOB_X_SPLIT);
    if (S_ <= 0) {
        // workgroups per CU to aim at: 2 (float64; 128^4: 1 range per unit
        // 786 us for the pair against 794-807 for 2-4); float32 with at least
        // one unit per CU: 1 for the pair, 4 for a single pass (128^4: pair
        // 394 us against 403 with 2 ranges; single 437-440 against 467-471)
        int target = 2;
        if (sizeof(S) == 4 && nvx * nq >= cu_count()) target = passes == 2 ? 1 : 4;
        S_ = 1;
        while (nvx * nq * S_ < target * cu_count() && S_ < static_cast<int>(ny)) S_ *= 2;
    }
    if (S_ > static_cast<int>(ny)) S_ = static_cast<int>(ny);
    const int64_t a_stride = nvy, row_stride = ny * nvx * nvy, loop_stride = nvx * nvy;
    const int nvel = static_cast<int>(nvx), lt = static_cast<int>(ny);
    int rc = VP_ERR_UNSUPPORTED;
    // float32 storage: neighbouring vy columns share the X drift's weights, so a thread can take
    // a pair of them (8-byte loads and stores, the float64 thread layout).  Measured at 128^4:
    // 4 % faster for the two-pass drift, 6 % slower for one pass; the knob: 1 never, 2 always.
    const int pk = knob(VP_KNOB_F32_PAIRS);
    const bool pairs = sizeof(S) == 4 && pk != 1 && (pk == 2 || passes == 2) &&
                       (reinterpret_cast<uintptr_t>(fout) & 7) == 0;
#define VP_X(N, PV)                                                                                                  \
    if (nx == N && (N) % KG == 0 && (N) / KG >= 1) {                                                                 \
        constexpr int RV = (N) / KG > 0 ? (N) / KG : 1;                                                              \
        constexpr int RV2 = (N) / (2 * KG) > 0 ? (N) / (2 * KG) : 1;                                                 \
        if (pairs && (N) >= 64)                                                                                      \
            rc = launch_drift<S, RV2, PV, false, false, 2>(fin, fout, nq, S_, static_cast<int>(nvx), 0, 0, a_stride, \
                                                           row_stride, loop_stride, lt, L.Wx, L.ILx, nvel, nullptr, s); \
        else                                                                                                         \
            rc = launch_drift<S, RV, PV, false, false>(fin, fout, nq, S_, static_cast<int>(nvx), 0, 0, a_stride,     \
                                                       row_stride, loop_stride, lt, L.Wx, L.ILx, nvel, nullptr, s);  \
    }
    if (passes == 2) {
        VP_X(128, 2) else VP_X(64, 2) else VP_X(32, 2)
    } else {
        VP_X(128, 1) else VP_X(64, 1) else VP_X(32, 1)
    }
#undef VP_X
    if (rc != VP_OK) return rc;
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// Y drift (+ rho partials when rho != 0) from fin to fout (may alias).
template <typename S>
static int ydrift(const S* fin, S* fout, const int64_t shape[4], const Fast2DWs& L, bool rho, hipStream_t s) {
    constexpr int KC = RingGeom<S>::col, KG = RingGeom<S>::groups;
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    if (nvy % KC != 0) return VP_ERR_UNSUPPORTED;
    const int nq = static_cast<int>(nvy / KC);
    const YPlan plan = y_plan(nx, ny, nvx, nvy, static_cast<int>(sizeof(S)));
    const int S_ = plan.nsplit;
    if (rho && nq * S_ != L.nparts) return VP_ERR_WORKSPACE;  // the layout was computed for another split
    const int64_t a_stride = ny * nvx * nvy, row_stride = nvx * nvy, loop_stride = nvy;
    const int nvel = static_cast<int>(nvy), lt = static_cast<int>(nvx);
    int rc = VP_ERR_UNSUPPORTED;
    // float32 storage: a thread takes a pair of vy columns (8-byte loads and
    // stores, a weight set per column) unless the pairs knob is 1
    const bool pairs = sizeof(S) == 4 && knob(VP_KNOB_F32_PAIRS) != 1 &&
                       ((reinterpret_cast<uintptr_t>(fin) | reinterpret_cast<uintptr_t>(fout)) & 7) == 0;
#define VP_Y(N, RH)                                                                                              \
    if (ny == N && (N) % KG == 0) {                                                                              \
        constexpr int RV = (N) / KG > 0 ? (N) / KG : 1;                                                          \
        constexpr int RV2 = (N) / (2 * KG) > 0 ? (N) / (2 * KG) : 1;                                             \
        if constexpr (sizeof(S) == 4 && (N) >= 64) {                                                             \
            if (pairs)                                                                                           \
                rc = launch_drift<S, RV2, 1, RH, true, 2>(fin, fout, nq, S_, static_cast<int>(nx), plan.groups,  \
                                                          plan.rot, a_stride, row_stride, loop_stride, lt, L.Wy, \
                                                          L.ILy, nvel, L.partial, s);                            \
        }                                                                                                        \
        if (rc == VP_ERR_UNSUPPORTED)                                                                            \
            rc = launch_drift<S, RV, 1, RH, true>(fin, fout, nq, S_, static_cast<int>(nx), plan.groups, plan.rot, \
                                                  a_stride, row_stride, loop_stride, lt, L.Wy, L.ILy, nvel,       \
                                                  L.partial, s);                                                  \
    }
    if (rho) {
        VP_Y(128, true) else VP_Y(64, true) else VP_Y(32, true)
    } else {
        VP_Y(128, false) else VP_Y(64, false) else VP_Y(32, false)
    }
#undef VP_Y
    if (rc != VP_OK) return rc;
    VP_CHECK_LAUNCH();
    return VP_OK;
}

int fast2d_xdrift(const double* fin, double* fout, const int64_t shape[4], const Fast2DWs& L, int passes,
                  hipStream_t s) {
    return xdrift<double>(fin, fout, shape, L, passes, s);
}
int fast2d_xdrift(const float* fin, float* fout, const int64_t shape[4], const Fast2DWs& L, int passes,
                  hipStream_t s) {
    return xdrift<float>(fin, fout, shape, L, passes, s);
}
int fast2d_ydrift(const double* fin, double* fout, const int64_t shape[4], const Fast2DWs& L, bool rho,
                  hipStream_t s) {
    return ydrift<double>(fin, fout, shape, L, rho, s);
}
int fast2d_ydrift(const float* fin, float* fout, const int64_t shape[4], const Fast2DWs& L, bool rho,
                  hipStream_t s) {
    return ydrift<float>(fin, fout, shape, L, rho, s);
}

template <int NVX, int NVY, bool PF>
static int launch_kick(const double* fin, double* fout, int nplanes, const double* Ex, const double* Ey, double dt,
                       double dvx, double dvy, hipStream_t s) {
    const size_t lds = static_cast<size_t>(NVX) * NVY * sizeof(double) + (NVX + NVY) * 6 * sizeof(double) +
                       (NVX + NVY) * sizeof(int);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&kick_kernel<NVX, NVY, PF>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return static_cast<int>(e);
        attr_set = true;
    }
    const int per_cu = lds > 81920 ? 1 : (lds > 54000 ? 2 : (lds > 40000 ? 3 : 4));
    int grid = cu_count() * per_cu;
    if (grid > nplanes) grid = nplanes;
    hipLaunchKernelGGL((kick_kernel<NVX, NVY, PF>), dim3(grid), dim3(1024), lds, s, fin, fout, nplanes, Ex, Ey, dt, dvx,
                       dvy);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

template <typename S, int NVX, int NVY>
static int launch_kick_stream(const S* fin, S* fout, int nplanes, const double* Ex, const double* Ey, double dt,
                              double dvx, double dvy, hipStream_t s) {
    constexpr int PPW = 128 / NVY;
    if (nplanes % PPW != 0) return VP_ERR_UNSUPPORTED;
    const int grid = nplanes / PPW;
    const bool pf16 = knob(VP_KNOB_KICK_PREFETCH) == 2;
    const bool nt = (nt_mask() & 4) != 0;
#define VP_K(PF, NT)                                                                                              \
    hipLaunchKernelGGL((kick_stream_kernel<S, NVX, NVY, PF, NT>), dim3(grid), dim3(128), 0, s, fin, fout, nplanes, \
                       Ex, Ey, dt, dvx, dvy)
    if (pf16) {
        if (nt) VP_K(16, true); else VP_K(16, false);
    } else {
        if (nt) VP_K(8, true); else VP_K(8, false);
    }
#undef VP_K
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// Row-split kick, 128 vy columns, WPP waves per plane, pf8: 8 rows
// prefetched (else 4), at most the rows per wave; feq != nullptr selects
// deviation storage (float only).
template <typename S, int NVX, int WPP, int NVY>
static int launch_kick_rows(bool pf8, const S* fin, S* fout, const double* feq, int nplanes, const double* Ex,
                            const double* Ey, double dt, double dvx, double dvy, hipStream_t s) {
    constexpr int RPW = NVX / WPP;
    constexpr int PF8 = RPW < 8 ? RPW : 8;
    constexpr int PF4 = RPW < 4 ? RPW : 4;
    constexpr bool F32 = sizeof(S) == 4;
    const bool nt = (nt_mask() & 4) != 0;
#define VP_K(PF, NT, DEV)                                                                                     \
    hipLaunchKernelGGL((kick_rows_kernel<S, NVX, WPP, PF, NT, DEV, NVY>), dim3(nplanes), dim3(64 * WPP), 0, s, \
                       fin, fout, feq, Ex, Ey, dt, dvx, dvy)
#define VP_K2(PF, DEV)                      \
    if (nt) VP_K(PF, true, DEV); else VP_K(PF, false, DEV)
    if (feq) {
        if constexpr (F32 && NVY == 128) {
            if (pf8) { VP_K2(PF8, true); } else { VP_K2(PF4, true); }
        } else {
            return VP_ERR_UNSUPPORTED;
        }
    } else {
        if (pf8) { VP_K2(PF8, false); } else { VP_K2(PF4, false); }
    }
#undef VP_K2
#undef VP_K
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// Kick kernel choice at 128 vy columns (and 64 for float64): the kick-kernel
// knob 3, 4, 5, 6 picks the row-split kernel with 2, 4, 8, 16 waves per
// plane, 2 the two-wave row-streaming kernel; the kick-prefetch knob 1 or 3
// picks 4 or 8 rows prefetched.  Automatic (knobs 0 and 2): 8 waves per
// plane, 4 rows prefetched for double and 8 for float (measured fastest at
// 128^4).  Deviation storage takes a row-split kernel, automatic = 4 waves
// per plane and 8 rows prefetched (at 128^4: 0.39 ms against 0.41 with 2
// waves and 0.58 with 8: its per-plane start reads the profiles' taps).
template <typename S, int NVX, int NVY = 128>
static int kick_rows_dispatch(int kk, int pfk, const S* fin, S* fout, const double* feq, int nplanes,
                              const double* Ex, const double* Ey, double dt, double dvx, double dvy, hipStream_t s) {
    if (feq && (kk < 3 || kk > 6)) kk = 4;
    // default: 8 waves per plane for float64, 4 for float32 (measured at
    // 128^4 on a regular field: float64 726 us with 8 against 755-771 with 4;
    // float32 371-379 us with 4 against 419-447 with 8)
    if (kk == 0) kk = sizeof(S) == 4 ? 4 : 5;
    const bool pf8 = pfk == 3 || (pfk != 1 && sizeof(S) == 4);
#define VP_R(W) return launch_kick_rows<S, NVX, W, NVY>(pf8, fin, fout, feq, nplanes, Ex, Ey, dt, dvx, dvy, s)
    if (kk == 3) VP_R(2);
    if (kk == 4) VP_R(4);
    if (kk == 6 && NVX >= 64) VP_R(16);
    if (kk == 5 || kk == 6) VP_R(8);
#undef VP_R
    return VP_ERR_UNSUPPORTED;  // row-streaming kernel
}

template <typename S>
static int kick_stream(const S* fin, S* fout, const int64_t shape[4], const double* Ex, const double* Ey, double dt,
                       double dvx, double dvy, hipStream_t s) {
    const int64_t nvx = shape[2], nvy = shape[3];
    const int nplanes = static_cast<int>(shape[0] * shape[1]);
    if (nvy == 128 || (nvy == 64 && sizeof(S) == 8)) {
        const int kk = knob(VP_KNOB_KICK_KERNEL), pfk = knob(VP_KNOB_KICK_PREFETCH);
        int rc = VP_ERR_UNSUPPORTED;
#define VP_RD(N, M) rc = kick_rows_dispatch<S, N, M>(kk, pfk, fin, fout, nullptr, nplanes, Ex, Ey, dt, dvx, dvy, s)
        if (nvy == 128) {
            if (nvx == 128) VP_RD(128, 128);
            if (nvx == 64) VP_RD(64, 128);
            if (nvx == 32) VP_RD(32, 128);
        } else {
            if (nvx == 128) VP_RD(128, 64);
            if (nvx == 64) VP_RD(64, 64);
            if (nvx == 32) VP_RD(32, 64);
        }
#undef VP_RD
        if (rc != VP_ERR_UNSUPPORTED) return rc;
    }
#define VP_KS(A, B) \
    if (nvx == A && nvy == B) return launch_kick_stream<S, A, B>(fin, fout, nplanes, Ex, Ey, dt, dvx, dvy, s);
    VP_KS(128, 128)
    VP_KS(128, 64)
    VP_KS(128, 32)
    VP_KS(64, 128)
    VP_KS(64, 64)
    VP_KS(64, 32)
    VP_KS(32, 128)
    VP_KS(32, 64)
    VP_KS(32, 32)
#undef VP_KS
    return VP_ERR_UNSUPPORTED;
}

int fast2d_kick(const double* fin, double* fout, const int64_t shape[4], const double* Ex, const double* Ey,
                double dt, double dvx, double dvy, hipStream_t s) {
    const int64_t nvx = shape[2], nvy = shape[3];
    const int nplanes = static_cast<int>(shape[0] * shape[1]);
    const int64_t cells = shape[0] * shape[1] * nvx * nvy;
    const bool disjoint = fout + cells <= fin || fin + cells <= fout;
    if (disjoint && knob(VP_KNOB_KICK_KERNEL) != 1) {
        const int rc = kick_stream<double>(fin, fout, shape, Ex, Ey, dt, dvx, dvy, s);
        if (rc != VP_ERR_UNSUPPORTED) return rc;
    }
    const bool pf = knob(VP_KNOB_KICK_PREFETCH) != 0;
#define VP_K(A, B)                                                                                  \
    if (nvx == A && nvy == B)                                                                       \
        return pf ? launch_kick<A, B, true>(fin, fout, nplanes, Ex, Ey, dt, dvx, dvy, s)            \
                  : launch_kick<A, B, false>(fin, fout, nplanes, Ex, Ey, dt, dvx, dvy, s);
    VP_K(128, 128)
    VP_K(128, 64)
    VP_K(128, 32)
    VP_K(64, 128)
    VP_K(64, 64)
    VP_K(64, 32)
    VP_K(32, 128)
    VP_K(32, 64)
    VP_K(32, 32)
#undef VP_K
    return VP_ERR_UNSUPPORTED;
}

// FP32 deviation storage: g = f - feq in float32 with the separable feq = fx fy
// (feq = fx | fy, float64): the float32 kick of g plus a float64 correction
// (row-split kernel; 128 vy columns; g_in and g_out must not overlap).
int fast2d_kick_dev(const float* gin, float* gout, const int64_t shape[4], const double* feq, const double* Ex,
                    const double* Ey, double dt, double dvx, double dvy, hipStream_t s) {
    const int64_t nvx = shape[2], nvy = shape[3];
    const int64_t cells = shape[0] * shape[1] * nvx * nvy;
    const bool disjoint = gout + cells <= gin || gin + cells <= gout;
    if (!disjoint || nvy != 128 || !feq) return VP_ERR_UNSUPPORTED;
    const int nplanes = static_cast<int>(shape[0] * shape[1]);
    const int kk = knob(VP_KNOB_KICK_KERNEL), pfk = knob(VP_KNOB_KICK_PREFETCH);
#define VP_RD(N) return kick_rows_dispatch<float, N>(kk, pfk, gin, gout, feq, nplanes, Ex, Ey, dt, dvx, dvy, s)
    if (nvx == 128) VP_RD(128);
    if (nvx == 64) VP_RD(64);
    if (nvx == 32) VP_RD(32);
#undef VP_RD
    return VP_ERR_UNSUPPORTED;
}

// FP32 storage: row-streaming kick only (f_in and f_out must not overlap).
int fast2d_kick(const float* fin, float* fout, const int64_t shape[4], const double* Ex, const double* Ey, double dt,
                double dvx, double dvy, hipStream_t s) {
    const int64_t cells = shape[0] * shape[1] * shape[2] * shape[3];
    const bool disjoint = fout + cells <= fin || fin + cells <= fout;
    if (!disjoint) return VP_ERR_UNSUPPORTED;
    return kick_stream<float>(fin, fout, shape, Ex, Ey, dt, dvx, dvy, s);
}

}  // namespace vp

extern "C" int vp_tune_set(int knob, int value) {
    if (knob < 0 || knob >= VP_KNOB_COUNT) return VP_ERR_ARG;
    vp::g_knobs[knob] = value;
    return VP_OK;
}

extern "C" int vp_tune_get(int knob) { return vp::knob(knob); }
