// Ensemble of independent 1D1V Vlasov-Poisson instances of different sizes
// (f_i[nx_i][nv_i] in one pool), stepped together: every launch covers all
// instances, so a 500-step run of 512 instances is 3 launches per step
// instead of ~5 per step per instance.
//
// Each instance follows the reference step (vlasov_step, vlasov1D1V.py:574-622):
// X(dt/2) periodic along x with velocity v_j, rho = sum_v f dv - 1, E from
// the 1D Poisson solve, V(dt) along v with E(x) (zero outside [0, nv)),
// X(dt/2).  The sweeps evaluate the reference arithmetic element by element
// (departure, power-form weights, left-to-right six-tap sum), so they are
// bit-identical to the single-instance path; only the charge-density sum
// order differs (compensated here).  The closing X(dt/2) of a step and the
// opening X(dt/2) of the next run as one launch (two passes through LDS).
//
// Launch shapes: grid.y = instance; grid.x = column chunk (X drift) or row
// block (kick); workgroups past an instance's extent exit at once.
#include "binade.hpp"
#include "field_dev.hpp"

namespace vp {

constexpr int kEnsThreads = 256;
constexpr int kEnsMaxNx = 256;
constexpr int kEnsMaxNv = 512;
// Binade codes tabulated per line (a line of at most 512 points reaches
// about 10; departure points with other codes compute their weights).
constexpr int kEnsCodes = 16;
constexpr int kEnsTabStride = kEnsCodes * 6 + 1;  // odd: lines' tables in different LDS banks

// Weight table of one line of n points with displacement disp: codes
// [q[0], q[0] + q[1]) of tab[code slot * 6 + k], filled by fill_line_tab.
__device__ __forceinline__ void line_codes(double disp, int n, int* q) {
    int q0, nq;
    segment_codes(-disp, static_cast<double>(n - 1) - disp, q0, nq);  // departure points of 0 and n - 1
    q[0] = q0;
    q[1] = nq < kEnsCodes ? nq : kEnsCodes;
}

// Every workgroup thread fills entries of `lines` tables of lines of n
// points (line l at tab + l * kEnsTabStride, displacement disp_of(l), which
// each thread evaluates for the lines it fills: no barrier is needed before
// the fill, so loads issued before it stay in flight).
template <typename DispOf>
__device__ __forceinline__ void fill_line_tabs(double* tab, int lines, int n, DispOf disp_of, int nt) {
    for (int e = threadIdx.x; e < lines * kEnsCodes; e += nt) {
        const int l = e / kEnsCodes, slot = e % kEnsCodes;
        const double disp = disp_of(l);
        int q[2];
        line_codes(disp, n, q);
        double w[6];
        if (slot < q[1]) {
            code_weights<double>(disp, q[0] + slot, w);
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) tab[l * kEnsTabStride + slot * 6 + k] = w[k];
    }
}

// One periodic X-drift value: row r of column c of a tile of n (>= 6) rows x
// C columns, the departure form with the column's displacement and weight
// table.  One int64 remainder per element, the six taps by stepping.
template <int C>
__device__ __forceinline__ double ens_periodic_sum(const double* tile, int r, double disp, const double* tab,
                                                   const int* q, int n, int c) {
    const double dep = static_cast<double>(r) - disp;
    const double fl = floor(dep);
    double w[6];
    departure_weights<double>(dep, fl, tab, q[0], q[1], w);
    int jr = wrap_index(static_cast<int64_t>(fl) - 2, n);
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        v[k] = tile[jr * C + c];
        jr += 1;
        jr -= (jr >= n) ? n : 0;
    }
    return tap_sum<double>(w, v);
}

// X drift of one column chunk (C velocities, all nx rows) of one instance,
// PASSES half-drifts; with RHO the row sums of the result over the chunk go
// to partial[(inst * maxch + chunk) * kEnsMaxNx + x] as (sum, compensation).
template <int C, int NT, int PASSES, bool RHO>
__global__ __launch_bounds__(NT) void ens_xdrift_kernel(double* __restrict__ pool,
                                                                 const vp_ens_inst* __restrict__ inst,
                                                                 const double* __restrict__ tables, double dt,
                                                                 int maxch, double2* __restrict__ partial) {
    __shared__ double bufA[kEnsMaxNx * C];
    __shared__ double bufB[kEnsMaxNx * C];
    __shared__ double tab[C * kEnsTabStride];
    const vp_ens_inst I = inst[blockIdx.y];
    const int nx = I.nx, nv = I.nv;
    const int c0 = blockIdx.x * C;
    if (c0 >= nv) return;  // uniform over the workgroup
    constexpr int NG = NT / C;
    const int t = threadIdx.x;
    const int c = t % C;
    const int g = t / C;
    const int j = c0 + c;
    const bool valid = j < nv;
    double* f = pool + I.f_off;
    // departure(), vlasov1D1V.py:229-238: the column's displacement
    auto disp_of = [&](int cc) {
        const double v = c0 + cc < nv ? tables[I.v_off + c0 + cc] : 0.0;
        return (v * dt) / I.dx;
    };
    const double disp = disp_of(c);
    int myq[2];
    line_codes(disp, nx, myq);
    // this thread's first rows load while the tables are computed
    constexpr int PB = kEnsMaxNx / NG < 8 ? kEnsMaxNx / NG : 8;
    double pre[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        const int r = g + k * NG;
        pre[k] = (valid && r < nx) ? f[static_cast<int64_t>(r) * nv + j] : 0.0;
    }
    fill_line_tabs(tab, C, nx, disp_of, NT);
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        const int r = g + k * NG;
        if (r < nx) bufA[r * C + c] = pre[k];
    }
    for (int r = g + PB * NG; r < nx; r += NG) bufA[r * C + c] = valid ? f[static_cast<int64_t>(r) * nv + j] : 0.0;
    __syncthreads();
    const double* mytab = tab + c * kEnsTabStride;
    for (int r = g; r < nx; r += NG) bufB[r * C + c] = ens_periodic_sum<C>(bufA, r, disp, mytab, myq, nx, c);
    const double* fin = bufB;  // result of the last pass
    if (PASSES == 2) {
        __syncthreads();
        for (int r = g; r < nx; r += NG) bufA[r * C + c] = ens_periodic_sum<C>(bufB, r, disp, mytab, myq, nx, c);
        fin = bufA;
    }
    __syncthreads();
    if (valid)
        for (int r = g; r < nx; r += NG) f[static_cast<int64_t>(r) * nv + j] = fin[r * C + c];
    if (RHO) {
        // invalid columns hold exact zeros (their taps read zeros), so they add nothing
        for (int r = t; r < nx; r += NT) {
            KSum k{0.0, 0.0};
            for (int cc = 0; cc < C; ++cc) ksum_add(k, fin[r * C + cc]);
            partial[(static_cast<int64_t>(blockIdx.y) * maxch + blockIdx.x) * kEnsMaxNx + r] = make_double2(k.s, k.c);
        }
    }
}

// Field solve of one instance per workgroup: rho from the chunk partials,
// E by poisson1d_block, E to Eout[inst * kEnsMaxNx + x], field energy
// ((1/2 sum E^2) dx) to energy[inst] (if energy is not null).
__global__ __launch_bounds__(kEnsThreads) void ens_poisson_kernel(const vp_ens_inst* __restrict__ inst,
                                                                  const double* __restrict__ tables,
                                                                  const double2* __restrict__ partial, int maxch,
                                                                  int C, int physics_mode, double* __restrict__ Eout,
                                                                  double* __restrict__ energy) {
    __shared__ double rho[kEnsMaxNx];
    __shared__ double El[kEnsMaxNx];
    __shared__ double2 tw[kEnsMaxNx];
    __shared__ double2 Ek[kEnsMaxNx / 2 + 1];
    __shared__ double2 Pk[kEnsMaxNx / 2 + 1];
    __shared__ double scratch[2 * kEnsThreads / kWave];
    const vp_ens_inst I = inst[blockIdx.x];
    const int nx = I.nx;
    const int nch = (I.nv + C - 1) / C;
    double* E = Eout + static_cast<int64_t>(blockIdx.x) * kEnsMaxNx;
    if (!physics_mode) {
        for (int x = threadIdx.x; x < nx; x += blockDim.x) E[x] = 0.0;
        if (energy && threadIdx.x == 0) energy[blockIdx.x] = 0.0;
        return;
    }
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        KSum k{0.0, 0.0};
        for (int ch = 0; ch < nch; ++ch) {
            const double2 p = partial[(static_cast<int64_t>(blockIdx.x) * maxch + ch) * kEnsMaxNx + x];
            ksum_merge(k, KSum{p.x, p.y});
        }
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

// V kick (dt, velocity E(x)) of RK rows of one instance, taps outside
// [0, nv) read as zero (clamp + mask, vlasov1D1V.py:248-253).
template <int RK>
__global__ __launch_bounds__(kEnsThreads) void ens_kick_kernel(double* __restrict__ pool,
                                                               const vp_ens_inst* __restrict__ inst,
                                                               const double* __restrict__ Ein, double dt) {
    __shared__ double tile[RK * kEnsMaxNv];
    __shared__ double tab[RK * kEnsTabStride];
    __shared__ double disps[RK];
    __shared__ int codes[2 * RK];
    const vp_ens_inst I = inst[blockIdx.y];
    const int nx = I.nx, nv = I.nv;
    const int x0 = blockIdx.x * RK;
    if (x0 >= nx) return;  // uniform over the workgroup
    const int rows = nx - x0 < RK ? nx - x0 : RK;
    double* f = pool + I.f_off + static_cast<int64_t>(x0) * nv;
    const int cnt = rows * nv;
    const double* E = Ein + static_cast<int64_t>(blockIdx.y) * kEnsMaxNx + x0;
    auto disp_of = [&](int rr) { return (E[rr] * dt) / I.dv; };  // departure(), vlasov1D1V.py:229-238
    if (static_cast<int>(threadIdx.x) < rows) {
        const double disp = disp_of(threadIdx.x);
        disps[threadIdx.x] = disp;
        line_codes(disp, nv, codes + 2 * threadIdx.x);
    }
    // the first elements of this thread load while the tables are computed
    constexpr int PB = 8;
    double pre[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        const int idx = static_cast<int>(threadIdx.x) + k * kEnsThreads;
        pre[k] = idx < cnt ? f[idx] : 0.0;
    }
    fill_line_tabs(tab, rows, nv, disp_of, kEnsThreads);
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        const int idx = static_cast<int>(threadIdx.x) + k * kEnsThreads;
        if (idx < cnt) tile[idx] = pre[k];
    }
    for (int idx = static_cast<int>(threadIdx.x) + PB * kEnsThreads; idx < cnt; idx += kEnsThreads) tile[idx] = f[idx];
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt; idx += kEnsThreads) {
        const int rr = idx / nv;
        const int j = idx - rr * nv;
        const double dep = static_cast<double>(j) - disps[rr];
        const double fl = floor(dep);
        double w[6];
        departure_weights<double>(dep, fl, tab + rr * kEnsTabStride, codes[2 * rr], codes[2 * rr + 1], w);
        const int64_t il = static_cast<int64_t>(fl);
        const double* row = tile + rr * nv;
        double v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int64_t jj = il + (k - 2);
            const int jc = static_cast<int>(jj < 0 ? 0 : (jj > nv - 1 ? nv - 1 : jj));
            const double m = (jj >= 0 && jj < nv) ? 1.0 : 0.0;
            v[k] = row[jc] * m;
        }
        f[idx] = tap_sum<double>(w, v);
    }
}

static int ens_cols() {
    const int c = knob(VP_KNOB_ENS_COLS);
    return c == 16 ? 16 : 32;
}

static int ens_rows() {
    const int r = knob(VP_KNOB_ENS_ROWS);
    return (r == 2 || r == 8 || r == 16) ? r : 4;
}

static size_t a256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

static size_t ens_ws(int ninst, int max_nv, int C) {
    const int maxch = (max_nv + C - 1) / C;
    return a256(static_cast<size_t>(ninst) * maxch * kEnsMaxNx * sizeof(double2)) +
           a256(static_cast<size_t>(ninst) * kEnsMaxNx * sizeof(double));
}

template <int C, int NT>
static void launch_x_nt(double* pool, const vp_ens_inst* inst, const double* tables, int ninst, int maxch, double dt,
                        int passes, bool rho, double2* partial, hipStream_t s) {
    const dim3 grid(maxch, ninst);
    if (passes == 2) {
        if (rho)
            hipLaunchKernelGGL((ens_xdrift_kernel<C, NT, 2, true>), grid, dim3(NT), 0, s, pool, inst, tables, dt,
                               maxch, partial);
        else
            hipLaunchKernelGGL((ens_xdrift_kernel<C, NT, 2, false>), grid, dim3(NT), 0, s, pool, inst, tables, dt,
                               maxch, partial);
    } else {
        if (rho)
            hipLaunchKernelGGL((ens_xdrift_kernel<C, NT, 1, true>), grid, dim3(NT), 0, s, pool, inst, tables, dt,
                               maxch, partial);
        else
            hipLaunchKernelGGL((ens_xdrift_kernel<C, NT, 1, false>), grid, dim3(NT), 0, s, pool, inst, tables, dt,
                               maxch, partial);
    }
}

template <int C>
static int launch_x(double* pool, const vp_ens_inst* inst, const double* tables, int ninst, int maxch, double dt,
                    int passes, bool rho, double2* partial, hipStream_t s) {
    const int nt = knob(VP_KNOB_ENS_XTHREADS);
    if (nt == 256)
        launch_x_nt<C, 256>(pool, inst, tables, ninst, maxch, dt, passes, rho, partial, s);
    else if (nt == 1024)
        launch_x_nt<C, 1024>(pool, inst, tables, ninst, maxch, dt, passes, rho, partial, s);
    else
        launch_x_nt<C, 512>(pool, inst, tables, ninst, maxch, dt, passes, rho, partial, s);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

static int launch_kick(double* pool, const vp_ens_inst* inst, const double* E, int ninst, int max_nx, double dt,
                       hipStream_t s) {
    const int RK = ens_rows();
    const dim3 grid((max_nx + RK - 1) / RK, ninst);
    switch (RK) {
        case 2: hipLaunchKernelGGL(ens_kick_kernel<2>, grid, dim3(kEnsThreads), 0, s, pool, inst, E, dt); break;
        case 8: hipLaunchKernelGGL(ens_kick_kernel<8>, grid, dim3(kEnsThreads), 0, s, pool, inst, E, dt); break;
        case 16: hipLaunchKernelGGL(ens_kick_kernel<16>, grid, dim3(kEnsThreads), 0, s, pool, inst, E, dt); break;
        default: hipLaunchKernelGGL(ens_kick_kernel<4>, grid, dim3(kEnsThreads), 0, s, pool, inst, E, dt); break;
    }
    VP_CHECK_LAUNCH();
    return VP_OK;
}

}  // namespace vp

extern "C" size_t vp_ens_workspace(int ninst, int max_nx, int max_nv) {
    if (ninst < 1 || max_nx < 1 || max_nx > vp::kEnsMaxNx || max_nv < 1 || max_nv > vp::kEnsMaxNv) return 0;
    const size_t a = vp::ens_ws(ninst, max_nv, 16), b = vp::ens_ws(ninst, max_nv, 32);
    return a > b ? a : b;  // valid for either column-chunk width
}

extern "C" int vp_ens_run_f64(double* pool, const vp_ens_inst* inst, int ninst, int max_nx, int max_nv,
                              const double* tables, int nsteps, double dt, int physics_mode, double* energy,
                              void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_RETURN_IF(!pool || !inst || !tables || !ws || nsteps < 0, VP_ERR_ARG);
    VP_RETURN_IF(ninst < 1 || ninst > 65535, VP_ERR_SHAPE);
    VP_RETURN_IF(max_nx < 6 || max_nx > vp::kEnsMaxNx || max_nv < 6 || max_nv > vp::kEnsMaxNv, VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < vp_ens_workspace(ninst, max_nx, max_nv), VP_ERR_WORKSPACE);
    if (nsteps == 0) return VP_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int C = vp::ens_cols();
    const int maxch = (max_nv + C - 1) / C;
    double2* partial = static_cast<double2*>(ws);
    double* E = reinterpret_cast<double*>(static_cast<unsigned char*>(ws) +
                                          vp::a256(static_cast<size_t>(ninst) * maxch * vp::kEnsMaxNx *
                                                   sizeof(double2)));
    const double half = dt / 2.0;
    auto xdrift = [&](int passes, bool rho) {
        return C == 16 ? vp::launch_x<16>(pool, inst, tables, ninst, maxch, half, passes, rho, partial, s)
                       : vp::launch_x<32>(pool, inst, tables, ninst, maxch, half, passes, rho, partial, s);
    };
    VP_TRY(xdrift(1, physics_mode != 0));
    for (int st = 0; st < nsteps; ++st) {
        hipLaunchKernelGGL(vp::ens_poisson_kernel, dim3(ninst), dim3(vp::kEnsThreads), 0, s, inst, tables, partial,
                           maxch, C, physics_mode, E, energy ? energy + static_cast<int64_t>(st) * ninst : nullptr);
        VP_CHECK_LAUNCH();
        VP_TRY(vp::launch_kick(pool, inst, E, ninst, max_nx, dt, s));
        if (st + 1 < nsteps)
            VP_TRY(xdrift(2, physics_mode != 0));
        else
            VP_TRY(xdrift(1, false));
    }
    return VP_OK;
}
