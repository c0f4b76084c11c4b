// Velocity moments (charge density) and the FFT Poisson / E-field solves.
//
// rho: one workgroup per spatial point sums its velocity row with Neumaier
// compensation (wave shuffle + LDS tree), replacing torch.sum(f, dim=...)
// of vlasov1D1V.py:559 and vlasov2D2V.py:416-431.
//
// Poisson: the spatial grids are small (<= 512 per axis, 128 in the
// benchmark configs), so the transforms are exact-twiddle float64 DFTs,
// one line per workgroup, twiddles from sincospi; the 1D solve of a
// power-of-two grid uses radix-2 FFTs in one workgroup.  Conventions follow
// torch.fft / numpy.fft: rfft, then c2c along x; the inverse is c2c along x
// (scaled 1/nx), then c2r along y (scaled 1/ny) reading only the real parts
// of the DC and Nyquist bins.
#include "field_dev.hpp"

namespace vp {

template <typename T>
__global__ __launch_bounds__(256) void row_sum_kernel(const T* __restrict__ f, int64_t len, double s1,
                                                      double s2, double shift, double* __restrict__ out) {
    __shared__ double scratch[2 * 256 / kWave];
    const T* row = f + static_cast<int64_t>(blockIdx.x) * len;
    KSum k{0.0, 0.0};
    for (int64_t j = threadIdx.x; j < len; j += blockDim.x) ksum_add(k, static_cast<double>(row[j]));
    const double total = block_ksum(k, scratch);
    if (threadIdx.x == 0) out[blockIdx.x] = ((total * s1) * s2) - shift;
}

// ---------------------------------------------------------------------------
// 1D: solve_poisson_real (vlasov1D1V.py:538-572), single workgroup.
// ws: tw[nx], Ek[nx/2+1], Pk[nx/2+1]  (double2)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void poisson1d_kernel(const double* __restrict__ rho, int nx,
                                                         const double* __restrict__ inv_kx,
                                                         const double* __restrict__ inv_kx_sq,
                                                         double* __restrict__ E, double* __restrict__ phi,
                                                         double2* __restrict__ ws) {
    poisson1d_block(rho, nx, inv_kx, inv_kx_sq, E, phi, ws, ws + nx, ws + nx + (nx / 2 + 1));
}

// 1D, many workgroups (nx from 512 to 4096): the DFT sums of
// poisson1d_block split over 16 lanes each (16 outputs per workgroup of 256
// threads, each lane one sixteenth of the sum, merged by a lane-shuffle
// tree), with the twiddles, rho and the half spectra staged in LDS.  Same
// terms, another summation order: agrees with the one-workgroup solve to
// rounding.  ws: Ek, Pk after the twiddle slot of the one-workgroup layout.
constexpr int kP1dParts = 16;

__device__ __forceinline__ double p1d_merge(double v) {
#pragma unroll
    for (int off = kP1dParts / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kP1dParts);
    return v;
}

__global__ __launch_bounds__(256) void p1d_fwd_kernel(const double* __restrict__ rho, int nx,
                                                      const double* __restrict__ inv_kx,
                                                      const double* __restrict__ inv_kx_sq, double2* __restrict__ Ek,
                                                      double2* __restrict__ Pk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);
    double* r = reinterpret_cast<double*>(tw + nx);
    fill_twiddles(tw, nx);
    for (int x = threadIdx.x; x < nx; x += blockDim.x) r[x] = rho[x];
    __syncthreads();
    const int nk = nx / 2 + 1;
    const int part = threadIdx.x % kP1dParts;
    const int k = blockIdx.x * (256 / kP1dParts) + threadIdx.x / kP1dParts;
    const int kk = k < nk ? k : 0;  // lanes past nk compute a dummy sum (every lane joins the shuffles)
    const int x0 = static_cast<int>(static_cast<int64_t>(part) * nx / kP1dParts);
    const int x1 = static_cast<int>(static_cast<int64_t>(part + 1) * nx / kP1dParts);
    double re = 0.0, im = 0.0;
    int ti = static_cast<int>((static_cast<int64_t>(kk) * x0) % nx);  // (k * x) mod nx, stepped
    for (int x = x0; x < x1; ++x) {
        const double2 t = tw[ti];
        re += r[x] * t.x;
        im -= r[x] * t.y;
        ti += kk;
        ti -= (ti >= nx) ? nx : 0;
    }
    re = p1d_merge(re);
    im = p1d_merge(im);
    if (part == 0 && k < nk) {
        Ek[k] = make_double2(im * inv_kx[k], -re * inv_kx[k]);
        Pk[k] = make_double2(re * inv_kx_sq[k], im * inv_kx_sq[k]);
    }
}

__global__ __launch_bounds__(256) void p1d_inv_kernel(const double2* __restrict__ Ek_g,
                                                      const double2* __restrict__ Pk_g, int nx,
                                                      double* __restrict__ E, double* __restrict__ phi) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nk = nx / 2 + 1;
    double2* tw = reinterpret_cast<double2*>(smem_raw);
    double2* Ek = tw + nx;
    double2* Pk = Ek + nk;
    fill_twiddles(tw, nx);
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        Ek[k] = Ek_g[k];
        Pk[k] = Pk_g[k];
    }
    __syncthreads();
    const int part = threadIdx.x % kP1dParts;
    const int x = blockIdx.x * (256 / kP1dParts) + threadIdx.x / kP1dParts;
    const int xx = x < nx ? x : 0;
    const int khalf = (nx + 1) / 2;  // bins 1..khalf-1 are doubled
    const int k0 = 1 + static_cast<int>(static_cast<int64_t>(part) * (khalf - 1) / kP1dParts);
    const int k1 = 1 + static_cast<int>(static_cast<int64_t>(part + 1) * (khalf - 1) / kP1dParts);
    double se = 0.0, sp = 0.0;
    int ti = static_cast<int>((static_cast<int64_t>(xx) * k0) % nx);  // (x * k) mod nx, stepped
    for (int k = k0; k < k1; ++k) {
        const double2 t = tw[ti];
        se += 2.0 * (Ek[k].x * t.x - Ek[k].y * t.y);
        sp += 2.0 * (Pk[k].x * t.x - Pk[k].y * t.y);
        ti += xx;
        ti -= (ti >= nx) ? nx : 0;
    }
    se = Ek[0].x + p1d_merge(se);
    sp = Pk[0].x + p1d_merge(sp);
    if (part != 0 || x >= nx) return;
    if ((nx & 1) == 0) {
        const double sgn = (x & 1) ? -1.0 : 1.0;
        se += Ek[nx / 2].x * sgn;
        sp += Pk[nx / 2].x * sgn;
    }
    const double inv_n = 1.0 / static_cast<double>(nx);
    E[x] = se * inv_n;
    phi[x] = sp * inv_n;
}

// 1D, power-of-two nx (8 to 4096): one workgroup, radix-2 FFTs in LDS (up
// to three stages per pass through LDS).
// Forward: the complex FFT of rho (decimation in frequency, bit-reversed
// output) gives X_k; E_k = -i X_k inv_kx[k], P_k = X_k inv_kx_sq[k] for
// k <= nx/2.  Both inverses at once: E and phi are real, so their
// Hermitian spectra combine into Z = E + i P (Z_k = E_k + i P_k,
// Z_{nx-k} = conj(E_k) + i conj(P_k), and only the real parts of the DC and
// Nyquist bins, as irfft reads them), whose inverse FFT is e + i p.  Same
// transform as the DFT kernels, O(nx log nx) instead of O(nx^2); agrees to
// rounding.  Twiddles from sincospi, one contiguous table per stage: the
// stage of half-length h reads exp(-i pi p / h), p < h, at tw[h - 1 + p]
// (n - 1 entries), so the lanes of a butterfly stage read consecutive
// entries instead of one entry every n / (2 h) (which put every lane of a
// wave in the same LDS bank).
constexpr int kFftMaxN = 4096;

// One pass of S consecutive radix-2 stages (S <= 3) done in registers: the
// 2^S elements b + m d (m < 2^S) are closed under those stages' butterflies,
// so each thread loads them once, applies the same butterflies with the same
// twiddles in the same order as stage-by-stage passes (identical results),
// and stores them once.  d is the smallest half-length of the pass.
// Decimation in frequency (DIF): halves d 2^(S-1), ..., d, butterfly
// (u + b, (u - b) w).  Decimation in time: halves d, ..., d 2^(S-1),
// butterfly (u + b w, u - b w), conjugate twiddles for the inverse.
// LDS slot of element i of the transform: i with its low 4 bits XORed by a
// fold of the higher bits (a bijection within each block of 16), so that the
// strided element sets of the passes and the bit-reversed bins of the
// spectral step fall in different banks (SW; without it, slot i).
template <bool SW>
__device__ __forceinline__ int fft_slot(int i) {
    if (!SW) return i;
    const int h = i >> 4;
    return i ^ ((h ^ (h >> 4) ^ (h >> 8)) & 15);
}

template <int S, bool DIF, bool INV, bool SW>
__device__ __forceinline__ void fft_pass(double2* a, const double2* tw, int nx, int d) {
    constexpr int M = 1 << S;
    const int groups = nx >> S;
    for (int t = threadIdx.x; t < groups; t += blockDim.x) {
        const int p = t & (d - 1);
        const int b = ((t - p) << S) + p;
        double2 x[M];
#pragma unroll
        for (int m = 0; m < M; ++m) x[m] = a[fft_slot<SW>(b + m * d)];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int sh = DIF ? S - 1 - k : k;  // pairs (m, m + 2^sh)
            const int hk = d << sh;              // this stage's half-length
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if ((m >> sh) & 1) continue;
                double2 w = tw[hk - 1 + (m & ((1 << sh) - 1)) * d + p];
                if (INV) w.y = -w.y;
                const double2 u = x[m];
                const double2 v = x[m + (1 << sh)];
                if (DIF) {
                    const double2 df = make_double2(u.x - v.x, u.y - v.y);
                    x[m] = make_double2(u.x + v.x, u.y + v.y);
                    x[m + (1 << sh)] = make_double2(df.x * w.x - df.y * w.y, df.x * w.y + df.y * w.x);
                } else {
                    const double2 bw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
                    x[m] = make_double2(u.x + bw.x, u.y + bw.y);
                    x[m + (1 << sh)] = make_double2(u.x - bw.x, u.y - bw.y);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) a[fft_slot<SW>(b + m * d)] = x[m];
    }
    __syncthreads();
}

// Decimation in time, conjugate twiddles: the inverse transform (without the
// 1/nx) of a bit-reversed spectrum, natural-order output; at most MAXS
// stages per pass.
template <int MAXS>
__device__ __forceinline__ void fft_inverse_dit(double2* a, const double2* tw, int nx, int lg) {
    int h = 1;
    for (int left = lg; left > 0;) {
        const int st = left >= MAXS ? MAXS : left;
        if (st == 3)
            fft_pass<3, false, true, (MAXS > 1)>(a, tw, nx, h);
        else if (st == 2)
            fft_pass<2, false, true, (MAXS > 1)>(a, tw, nx, h);
        else
            fft_pass<1, false, true, (MAXS > 1)>(a, tw, nx, h);
        h <<= st;
        left -= st;
    }
}

// Decimation in frequency: natural-order input, bit-reversed output (the
// order the decimation-in-time inverse reads), so neither transform needs a
// permutation pass (scattered LDS accesses).
template <int MAXS>
__device__ __forceinline__ void fft_forward_dif(double2* a, const double2* tw, int nx, int lg) {
    int H = nx >> 1;  // largest half-length left
    for (int left = lg; left > 0;) {
        const int st = left >= MAXS ? MAXS : left;
        const int d = H >> (st - 1);
        if (st == 3)
            fft_pass<3, true, false, (MAXS > 1)>(a, tw, nx, d);
        else if (st == 2)
            fft_pass<2, true, false, (MAXS > 1)>(a, tw, nx, d);
        else
            fft_pass<1, true, false, (MAXS > 1)>(a, tw, nx, d);
        H = d >> 1;
        left -= st;
    }
}

constexpr int kFftThreads = 1024;
constexpr int kFftPer = kFftMaxN / kFftThreads;           // points per thread
constexpr int kFftBins = kFftMaxN / 2 / kFftThreads + 1;  // bins k <= nx/2 per thread

// MAXS: radix-2 stages per LDS pass (3 from 1024 points up; below that the
// few active threads of a three-stage pass make it slower than single stages).
template <int MAXS>
__global__ __launch_bounds__(kFftThreads) void poisson1d_fft_kernel(const double* __restrict__ rho, int nx, int lg,
                                                                    const double* __restrict__ inv_kx,
                                                                    const double* __restrict__ inv_kx_sq,
                                                                    double* __restrict__ E, double* __restrict__ phi) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* a = reinterpret_cast<double2*>(smem_raw);  // [nx]
    double2* tw = a + nx;                                 // [nx - 1]: per-stage tables
    const int half_n = nx >> 1;
    const int t0 = static_cast<int>(threadIdx.x);
    // this thread's rho values and spectral factors load while the twiddles
    // are computed
    double rv[kFftPer], ik[kFftBins], ik2[kFftBins];
#pragma unroll
    for (int i = 0; i < kFftPer; ++i) {
        const int x = t0 + i * kFftThreads;
        rv[i] = x < nx ? rho[x] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kFftBins; ++i) {
        const int k = t0 + i * kFftThreads;
        ik[i] = k <= half_n ? inv_kx[k] : 0.0;
        ik2[i] = k <= half_n ? inv_kx_sq[k] : 0.0;
    }
    const int lh = lg - 1;
    if (nx - 1 <= kFftThreads) {
        // one entry per thread: entry e of the stage h' = 2^le is exp(-i pi pos / h')
        if (t0 < nx - 1) {
            const int le = 31 - __clz(t0 + 1);  // h' - 1 <= e < 2 h' - 1
            double sn, cs;
            sincospi(ldexp(static_cast<double>(t0 - ((1 << le) - 1)), -le), &sn, &cs);
            tw[t0] = make_double2(cs, -sn);
        }
    } else {
        // the largest stage's table (h = nx/2) from sincospi over its first half
        // (exp(-i pi (p + h/2) / h) = -i exp(-i pi p / h)); every smaller stage's
        // entry exp(-i pi pos / h') is the entry pos * (nx/2) / h' of that table
        double2* big = tw + (half_n - 1);
        for (int p = t0; p < half_n / 2; p += kFftThreads) {
            double sn, cs;
            sincospi(ldexp(static_cast<double>(p), -lh), &sn, &cs);  // p / (nx/2), exact
            big[p] = make_double2(cs, -sn);
            big[p + half_n / 2] = make_double2(-sn, -cs);
        }
        __syncthreads();
        for (int e = t0; e < half_n - 1; e += kFftThreads) {
            const int le = 31 - __clz(e + 1);  // stage of entry e: h' = 2^le, h' - 1 <= e < 2 h' - 1
            tw[e] = big[(e - ((1 << le) - 1)) << (lh - le)];
        }
    }
#pragma unroll
    for (int i = 0; i < kFftPer; ++i) {
        const int x = t0 + i * kFftThreads;
        if (x < nx) a[fft_slot<(MAXS > 1)>(x)] = make_double2(rv[i], 0.0);
    }
    __syncthreads();
    fft_forward_dif<MAXS>(a, tw, nx, lg);  // X_k at a[brev(k)]
    auto br = [lg](int k) { return fft_slot<(MAXS > 1)>(static_cast<int>(__brev(static_cast<unsigned>(k)) >> (32 - lg))); };
    // each thread reads bins k <= nx/2 it alone writes; bins above nx/2 are not
    // read (br gives the LDS slot of bin k)
#pragma unroll
    for (int i = 0; i < kFftBins; ++i) {
        const int k = t0 + i * kFftThreads;
        if (k > half_n) continue;
        const double2 X = a[br(k)];
        const double er = X.y * ik[i], ei = -X.x * ik[i];
        const double pr = X.x * ik2[i], pi = X.y * ik2[i];
        if (k == 0 || k == half_n) {
            a[br(k)] = make_double2(er, pr);
        } else {
            a[br(k)] = make_double2(er - pi, ei + pr);
            a[br(nx - k)] = make_double2(er + pi, pr - ei);
        }
    }
    __syncthreads();
    fft_inverse_dit<MAXS>(a, tw, nx, lg);  // bit-reversed spectrum in, natural order out
    const double inv_n = 1.0 / static_cast<double>(nx);  // exact for a power of two
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        const double2 z = a[fft_slot<(MAXS > 1)>(x)];
        E[x] = z.x * inv_n;
        phi[x] = z.y * inv_n;
    }
}

// ---------------------------------------------------------------------------
// 2D: compute_poisson_real (vlasov2D2V.py:488-517).
// ws: R[nx][nky], T[3][nx][nky]  (double2), nky = ny/2 + 1
// ---------------------------------------------------------------------------
// rho comes either from `rho` or, when `partial` is given, from nparts
// compensated partial sums per point (layout [part][x * ny + y], merged in
// part order) times dvx, then dvy.
__global__ __launch_bounds__(256) void p2d_rows_fwd(const double* __restrict__ rho, const double2* __restrict__ partial,
                                                    int nparts, double dvx, double dvy, int nx, int ny,
                                                    const double* __restrict__ rho_shift,
                                                    double* __restrict__ delta_rho, double2* __restrict__ R) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);
    double* row = reinterpret_cast<double*>(tw + ny);
    const int x = blockIdx.x;
    fill_twiddles(tw, ny);
    for (int y = threadIdx.x; y < ny; y += blockDim.x) {
        double rv;
        if (partial) {
            const int64_t npoints = static_cast<int64_t>(nx) * ny;
            const double2* pp = partial + static_cast<int64_t>(x) * ny + y;  // layout [part][x * ny + y]
            KSum k{0.0, 0.0};
            for (int i = 0; i < nparts; ++i) ksum_merge(k, KSum{pp[i * npoints].x, pp[i * npoints].y});
            rv = ((k.s + k.c) * dvx) * dvy;
        } else {
            rv = rho[static_cast<int64_t>(x) * ny + y];
        }
        // delta_rho = rho - 1; with deviation storage the partial sums hold
        // sum g dv and rho_shift[0] = sum feq dv - 1
        const double d = rho_shift ? rv + rho_shift[0] : rv - 1.0;
        row[y] = d;
        if (delta_rho) delta_rho[static_cast<int64_t>(x) * ny + y] = d;
    }
    __syncthreads();
    const int nky = ny / 2 + 1;
    for (int k = threadIdx.x; k < nky; k += blockDim.x) {
        double re = 0.0, im = 0.0;
        int ti = 0;  // (k * y) mod ny, stepped
        for (int y = 0; y < ny; ++y) {
            const double2 t = tw[ti];
            re += row[y] * t.x;
            im -= row[y] * t.y;
            ti += k;
            ti -= (ti >= ny) ? ny : 0;
        }
        R[static_cast<int64_t>(x) * nky + k] = make_double2(re, im);
    }
}

__global__ __launch_bounds__(256) void p2d_cols(const double2* __restrict__ R, int nx, int ny,
                                                const double* __restrict__ kx, const double* __restrict__ ky,
                                                const double* __restrict__ inv_ksq, double2* __restrict__ T) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);
    double2* col = tw + nx;
    double2* Q = col + nx;  // [3][nx]
    const int nky = ny / 2 + 1;
    const int ky_i = blockIdx.x;
    fill_twiddles(tw, nx);
    for (int x = threadIdx.x; x < nx; x += blockDim.x) col[x] = R[static_cast<int64_t>(x) * nky + ky_i];
    __syncthreads();
    const double kyv = ky[ky_i];
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
        double re = 0.0, im = 0.0;
        int ti = 0;  // (k * x) mod nx, stepped
        for (int x = 0; x < nx; ++x) {
            const double2 t = tw[ti];
            // (a + ib)(c - is) = (ac + bs) + i(bc - as)
            re += col[x].x * t.x + col[x].y * t.y;
            im += col[x].y * t.x - col[x].x * t.y;
            ti += k;
            ti -= (ti >= nx) ? nx : 0;
        }
        const double iks = inv_ksq[static_cast<int64_t>(k) * nky + ky_i];
        const double pr = re * iks, pi = im * iks;
        const double kxv = kx[k];
        Q[k] = make_double2(pr, pi);                      // phi_k
        Q[nx + k] = make_double2(pi * kxv, -pr * kxv);    // Ex_k = -i kx phi_k
        Q[2 * nx + k] = make_double2(pi * kyv, -pr * kyv);  // Ey_k = -i ky phi_k
    }
    __syncthreads();
    const double inv_n = 1.0 / static_cast<double>(nx);
    for (int idx = threadIdx.x; idx < 3 * nx; idx += blockDim.x) {
        const int q = idx / nx;
        const int x = idx - q * nx;
        const double2* Qq = Q + q * nx;
        double re = 0.0, im = 0.0;
        int ti = 0;  // (x * k) mod nx, stepped
        for (int k = 0; k < nx; ++k) {
            const double2 t = tw[ti];
            // (a + ib)(c + is) = (ac - bs) + i(bc + as)
            re += Qq[k].x * t.x - Qq[k].y * t.y;
            im += Qq[k].y * t.x + Qq[k].x * t.y;
            ti += x;
            ti -= (ti >= nx) ? nx : 0;
        }
        T[(static_cast<int64_t>(q) * nx + x) * nky + ky_i] = make_double2(re * inv_n, im * inv_n);
    }
}

__global__ __launch_bounds__(256) void p2d_rows_inv(const double2* __restrict__ T, int nx, int ny,
                                                    double* __restrict__ phi, double* __restrict__ Ex,
                                                    double* __restrict__ Ey) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);
    double2* row = tw + ny;  // [nky]
    const int nky = ny / 2 + 1;
    const int x = blockIdx.x % nx;
    const int q = blockIdx.x / nx;
    double* out = q == 0 ? phi : (q == 1 ? Ex : Ey);
    if (!out) return;
    fill_twiddles(tw, ny);
    for (int k = threadIdx.x; k < nky; k += blockDim.x)
        row[k] = T[(static_cast<int64_t>(q) * nx + x) * nky + k];
    __syncthreads();
    const int khalf = (ny + 1) / 2;
    const double inv_n = 1.0 / static_cast<double>(ny);
    for (int y = threadIdx.x; y < ny; y += blockDim.x) {
        double s = row[0].x;
        int ti = y % ny;  // (y * k) mod ny, stepped
        for (int k = 1; k < khalf; ++k) {
            const double2 t = tw[ti];
            s += 2.0 * (row[k].x * t.x - row[k].y * t.y);
            ti += y;
            ti -= (ti >= ny) ? ny : 0;
        }
        if ((ny & 1) == 0) s += row[ny / 2].x * ((y & 1) ? -1.0 : 1.0);
        out[static_cast<int64_t>(x) * ny + y] = s * inv_n;
    }
}

// ---------------------------------------------------------------------------
// 2D, power-of-two nx and ny (8..512): the same transforms as radix-2 FFTs.
// Three kernels as above (rows forward, columns, rows inverse); each
// workgroup holds several lines in LDS and runs up to three butterfly stages
// per pass through LDS on all of them at once.  Forward transforms are
// decimation in frequency (natural order in, bit-reversed out), inverses
// decimation in time with conjugate twiddles (bit-reversed in, natural
// out), so no permutation pass is needed; twiddles come from one table per
// stage, tw[h - 1 + p] = exp(-i pi p / h), h the stage's half-length
// (sincospi of exact arguments).  Agrees with the DFT kernels to rounding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fill_stage_twiddles(double2* tw, int n) {
    for (int e = threadIdx.x; e < n - 1; e += blockDim.x) {
        const int le = 31 - __clz(e + 1);  // stage h = 2^le holds entries h - 1 .. 2h - 2
        double sn, cs;
        sincospi(ldexp(static_cast<double>(e - ((1 << le) - 1)), -le), &sn, &cs);
        tw[e] = make_double2(cs, -sn);
    }
}

// S consecutive stages on `lines` lines of n points stored back to back at
// a (see fft_pass: the 2^S elements b + m d of a line are closed under
// those stages), then a barrier.
template <int S, bool DIF, bool INV>
__device__ __forceinline__ void lines_pass(double2* a, int lines, int n, int d, const double2* tw) {
    constexpr int M = 1 << S;
    const int gshift = 31 - __clz(n >> S);  // groups per line = n >> S, a power of two
    const int total = lines << gshift;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int line = idx >> gshift;
        const int t = idx - (line << gshift);
        const int p = t & (d - 1);
        const int b = ((t - p) << S) + p;
        double2* base = a + line * n;
        double2 x[M];
#pragma unroll
        for (int m = 0; m < M; ++m) x[m] = base[b + m * d];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int sh = DIF ? S - 1 - k : k;
            const int hk = d << sh;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if ((m >> sh) & 1) continue;
                double2 w = tw[hk - 1 + (m & ((1 << sh) - 1)) * d + p];
                if (INV) w.y = -w.y;
                const double2 u = x[m];
                const double2 v = x[m + (1 << sh)];
                if (DIF) {
                    const double2 df = make_double2(u.x - v.x, u.y - v.y);
                    x[m] = make_double2(u.x + v.x, u.y + v.y);
                    x[m + (1 << sh)] = make_double2(df.x * w.x - df.y * w.y, df.x * w.y + df.y * w.x);
                } else {
                    const double2 bw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
                    x[m] = make_double2(u.x + bw.x, u.y + bw.y);
                    x[m + (1 << sh)] = make_double2(u.x - bw.x, u.y - bw.y);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) base[b + m * d] = x[m];
    }
    __syncthreads();
}

// Forward transform of every line (natural order in, X_k at brev(k) out).
__device__ __forceinline__ void lines_forward(double2* a, int lines, int n, int lg, const double2* tw) {
    int H = n >> 1;
    for (int left = lg; left > 0;) {
        const int st = left >= 3 ? 3 : left;
        const int d = H >> (st - 1);
        if (st == 3)
            lines_pass<3, true, false>(a, lines, n, d, tw);
        else if (st == 2)
            lines_pass<2, true, false>(a, lines, n, d, tw);
        else
            lines_pass<1, true, false>(a, lines, n, d, tw);
        H = d >> 1;
        left -= st;
    }
}

// Inverse transform without the 1/n (bit-reversed in, natural order out).
__device__ __forceinline__ void lines_inverse(double2* a, int lines, int n, int lg, const double2* tw) {
    int h = 1;
    for (int left = lg; left > 0;) {
        const int st = left >= 3 ? 3 : left;
        if (st == 3)
            lines_pass<3, false, true>(a, lines, n, h, tw);
        else if (st == 2)
            lines_pass<2, false, true>(a, lines, n, h, tw);
        else
            lines_pass<1, false, true>(a, lines, n, h, tw);
        h <<= st;
        left -= st;
    }
}

__device__ __forceinline__ int brev_lg(int k, int lg) {
    return static_cast<int>(__brev(static_cast<unsigned>(k)) >> (32 - lg));
}

// rows x0 .. x0 + rows - 1: rho (from `rho` or merged partial sums), delta_rho,
// rfft along y -> R[x][k], k <= ny/2.  TPP = blockDim.x / (rows * ny)
// threads per point (a power of two, at most 8; 1 when points outnumber
// threads): thread part h of a point loads the partial sums h, h + TPP, ...
// all at once (up to kMergeLoads in flight), the stage twiddles are computed
// while they land, then the thread merges them in order (compensated) and the
// TPP threads of the point merge their sums by lane shuffles.
constexpr int kMergeLoads = 24;

__global__ __launch_bounds__(256) void p2d_fft_rows_fwd(const double* __restrict__ rho,
                                                        const double2* __restrict__ partial, int nparts, double dvx,
                                                        double dvy, int nx, int ny, int lgy, int rows,
                                                        const double* __restrict__ rho_shift,
                                                        double* __restrict__ delta_rho, double2* __restrict__ R) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);  // [ny - 1]
    double2* a = tw + ny;                                  // [rows][ny]
    const int x0 = static_cast<int>(blockIdx.x) * rows;
    const int64_t npoints = static_cast<int64_t>(nx) * ny;
    const int npts = rows * ny;
    int tpp = static_cast<int>(blockDim.x) / npts;
    tpp = tpp >= 8 ? 8 : (tpp >= 4 ? 4 : (tpp >= 2 ? 2 : 1));
    if (partial && tpp > 1 && nparts <= tpp * kMergeLoads) {
        // one point per TPP consecutive threads; the threads past npts * tpp idle
        const int t = threadIdx.x;
        const int idx = t / tpp, h = t - idx * tpp;
        const bool active = idx < npts;
        const int64_t pt = static_cast<int64_t>(x0) * ny + (active ? idx : 0);
        const double2* pp = partial + pt;
        double2 v[kMergeLoads];
#pragma unroll
        for (int u = 0; u < kMergeLoads; ++u) {
            const int i = h + u * tpp;
            v[u] = (active && i < nparts) ? pp[i * npoints] : make_double2(0.0, 0.0);
        }
        fill_stage_twiddles(tw, ny);
        KSum k{0.0, 0.0};
#pragma unroll
        for (int u = 0; u < kMergeLoads; ++u) ksum_merge(k, KSum{v[u].x, v[u].y});
        for (int off = 1; off < tpp; off <<= 1) {
            KSum o;
            o.s = __shfl_xor(k.s, off, kWave);
            o.c = __shfl_xor(k.c, off, kWave);
            ksum_merge(k, o);
        }
        if (active && h == 0) {
            const double rv = ((k.s + k.c) * dvx) * dvy;
            const double d = rho_shift ? rv + rho_shift[0] : rv - 1.0;
            a[idx] = make_double2(d, 0.0);
            if (delta_rho) delta_rho[pt] = d;
        }
    } else {
        fill_stage_twiddles(tw, ny);
        for (int idx = threadIdx.x; idx < npts; idx += blockDim.x) {
            const int64_t pt = static_cast<int64_t>(x0) * ny + idx;  // x * ny + y
            double rv;
            if (partial) {
                const double2* pp = partial + pt;  // layout [part][x * ny + y]
                KSum k{0.0, 0.0};
                int i = 0;
                for (; i + 8 <= nparts; i += 8) {
                    double2 v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) v[u] = pp[(i + u) * npoints];
#pragma unroll
                    for (int u = 0; u < 8; ++u) ksum_merge(k, KSum{v[u].x, v[u].y});
                }
                for (; i < nparts; ++i) ksum_merge(k, KSum{pp[i * npoints].x, pp[i * npoints].y});
                rv = ((k.s + k.c) * dvx) * dvy;
            } else {
                rv = rho[pt];
            }
            const double d = rho_shift ? rv + rho_shift[0] : rv - 1.0;
            a[idx] = make_double2(d, 0.0);
            if (delta_rho) delta_rho[pt] = d;
        }
    }
    __syncthreads();
    lines_forward(a, rows, ny, lgy, tw);
    const int nky = ny / 2 + 1;
    for (int idx = threadIdx.x; idx < rows * nky; idx += blockDim.x) {
        const int r = idx / nky;
        const int k = idx - r * nky;
        R[static_cast<int64_t>(x0 + r) * nky + k] = a[r * ny + brev_lg(k, lgy)];
    }
}

// column ky_i: c2c FFT along x, phi_k = R_k inv_ksq, Ex_k = -i kx phi_k,
// Ey_k = -i ky phi_k, three inverse FFTs along x (scaled 1/nx) -> T[q][x][ky_i].
// blockDim.x == 256 and nx <= 512: each thread owns slots threadIdx.x and
// threadIdx.x + 256; the column and the inv_ksq and kx values of a thread's
// slots (bins in bit-reversed order) are loaded before the twiddles are
// computed.
__global__ __launch_bounds__(256) void p2d_fft_cols(const double2* __restrict__ R, int nx, int ny, int lgx,
                                                    const double* __restrict__ kx, const double* __restrict__ ky,
                                                    const double* __restrict__ inv_ksq, double2* __restrict__ T) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);  // [nx - 1]
    double2* a = tw + nx;                                  // [4][nx]: the column, then phi_k, Ex_k, Ey_k
    const int ky_i = static_cast<int>(blockIdx.x);
    const int nky = ny / 2 + 1;
    double iks[2] = {0.0, 0.0}, kxs[2] = {0.0, 0.0};
    double2 col[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int s = threadIdx.x + u * blockDim.x;
        if (s < nx) {
            const int k = brev_lg(s, lgx);  // slot s holds bin k
            iks[u] = inv_ksq[static_cast<int64_t>(k) * nky + ky_i];
            kxs[u] = kx[k];
            col[u] = R[static_cast<int64_t>(s) * nky + ky_i];
        }
    }
    const double kyv = ky[ky_i];
    fill_stage_twiddles(tw, nx);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int x = threadIdx.x + u * blockDim.x;
        if (x < nx) a[x] = col[u];
    }
    __syncthreads();
    lines_forward(a, 1, nx, lgx, tw);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int s = threadIdx.x + u * blockDim.x;
        if (s < nx) {
            const double2 X = a[s];
            const double pr = X.x * iks[u], pi = X.y * iks[u];
            a[nx + s] = make_double2(pr, pi);                        // phi_k
            a[2 * nx + s] = make_double2(pi * kxs[u], -pr * kxs[u]);  // Ex_k = -i kx phi_k
            a[3 * nx + s] = make_double2(pi * kyv, -pr * kyv);        // Ey_k = -i ky phi_k
        }
    }
    __syncthreads();
    lines_inverse(a + nx, 3, nx, lgx, tw);
    const double inv_n = 1.0 / static_cast<double>(nx);  // exact: nx is a power of two
    for (int idx = threadIdx.x; idx < 3 * nx; idx += blockDim.x) {
        const int q = idx / nx;
        const int x = idx - q * nx;
        const double2 v = a[nx + idx];
        T[(static_cast<int64_t>(q) * nx + x) * nky + ky_i] = make_double2(v.x * inv_n, v.y * inv_n);
    }
}

// rows x0 .. x0 + rows - 1: irfft along y of phi_k, Ex_k, Ey_k (bins k <=
// ny/2; the imaginary parts of the DC and Nyquist bins dropped, as irfft
// does).  Ex and Ey are real, so one inverse FFT of the Hermitian extension
// of Ex_k + i Ey_k gives Ex + i Ey; phi takes a transform of its own.
// rows * (ny/2 + 1) <= 2 * blockDim.x: each thread loads at most two bins,
// before the twiddles are computed.
__global__ __launch_bounds__(256) void p2d_fft_rows_inv(const double2* __restrict__ T, int nx, int ny, int lgy,
                                                        int rows, double* __restrict__ phi, double* __restrict__ Ex,
                                                        double* __restrict__ Ey) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* tw = reinterpret_cast<double2*>(smem_raw);  // [ny - 1]
    double2* a = tw + ny;                                  // [rows][2][ny]: Ex + i Ey, then phi
    const int nky = ny / 2 + 1;
    const int half = ny / 2;
    const int nb = rows * nky;
    const int x0 = static_cast<int>(blockIdx.x) * rows;
    double2 P[2], E[2], Y[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int idx = threadIdx.x + u * blockDim.x;
        if (idx < nb) {
            const int r = idx / nky;
            const int k = idx - r * nky;
            const int x = x0 + r;
            P[u] = T[(static_cast<int64_t>(0) * nx + x) * nky + k];
            E[u] = T[(static_cast<int64_t>(1) * nx + x) * nky + k];
            Y[u] = T[(static_cast<int64_t>(2) * nx + x) * nky + k];
        }
    }
    fill_stage_twiddles(tw, ny);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int idx = threadIdx.x + u * blockDim.x;
        if (idx < nb) {
            const int r = idx / nky;
            const int k = idx - r * nky;
            const double2 p = P[u], e = E[u], y = Y[u];
            double2* z = a + (2 * r) * ny;
            double2* w = z + ny;
            if (k == 0 || k == half) {
                z[brev_lg(k, lgy)] = make_double2(e.x, y.x);
                w[brev_lg(k, lgy)] = make_double2(p.x, 0.0);
            } else {
                z[brev_lg(k, lgy)] = make_double2(e.x - y.y, e.y + y.x);       // E_k + i Y_k
                z[brev_lg(ny - k, lgy)] = make_double2(e.x + y.y, y.x - e.y);  // conj(E_k) + i conj(Y_k)
                w[brev_lg(k, lgy)] = p;
                w[brev_lg(ny - k, lgy)] = make_double2(p.x, -p.y);
            }
        }
    }
    __syncthreads();
    lines_inverse(a, 2 * rows, ny, lgy, tw);
    const double inv_n = 1.0 / static_cast<double>(ny);  // exact: ny is a power of two
    for (int idx = threadIdx.x; idx < rows * ny; idx += blockDim.x) {
        const int r = idx / ny;
        const int y = idx - r * ny;
        const int64_t o = static_cast<int64_t>(x0 + r) * ny + y;
        const double2 z = a[(2 * r) * ny + y];
        if (Ex) Ex[o] = z.x * inv_n;
        if (Ey) Ey[o] = z.y * inv_n;
        if (phi) phi[o] = a[(2 * r + 1) * ny + y].x * inv_n;
    }
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int fill(double* p, int64_t n, double v, hipStream_t s) {
    if (!p || n <= 0) return VP_OK;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

template <typename T>
int row_sum(const T* f, int64_t rows, int64_t len, double s1, double s2, double shift, double* out,
            hipStream_t s) {
    VP_RETURN_IF(!f || !out, VP_ERR_ARG);
    VP_RETURN_IF(rows < 1 || len < 1 || rows > 0x7fffffffLL, VP_ERR_SHAPE);
    hipLaunchKernelGGL((row_sum_kernel<T>), dim3(rows), dim3(256), 0, s, f, len, s1, s2, shift, out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}
template int row_sum<double>(const double*, int64_t, int64_t, double, double, double, double*, hipStream_t);
template int row_sum<float>(const float*, int64_t, int64_t, double, double, double, double*, hipStream_t);

size_t poisson1d_ws(int64_t nx) { return static_cast<size_t>(nx + 2 * (nx / 2 + 1)) * sizeof(double2); }

int poisson1d(const double* rho, int64_t nx, const double* inv_kx, const double* inv_kx_sq, double* E,
              double* phi, void* ws, size_t ws_bytes, hipStream_t s) {
    VP_RETURN_IF(!rho || !inv_kx || !inv_kx_sq || !E || !phi || !ws, VP_ERR_ARG);
    VP_RETURN_IF(nx < 1 || nx > 8192, VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < poisson1d_ws(nx), VP_ERR_WORKSPACE);
    if (nx >= 8 && nx <= kFftMaxN && (nx & (nx - 1)) == 0) {
        const int n = static_cast<int>(nx);
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        static bool attr = false;
        if (!attr) {
            VP_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&poisson1d_fft_kernel<1>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(2 * kFftMaxN * sizeof(double2))));
            VP_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&poisson1d_fft_kernel<3>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(2 * kFftMaxN * sizeof(double2))));
            attr = true;
        }
        const size_t lds = static_cast<size_t>(2 * n - 1) * sizeof(double2);
        if (n >= 1024)
            hipLaunchKernelGGL(poisson1d_fft_kernel<3>, dim3(1), dim3(kFftThreads), lds, s, rho, n, lg, inv_kx,
                               inv_kx_sq, E, phi);
        else
            hipLaunchKernelGGL(poisson1d_fft_kernel<1>, dim3(1), dim3(kFftThreads), lds, s, rho, n, lg, inv_kx,
                               inv_kx_sq, E, phi);
        VP_CHECK_LAUNCH();
        return VP_OK;
    }
    if (nx >= 512 && nx <= 4096) {
        // many workgroups; LDS: fwd tw + rho (96 KB at 4096), inv tw + Ek + Pk (128 KB)
        const int n = static_cast<int>(nx), nk = n / 2 + 1;
        const size_t lds_f = static_cast<size_t>(n) * (sizeof(double2) + sizeof(double));
        const size_t lds_i = static_cast<size_t>(n + 2 * nk) * sizeof(double2);
        static bool attr = false;
        if (!attr) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&p1d_fwd_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 4096 * 24);
            if (e == hipSuccess)
                e = hipFuncSetAttribute(reinterpret_cast<const void*>(&p1d_inv_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (4096 + 2 * 2049) * 16);
            if (e != hipSuccess) return static_cast<int>(e);
            attr = true;
        }
        double2* Ek = static_cast<double2*>(ws) + n;
        double2* Pk = Ek + nk;
        constexpr int per = 256 / kP1dParts;  // outputs per workgroup
        hipLaunchKernelGGL(p1d_fwd_kernel, dim3((nk + per - 1) / per), dim3(256), lds_f, s, rho, n, inv_kx, inv_kx_sq,
                           Ek, Pk);
        VP_CHECK_LAUNCH();
        hipLaunchKernelGGL(p1d_inv_kernel, dim3((n + per - 1) / per), dim3(256), lds_i, s, Ek, Pk, n, E, phi);
        VP_CHECK_LAUNCH();
        return VP_OK;
    }
    hipLaunchKernelGGL(poisson1d_kernel, dim3(1), dim3(1024), 0, s, rho, static_cast<int>(nx), inv_kx,
                       inv_kx_sq, E, phi, static_cast<double2*>(ws));
    VP_CHECK_LAUNCH();
    return VP_OK;
}

size_t poisson2d_ws(int64_t nx, int64_t ny) {
    return static_cast<size_t>(4) * nx * (ny / 2 + 1) * sizeof(double2);
}

static bool fft2d_enabled() { return knob(VP_KNOB_POISSON_DFT) != 1; }

int poisson2d_from(const double* rho, const double2* partial, int nparts, double dvx, double dvy, int64_t nx,
                   int64_t ny, const double* kx, const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                   double* delta_rho, double* phi, void* ws, size_t ws_bytes, hipStream_t s,
                   const double* rho_shift) {
    VP_RETURN_IF((!rho && !partial) || !kx || !ky || !inv_ksq || !ws, VP_ERR_ARG);
    VP_RETURN_IF(nx < 1 || ny < 1 || nx > 512 || ny > 512, VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < poisson2d_ws(nx, ny), VP_ERR_WORKSPACE);
    const int nxi = static_cast<int>(nx), nyi = static_cast<int>(ny);
    const int nky = nyi / 2 + 1;
    double2* R = static_cast<double2*>(ws);
    double2* T = R + static_cast<int64_t>(nx) * nky;
    auto pow2 = [](int n) { return n >= 8 && (n & (n - 1)) == 0; };
    if (pow2(nxi) && pow2(nyi) && fft2d_enabled()) {
        const int lgx = 31 - __builtin_clz(static_cast<unsigned>(nxi));
        const int lgy = 31 - __builtin_clz(static_cast<unsigned>(nyi));
        // rows per workgroup: 256 / ny (at least 1, at most nx): nx is a
        // multiple; merging partial sums, 128 / ny (two threads per point).
        // The inverse rows take at most 2 * 256 bins per workgroup.
        int rows = (partial ? 128 : 256) / nyi;
        rows = rows < 1 ? 1 : (rows > nxi ? nxi : rows);
        int rows_i = 256 / nyi;
        rows_i = rows_i < 1 ? 1 : (rows_i > nxi ? nxi : rows_i);
        const size_t lds_f = (ny + rows * ny) * sizeof(double2);
        const size_t lds_c = 5 * nx * sizeof(double2);
        const size_t lds_i = (ny + 2 * rows_i * ny) * sizeof(double2);
        hipLaunchKernelGGL(p2d_fft_rows_fwd, dim3(nxi / rows), dim3(256), lds_f, s, rho, partial, nparts, dvx, dvy,
                           nxi, nyi, lgy, rows, rho_shift, delta_rho, R);
        VP_CHECK_LAUNCH();
        hipLaunchKernelGGL(p2d_fft_cols, dim3(nky), dim3(256), lds_c, s, R, nxi, nyi, lgx, kx, ky, inv_ksq, T);
        VP_CHECK_LAUNCH();
        hipLaunchKernelGGL(p2d_fft_rows_inv, dim3(nxi / rows_i), dim3(256), lds_i, s, T, nxi, nyi, lgy, rows_i, phi,
                           Ex, Ey);
        VP_CHECK_LAUNCH();
        return VP_OK;
    }
    hipLaunchKernelGGL(p2d_rows_fwd, dim3(nxi), dim3(256), ny * (sizeof(double2) + sizeof(double)), s, rho, partial,
                       nparts, dvx, dvy, nxi, nyi, rho_shift, delta_rho, R);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(p2d_cols, dim3(nky), dim3(256), 5 * nx * sizeof(double2), s, R, nxi, nyi, kx, ky,
                       inv_ksq, T);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(p2d_rows_inv, dim3(3 * nxi), dim3(256), (ny + nky) * sizeof(double2), s, T, nxi, nyi,
                       phi, Ex, Ey);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

int poisson2d(const double* rho, int64_t nx, int64_t ny, const double* kx, const double* ky,
              const double* inv_ksq, double* Ex, double* Ey, double* delta_rho, double* phi, void* ws,
              size_t ws_bytes, hipStream_t s) {
    VP_RETURN_IF(!rho, VP_ERR_ARG);
    return poisson2d_from(rho, nullptr, 0, 1.0, 1.0, nx, ny, kx, ky, inv_ksq, Ex, Ey, delta_rho, phi, ws, ws_bytes,
                          s, nullptr);
}

// out[0] = ((0.5 * sum(a^2 + b^2)) * s1) * s2  (b may be NULL), compensated.
__global__ __launch_bounds__(1024) void energy_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                      int64_t n, double s1, double s2, double* __restrict__ out) {
    __shared__ double scratch[2 * 1024 / kWave];
    KSum k{0.0, 0.0};
    constexpr int kU = 8;  // loads in flight per thread; each thread adds in index order
    const int64_t st = blockDim.x;
    int64_t i = threadIdx.x;
    for (; i + (kU - 1) * st < n; i += kU * st) {
        double av[kU], bv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) av[u] = a[i + u * st];
        if (b) {
#pragma unroll
            for (int u = 0; u < kU; ++u) bv[u] = b[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            double e = av[u] * av[u];
            if (b) e = e + bv[u] * bv[u];
            ksum_add(k, e);
        }
    }
    for (; i < n; i += st) {
        const double av = a[i];
        double e = av * av;
        if (b) {
            const double bv = b[i];
            e = e + bv * bv;
        }
        ksum_add(k, e);
    }
    const double total = block_ksum(k, scratch);
    if (threadIdx.x == 0) out[0] = ((0.5 * total) * s1) * s2;
}

int field_energy(const double* a, const double* b, int64_t n, double s1, double s2, double* out, hipStream_t s) {
    VP_RETURN_IF(!a || !out || n < 1, VP_ERR_ARG);
    hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, s, a, b, n, s1, s2, out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

// Moments of f [rows][len]: mass sum f, kinetic sum f v2 and L2 sum f^2 with
// v2 of column j = vx[j / nvy]^2 + vy[j % nvy]^2 (2D2V; vy null: vx[j]^2).
// Grid-stride compensated sums; partial[block][3] as (sum, compensation).
template <typename T>
__global__ __launch_bounds__(256) void moments_partial_kernel(const T* __restrict__ f, int64_t rows, int64_t len,
                                                              const double* __restrict__ vx,
                                                              const double* __restrict__ vy, int64_t nvy,
                                                              double2* __restrict__ partial) {
    __shared__ double scratch[2 * 256 / kWave];
    KSum m{0.0, 0.0}, k{0.0, 0.0}, l{0.0, 0.0};
    const int64_t n = rows * len;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const double v = static_cast<double>(f[i]);
        const int64_t j = i % len;
        double v2;
        if (vy) {
            const double a = vx[j / nvy], b = vy[j % nvy];
            v2 = a * a + b * b;
        } else {
            const double a = vx[j];
            v2 = a * a;
        }
        ksum_add(m, v);
        ksum_add(k, v * v2);
        ksum_add(l, v * v);
    }
    const double tm = block_ksum(m, scratch);
    const double tk = block_ksum(k, scratch);
    const double tl = block_ksum(l, scratch);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = make_double2(tm, 0.0);
        partial[3 * blockIdx.x + 1] = make_double2(tk, 0.0);
        partial[3 * blockIdx.x + 2] = make_double2(tl, 0.0);
    }
}

// out = (((sum) * s1) * s2, ((0.5 * sum f v2) * s1) * s2, ((sum f^2) * s1) * s2)
__global__ __launch_bounds__(256) void moments_final_kernel(const double2* __restrict__ partial, int nblocks,
                                                            double s1, double s2, double* __restrict__ out) {
    __shared__ double scratch[2 * 256 / kWave];
    double r[3];
    for (int q = 0; q < 3; ++q) {
        KSum k{0.0, 0.0};
        for (int b = threadIdx.x; b < nblocks; b += 256) ksum_add(k, partial[3 * b + q].x);
        r[q] = block_ksum(k, scratch);
    }
    if (threadIdx.x == 0) {
        out[0] = (r[0] * s1) * s2;
        out[1] = ((0.5 * r[1]) * s1) * s2;
        out[2] = (r[2] * s1) * s2;
    }
}

constexpr int kMomentBlocks = 1024;

template <typename T>
int moments(const T* f, int64_t rows, int64_t len, const double* vx, const double* vy, int64_t nvy, double s1,
            double s2, double* out, void* ws, size_t ws_bytes, hipStream_t s) {
    VP_RETURN_IF(!f || !vx || !out || !ws, VP_ERR_ARG);
    VP_RETURN_IF(rows < 1 || len < 1 || (vy && (nvy < 1 || len % nvy != 0)), VP_ERR_SHAPE);
    VP_RETURN_IF(ws_bytes < static_cast<size_t>(kMomentBlocks) * 3 * sizeof(double2), VP_ERR_WORKSPACE);
    const int64_t n = rows * len;
    int blocks = static_cast<int>((n + 255) / 256);
    if (blocks > kMomentBlocks) blocks = kMomentBlocks;
    double2* part = static_cast<double2*>(ws);
    hipLaunchKernelGGL((moments_partial_kernel<T>), dim3(blocks), dim3(256), 0, s, f, rows, len, vx, vy, nvy, part);
    VP_CHECK_LAUNCH();
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, s, part, blocks, s1, s2, out);
    VP_CHECK_LAUNCH();
    return VP_OK;
}

}  // namespace vp

extern "C" int vp_row_sum_f64(const double* f, int64_t rows, int64_t len, double scale1, double scale2,
                              double shift, double* out, vp_stream_t stream) {
    return vp::row_sum<double>(f, rows, len, scale1, scale2, shift, out, static_cast<hipStream_t>(stream));
}

extern "C" int vp_row_sum_f32(const float* f, int64_t rows, int64_t len, double scale1, double scale2,
                              double shift, double* out, vp_stream_t stream) {
    return vp::row_sum<float>(f, rows, len, scale1, scale2, shift, out, static_cast<hipStream_t>(stream));
}

extern "C" size_t vp_poisson_1d_workspace(int64_t nx) { return vp::poisson1d_ws(nx); }

extern "C" int vp_poisson_1d_f64(const double* rho, int64_t nx, const double* inv_kx, const double* inv_kx_sq,
                                 double* E, double* phi, void* ws, size_t ws_bytes, vp_stream_t stream) {
    return vp::poisson1d(rho, nx, inv_kx, inv_kx_sq, E, phi, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" size_t vp_poisson_2d_workspace(int64_t nx, int64_t ny) { return vp::poisson2d_ws(nx, ny); }

extern "C" int vp_poisson_2d_f64(const double* rho, int64_t nx, int64_t ny, const double* kx, const double* ky,
                                 const double* inv_ksq, double* Ex, double* Ey, double* delta_rho, double* phi,
                                 void* ws, size_t ws_bytes, vp_stream_t stream) {
    return vp::poisson2d(rho, nx, ny, kx, ky, inv_ksq, Ex, Ey, delta_rho, phi, ws, ws_bytes,
                         static_cast<hipStream_t>(stream));
}

extern "C" size_t vp_moments_workspace(void) { return static_cast<size_t>(vp::kMomentBlocks) * 3 * sizeof(double2); }

extern "C" int vp_moments_f64(const double* f, int64_t rows, int64_t len, const double* vx, const double* vy,
                              int64_t nvy, double s1, double s2, double* out, void* ws, size_t ws_bytes,
                              vp_stream_t stream) {
    return vp::moments<double>(f, rows, len, vx, vy, nvy, s1, s2, out, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" int vp_moments_f32(const float* f, int64_t rows, int64_t len, const double* vx, const double* vy,
                              int64_t nvy, double s1, double s2, double* out, void* ws, size_t ws_bytes,
                              vp_stream_t stream) {
    return vp::moments<float>(f, rows, len, vx, vy, nvy, s1, s2, out, ws, ws_bytes, static_cast<hipStream_t>(stream));
}
