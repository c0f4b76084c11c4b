// Workgroup-level pieces of the field solve shared by the single-line
// Poisson kernels (field.hip) and the ensemble kernels (ensemble.hip), so
// both evaluate exactly the same float64 operations.
#pragma once

#include "vp_common.hpp"

namespace vp {

// Twiddle table tw[m] = (cos(2 pi m / n), sin(2 pi m / n)).
__device__ __forceinline__ void fill_twiddles(double2* tw, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * static_cast<double>(m) / static_cast<double>(n), &s, &c);
        tw[m] = make_double2(c, s);
    }
}

// solve_poisson_real (vlasov1D1V.py:538-572) by one workgroup, from
// rho = n - 1: rho_k = rfft(rho); E_k = -i rho_k inv_kx; phi_k = rho_k
// inv_kx_sq; E, phi = irfft(., n = nx) (imaginary parts of the DC and
// Nyquist bins dropped).  tw [nx], Ek and Pk [nx/2+1] are scratch (LDS or
// global); phi may be null.  Ends with the outputs written (no barrier).
__device__ __forceinline__ void poisson1d_block(const double* rho, int nx, const double* inv_kx,
                                                const double* inv_kx_sq, double* E, double* phi, double2* tw,
                                                double2* Ek, double2* Pk) {
    fill_twiddles(tw, nx);
    __syncthreads();
    const int nk = nx / 2 + 1;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        double re = 0.0, im = 0.0;
        int ti = 0;  // (k * x) mod nx, stepped
        for (int x = 0; x < nx; ++x) {
            const double2 t = tw[ti];
            re += rho[x] * t.x;
            im -= rho[x] * t.y;
            ti += k;
            ti -= (ti >= nx) ? nx : 0;
        }
        // E_k = -i rho_k inv_kx ; phi_k = rho_k inv_kx^2   (vlasov1D1V.py:556-567)
        Ek[k] = make_double2(im * inv_kx[k], -re * inv_kx[k]);
        Pk[k] = make_double2(re * inv_kx_sq[k], im * inv_kx_sq[k]);
    }
    __syncthreads();
    const int khalf = (nx + 1) / 2;  // bins 1..khalf-1 are doubled
    const double inv_n = 1.0 / static_cast<double>(nx);
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        double se = Ek[0].x, sp = Pk[0].x;
        int ti = x % nx;  // (x * k) mod nx, stepped
        for (int k = 1; k < khalf; ++k) {
            const double2 t = tw[ti];
            se += 2.0 * (Ek[k].x * t.x - Ek[k].y * t.y);
            if (phi) sp += 2.0 * (Pk[k].x * t.x - Pk[k].y * t.y);
            ti += x;
            ti -= (ti >= nx) ? nx : 0;
        }
        if ((nx & 1) == 0) {
            const double sgn = (x & 1) ? -1.0 : 1.0;
            se += Ek[nx / 2].x * sgn;
            sp += Pk[nx / 2].x * sgn;
        }
        E[x] = se * inv_n;
        if (phi) phi[x] = sp * inv_n;
    }
}

}  // namespace vp
