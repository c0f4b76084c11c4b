/*
 * vlasov_amd.h -- C ABI of the MI355X (gfx950) Vlasov-Poisson library.
 *
 * Plain pointers and sizes only.  Every pointer argument named f / vel / E /
 * rho / phi / k* / ws is a DEVICE pointer.  Every entry point is
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream), never
 * synchronises and never allocates device memory: scratch space is passed
 * in as `ws`/`ws_bytes`, sized by the matching *_workspace() query.
 *
 * Return value: 0 on success, a negative VP_ERR_* code for a rejected
 * argument (nothing was launched), or a positive hipError_t from the launch.
 * vp_error_string() turns either into text.
 *
 * Array layouts are C order (last index fastest), as in the reference
 * PyTorch solver:
 *   1D1V  f[Nx][Nv]                  (simulations/vlasov1d_1v/pytorch/src/vlasov1D1V.py)
 *   2D2V  f[Nx][Ny][Nvx][Nvy]        (simulations/vlasov2d_2v/pytorch/src/vlasov2D2V.py)
 * Below, P1 and P2 abbreviate those two files.
 */
#ifndef VLASOV_AMD_H
#define VLASOV_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* vp_stream_t;

enum {
    VP_OK = 0,
    VP_ERR_ARG = -1,        /* null pointer, bad axis / flag                 */
    VP_ERR_SHAPE = -2,      /* size out of the supported range              */
    VP_ERR_WORKSPACE = -3,  /* ws too small (see the *_workspace query)     */
    VP_ERR_UNSUPPORTED = -4, /* shape/variant this library has no kernel for */
    VP_ERR_SCRATCH = -5      /* problem larger than VP_STATIC_SCRATCH_BYTES  */
};

const char* vp_error_string(int code);
int vp_abi_version(void);

/* ---------------------------------------------------------------------
 * Semi-Lagrangian sweep along one axis of a 4D array.
 *
 * Replaces advect_1d (P2:356-413) and advect_1d_vectorized (P1:212-256,
 * P2:287-353) without the permute/reshape copies: for every line along
 * `axis`, out[i] = sum_k w_k(alpha_i) * in[(il_i + k - 2)], with
 * dep_i = i - (vel*dt)/dx, il_i = floor(dep_i), alpha_i = dep_i - il_i and
 * the 6-point Lagrange weights in the reference's power form
 * (lagrange5_interp_sum, P1:163-189).  periodic != 0 wraps the taps
 * (remainder); periodic == 0 zeroes taps outside [0, n) (Dirichlet clamp +
 * mask, P1:248-253).
 *
 * The velocity of element (i0,i1,i2,i3) is vel[sum_d i_d * vel_strides[d]]
 * (stride 0 = broadcast).  vel_strides[axis] != 0 gives a per-point
 * velocity, the [B, N] form accepted by advect_1d_vectorized.
 * A 2D array f[B][N] is shape {1,1,B,N}.
 * f_in and f_out must not overlap.
 * --------------------------------------------------------------------- */
int vp_advect_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                  const double* vel, const int64_t vel_strides[4],
                  double dt, double dx, int periodic, vp_stream_t stream);

/* Kokkos-semantics sweep along `axis` (advect_1d_periodic / _dirichlet,
 * simulations/vlasov2d_2v/kokkos/src/vlasov_poisson.hpp:293-339, with
 * lagrange_interp_5_uniform :244-291), for cross-checking against Kokkos
 * runs: departure point grid[i] - vel*dt in physical coordinates (grid has
 * shape[axis] points; vel indexed as in vp_advect_f64); periodic: wrapped
 * by fmod(dp - xmin, L), index floor(. * inv_dx); Dirichlet: 0 outside
 * [xmin, xmin + L], taps outside [0, n) read as 0; Horner-form weights.
 * Not the parity target of the rest of the library (which follows the
 * PyTorch solvers); f_in and f_out must not overlap. */
int vp_advect_kokkos_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                         const double* grid, const double* vel, const int64_t vel_strides[4],
                         double dt, double L, double xmin, double inv_dx, int periodic,
                         vp_stream_t stream);

/* FP32-storage form: f in float32, departure/weights in float64, the
 * 6-tap sum in float32.  No reference counterpart (the reference's float32
 * mode computes everything in float32). */
int vp_advect_f32(const float* f_in, float* f_out, const int64_t shape[4], int axis,
                  const double* vel, const int64_t vel_strides[4],
                  double dt, double dx, int periodic, vp_stream_t stream);

/* Reflected (shift) form of the sweep, used by the reference's solver 1
 * (advect_1d_matrix_diag, P1:258-337, P2:557-641) and solver 2
 * (advect_v_static, P1:479-501, P2:868-913): d = (vel*dt)/dx,
 * s = floor(d), alpha = d - s, out[i] = sum_k w_k(alpha) in[i - s - o_k]
 * with o_k = k - 2 (taps in that order, wrapped or zero outside).  Agrees
 * with vp_advect_f64 to roundoff (the reference's tests allow 1e-10).
 * window >= 0 (Dirichlet only) keeps just the taps with |s + o_k| <=
 * window + 2, the static window of solver 2 (exact while |d| < window);
 * window < 0 keeps all taps (solver 1). */
int vp_advect_reflected_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                            const double* vel, const int64_t vel_strides[4], double dt, double dx,
                            int periodic, int window, vp_stream_t stream);

/* advect_x_static (P1:408-448, P2:756-819): periodic sweep with the shift
 * and the six weights of every line given: out[i] = sum_k W[k][p]
 * in[(i - shift[p] - o_k) mod n], p = sum_d i_d * param_strides[d] (the
 * line's parameter index), W[k][p] at W[k * w_tap_stride + p]. */
int vp_advect_static_f64(const double* f_in, double* f_out, const int64_t shape[4], int axis,
                         const int64_t* shift, const double* W, int64_t w_tap_stride,
                         const int64_t param_strides[4], vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Velocity moments and field solves.
 * --------------------------------------------------------------------- */

/* Row sums with compensated (Neumaier) float64 accumulation:
 *   out[r] = ((sum_j f[r][j]) * scale1) * scale2 - shift
 * 1D rho (P1:559-561):  rows=Nx, len=Nv, scale1=dv,  scale2=1, shift=1
 * 2D rho (P2:416-431):  rows=Nx*Ny, len=Nvx*Nvy, scale1=dvx, scale2=dvy, shift=0 */
int vp_row_sum_f64(const double* f, int64_t rows, int64_t len, double scale1, double scale2,
                   double shift, double* out, vp_stream_t stream);
int vp_row_sum_f32(const float* f, int64_t rows, int64_t len, double scale1, double scale2,
                   double shift, double* out, vp_stream_t stream);

/* solve_poisson_real (P1:538-572) from rho = n - 1 (already reduced):
 * E = irfft(-i rho_k inv_kx), phi = irfft(rho_k inv_kx_sq).  nx <= 8192. */
size_t vp_poisson_1d_workspace(int64_t nx);
int vp_poisson_1d_f64(const double* rho, int64_t nx, const double* inv_kx,
                      const double* inv_kx_sq, double* E, double* phi,
                      void* ws, size_t ws_bytes, vp_stream_t stream);

/* compute_poisson_real (P2:488-517): delta_rho = rho - 1, phi_k =
 * rfft2(delta_rho) * inv_ksq, E*_k = -i k* phi_k, three irfft2(s=(nx,ny)).
 * kx is [nx] (fftfreq order), ky is [ny/2+1], inv_ksq is [nx][ny/2+1].
 * Any output pointer may be NULL.  nx, ny <= 512. */
size_t vp_poisson_2d_workspace(int64_t nx, int64_t ny);
int vp_poisson_2d_f64(const double* rho, int64_t nx, int64_t ny, const double* kx,
                      const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                      double* delta_rho, double* phi, void* ws, size_t ws_bytes,
                      vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Full Strang steps (one time step, functional: f_in is not modified).
 * --------------------------------------------------------------------- */

/* vlasov_step (P1:574-622): X(dt/2), Poisson, V(dt, E), X(dt/2).
 * E, rho, phi: [nx] outputs (mid-step field); zero when physics_mode == 0. */
size_t vp_step_1d1v_workspace(int64_t nx, int64_t nv);
int vp_step_1d1v_f64(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                     const double* v, const double* inv_kx, const double* inv_kx_sq,
                     double dx, double dv, int physics_mode, double* E, double* rho,
                     double* phi, void* ws, size_t ws_bytes, vp_stream_t stream);

/* FP32-storage form of vp_step_1d1v_f64 (no reference counterpart: the
 * reference's float32 mode, P1:782, computes everything in float32): f in
 * float32, the sweeps as vp_advect_f32 (float64 departure points and
 * weights, float32 6-tap sums), rho summed in float64 from the float32
 * values, Poisson, E, rho and phi in float64.  Workspace:
 * vp_step_1d1v_workspace. */
int vp_step_1d1v_f32(const float* f_in, float* f_out, int64_t nx, int64_t nv, double dt,
                     const double* v, const double* inv_kx, const double* inv_kx_sq,
                     double dx, double dv, int physics_mode, double* E, double* rho,
                     double* phi, void* ws, size_t ws_bytes, vp_stream_t stream);

/* vp_step_1d1v_f64 without the workspace argument and without allocating:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The scratch is a static device array of
 * VP_STATIC_SCRATCH_BYTES in the library, one per GPU, held until the
 * process ends; problems needing more (vp_step_1d1v_workspace(nx, nv) >
 * VP_STATIC_SCRATCH_BYTES: up to nx * nv ~ 1.1 M cells) return
 * VP_ERR_SCRATCH and launch nothing.  The scratch holds only intermediates
 * that each call writes before reading them.  Calls on one GPU share its
 * array: calls that may overlap (on different streams of the same GPU, or
 * from several threads without ordering) race; calls ordered on one stream
 * do not.  The device is the one current at the call; the first call on a
 * device looks the array's address up (make it before a graph capture
 * starts).  Results are those of vp_step_1d1v_f64. */
#define VP_STATIC_SCRATCH_BYTES ((size_t)9 << 20)
int vp_step_1d1v_f64_noalloc(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                             const double* v, const double* inv_kx, const double* inv_kx_sq,
                             double dx, double dv, int physics_mode, double* E, double* rho,
                             double* phi, vp_stream_t stream);

/* vlasov_step_matrix_diag (P1:625-665, solver 1) and
 * vlasov_step_gather_free_impl (P1:668-701, solver 2) in one call:
 * X(dt/2), Poisson, V(dt, E), X(dt/2) with the shift forms.  x_shift and W_x
 * both null: X sweeps as advect_1d_matrix_diag (P1:302-337) from v; both
 * given: X sweeps as advect_x_static (P1:408-448) with shift x_shift[j] and
 * weights W_x[k * w_stride + j] for v column j.  The V sweep is
 * advect_1d_matrix_diag (window < 0) or advect_v_static (P1:479-501, window
 * S >= 0).  Workspace: vp_step_1d1v_workspace. */
int vp_step_1d1v_shift_f64(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                           const double* v, const int64_t* x_shift, const double* W_x, int64_t w_stride,
                           const double* inv_kx, const double* inv_kx_sq, double dx, double dv,
                           int physics_mode, int window, double* E, double* rho, double* phi, void* ws,
                           size_t ws_bytes, vp_stream_t stream);

/* vlasov_step_impl (P2:943-998): X, Y (dt/2), Poisson, Vx, Vy (dt), Y, X (dt/2).
 * Ex, Ey, delta_rho, phi: [nx][ny] outputs; zero when physics_mode == 0.
 * For nx, ny, nvx, nvy in {32, 64, 128} the step runs the fused in-place kernels
 * (X drift, Y drift + rho, Vx+Vy kick per (x,y) plane); other shapes run the
 * generic sweeps.  Both give identical f. */
size_t vp_step_2d2v_workspace(const int64_t shape[4]);
int vp_step_2d2v_f64(const double* f_in, double* f_out, const int64_t shape[4], double dt,
                     const double* vx, const double* vy, const double* kx, const double* ky,
                     const double* inv_ksq, double dx, double dy, double dvx, double dvy,
                     int physics_mode, double* Ex, double* Ey, double* delta_rho,
                     double* phi, void* ws, size_t ws_bytes, vp_stream_t stream);

/* nsteps calls of vlasov_step_impl (the loop of Vlasov.run, P2:1392-1435)
 * in place on f, with the X(dt/2) that closes one step and the X(dt/2) that
 * opens the next fused into one pass; f after the call equals nsteps
 * separate steps.  energy (optional, [nsteps]) receives the field energy
 * 1/2 sum(Ex^2 + Ey^2) dx dy of every step's mid-step field; Ex, Ey,
 * delta_rho, phi hold the last step's field.  Fused shapes only
 * (vp_run_2d2v_workspace returns 0 otherwise). */
size_t vp_run_2d2v_workspace(const int64_t shape[4]);
int vp_run_2d2v_f64(double* f, const int64_t shape[4], int nsteps, double dt, const double* vx,
                    const double* vy, const double* kx, const double* ky, const double* inv_ksq,
                    double dx, double dy, double dvx, double dvy, int physics_mode, double* Ex,
                    double* Ey, double* delta_rho, double* phi, double* energy, void* ws,
                    size_t ws_bytes, vp_stream_t stream);

/* Kernel-level pieces of vp_run_2d2v_f64 / the fused vp_step_2d2v_f64, in
 * the order they run (ws laid out as vp_fused_workspace(shape, 8): weight
 * tables, rho partial sums, Poisson scratch; the caller owns the second
 * f-sized buffer the kick writes); callers that time or interleave the
 * kernels issue them one by one:
 *   tables(dt/2)                                      once per (grid, dt)
 *   xdrift(passes=1)                                  X(dt/2)
 *   per step: ydrift(with_rho=1), poisson, kick(dt) into a second buffer,
 *             ydrift(0) back, xdrift(passes=2 between steps, 1 after the last)
 * xdrift/ydrift may run in place (f_in == f_out).  kick streams the plane
 * rows when f_in and f_out do not overlap; in place it stages each whole
 * (vx, vy) plane in LDS instead (slower). */
size_t vp_fused_workspace(const int64_t shape[4], int elem_bytes); /* 8: f64 pieces, 4: f32; 0 = none */
int vp_fused_tables_f64(const int64_t shape[4], const double* vx, const double* vy, double half_dt,
                        double dx, double dy, void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_xdrift_f64(const double* f_in, double* f_out, const int64_t shape[4], int passes,
                        void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_ydrift_f64(const double* f_in, double* f_out, const int64_t shape[4], int with_rho,
                        void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_poisson_f64(const int64_t shape[4], double dvx, double dvy, const double* kx,
                         const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                         double* delta_rho, double* phi, void* ws, size_t ws_bytes,
                         vp_stream_t stream);
int vp_fused_kick_f64(const double* f_in, double* f_out, const int64_t shape[4], const double* Ex,
                      const double* Ey, double dt, double dvx, double dvy, vp_stream_t stream);

/* ---------------------------------------------------------------------
 * FP32-storage variant of the fused run (no reference counterpart: the
 * reference's float32 mode computes everything in float32).  f is float32
 * in HBM and interpolated in float32 with weights computed in float64 and
 * rounded once (the smallest one balanced so the six sum to 1 within half
 * its ulp); the charge density is summed in float64 from the stored
 * values, and the Poisson solve, fields and field energy stay float64.
 * Same arguments and step order as vp_run_2d2v_f64.  Shapes: nx, ny, nvy in
 * {64, 128}, nvx in {32, 64, 128} (vp_run_2d2v_f32_workspace returns 0
 * otherwise).  The pieces take ws laid out as vp_fused_workspace(shape, 4);
 * the kick must run out of place (f_in and f_out disjoint).
 * --------------------------------------------------------------------- */
size_t vp_run_2d2v_f32_workspace(const int64_t shape[4]);
int vp_run_2d2v_f32(float* f, const int64_t shape[4], int nsteps, double dt, const double* vx,
                    const double* vy, const double* kx, const double* ky, const double* inv_ksq,
                    double dx, double dy, double dvx, double dvy, int physics_mode, double* Ex,
                    double* Ey, double* delta_rho, double* phi, double* energy, void* ws,
                    size_t ws_bytes, vp_stream_t stream);
int vp_fused_tables_f32(const int64_t shape[4], const double* vx, const double* vy, double half_dt,
                        double dx, double dy, void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_xdrift_f32(const float* f_in, float* f_out, const int64_t shape[4], int passes,
                        void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_ydrift_f32(const float* f_in, float* f_out, const int64_t shape[4], int with_rho,
                        void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_poisson_f32(const int64_t shape[4], double dvx, double dvy, const double* kx,
                         const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                         double* delta_rho, double* phi, void* ws, size_t ws_bytes,
                         vp_stream_t stream);
int vp_fused_kick_f32(const float* f_in, float* f_out, const int64_t shape[4], const double* Ex,
                      const double* Ey, double dt, double dvx, double dvy, vp_stream_t stream);

/* ---------------------------------------------------------------------
 * FP32 deviation storage (no reference counterpart): f = feq(vx, vy) +
 * g(x, y, vx, vy) with g in float32 and a separable equilibrium
 * feq(vx, vy) = fx(vx) fy(vy) (the float64 product) from float64 profiles:
 * the rank-one part of the (x, y) mean of the initial f (fx = its sums over
 * vy, fy = its sums over vx over its total; for a mean a(vx) b(vy), as for
 * the reference's initial conditions, fx fy is that mean).  feq is constant
 * in x and y, so the drifts leave it unchanged and run on g (the
 * vp_fused_*drift_f32 kernels); the kick, linear, maps fx fy to
 * (Kvx fx)(Kvy fy), so it stores float(K32(g) + (Kvx fx)(Kvy fy) - fx fy):
 * the float32 kick of g plus a float64 correction from the two profiles'
 * 1D kicks per (x, y) plane; the charge density is sum(g) dv + sum(feq) dv.
 * Storage and arithmetic rounding then scale with the perturbation |g|
 * instead of |f|.  Shapes: those of vp_run_2d2v_f32 with nvy == 128.
 * feq arrays: nvx + nvy float64 values, fx then fy.
 *
 * vp_dev_split_f32: the profiles (compensated sums) into feq,
 *   g = float(f - fx fy), rho_shift[0] = ((sum fx fy) * dvx) * dvy - 1
 *   (device scalar); ws of vp_dev_split_workspace(shape) bytes.
 * vp_dev_join_f32: f = fx fy + g into f64_out and/or (rounded) f32_out.
 * vp_run_2d2v_dev_f32: vp_run_2d2v_f32 on g (same arguments and ws layout).
 * vp_fused_kick_dev_f32 / vp_fused_poisson_dev_f32: the two pieces that
 *   differ from the vp_fused_*_f32 ones (the drifts and tables are shared).
 * --------------------------------------------------------------------- */
size_t vp_dev_split_workspace(const int64_t shape[4]);
int vp_dev_split_f32(const double* f, const int64_t shape[4], double dvx, double dvy, double* feq,
                     double* rho_shift, float* g, void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_dev_join_f32(const float* g, const double* feq, const int64_t shape[4], float* f32_out,
                    double* f64_out, vp_stream_t stream);
int vp_run_2d2v_dev_f32(float* g, const double* feq, const double* rho_shift, const int64_t shape[4],
                        int nsteps, double dt, const double* vx, const double* vy, const double* kx,
                        const double* ky, const double* inv_ksq, double dx, double dy, double dvx,
                        double dvy, int physics_mode, double* Ex, double* Ey, double* delta_rho,
                        double* phi, double* energy, void* ws, size_t ws_bytes, vp_stream_t stream);
int vp_fused_kick_dev_f32(const float* g_in, float* g_out, const int64_t shape[4], const double* feq,
                          const double* Ex, const double* Ey, double dt, double dvx, double dvy,
                          vp_stream_t stream);
int vp_fused_poisson_dev_f32(const int64_t shape[4], double dvx, double dvy, const double* rho_shift,
                             const double* kx, const double* ky, const double* inv_ksq, double* Ex,
                             double* Ey, double* delta_rho, double* phi, void* ws, size_t ws_bytes,
                             vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Diagnostics (no reference counterpart; the reference only writes
 * snapshots).
 * --------------------------------------------------------------------- */

/* out[0] = ((1/2 * sum_i (Ex[i]^2 + Ey[i]^2)) * dx) * dy, compensated;
 * Ey may be NULL (1D: dy = 1). */
int vp_field_energy_f64(const double* Ex, const double* Ey, int64_t n, double dx, double dy,
                        double* out, vp_stream_t stream);

/* Moments of f viewed as [rows][len] (compensated float64 sums):
 *   out[0] = ((sum f) * s1) * s2                       mass
 *   out[1] = ((1/2 sum f v^2) * s1) * s2               kinetic energy
 *   out[2] = ((sum f^2) * s1) * s2                     L2 norm squared
 * 1D1V: rows = Nx, len = Nv, vx = v, vy = NULL, s1 = dx, s2 = dv.
 * 2D2V: rows = Nx*Ny, len = Nvx*Nvy, v^2 = vx[j / nvy]^2 + vy[j % nvy]^2,
 *       s1 = dx*dy*dvx*dvy, s2 = 1.  ws: vp_moments_workspace() bytes. */
size_t vp_moments_workspace(void);
int vp_moments_f64(const double* f, int64_t rows, int64_t len, const double* vx, const double* vy,
                   int64_t nvy, double s1, double s2, double* out, void* ws, size_t ws_bytes,
                   vp_stream_t stream);
int vp_moments_f32(const float* f, int64_t rows, int64_t len, const double* vx, const double* vy,
                   int64_t nvy, double s1, double s2, double* out, void* ws, size_t ws_bytes,
                   vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Ensemble of independent 1D1V instances (no reference counterpart: the
 * reference runs one simulation per process).  Each instance follows
 * vlasov_step (P1:574-622) on its own f[nx][nv] (C order) inside one pool;
 * all instances advance together, three launches per step (X drift + rho
 * partial sums, field solve + field energy, V kick; persistent sweep
 * kernels walking work lists of equal-sized tiles that a set-up launch
 * builds from `inst`), the closing X(dt/2) of a step fused with the opening
 * X(dt/2) of the next.  Sweeps are bit-identical to the single-instance
 * path.
 *
 * inst (device array, ninst entries): per instance the offsets of its f in
 * `pool` and of its tables in `tables` (v[nv] at v_off; inv_kx[nx/2+1] then
 * inv_kx_sq[nx/2+1] at k_off; both as built by Grid.create, P1:94-122), its
 * sizes and dx, dv.  Every nx, nv must be >= 6 and <= max_nx <= 256,
 * max_nv <= 512.  energy (optional, device, [nsteps][ninst]) receives each
 * step's field energy 1/2 sum(E^2) dx.  Instances may not overlap in pool.
 * --------------------------------------------------------------------- */
typedef struct {
    int64_t f_off;
    int64_t v_off;
    int64_t k_off;
    int32_t nx;
    int32_t nv;
    double dx;
    double dv;
} vp_ens_inst;

size_t vp_ens_workspace(int ninst, int max_nx, int max_nv);
int vp_ens_run_f64(double* pool, const vp_ens_inst* inst, int ninst, int max_nx, int max_nv,
                   const double* tables, int nsteps, double dt, int physics_mode, double* energy,
                   void* ws, size_t ws_bytes, vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Surrogate network kernels (no reference counterpart): the Fourier-operator
 * surrogate of the field-energy history is trained and evaluated with these.
 * All tensors are float32 device arrays.
 *
 * vp_nn_gemm: for b < batch, C_b(i,j) = alpha sum_k A_b(i,k) B_b(k,j)
 *   + beta C_b(i,j), i < M, j < N, k < K, with X_b(r,s) = X[b*sXb + r*sXr +
 *   s*sXs] (a transpose is a stride swap).  A and B are rounded to bfloat16
 *   (nearest even) and multiplied on the matrix cores with float32
 *   accumulation.  beta == 0 does not read C.
 * --------------------------------------------------------------------- */
int vp_nn_gemm(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
               int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
               int64_t sCb, int64_t sCi, int64_t sCj, vp_stream_t stream);
/* y = x Phi(x) (exact GELU) and its backward dx = dy (Phi(x) + x phi(x)) */
int vp_nn_gelu(const float* x, float* y, int64_t n, vp_stream_t stream);
int vp_nn_gelu_bwd(const float* x, const float* dy, float* dx, int64_t n, vp_stream_t stream);
/* y[r][c] += bias[c];  db[c] = sum_r dy[r][c] */
int vp_nn_bias_add(float* y, const float* bias, int64_t rows, int cols, vp_stream_t stream);
int vp_nn_bias_grad(const float* dy, float* db, int64_t rows, int cols, vp_stream_t stream);
/* grad = 2 (pred - target) / n (grad may be NULL); partial[ceil(n/256)]
 * receives per-block sums of squared errors */
int vp_nn_mse(const float* pred, const float* target, float* grad, float* partial, int64_t n,
              vp_stream_t stream);
/* weighted squared error: partial[b] = sum over block b (256 elements) of
 * w (pred - target)^2, grad = 2 w (pred - target) (grad may be NULL) */
int vp_nn_wmse(const float* pred, const float* target, const float* w, float* grad, float* partial,
               int64_t n, vp_stream_t stream);
/* Adam with bias correction at step >= 1 (float32 master weights) */
int vp_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
               float eps, int step, vp_stream_t stream);
/* Encoding of the surrogate's input f0 [n][nx][nv] (float64 device arrays)
 * on each run's grid v_j = -vmax[r] + j dv, dv = 2 vmax[r]/(nv - 1):
 * out[r][6] = ((1/nx) sum_i rho_i, (1/nx) sum_i rho_i cos(2 pi i/nx),
 * -(1/nx) sum_i rho_i sin(2 pi i/nx), M0, M2, M4) with rho_i = sum_j f_ij dv
 * and M_p = sum_j fbar_j v_j^p dv, fbar_j = (1/nx) sum_i f_ij (float64,
 * compensated sums). */
int vp_nn_f0_moments(const double* f0, int n, int nx, int nv, const double* vmax, double* out,
                     vp_stream_t stream);

/* ---------------------------------------------------------------------
 * Launch-shape knobs of the fused kernels (no reference counterpart).
 * 0 = automatic.  Process-wide; used by the autotuner and the benchmarks.
 * --------------------------------------------------------------------- */
enum {
    VP_KNOB_X_SPLIT = 0,       /* y-range splits of the X drift grid            */
    VP_KNOB_Y_SPLIT = 1,       /* vx-range splits of the Y drift grid           */
    VP_KNOB_KICK_PREFETCH = 2, /* row-streaming kick: 2 = 16 rows ahead, else 8;
                                  row-split kick: 1 = 4 rows, 3 = 8 rows
                                  (else 4 for double, 8 for float);
                                  in-place plane kick: != 0 prefetches the next
                                  plane (default 2)                            */
    VP_KNOB_KICK_KERNEL = 3,   /* 1: whole-plane LDS kick even out of place;
                                  at 128 vy columns: 2 = two-wave
                                  row-streaming kick, 3/4/5/6 = row-split
                                  kick with 2/4/8/16 waves per plane
                                  (0: 8 waves per plane for float64, 4 for
                                  float32)                                     */
    VP_KNOB_DRIFT_DMA = 4,     /* drifts at 64/128 rows: 1 = register prefetch,
                                  2 or 3 = direct-to-LDS ring of that many
                                  tiles (0 = 2)                                */
    VP_KNOB_NONTEMPORAL = 5,   /* sweeps whose f loads/stores carry the
                                  streaming (non-temporal) hint, a bit mask:
                                  1 X drift, 2 Y drift, 4 kick; 8 = none
                                  (0 = 7)                                      */
    VP_KNOB_ENS_XTILE = 6,     /* ensemble X drift: elements per work item,
                                  2048 or 4096 (0 = 4096)                      */
    VP_KNOB_ENS_KTILE = 7,     /* ensemble kick: elements per work item, 2048
                                  or 4096 (0 = 4096)                           */
    VP_KNOB_ENS_WGS = 8,       /* ensemble sweeps: persistent workgroups per
                                  CU, 1 or 2 (0 = 2)                           */
    VP_KNOB_SEG_ROWS = 9,      /* generic sweeps of strided lines longer than
                                  256: 16..128 rows per workgroup segment
                                  (0 = 112, or 64 / 32 when that gives under
                                  512 workgroups)                              */
    VP_KNOB_SEG_CONTIG = 10,   /* generic sweeps of contiguous lines longer
                                  than 256: 256..4096 indices per workgroup
                                  segment (0 = 2048 for lines of 2048 or
                                  more, else 1024)                             */
    VP_KNOB_F32_PAIRS = 11,    /* float32 drifts: pairs of vy columns moved
                                  as 8-byte elements, 1 = never, 2 = always
                                  (0 = the Y drifts and the two-pass X drift)  */
    VP_KNOB_POISSON_DFT = 12,  /* 1: 2D Poisson by the exact-twiddle DFT kernels
                                  also for power-of-two grids (default: radix-2
                                  FFTs there); results agree to rounding      */
    VP_KNOB_Y_UNITS = 13,      /* Y drift over 128-row lines: 1 = one workgroup
                                  per (x, vx range); 2 = persistent workgroups
                                  (one per CU) walking ranges of them with
                                  rotated starts; 3 = 2 with the vy chunks
                                  spread over the XCDs; 4 = 3 unrotated
                                  (0: 2 for float32 storage, else 1)          */
    VP_KNOB_COUNT = 14
};
int vp_tune_set(int knob, int value);
int vp_tune_get(int knob);

/* ---------------------------------------------------------------------
 * HBM calibration (STREAM-style; the reference ships the same kind of
 * benchmark, simulations/stream/pytorch/src/stream.py:92-105).
 * kind 0: copy a -> b; 1: read a (b receives `grid` partial sums);
 * 2: write b.  n doubles (even), 16-byte aligned pointers, unroll in
 * {1,2,4,8}, grid = number of 256-thread workgroups.
 * --------------------------------------------------------------------- */
int vp_stream_f64(int kind, const double* a, double* b, int64_t n, int unroll, int nontemporal,
                  int grid, vp_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* VLASOV_AMD_H */
