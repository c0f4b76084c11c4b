// Full Strang steps composed from the sweep and field kernels, plus the
// small C-ABI utilities (error strings, version).
#include <mutex>

#include "fast2d.hpp"

namespace vp {
int fill(double* p, int64_t n, double v, hipStream_t s);
template <typename T>
int row_sum(const T* f, int64_t rows, int64_t len, double s1, double s2, double shift, double* out,
            hipStream_t s);
size_t poisson1d_ws(int64_t nx);
int poisson1d(const double* rho, int64_t nx, const double* inv_kx, const double* inv_kx_sq, double* E,
              double* phi, void* ws, size_t ws_bytes, hipStream_t s);
size_t poisson2d_ws(int64_t nx, int64_t ny);
int poisson2d(const double* rho, int64_t nx, int64_t ny, const double* kx, const double* ky,
              const double* inv_ksq, double* Ex, double* Ey, double* delta_rho, double* phi, void* ws,
              size_t ws_bytes, hipStream_t s);

int poisson2d_from(const double* rho, const double2* partial, int nparts, double dvx, double dvy, int64_t nx,
                   int64_t ny, const double* kx, const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                   double* delta_rho, double* phi, void* ws, size_t ws_bytes, hipStream_t s,
                   const double* rho_shift = nullptr);
int field_energy(const double* a, const double* b, int64_t n, double s1, double s2, double* out, hipStream_t s);

static inline size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }


}  // namespace vp

extern "C" const char* vp_error_string(int code) {
    switch (code) {
        case VP_OK: return "ok";
        case VP_ERR_ARG: return "invalid argument (null pointer, axis or flag)";
        case VP_ERR_SHAPE: return "array size outside the supported range";
        case VP_ERR_WORKSPACE: return "workspace too small";
        case VP_ERR_UNSUPPORTED: return "shape or variant not supported by this library";
        case VP_ERR_SCRATCH:
            return "problem needs more scratch than the allocation-free entry point holds "
                   "(VP_STATIC_SCRATCH_BYTES = 9 MiB: vp_step_1d1v_workspace(nx, nv) must fit)";
        default: break;
    }
    if (code > 0) return hipGetErrorString(static_cast<hipError_t>(code));
    return "unknown error";
}

extern "C" int vp_abi_version(void) { return 1; }

// ---------------------------------------------------------------------------
// 1D1V: vlasov_step (vlasov1D1V.py:574-622)
// ws: [nx*nv] f scratch | [nx] rho scratch | poisson ws
// ---------------------------------------------------------------------------
extern "C" size_t vp_step_1d1v_workspace(int64_t nx, int64_t nv) {
    return vp::align256(static_cast<size_t>(nx) * nv * sizeof(double)) +
           vp::align256(static_cast<size_t>(nx) * sizeof(double)) + vp::poisson1d_ws(nx);
}

namespace {
int advect_s(const double* a, double* b, const int64_t shape[4], int axis, const double* vel, const int64_t vs[4],
             double dt, double dx, int periodic, vp_stream_t stream) {
    return vp_advect_f64(a, b, shape, axis, vel, vs, dt, dx, periodic, stream);
}
int advect_s(const float* a, float* b, const int64_t shape[4], int axis, const double* vel, const int64_t vs[4],
             double dt, double dx, int periodic, vp_stream_t stream) {
    return vp_advect_f32(a, b, shape, axis, vel, vs, dt, dx, periodic, stream);
}

// T = double: the reference step; T = float: FP32 storage (float64 rho, fields).
template <typename T>
int step_1d1v(const T* f_in, T* f_out, int64_t nx, int64_t nv, double dt, const double* v, const double* inv_kx,
              const double* inv_kx_sq, double dx, double dv, int physics_mode, double* E, double* rho, double* phi,
              void* ws, size_t ws_bytes, vp_stream_t stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    VP_RETURN_IF(!f_in || !f_out || !v || !E || !rho || !phi || !ws, VP_ERR_ARG);
    VP_RETURN_IF(physics_mode && (!inv_kx || !inv_kx_sq), VP_ERR_ARG);
    VP_RETURN_IF(ws_bytes < vp_step_1d1v_workspace(nx, nv), VP_ERR_WORKSPACE);
    unsigned char* wb = static_cast<unsigned char*>(ws);
    T* tmp = reinterpret_cast<T*>(wb);
    wb += vp::align256(static_cast<size_t>(nx) * nv * sizeof(double));
    wb += vp::align256(static_cast<size_t>(nx) * sizeof(double));
    void* pws = wb;
    const int64_t shape[4] = {1, 1, nx, nv};
    const int64_t vs_x[4] = {0, 0, 0, 1};  // drift along x with velocity v[j]
    const int64_t vs_v[4] = {0, 0, 1, 0};  // kick along v with field E[i]
    const double half = dt / 2.0;
    VP_TRY(advect_s(f_in, f_out, shape, 2, v, vs_x, half, dx, 1, stream));
    if (physics_mode) {
        VP_TRY(vp::row_sum<T>(f_out, nx, nv, dv, 1.0, 1.0, rho, s));
        VP_TRY(vp::poisson1d(rho, nx, inv_kx, inv_kx_sq, E, phi, pws, vp::poisson1d_ws(nx), s));
    } else {
        VP_TRY(vp::fill(E, nx, 0.0, s));
        VP_TRY(vp::fill(rho, nx, 0.0, s));
        VP_TRY(vp::fill(phi, nx, 0.0, s));
    }
    VP_TRY(advect_s(f_out, tmp, shape, 3, E, vs_v, dt, dv, 0, stream));
    VP_TRY(advect_s(tmp, f_out, shape, 2, v, vs_x, half, dx, 1, stream));
    return VP_OK;
}
}  // namespace

extern "C" int vp_step_1d1v_f64(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                                const double* v, const double* inv_kx, const double* inv_kx_sq, double dx,
                                double dv, int physics_mode, double* E, double* rho, double* phi, void* ws,
                                size_t ws_bytes, vp_stream_t stream) {
    return step_1d1v<double>(f_in, f_out, nx, nv, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode, E, rho, phi, ws,
                             ws_bytes, stream);
}

extern "C" int vp_step_1d1v_f32(const float* f_in, float* f_out, int64_t nx, int64_t nv, double dt,
                                const double* v, const double* inv_kx, const double* inv_kx_sq, double dx,
                                double dv, int physics_mode, double* E, double* rho, double* phi, void* ws,
                                size_t ws_bytes, vp_stream_t stream) {
    return step_1d1v<float>(f_in, f_out, nx, nv, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode, E, rho, phi, ws,
                            ws_bytes, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free 1D1V step: the workspace of vp_step_1d1v_f64 lives in a
// static device array (one per GPU, for the life of the process).  In host
// code `scratch` names a host-side stand-in, never the device memory: its
// device address is looked up once per device with hipGetSymbolAddress and
// only that address is passed on.
// ---------------------------------------------------------------------------
namespace {
__device__ __align__(256) unsigned char scratch[VP_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with it current).
int static_scratch(void** out) {
    int dev = -1;
    VP_TRY_HIP(hipGetDevice(&dev));
    VP_RETURN_IF(dev < 0 || dev >= kMaxDevices, VP_ERR_ARG);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        VP_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VP_RETURN_IF(p == nullptr, VP_ERR_ARG);
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return VP_OK;
}
}  // namespace

extern "C" int vp_step_1d1v_f64_noalloc(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                                        const double* v, const double* inv_kx, const double* inv_kx_sq, double dx,
                                        double dv, int physics_mode, double* E, double* rho, double* phi,
                                        vp_stream_t stream) {
    VP_RETURN_IF(nx < 1 || nv < 1 || nx > (1 << 24) || nv > (1 << 24), VP_ERR_SHAPE);
    VP_RETURN_IF(vp_step_1d1v_workspace(nx, nv) > VP_STATIC_SCRATCH_BYTES, VP_ERR_SCRATCH);
    void* ws = nullptr;
    VP_TRY(static_scratch(&ws));
    // every scratch region (the f copy between the kick and the closing
    // drift, the Poisson intermediates) is written before it is read within
    // the call, so nothing has to be cleared between calls
    return vp_step_1d1v_f64(f_in, f_out, nx, nv, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode, E, rho, phi, ws,
                            VP_STATIC_SCRATCH_BYTES, stream);
}

// Solver-1 and solver-2 steps in one call: the X sweeps are the reflected
// periodic form from v at dt/2 (x_shift == W_x == null, solver 1) or the
// static form with the given per-column shifts and weights W_x[k * w_stride
// + j] (solver 2); the V sweep is the reflected Dirichlet form from E at dt
// with the static window `window` (< 0: every tap, solver 1).  Same
// workspace as vp_step_1d1v_f64.
extern "C" int vp_step_1d1v_shift_f64(const double* f_in, double* f_out, int64_t nx, int64_t nv, double dt,
                                      const double* v, const int64_t* x_shift, const double* W_x, int64_t w_stride,
                                      const double* inv_kx, const double* inv_kx_sq, double dx, double dv,
                                      int physics_mode, int window, double* E, double* rho, double* phi, void* ws,
                                      size_t ws_bytes, vp_stream_t stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    VP_RETURN_IF(!f_in || !f_out || !E || !rho || !phi || !ws, VP_ERR_ARG);
    VP_RETURN_IF((x_shift == nullptr) != (W_x == nullptr), VP_ERR_ARG);
    VP_RETURN_IF(!x_shift && !v, VP_ERR_ARG);
    VP_RETURN_IF(x_shift && w_stride < nv, VP_ERR_ARG);
    VP_RETURN_IF(physics_mode && (!inv_kx || !inv_kx_sq), VP_ERR_ARG);
    VP_RETURN_IF(ws_bytes < vp_step_1d1v_workspace(nx, nv), VP_ERR_WORKSPACE);
    unsigned char* wb = static_cast<unsigned char*>(ws);
    double* tmp = reinterpret_cast<double*>(wb);
    wb += vp::align256(static_cast<size_t>(nx) * nv * sizeof(double));
    wb += vp::align256(static_cast<size_t>(nx) * sizeof(double));
    void* pws = wb;
    const int64_t shape[4] = {1, 1, nx, nv};
    const int64_t vs_x[4] = {0, 0, 0, 1};  // X sweep: parameters per v column
    const int64_t vs_v[4] = {0, 0, 1, 0};  // V sweep: field per x row
    const double half = dt / 2.0;
    auto xsweep = [&](const double* a, double* b) -> int {
        if (x_shift) return vp_advect_static_f64(a, b, shape, 2, x_shift, W_x, w_stride, vs_x, stream);
        return vp_advect_reflected_f64(a, b, shape, 2, v, vs_x, half, dx, 1, -1, stream);
    };
    VP_TRY(xsweep(f_in, f_out));
    if (physics_mode) {
        VP_TRY(vp::row_sum<double>(f_out, nx, nv, dv, 1.0, 1.0, rho, s));
        VP_TRY(vp::poisson1d(rho, nx, inv_kx, inv_kx_sq, E, phi, pws, vp::poisson1d_ws(nx), s));
    } else {
        VP_TRY(vp::fill(E, nx, 0.0, s));
        VP_TRY(vp::fill(rho, nx, 0.0, s));
        VP_TRY(vp::fill(phi, nx, 0.0, s));
    }
    VP_TRY(vp_advect_reflected_f64(f_out, tmp, shape, 3, E, vs_v, dt, dv, 0, window, stream));
    VP_TRY(xsweep(tmp, f_out));
    return VP_OK;
}

// ---------------------------------------------------------------------------
// 2D2V: vlasov_step_impl (vlasov2D2V.py:943-998)
// generic ws: [N^4] f scratch | [nx*ny] rho | poisson ws
// fused ws:   weight tables | rho partials | f-sized kick buffer | poisson ws
// ---------------------------------------------------------------------------
static size_t generic_2d_ws(const int64_t shape[4]) {
    const size_t cells = static_cast<size_t>(shape[0]) * shape[1] * shape[2] * shape[3];
    return vp::align256(cells * sizeof(double)) +
           vp::align256(static_cast<size_t>(shape[0]) * shape[1] * sizeof(double)) +
           vp::poisson2d_ws(shape[0], shape[1]);
}

// with_ftmp: the run / full-step layout (adds the kick's f-sized buffer);
// without: the kernel-level pieces (the caller owns the second f buffer).
static size_t fused_2d_ws(const int64_t shape[4], int elem_bytes, bool with_ftmp) {
    return vp::fast2d_layout(shape, nullptr, elem_bytes, with_ftmp).bytes;
}

extern "C" size_t vp_step_2d2v_workspace(const int64_t shape[4]) {
    if (!shape) return 0;
    return vp::fast2d_supported(shape) ? fused_2d_ws(shape, 8, true) : generic_2d_ws(shape);
}

extern "C" size_t vp_run_2d2v_workspace(const int64_t shape[4]) {
    if (!shape || !vp::fast2d_supported(shape)) return 0;
    return fused_2d_ws(shape, 8, true);
}

extern "C" size_t vp_run_2d2v_f32_workspace(const int64_t shape[4]) {
    if (!shape || !vp::fast2d_supported_f32(shape)) return 0;
    return fused_2d_ws(shape, 4, true);
}

extern "C" size_t vp_fused_workspace(const int64_t shape[4], int elem_bytes) {
    if (!shape) return 0;
    if (elem_bytes == 8) return vp::fast2d_supported(shape) ? fused_2d_ws(shape, 8, false) : 0;
    if (elem_bytes == 4) return vp::fast2d_supported_f32(shape) ? fused_2d_ws(shape, 4, false) : 0;
    return 0;
}

extern "C" int vp_field_energy_f64(const double* Ex, const double* Ey, int64_t n, double dx, double dy,
                                   double* out, vp_stream_t stream) {
    return vp::field_energy(Ex, Ey, n, dx, dy, out, static_cast<hipStream_t>(stream));
}

// Fused step: X (f_in -> f_out), then Y+rho, Poisson, kick, Y, X in place on f_out.
static int fused_step_2d(const double* f_in, double* f_out, const int64_t shape[4], double dt, const double* vx,
                         const double* vy, const double* kx, const double* ky, const double* inv_ksq, double dx,
                         double dy, double dvx, double dvy, int physics_mode, double* Ex, double* Ey,
                         double* delta_rho, double* phi, void* ws, hipStream_t s) {
    const int64_t nx = shape[0], ny = shape[1];
    vp::Fast2DWs L = vp::fast2d_layout(shape, ws, 8, true);
    void* pws = L.pws;
    const double half = dt / 2.0;
    VP_TRY(vp::fast2d_tables(shape, L, vx, vy, half, dx, dy, s));
    VP_TRY(vp::fast2d_xdrift(f_in, f_out, shape, L, 1, s));
    VP_TRY(vp::fast2d_ydrift(f_out, f_out, shape, L, physics_mode != 0, s));
    if (physics_mode) {
        VP_TRY(vp::poisson2d_from(nullptr, L.partial, L.nparts, dvx, dvy, nx, ny, kx, ky, inv_ksq, Ex, Ey,
                                  delta_rho, phi, pws, vp::poisson2d_ws(nx, ny), s));
    } else {
        VP_TRY(vp::fill(Ex, nx * ny, 0.0, s));
        VP_TRY(vp::fill(Ey, nx * ny, 0.0, s));
        VP_TRY(vp::fill(delta_rho, nx * ny, 0.0, s));
        VP_TRY(vp::fill(phi, nx * ny, 0.0, s));
    }
    double* tmp = static_cast<double*>(L.ftmp);
    VP_TRY(vp::fast2d_kick(f_out, tmp, shape, Ex, Ey, dt, dvx, dvy, s));
    VP_TRY(vp::fast2d_ydrift(tmp, f_out, shape, L, false, s));
    VP_TRY(vp::fast2d_xdrift(f_out, f_out, shape, L, 1, s));
    return VP_OK;
}

// Kernel-level pieces of the fused step, in the order vp_run_2d2v_f64 issues
// them; ws is laid out as for vp_fused_workspace (a prefix of the run layout).
#define VP_FUSED_GUARD(ws, ws_bytes)                                   \
    VP_RETURN_IF(!shape || !ws, VP_ERR_ARG);                           \
    VP_RETURN_IF(!vp::fast2d_supported(shape), VP_ERR_UNSUPPORTED);    \
    VP_RETURN_IF(ws_bytes < fused_2d_ws(shape, 8, false), VP_ERR_WORKSPACE)
#define VP_FUSED_GUARD_F32(ws, ws_bytes)                                \
    VP_RETURN_IF(!shape || !ws, VP_ERR_ARG);                            \
    VP_RETURN_IF(!vp::fast2d_supported_f32(shape), VP_ERR_UNSUPPORTED); \
    VP_RETURN_IF(ws_bytes < fused_2d_ws(shape, 4, false), VP_ERR_WORKSPACE)

extern "C" int vp_fused_tables_f64(const int64_t shape[4], const double* vx, const double* vy, double half_dt,
                                   double dx, double dy, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD(ws, ws_bytes);
    VP_RETURN_IF(!vx || !vy, VP_ERR_ARG);
    return vp::fast2d_tables(shape, vp::fast2d_layout(shape, ws, 8, false), vx, vy, half_dt, dx, dy,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_xdrift_f64(const double* f_in, double* f_out, const int64_t shape[4], int passes, void* ws,
                                   size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD(ws, ws_bytes);
    VP_RETURN_IF(!f_in || !f_out || (passes != 1 && passes != 2), VP_ERR_ARG);
    return vp::fast2d_xdrift(f_in, f_out, shape, vp::fast2d_layout(shape, ws, 8, false), passes,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_ydrift_f64(const double* f_in, double* f_out, const int64_t shape[4], int with_rho,
                                   void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD(ws, ws_bytes);
    VP_RETURN_IF(!f_in || !f_out, VP_ERR_ARG);
    return vp::fast2d_ydrift(f_in, f_out, shape, vp::fast2d_layout(shape, ws, 8, false), with_rho != 0,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_poisson_f64(const int64_t shape[4], double dvx, double dvy, const double* kx,
                                    const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                                    double* delta_rho, double* phi, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD(ws, ws_bytes);
    vp::Fast2DWs L = vp::fast2d_layout(shape, ws, 8, false);
    return vp::poisson2d_from(nullptr, L.partial, L.nparts, dvx, dvy, shape[0], shape[1], kx, ky, inv_ksq, Ex, Ey,
                              delta_rho, phi, L.pws, L.pws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_kick_f64(const double* f_in, double* f_out, const int64_t shape[4], const double* Ex,
                                 const double* Ey, double dt, double dvx, double dvy, vp_stream_t stream) {
    VP_RETURN_IF(!shape || !f_in || !f_out || !Ex || !Ey, VP_ERR_ARG);
    VP_RETURN_IF(shape[0] < 1 || shape[1] < 1 || shape[0] * shape[1] > 0x7fffffffLL, VP_ERR_SHAPE);
    VP_RETURN_IF(!vp::fast2d_kick_supported(shape), VP_ERR_UNSUPPORTED);
    return vp::fast2d_kick(f_in, f_out, shape, Ex, Ey, dt, dvx, dvy, static_cast<hipStream_t>(stream));
}

// nsteps Strang steps in place on f.  Between two steps the closing X(dt/2)
// and the opening X(dt/2) run in one pass (applied twice in LDS): the
// result is identical to nsteps separate steps.
// feq / rho_shift (float only): deviation storage, f holds g = f - feq.
template <typename S>
static int run_2d2v(S* f, const int64_t shape[4], int nsteps, double dt, const double* vx, const double* vy,
                    const double* kx, const double* ky, const double* inv_ksq, double dx, double dy, double dvx,
                    double dvy, int physics_mode, double* Ex, double* Ey, double* delta_rho, double* phi,
                    double* energy, void* ws, hipStream_t s, const double* feq = nullptr,
                    const double* rho_shift = nullptr) {
    const int64_t nx = shape[0], ny = shape[1];
    vp::Fast2DWs L = vp::fast2d_layout(shape, ws, static_cast<int>(sizeof(S)), true);
    void* pws = L.pws;
    S* tmp = static_cast<S*>(L.ftmp);
    const double half = dt / 2.0;
    VP_TRY(vp::fast2d_tables(shape, L, vx, vy, half, dx, dy, s));
    if (!physics_mode) {
        VP_TRY(vp::fill(Ex, nx * ny, 0.0, s));
        VP_TRY(vp::fill(Ey, nx * ny, 0.0, s));
        VP_TRY(vp::fill(delta_rho, nx * ny, 0.0, s));
        VP_TRY(vp::fill(phi, nx * ny, 0.0, s));
    }
    VP_TRY(vp::fast2d_xdrift(f, f, shape, L, 1, s));
    for (int st = 0; st < nsteps; ++st) {
        VP_TRY(vp::fast2d_ydrift(f, f, shape, L, physics_mode != 0, s));
        if (physics_mode)
            VP_TRY(vp::poisson2d_from(nullptr, L.partial, L.nparts, dvx, dvy, nx, ny, kx, ky, inv_ksq, Ex, Ey,
                                      delta_rho, phi, pws, vp::poisson2d_ws(nx, ny), s, rho_shift));
        if (energy) VP_TRY(vp::field_energy(Ex, Ey, nx * ny, dx, dy, energy + st, s));
        if constexpr (sizeof(S) == 4) {
            if (feq)
                VP_TRY(vp::fast2d_kick_dev(f, tmp, shape, feq, Ex, Ey, dt, dvx, dvy, s));
            else
                VP_TRY(vp::fast2d_kick(f, tmp, shape, Ex, Ey, dt, dvx, dvy, s));
        } else {
            VP_TRY(vp::fast2d_kick(f, tmp, shape, Ex, Ey, dt, dvx, dvy, s));
        }
        VP_TRY(vp::fast2d_ydrift(tmp, f, shape, L, false, s));
        VP_TRY(vp::fast2d_xdrift(f, f, shape, L, st == nsteps - 1 ? 1 : 2, s));
    }
    return VP_OK;
}

#define VP_RUN_CHECKS(SUPPORTED, WS)                                                                           \
    VP_RETURN_IF(!f || !shape || !vx || !vy || !Ex || !Ey || !delta_rho || !phi || !ws, VP_ERR_ARG);          \
    VP_RETURN_IF(physics_mode && (!kx || !ky || !inv_ksq), VP_ERR_ARG);                                       \
    VP_RETURN_IF(nsteps < 0, VP_ERR_ARG);                                                                    \
    VP_RETURN_IF(!(SUPPORTED), VP_ERR_UNSUPPORTED);                                                           \
    VP_RETURN_IF(ws_bytes < (WS), VP_ERR_WORKSPACE);                                                         \
    if (nsteps == 0) return VP_OK

extern "C" int vp_run_2d2v_f64(double* f, const int64_t shape[4], int nsteps, double dt, const double* vx,
                               const double* vy, const double* kx, const double* ky, const double* inv_ksq,
                               double dx, double dy, double dvx, double dvy, int physics_mode, double* Ex,
                               double* Ey, double* delta_rho, double* phi, double* energy, void* ws,
                               size_t ws_bytes, vp_stream_t stream) {
    VP_RUN_CHECKS(vp::fast2d_supported(shape), fused_2d_ws(shape, 8, true));
    return run_2d2v<double>(f, shape, nsteps, dt, vx, vy, kx, ky, inv_ksq, dx, dy, dvx, dvy, physics_mode, Ex, Ey,
                            delta_rho, phi, energy, ws, static_cast<hipStream_t>(stream));
}

extern "C" int vp_run_2d2v_f32(float* f, const int64_t shape[4], int nsteps, double dt, const double* vx,
                               const double* vy, const double* kx, const double* ky, const double* inv_ksq,
                               double dx, double dy, double dvx, double dvy, int physics_mode, double* Ex,
                               double* Ey, double* delta_rho, double* phi, double* energy, void* ws,
                               size_t ws_bytes, vp_stream_t stream) {
    VP_RUN_CHECKS(vp::fast2d_supported_f32(shape), fused_2d_ws(shape, 4, true));
    return run_2d2v<float>(f, shape, nsteps, dt, vx, vy, kx, ky, inv_ksq, dx, dy, dvx, dvy, physics_mode, Ex, Ey,
                           delta_rho, phi, energy, ws, static_cast<hipStream_t>(stream));
}

// FP32 deviation storage: g (float32) = f - feq, feq = fx | fy (float64
// profiles of the separable feq = fx(vx) fy(vy)), rho_shift[0] = sum feq dvx
// dvy - 1 (all from vp_dev_split_f32).
static bool dev_supported(const int64_t shape[4]) {
    return shape && vp::fast2d_supported_f32(shape) && shape[3] == 128;
}

extern "C" int vp_run_2d2v_dev_f32(float* g, const double* feq, const double* rho_shift, const int64_t shape[4],
                                   int nsteps, double dt, const double* vx, const double* vy, const double* kx,
                                   const double* ky, const double* inv_ksq, double dx, double dy, double dvx,
                                   double dvy, int physics_mode, double* Ex, double* Ey, double* delta_rho,
                                   double* phi, double* energy, void* ws, size_t ws_bytes, vp_stream_t stream) {
    float* f = g;
    VP_RETURN_IF(!feq || !rho_shift, VP_ERR_ARG);
    VP_RUN_CHECKS(dev_supported(shape), fused_2d_ws(shape, 4, true));
    return run_2d2v<float>(f, shape, nsteps, dt, vx, vy, kx, ky, inv_ksq, dx, dy, dvx, dvy, physics_mode, Ex, Ey,
                           delta_rho, phi, energy, ws, static_cast<hipStream_t>(stream), feq, rho_shift);
}

extern "C" int vp_fused_poisson_dev_f32(const int64_t shape[4], double dvx, double dvy, const double* rho_shift,
                                        const double* kx, const double* ky, const double* inv_ksq, double* Ex,
                                        double* Ey, double* delta_rho, double* phi, void* ws, size_t ws_bytes,
                                        vp_stream_t stream) {
    VP_FUSED_GUARD_F32(ws, ws_bytes);
    VP_RETURN_IF(!rho_shift, VP_ERR_ARG);
    vp::Fast2DWs L = vp::fast2d_layout(shape, ws, 4, false);
    return vp::poisson2d_from(nullptr, L.partial, L.nparts, dvx, dvy, shape[0], shape[1], kx, ky, inv_ksq, Ex, Ey,
                              delta_rho, phi, L.pws, L.pws_bytes, static_cast<hipStream_t>(stream), rho_shift);
}

extern "C" int vp_fused_kick_dev_f32(const float* g_in, float* g_out, const int64_t shape[4], const double* feq,
                                     const double* Ex, const double* Ey, double dt, double dvx, double dvy,
                                     vp_stream_t stream) {
    VP_RETURN_IF(!shape || !g_in || !g_out || !feq || !Ex || !Ey, VP_ERR_ARG);
    VP_RETURN_IF(shape[0] < 1 || shape[1] < 1 || shape[0] * shape[1] > 0x7fffffffLL, VP_ERR_SHAPE);
    VP_RETURN_IF(!vp::fast2d_kick_supported(shape) || shape[3] != 128, VP_ERR_UNSUPPORTED);
    return vp::fast2d_kick_dev(g_in, g_out, shape, feq, Ex, Ey, dt, dvx, dvy, static_cast<hipStream_t>(stream));
}
#undef VP_RUN_CHECKS

// Pieces of vp_run_2d2v_f32; ws laid out as vp_fused_workspace(shape, 4).
extern "C" int vp_fused_tables_f32(const int64_t shape[4], const double* vx, const double* vy, double half_dt,
                                   double dx, double dy, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD_F32(ws, ws_bytes);
    VP_RETURN_IF(!vx || !vy, VP_ERR_ARG);
    return vp::fast2d_tables(shape, vp::fast2d_layout(shape, ws, 4, false), vx, vy, half_dt, dx, dy,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_xdrift_f32(const float* f_in, float* f_out, const int64_t shape[4], int passes, void* ws,
                                   size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD_F32(ws, ws_bytes);
    VP_RETURN_IF(!f_in || !f_out || (passes != 1 && passes != 2), VP_ERR_ARG);
    return vp::fast2d_xdrift(f_in, f_out, shape, vp::fast2d_layout(shape, ws, 4, false), passes,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_ydrift_f32(const float* f_in, float* f_out, const int64_t shape[4], int with_rho,
                                   void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD_F32(ws, ws_bytes);
    VP_RETURN_IF(!f_in || !f_out, VP_ERR_ARG);
    return vp::fast2d_ydrift(f_in, f_out, shape, vp::fast2d_layout(shape, ws, 4, false), with_rho != 0,
                             static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_poisson_f32(const int64_t shape[4], double dvx, double dvy, const double* kx,
                                    const double* ky, const double* inv_ksq, double* Ex, double* Ey,
                                    double* delta_rho, double* phi, void* ws, size_t ws_bytes, vp_stream_t stream) {
    VP_FUSED_GUARD_F32(ws, ws_bytes);
    vp::Fast2DWs L = vp::fast2d_layout(shape, ws, 4, false);
    return vp::poisson2d_from(nullptr, L.partial, L.nparts, dvx, dvy, shape[0], shape[1], kx, ky, inv_ksq, Ex, Ey,
                              delta_rho, phi, L.pws, L.pws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" int vp_fused_kick_f32(const float* f_in, float* f_out, const int64_t shape[4], const double* Ex,
                                 const double* Ey, double dt, double dvx, double dvy, vp_stream_t stream) {
    VP_RETURN_IF(!shape || !f_in || !f_out || !Ex || !Ey, VP_ERR_ARG);
    VP_RETURN_IF(shape[0] < 1 || shape[1] < 1 || shape[0] * shape[1] > 0x7fffffffLL, VP_ERR_SHAPE);
    VP_RETURN_IF(!vp::fast2d_kick_supported(shape), VP_ERR_UNSUPPORTED);
    return vp::fast2d_kick(f_in, f_out, shape, Ex, Ey, dt, dvx, dvy, static_cast<hipStream_t>(stream));
}

extern "C" int vp_step_2d2v_f64(const double* f_in, double* f_out, const int64_t shape[4], double dt,
                                const double* vx, const double* vy, const double* kx, const double* ky,
                                const double* inv_ksq, double dx, double dy, double dvx, double dvy,
                                int physics_mode, double* Ex, double* Ey, double* delta_rho, double* phi,
                                void* ws, size_t ws_bytes, vp_stream_t stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    VP_RETURN_IF(!f_in || !f_out || !shape || !vx || !vy || !Ex || !Ey || !delta_rho || !phi || !ws,
                 VP_ERR_ARG);
    VP_RETURN_IF(physics_mode && (!kx || !ky || !inv_ksq), VP_ERR_ARG);
    VP_RETURN_IF(ws_bytes < vp_step_2d2v_workspace(shape), VP_ERR_WORKSPACE);
    if (vp::fast2d_supported(shape))
        return fused_step_2d(f_in, f_out, shape, dt, vx, vy, kx, ky, inv_ksq, dx, dy, dvx, dvy, physics_mode, Ex, Ey,
                             delta_rho, phi, ws, s);
    const int64_t nx = shape[0], ny = shape[1], nvx = shape[2], nvy = shape[3];
    const size_t cells = static_cast<size_t>(nx) * ny * nvx * nvy;
    unsigned char* wb = static_cast<unsigned char*>(ws);
    double* tmp = reinterpret_cast<double*>(wb);
    wb += vp::align256(cells * sizeof(double));
    double* rho = reinterpret_cast<double*>(wb);
    wb += vp::align256(static_cast<size_t>(nx) * ny * sizeof(double));
    void* pws = wb;
    const double half = dt / 2.0;
    const int64_t vs_x[4] = {0, 0, 1, 0};  // vx[None, :, None]
    const int64_t vs_y[4] = {0, 0, 0, 1};  // vy[None, None, :]
    const int64_t vs_e[4] = {ny, 1, 0, 0};  // E[:, :, None]
    VP_TRY(vp_advect_f64(f_in, tmp, shape, 0, vx, vs_x, half, dx, 1, stream));
    VP_TRY(vp_advect_f64(tmp, f_out, shape, 1, vy, vs_y, half, dy, 1, stream));
    if (physics_mode) {
        VP_TRY(vp::row_sum<double>(f_out, nx * ny, nvx * nvy, dvx, dvy, 0.0, rho, s));
        VP_TRY(vp::poisson2d(rho, nx, ny, kx, ky, inv_ksq, Ex, Ey, delta_rho, phi, pws,
                             vp::poisson2d_ws(nx, ny), s));
    } else {
        VP_TRY(vp::fill(Ex, nx * ny, 0.0, s));
        VP_TRY(vp::fill(Ey, nx * ny, 0.0, s));
        VP_TRY(vp::fill(delta_rho, nx * ny, 0.0, s));
        VP_TRY(vp::fill(phi, nx * ny, 0.0, s));
    }
    VP_TRY(vp_advect_f64(f_out, tmp, shape, 2, Ex, vs_e, dt, dvx, 0, stream));
    VP_TRY(vp_advect_f64(tmp, f_out, shape, 3, Ey, vs_e, dt, dvy, 0, stream));
    VP_TRY(vp_advect_f64(f_out, tmp, shape, 1, vy, vs_y, half, dy, 1, stream));
    VP_TRY(vp_advect_f64(tmp, f_out, shape, 0, vx, vs_x, half, dx, 1, stream));
    return VP_OK;
}
