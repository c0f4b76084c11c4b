// Host interface of the fused in-place 2D2V sweeps (fast2d.hip).
#pragma once

#include "vp_common.hpp"

namespace vp {

// Workspace of the fused path: X/Y weight tables and the rho partial sums.
struct Fast2DWs {
    double* Wx;
    int* ILx;
    double* Wy;
    int* ILy;
    double2* partial;
    int nparts;
    void* pws;  // Poisson workspace
    size_t pws_bytes;
    void* ftmp;  // one f-sized buffer (the out-of-place kick writes here); null without it
    size_t bytes;
};

bool fast2d_supported(const int64_t shape[4]);
// FP32 storage (float f, float64 tables, rho, fields)
bool fast2d_supported_f32(const int64_t shape[4]);
// the kick only constrains the (vx, vy) plane
bool fast2d_kick_supported(const int64_t shape[4]);
// Carves the workspace (ws == nullptr: only computes .bytes); elem_bytes is
// the size of one f element (8: float64, 4: float32 storage).  Order: weight
// tables, rho partial sums, Poisson workspace, then (with_ftmp) one f-sized
// buffer, so the layout without it is a prefix of the layout with it.
Fast2DWs fast2d_layout(const int64_t shape[4], void* ws, int elem_bytes, bool with_ftmp);
size_t poisson2d_ws(int64_t nx, int64_t ny);
int fast2d_tables(const int64_t shape[4], const Fast2DWs& L, const double* vx, const double* vy, double half_dt,
                  double dx, double dy, hipStream_t s);
int fast2d_xdrift(const double* fin, double* fout, const int64_t shape[4], const Fast2DWs& L, int passes,
                  hipStream_t s);
int fast2d_ydrift(const double* fin, double* fout, const int64_t shape[4], const Fast2DWs& L, bool rho,
                  hipStream_t s);
int fast2d_kick(const double* fin, double* fout, const int64_t shape[4], const double* Ex, const double* Ey,
                double dt, double dvx, double dvy, hipStream_t s);
int fast2d_xdrift(const float* fin, float* fout, const int64_t shape[4], const Fast2DWs& L, int passes,
                  hipStream_t s);
int fast2d_ydrift(const float* fin, float* fout, const int64_t shape[4], const Fast2DWs& L, bool rho,
                  hipStream_t s);
// Deviation storage (float32 g = f - feq, feq = fx | fy: float64 profiles of
// the separable feq(vx, vy) = fx(vx) fy(vy)): the float32 kick of g plus the
// float64 correction (Kvx fx)(Kvy fy) - fx fy; nvy == 128, out of place.
int fast2d_kick_dev(const float* gin, float* gout, const int64_t shape[4], const double* feq, const double* Ex,
                    const double* Ey, double dt, double dvx, double dvy, hipStream_t s);
int fast2d_kick(const float* fin, float* fout, const int64_t shape[4], const double* Ex, const double* Ey, double dt,
                double dvx, double dvy, hipStream_t s);

}  // namespace vp
