// mr_kernels.h -- kernel argument blocks and launchers (host-visible).
#pragma once

#include "mr_device64.h"

namespace mr {

// The numeric field views live in device memory and are copied into
// per-lane registers at kernel start (see load_view); the field pointers
// travel as kernel arguments.
struct TraceArgs {
    const void *B;                     // BathyView in device memory
    const void *C;                     // CurrentView in device memory
    BathyPtrs bp;
    CurrentPtrs cp;
    int64_t n_rays, n_steps;
    double t0, dt, half, sixth;
    double *t_out;                     // n_steps + 1 times, may be null
    int ray_blocks;                    // workgroups of rays; one more writes t_out
    const double *x0, *y0, *kx0, *ky0;
    double *xo, *yo, *kxo, *kyo;       // step-major, may be null
    int32_t *len_out;                  // may be null
    double *end_out;                   // 4 per ray, may be null
    int solout;                        // 0 state, 1 next derivative
    int shared_cells;                  // gridded depth and current on one set of cells
    const SCell *scells;               // their combined records (with shared_cells)
};

// One scene of a batched trace, in device memory.
struct SceneDesc {
    const void *B, *C;                 // BathyView / CurrentView in device memory
    BathyPtrs bp;
    CurrentPtrs cp;
    double dt, half, sixth;
    int64_t n_steps;
    int64_t ray0;                      // first global ray of the scene
    const double *x0, *y0, *kx0, *ky0; // the scene's initial conditions
    int32_t *len_out;                  // the scene's outputs
    double *end_out;
    const SCell *scells;               // combined records (batches with shared_cells)
};

// A persistent batched trace: rays of every scene taken from one queue.
struct BatchArgs {
    const SceneDesc *scenes;
    const int32_t *ray_scene;          // scene of each global ray
    int64_t n_rays;                    // over all scenes
    unsigned long long *next;          // queue head, cleared before the launch
    int solout;
    int shared_cells;                  // as in TraceArgs, for every scene
};

struct RhsArgs {
    const void *B;
    const void *C;
    BathyPtrs bp;
    CurrentPtrs cp;
    int64_t n;
    const double *s;
    double *ds;
};

struct BathyBuildArgs {
    int64_t nx, ny;
    const float2 *xp, *yp;
    const double *raw;
    int64_t oob_from;                  // cells from here on are never read
    double xspace, yspace;
    int ntx;                           // (record tiles along x) << 4 | log2 tile edge
    BCell *cells;
};

// combined records of a gridded depth and current on shared cells
struct SharedBuildArgs {
    int64_t nx, ny;
    const float2 *xp, *yp;             // the depth's f32 axis pairs (a[i], a[i+1])
    const float2 *fx, *fy;             // per axis cell: the fast region's (lo, hi)
    const double *depth, *u, *v;       // raw values
    int64_t oob_from;                  // cells from here on have a corner past some array
    double bxspace, byspace;           // (f64)x[1] - (f64)x[0] of the f32 axes (depth slopes)
    double cxspace, cyspace;           // x[1] - x[0] of the f64 axes (current slopes)
    int ntx;                           // as in BathyBuildArgs
    SCell *cells;
};

struct CurrentBuildArgs {
    int64_t nx, ny;
    const float2 *xp, *yp;
    const double *u, *v;
    int64_t oob_from;
    double xspace, yspace;
    int ntx;                           // as in BathyBuildArgs
    CCell *cells;
    float2 *mtab;                      // per-cell change of basis
};

// fp64-interpolation trace (MR_PREC_FP64_INTERP), optionally writing the
// state at every stride-th step (checkpoints for the adjoint sweep).
struct Trace64Args {
    const void *B, *C;                 // Grid64View in device memory
    Grid64Ptrs bp, cp;
    Fields64 f;                        // field values (depth, u, v)
    int64_t n_rays, n_steps;
    double t0, dt, half, sixth;
    double *t_out;
    int ray_blocks;
    const double *x0, *y0, *kx0, *ky0;
    double *xo, *yo, *kxo, *kyo;
    int32_t *len_out;
    double *end_out;
    int solout;
    int stride;                        // checkpoint every stride steps
    double *ckpt;                      // [n_ckpt][4][n_rays], may be null
};

// Reverse sweep of the discrete RK4 adjoint.
struct AdjointArgs {
    const void *B, *C;
    Grid64Ptrs bp, cp;
    Fields64 f;
    int64_t n_rays, n_steps;
    double dt, half, sixth;
    int stride;
    const double *ckpt;                // [n_ckpt][4][n_rays]
    const int32_t *len;                // rows of each ray
    const double *end;                 // last row of each ray (4 per ray)
    const double *g_end;               // cotangent of the last row (4 per ray), may be null
    const double *g_x, *g_y, *g_kx, *g_ky;   // cotangents of the rows, step-major, may be null
    double *work;                      // [2][stride][8][work_ld] pairs of doubles
    int64_t work_ld;                   // >= blocks * block threads
    double *g_x0, *g_y0, *g_kx0, *g_ky0;     // may be null
    Grads64 grads;                     // gradient grids, cleared by the caller
};

struct Rhs64Args {
    const void *B, *C;
    Grid64Ptrs bp, cp;
    Fields64 f;
    int64_t n;
    const double *s;
    double *ds;
};

struct M64BuildArgs {
    int64_t nx, ny;
    const double *ax, *ay;
    double2 *m;
};

hipError_t launch_trace(const TraceArgs &a, int bk, int ck, int math, int block,
                        hipStream_t stream);
hipError_t launch_rhs(const RhsArgs &a, int bk, int ck, int math, hipStream_t stream);
hipError_t launch_fill_time(double t0, double dt, int64_t n, double *t, hipStream_t stream);
hipError_t launch_batch(const BatchArgs &a, int bk, int ck, int math, int block, int blocks,
                        hipStream_t stream);
hipError_t launch_clear_counter(unsigned long long *p, hipStream_t stream);
hipError_t launch_build_bathy(const BathyBuildArgs &a, hipStream_t stream);
hipError_t launch_build_current(const CurrentBuildArgs &a, hipStream_t stream);
hipError_t launch_build_shared(const SharedBuildArgs &a, hipStream_t stream);
hipError_t launch_trace64(const Trace64Args &a, int bk, int ck, int math, int block,
                          hipStream_t stream);
// waves per 64 rays: 2 (recompute | sweep) or 4 (recompute | two linearise |
// chain)
hipError_t launch_adjoint(const AdjointArgs &a, int bk, int ck, int math, int waves,
                          hipStream_t stream);
hipError_t launch_rhs64(const Rhs64Args &a, int bk, int ck, int math, hipStream_t stream);
hipError_t launch_build_m64(const M64BuildArgs &a, hipStream_t stream);
hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t stream);
// the device address of this device's MR_ADJOINT_SCRATCH_BYTES scratch array
// (looked up on the device's first call, then cached)
hipError_t adjoint_scratch(void **ptr);

}  // namespace mr
