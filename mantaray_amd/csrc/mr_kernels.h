// mr_kernels.h -- kernel argument blocks and launchers (host-visible).
#pragma once

#include "mr_device.h"

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
    int ntx;                           // 8 x 8 record tiles along x
    BCell *cells;
};

struct CurrentBuildArgs {
    int64_t nx, ny;
    const float2 *xp, *yp;
    const double *u, *v;
    int64_t oob_from;
    double xspace, yspace;
    int ntx;
    CCell *cells;
    float2 *mtab;                      // per-cell change of basis
};

hipError_t launch_trace(const TraceArgs &a, int bk, int ck, int math, int block,
                        hipStream_t stream);
hipError_t launch_rhs(const RhsArgs &a, int bk, int ck, int math, hipStream_t stream);
hipError_t launch_fill_time(double t0, double dt, int64_t n, double *t, hipStream_t stream);
hipError_t launch_build_bathy(const BathyBuildArgs &a, hipStream_t stream);
hipError_t launch_build_current(const CurrentBuildArgs &a, hipStream_t stream);

}  // namespace mr
