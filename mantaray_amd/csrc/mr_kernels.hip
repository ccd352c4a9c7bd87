// mr_kernels.hip -- gfx950 kernels of the many-ray integrator.
//
//   trace_kernel   ManyRays::trace_many (src/ray.rs:98-127): one ray per
//                  lane, the whole RK4 step loop in registers, rows written
//                  step-major so each wave stores 64 contiguous doubles per
//                  component and step.
//   rhs_kernel     WaveRayPath::system (src/wave_ray_path.rs:220-234) at a
//                  batch of states (parity tests of the right-hand side).
//   build_*        per-cell records of a gridded field (see mr_device.h).
//   fill_time      the time column, t += dt accumulated as ode_solvers does.
#include "mr_kernels.h"

namespace mr {

// Copy a field view from device memory into per-lane registers.  The loads
// go through a lane-varying (always zero) offset so they are vector loads
// and the view lives in VGPRs for the whole kernel: the scalar register file
// cannot hold both views next to the loop's own scalars, and re-reading
// them with scalar loads puts a load latency into every evaluation.
template <class View>
__device__ __forceinline__ View load_view(const void *p)
{
    static_assert(sizeof(View) % 4 == 0, "view size");
    int zero;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(zero));
    const int *src = reinterpret_cast<const int *>(p) + zero;
    int buf[sizeof(View) / 4];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(View) / 4); ++i) buf[i] = src[i];
    View v;
    __builtin_memcpy(&v, buf, sizeof(View));
    return v;
}

// ---------------------------------------------------------------------------
// ode_solvers 0.4.0 Rk4::step/integrate for one lane:
//   k1 = f(y), k2 = f(y + k1 (dt/2)), k3 = f(y + k2 (dt/2)), k4 = f(y + k3 dt)
//   y <- y + (((k1 + k2 2) + k3 2) + k4) (dt/6)
// with the stop test of WaveRayPath::solout (wave_ray_path.rs:236-246).
template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(256) trace_kernel(TraceArgs a)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = r < a.n_rays;
    const int64_t nr = a.n_rays;
    double x = 0.0, y = 0.0, kx = 0.0, ky = 0.0;
    if (active) {
        x = a.x0[r]; y = a.y0[r]; kx = a.kx0[r]; ky = a.ky0[r];
        if (a.xo) a.xo[r] = x;
        if (a.yo) a.yo[r] = y;
        if (a.kxo) a.kxo[r] = kx;
        if (a.kyo) a.kyo[r] = ky;
    }
    const BathyView B = load_view<BathyView>(a.B);
    const CurrentView C = load_view<CurrentView>(a.C);
    BathyCache bc;
    CurrentCache cc;
    bool alive = active;
    int64_t len = active ? 1 : 0;
    const double half = a.half, dt = a.dt, sixth = a.sixth;
    // the time column t_s = t_{s-1} + dt (ode_solvers accumulates), written
    // by the first lane of the grid alongside its own rows
    const bool timekeeper = a.t_out && r == 0;
    double tcur = a.t0;
    if (timekeeper) a.t_out[0] = tcur;
    int64_t s = 1;
    for (; s <= a.n_steps; ++s) {
        tcur = tcur + dt;
        if (timekeeper) a.t_out[s] = tcur;
        if (!__any(alive)) break;
        double ox = MR_NAN, oy = MR_NAN, okx = MR_NAN, oky = MR_NAN;
        if (alive) {
            double sx = x, sy = y, skx = kx, sky = ky;   // stage state
            double ax = 0.0, ay = 0.0, akx = 0.0, aky = 0.0;
            bool stop_before = false;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                double dx, dy, dkx, dky;
                rhs<BK, CK, MATH>(B, a.bp, bc, C, a.cp, cc, sx, sy, skx, sky, dx, dy, dkx, dky);
                if (st == 0) {
                    // With MR_SOLOUT_NEXT_DERIV, solout saw f(y) of the row
                    // just pushed: an all-NaN k1 here means the ray already
                    // ended on that row.
                    if (a.solout == 1 && all_nan4(dx, dy, dkx, dky)) {
                        stop_before = true;
                        break;
                    }
                    ax = dx; ay = dy; akx = dkx; aky = dky;
                } else if (st < 3) {
                    // k * 2 is exact, so fma(k, 2, a) rounds as a + k * 2
                    ax = __builtin_fma(dx, 2.0, ax); ay = __builtin_fma(dy, 2.0, ay);
                    akx = __builtin_fma(dkx, 2.0, akx); aky = __builtin_fma(dky, 2.0, aky);
                } else {
                    ax = ax + dx; ay = ay + dy; akx = akx + dkx; aky = aky + dky;
                }
                const double hs = st < 2 ? half : dt;
                sx = x + dx * hs; sy = y + dy * hs;
                skx = kx + dkx * hs; sky = ky + dky * hs;
            }
            if (stop_before) {
                alive = false;
            } else {
                ox = x + ax * sixth; oy = y + ay * sixth;
                okx = kx + akx * sixth; oky = ky + aky * sixth;
                x = ox; y = oy; kx = okx; ky = oky;
                len = s + 1;
                if (all_nan4(ox, oy, okx, oky)) alive = false;
            }
        }
        if (active) {
            const int64_t o = s * nr + r;
            if (a.xo) a.xo[o] = ox;
            if (a.yo) a.yo[o] = oy;
            if (a.kxo) a.kxo[o] = okx;
            if (a.kyo) a.kyo[o] = oky;
        }
    }
    if (active) {
        // rows past the wave's last live step (its time entry is written)
        bool first = true;
        for (; s <= a.n_steps; ++s) {
            if (!first) {
                tcur = tcur + dt;
                if (timekeeper) a.t_out[s] = tcur;
            }
            first = false;
            const int64_t o = s * nr + r;
            if (a.xo) a.xo[o] = MR_NAN;
            if (a.yo) a.yo[o] = MR_NAN;
            if (a.kxo) a.kxo[o] = MR_NAN;
            if (a.kyo) a.kyo[o] = MR_NAN;
        }
        if (a.len_out) a.len_out[r] = (int32_t)len;
        if (a.end_out) {
            a.end_out[4 * r + 0] = x; a.end_out[4 * r + 1] = y;
            a.end_out[4 * r + 2] = kx; a.end_out[4 * r + 3] = ky;
        }
    }
}

template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(256) rhs_kernel(RhsArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double dx, dy, dkx, dky;
    const BathyView B = load_view<BathyView>(a.B);
    const CurrentView C = load_view<CurrentView>(a.C);
    BathyCache bc;
    CurrentCache cc;
    rhs<BK, CK, MATH>(B, a.bp, bc, C, a.cp, cc, a.s[4 * i], a.s[4 * i + 1], a.s[4 * i + 2], a.s[4 * i + 3],
                      dx, dy, dkx, dky);
    a.ds[4 * i] = dx; a.ds[4 * i + 1] = dy; a.ds[4 * i + 2] = dkx; a.ds[4 * i + 3] = dky;
}

__global__ void fill_time_kernel(double t0, double dt, int64_t n, double *t)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double v = t0;
        t[0] = v;
        for (int64_t s = 1; s <= n; ++s) {
            v = v + dt;
            t[s] = v;
        }
    }
}

// Cell records (mr_device.h).  Cell (i, j) has corners A = (i, j),
// B = (i, j+1), C = (i+1, j+1), D = (i+1, j) in the reference's order
// (cartesian_netcdf3.rs:389); every value is produced by the reference's own
// f32 / f64 operations so the records are bitwise what it would compute.
// Cells with a corner past the value arrays (c >= oob_from) are never read.

// change of basis of interpolator.rs:59-72 for the cell; m01 = NaN marks
// det == 0, on which the reference errs
__device__ __forceinline__ float2 cell_basis(const float2 *xp, const float2 *yp, int64_t i,
                                             int64_t j)
{
    const float2 xq = xp[i], yq = yp[j];                   // (a[i], a[i+1])
    const float btx = xq.x - xq.x, bty = yq.y - yq.x;
    const float dtx = xq.y - xq.x, dty = yq.x - yq.x;
    const float det = (btx * dty) - (dtx * bty);
    if (det == 0.0f) return make_float2(MR_NANF, MR_NANF);
    return make_float2(-(dtx / det), -(bty / det));
}

__global__ void build_bathy_kernel(BathyBuildArgs a)
{
    const int64_t ncx = a.nx - 1;
    const int64_t ncell = ncx * (a.ny - 1);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int64_t i = c % ncx, j = c / ncx;
    BCell r;
    if (c >= a.oob_from) {
        r.a00 = r.a10 = r.a01 = r.a11 = MR_NANF;
        r.hx = r.hy = MR_NANF;
    } else {
        const int64_t iA = a.nx * j + i, iB = iA + a.nx, iC = iA + a.nx + 1, iD = iA + 1;
        const float vA = (float)a.raw[iA], vB = (float)a.raw[iB];
        const float vC = (float)a.raw[iC], vD = (float)a.raw[iD];
        r.a00 = vA;                                           // interpolator.rs:78-81
        r.a10 = vB - vA;
        r.a01 = vD - vA;
        r.a11 = ((vC - vA) - r.a10) - r.a01;
        r.hx = (float)((a.raw[iD] - a.raw[iA]) / a.xspace);  // cartesian_netcdf3.rs:119-134
        r.hy = (float)((a.raw[iB] - a.raw[iA]) / a.yspace);
    }
    const float2 m = cell_basis(a.xp, a.yp, i, j);   // m01 = NaN marks det == 0
    r.m01 = m.x;
    r.m10 = m.y;
    a.cells[tiled_cell((int)i, (int)j, a.ntx)] = r;
}

__global__ void build_current_kernel(CurrentBuildArgs a)
{
    const int64_t ncx = a.nx - 1;
    const int64_t ncell = ncx * (a.ny - 1);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int64_t i = c % ncx, j = c / ncx;
    CCell r;
    if (c >= a.oob_from) {
        r.ua00 = r.ua10 = r.ua01 = r.ua11 = MR_NANF;
        r.va00 = r.va10 = r.va01 = r.va11 = MR_NANF;
        r.ux = r.uy = r.vx = r.vy = MR_NAN;
    } else {
        const int64_t iA = a.nx * j + i, iB = iA + a.nx, iC = iA + a.nx + 1, iD = iA + 1;
        float vA = (float)a.u[iA], vB = (float)a.u[iB], vC = (float)a.u[iC], vD = (float)a.u[iD];
        r.ua00 = vA; r.ua10 = vB - vA; r.ua01 = vD - vA; r.ua11 = ((vC - vA) - r.ua10) - r.ua01;
        vA = (float)a.v[iA]; vB = (float)a.v[iB]; vC = (float)a.v[iC]; vD = (float)a.v[iD];
        r.va00 = vA; r.va10 = vB - vA; r.va01 = vD - vA; r.va11 = ((vC - vA) - r.va10) - r.va01;
        r.ux = (a.u[iD] - a.u[iA]) / a.xspace;                // cartesian_current.rs:515-536
        r.uy = (a.u[iB] - a.u[iA]) / a.yspace;
        r.vx = (a.v[iD] - a.v[iA]) / a.xspace;
        r.vy = (a.v[iB] - a.v[iA]) / a.yspace;
    }
    const uint32_t t = tiled_cell((int)i, (int)j, a.ntx);
    a.cells[t] = r;
    if (a.mtab) a.mtab[t] = cell_basis(a.xp, a.yp, i, j);
}

// ---------------------------------------------------------------------------
// launchers

#define MR_DISPATCH(KERNEL, BK, CK, MATH, GRID, BLOCK, STREAM, ARGS)                     \
    do {                                                                                  \
        if (MATH == MATH_FAST) {                                                          \
            DISPATCH_BC(KERNEL, BK, CK, MATH_FAST, GRID, BLOCK, STREAM, ARGS);            \
        } else {                                                                          \
            DISPATCH_BC(KERNEL, BK, CK, MATH_LITERAL, GRID, BLOCK, STREAM, ARGS);         \
        }                                                                                 \
    } while (0)

#define DISPATCH_BC(KERNEL, BK, CK, M, GRID, BLOCK, STREAM, ARGS)                           \
    do {                                                                                    \
        if (CK == CK_CONSTANT) {                                                            \
            switch (BK) {                                                                   \
            case BK_CONSTANT: hipLaunchKernelGGL((KERNEL<BK_CONSTANT, CK_CONSTANT, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            case BK_SLOPE: hipLaunchKernelGGL((KERNEL<BK_SLOPE, CK_CONSTANT, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            case BK_ARRAY: hipLaunchKernelGGL((KERNEL<BK_ARRAY, CK_CONSTANT, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            default: hipLaunchKernelGGL((KERNEL<BK_GRID, CK_CONSTANT, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            }                                                                               \
        } else {                                                                            \
            switch (BK) {                                                                   \
            case BK_CONSTANT: hipLaunchKernelGGL((KERNEL<BK_CONSTANT, CK_GRID, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            case BK_SLOPE: hipLaunchKernelGGL((KERNEL<BK_SLOPE, CK_GRID, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            case BK_ARRAY: hipLaunchKernelGGL((KERNEL<BK_ARRAY, CK_GRID, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            default: hipLaunchKernelGGL((KERNEL<BK_GRID, CK_GRID, M>), GRID, BLOCK, 0, STREAM, ARGS); break; \
            }                                                                               \
        }                                                                                   \
    } while (0)

hipError_t launch_trace(const TraceArgs &a, int bk, int ck, int math, int block,
                        hipStream_t stream)
{
    if (a.n_rays <= 0) return hipSuccess;
    const int64_t nblk = (a.n_rays + block - 1) / block;
    MR_DISPATCH(trace_kernel, bk, ck, math, dim3((unsigned)nblk), dim3(block),
                stream, a);
    return hipGetLastError();
}

hipError_t launch_rhs(const RhsArgs &a, int bk, int ck, int math, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t nblk = (a.n + block - 1) / block;
    MR_DISPATCH(rhs_kernel, bk, ck, math, dim3((unsigned)nblk), dim3(block),
                stream, a);
    return hipGetLastError();
}

hipError_t launch_fill_time(double t0, double dt, int64_t n, double *t, hipStream_t stream)
{
    hipLaunchKernelGGL(fill_time_kernel, dim3(1), dim3(64), 0, stream, t0, dt, n, t);
    return hipGetLastError();
}

hipError_t launch_build_bathy(const BathyBuildArgs &a, hipStream_t stream)
{
    const int64_t ncell = (a.nx - 1) * (a.ny - 1);
    if (ncell <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(build_bathy_kernel, dim3((unsigned)((ncell + block - 1) / block)),
                       dim3(block), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_build_current(const CurrentBuildArgs &a, hipStream_t stream)
{
    const int64_t ncell = (a.nx - 1) * (a.ny - 1);
    if (ncell <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(build_current_kernel, dim3((unsigned)((ncell + block - 1) / block)),
                       dim3(block), 0, stream, a);
    return hipGetLastError();
}

}  // namespace mr
