// mr_device.h -- device-side field views, field queries and the ray-equation
// right-hand side for gfx950.
//
// Citations are to files of mines-oceanography/mantaray 0.1.0.
//
// Data layout in HBM (built once per field and device by build_bathy_kernel
// and build_current_kernel in mr_kernels.hip):
//   gridded bathymetry: one 32-byte record per cell (i, j), i < nx-1,
//     j < ny-1, in 8 x 8-cell tiles (tiled_cell), holding the f32 bilinear
//     coefficients of the cell, the f32 cell gradient and the cell's
//     change-of-basis entries; the f32 axes as (a[i], a[i+1]) pairs; the raw
//     f64 depth (read only when the target is exactly a grid node).
//   gridded current: one 64-byte record per cell with the f32 bilinear
//     coefficients of u and v and the four f64 cell gradients; pairs of the
//     f32 casts of the f64 axes; the change of basis as two constants when
//     the f32 axes are uniform, else a per-cell float2 table; the raw f64 u
//     and v.
// A query loads one record per field and the two axis pairs around the cell
// instead of 12 scattered corner values, and performs no division:
// everything cell-constant is precomputed with the reference's own
// operations, so the interpolated values are bitwise those of the reference.
//
// Register layout in the kernels: the numeric parameters of a field (a
// *View) are held per lane in VGPRs; its device pointers (a *Ptrs) are kernel
// arguments in SGPRs, so record loads are global loads with a scalar base and
// a 32-bit lane offset.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mr {

#define MR_NANF_INIT __builtin_nanf("")
#define MR_NAN_INIT __builtin_nan("")
// a double whose two f32 halves are both -inf
#define MR_EMPTY_REGION __builtin_bit_cast(double, 0xFF800000FF800000ull)

constexpr double G = 9.8;                        // src/wave_ray_path.rs:23

enum BathyKind : int { BK_CONSTANT = 0, BK_SLOPE = 1, BK_ARRAY = 2, BK_GRID = 3 };
enum CurrentKind : int { CK_CONSTANT = 0, CK_GRID = 1 };
enum MathMode : int { MATH_FAST = 0, MATH_LITERAL = 1 };

struct BCell {                    // 32 B
    float a00, a10, a01, a11;     // interpolator.rs:78-81 over (f32)depth
    float hx, hy;                 // cartesian_netcdf3.rs:119-134
    float m01, m10;               // interpolator.rs:69-72 (m00 = m11 = 0/det)
};

struct CCell {                    // 64 B
    float ua00, ua10, ua01, ua11; // u coefficients (f32)
    float va00, va10, va01, va11; // v coefficients (f32)
    double ux, uy, vx, vy;        // cartesian_current.rs:515-536
};

// One cell of a gridded depth and a gridded current that share their cells
// (same node counts; the current's axes rounded to f32 are the depth's), in
// one 128-byte record, so that one load and one test serve both fields.
// [lo_x, hi_x] x [lo_y, hi_y] is the cell's fast region: a target whose f32
// rounding (px, py) lies in it is in this cell for the depth's f32 index
// rule, and, for every f64 point that rounds to (px, py), for the current's
// f64 index rule; and it is on no corner (px, py are strictly between the
// cell's node coordinates).  Such a target needs no other test: the depth
// and both current components are the cell's bilinear pieces at (px, py),
// with the cell's slopes.  Cells where the reference errs (a corner past a
// value array, det == 0) get an empty region.  Each bound is found on the
// host from the exact threshold tables of both index rules.
struct SCell {                    // 128 B
    float lo_x, hi_x, lo_y, hi_y; // fast region (f32, closed)
    float x0, y0, m01, m10;       // cell origin (f32 node coordinates); change of basis
    float a00, a10, a01, a11;     // depth, interpolator.rs:78-81 over (f32)depth
    double hx, hy;                // depth slopes (f32 values, cartesian_netcdf3.rs:119-134,
                                  // widened as the right-hand side widens them)
    float ua00, ua10, ua01, ua11; // u over (f32)u
    float va00, va10, va01, va11; // v over (f32)v
    double ux, uy, vx, vy;        // current slopes (f64), cartesian_current.rs:515-536
};

struct BathyView {
    int kind;
    float h;                              // ConstantDepth
    float h0, x0, y0, dhdx, dhdy;         // ConstantSlope
    int an;                               // ArrayDepth (an x an)
    int nx, ny, ncx;                      // CartesianNetcdf3; ncx = nx - 1
    float ax0, ay0, nxm1, nym1;           // a[0]; (f32)(n-1)
    double inv_sx, inv_sy;                // RN64(1/(double)|a1-a0|)
    uint32_t oob_from;                    // first cell (row-major) with a corner past the data
    int ntx;                              // (record tiles along x) << 4 | log2 tile edge
    float pinv_sx, pinv_sy;               // cell guess for prefetch: (f32) inv_s
    float pmax_x, pmax_y;                 // (f32)(n - 2)
};

struct BathyPtrs {
    const float *arr;                     // ArrayDepth values
    const float4 *xa, *ya;                // per axis cell i < n-1: (a[i], a[i+1], T[i], T[i+1])
    const BCell *cells;
    const double *raw;                    // raw depth (exact-corner values)
};

struct CurrentView {
    int kind;
    int nx, ny, ncx;                      // CartesianCurrent; ncx = nx - 1
    double u, v;                          // ConstantCurrent
    double cx0, cy0, sx, sy, inv_sx, inv_sy;
    double nxm1, nym1;                    // n - 1
    uint32_t oob_from;                    // first cell (row-major) with a corner past u or v
    int ntx;                              // (record tiles along x) << 4 | log2 tile edge
    float px0, py0, pinv_sx, pinv_sy;     // cell guess for prefetch, in f32
    float pmax_x, pmax_y;                 // (f32)(n - 2)
};

struct CurrentPtrs {
    const float2 *xp, *yp;                // pairs of the (f32) casts of the f64 axes
    const double2 *xt, *yt;               // per axis cell: (T[i], T[i+1]) in f64
    const CCell *cells;
    const float2 *mtab;                   // per-cell (m01, m10)
    const double *raw_u, *raw_v;          // raw values (exact-corner values)
};

// Per-lane copy of a cell record, with the cell's index range.  Along each
// axis the cell the reference chooses is a monotone step function of the
// coordinate (cartesian_netcdf3.rs:274-390, cartesian_current.rs:231-339
// round, divide and floor monotonically), so cell i is exactly the interval
// T[i] <= p < T[i+1] of a per-axis threshold table T (T[0] the first valid
// coordinate, T[n-1] the first one past the last node).  The thresholds are
// found on the host by bisection over the floating-point values with the
// reference's own arithmetic.  A query whose target lies in the slot's cell
// (four comparisons per field, and the slot valid) needs no index arithmetic
// and no load of its own.  valid is false for a cell with a corner past the
// data (an error cell) and for a slot never loaded; thresholds start as NaN.
struct BathySlot {
    float4 r0 = {0.f, 0.f, 0.f, 0.f}, r1 = {0.f, 0.f, 0.f, 0.f};
    float4 ax = {0.f, 0.f, MR_NANF_INIT, MR_NANF_INIT};   // (x_i, x_i+1, T_i, T_i+1)
    float4 ay = {0.f, 0.f, MR_NANF_INIT, MR_NANF_INIT};
    int i1 = 0, j1 = 0;
    bool valid = false;
};

struct CurrentSlot {
    float4 r0 = {0.f, 0.f, 0.f, 0.f}, r1 = {0.f, 0.f, 0.f, 0.f};
    double2 g0 = {0.0, 0.0}, g1 = {0.0, 0.0};
    float2 xp = {0.f, 0.f}, yp = {0.f, 0.f}, m = {0.f, 0.f};
    double2 tx = {MR_NAN_INIT, MR_NAN_INIT}, ty = {MR_NAN_INIT, MR_NAN_INIT};
    int i1 = 0, j1 = 0;
    bool valid = false;
};

// Loads from device (global) memory through the global address space, so
// that the compiler emits global (not flat) loads.
#define MR_GLOBAL __attribute__((address_space(1)))
typedef float mr_f2 __attribute__((ext_vector_type(2)));
typedef float mr_f4 __attribute__((ext_vector_type(4)));
typedef double mr_d2 __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ T gld(const T *p, uint32_t i)
{
    return *(const MR_GLOBAL T *)((const MR_GLOBAL char *)p + i * (uint32_t)sizeof(T));
}
__device__ __forceinline__ float4 gld4(const void *p, uint32_t byte_off)
{
    const mr_f4 v = *(const MR_GLOBAL mr_f4 *)((const MR_GLOBAL char *)p + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 gldd2(const void *p, uint32_t byte_off)
{
    const mr_d2 v = *(const MR_GLOBAL mr_d2 *)((const MR_GLOBAL char *)p + byte_off);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 gld2(const float2 *p, uint32_t i)
{
    const mr_f2 v = *(const MR_GLOBAL mr_f2 *)((const MR_GLOBAL char *)p + i * 8u);
    return make_float2(v.x, v.y);
}

#define MR_NAN __builtin_nan("")
#define MR_NANF __builtin_nanf("")

// Records are stored in square tiles of 2^tl x 2^tl cells (tile-major,
// row-major inside a tile; 8 x 8 by default), so that the neighbours a ray
// moves into lie on the same memory page as the cell it leaves.  ntx_tl =
// (number of tiles along x) << 4 | tl; i, j >= 0.
__device__ __forceinline__ uint32_t tiled_cell(int i, int j, int ntx_tl)
{
    const uint32_t tl = (uint32_t)ntx_tl & 15u, ntx = (uint32_t)ntx_tl >> 4;
    const uint32_t m = (1u << tl) - 1u;
    const uint32_t t = ((uint32_t)j >> tl) * ntx + ((uint32_t)i >> tl);
    return (t << (2u * tl)) | (((uint32_t)j & m) << tl) | ((uint32_t)i & m);
}

// ---------------------------------------------------------------------------
// Interpolation from cell records.
//
// interpolator::bilinear (interpolator.rs:39-84) on a cell A=(x_i, y_j),
// B=(x_i, y_j+1), C=(x_i+1, y_j+1), D=(x_i+1, y_j) has bt = B - A with an
// exactly zero x component and dt = D - A with an exactly zero y component,
// so m00 = m11 = 0/det and m00*ttx, m11*tty are exact zeros: u = m01*tty and
// w = m10*ttx reproduce its lines 74-75 (up to the sign of a zero, which no
// later operation can observe).  The reference's three exits are kept:
//   - a corner index past the value array (cartesian_netcdf3.rs:465-471):
//     Err.  Such cells form a suffix of the cell order (their largest
//     corner index, C, grows with the cell index), so one bound per field,
//     oob_from, identifies them;
//   - target equal to a corner (interpolator.rs:47-51): that corner's value,
//     first match in the order A, B, C, D, read from the raw values;
//   - det == 0 (interpolator.rs:65-67): Err.  The record build marks such a
//     cell with m01 = NaN; no other cell has a NaN m01 (det is finite and
//     nonzero there, so dtx/det is finite or infinite).
// A query reports Err by returning NaN for the depth (bathymetry) or by
// clearing its ok flag (current): the right-hand side then yields all-NaN
// derivatives, which is what the reference's Err produces
// (wave_ray_path.rs:222-227).

__device__ __forceinline__ float bilin_eval(float a00, float a10, float a01, float a11,
                                            float u, float w)
{
    return ((a00 + a10 * u) + a01 * w) + (a11 * u) * w;   // interpolator.rs:83
}

// true when the target is exactly on one of the cell's corners: t - a is 0
// only for t == a (subtraction of distinct finite floats is never zero)
__device__ __forceinline__ bool on_corner(float dxa, float dxd, float dya, float dyb)
{
    const float cx = __builtin_fminf(__builtin_fabsf(dxa), __builtin_fabsf(dxd));
    const float cy = __builtin_fminf(__builtin_fabsf(dya), __builtin_fabsf(dyb));
    return __builtin_fmaxf(cx, cy) == 0.0f;
}

// flat index (row stride nx) of the first corner, in the order A, B, C, D,
// equal to the target
__device__ __forceinline__ uint32_t corner_index(float dxa, float dya, float dyb, int nx,
                                                 int i1, int j1)
{
    uint32_t ic = (uint32_t)nx * (uint32_t)j1 + (uint32_t)i1;
    if (dxa == 0.0f) ic += dya == 0.0f ? 0u : (uint32_t)nx;
    else ic += dyb == 0.0f ? (uint32_t)nx + 1u : 1u;
    return ic;
}

// f32 fractional index RN32(t / sp) (cartesian_netcdf3.rs:285) as
// RN32(t * RN64(1/sp)): the exact quotient of two f32 values lies at least
// 2^-49 (relative) from any f32 rounding boundary and is never on one, while
// the f64 product is within 2^-52 of it, so the final rounding is the
// correctly rounded f32 quotient.
__device__ __forceinline__ float index32(float t, double inv_sp)
{
    return (float)((double)t * inv_sp);
}

// f64 fractional index RN64(t / sp) (cartesian_current.rs:244), reduced to
// what the cell choice needs: the bound checks and floor.  The product with
// the rounded reciprocal is within 2^-50 (relative) of the quotient; only
// when it lies that close to an integer is the exact division taken.  ok is
// cleared where the reference returns Err (a NaN index errs later, in
// bilinear, because all four corners collapse onto (0, 0)).  The returned
// floor is meaningful only when ok is set.
__device__ __forceinline__ double index64(double p, double a0, double sp, double inv_sp,
                                          double nm1, bool &ok)
{
    const double t = p - a0;
    double q = t * inv_sp;
    double fl = __builtin_floor(q);
    if (__builtin_fabs((q - fl) - 0.5) > 0.5 - q * 0x1p-50) {   // near an integer
        q = t / sp;
        fl = __builtin_floor(q);
    }
    ok = ok & (t >= 0.0) & (q <= nm1);
    return fl;
}

// Wait for every outstanding vector memory operation (s_waitcnt vmcnt(0),
// expcnt and lgkmcnt left alone).  Used at the end of a reload path so that
// the loads it issued are complete on that path: the compiler's own wait
// would otherwise sit after the join, where, because vmcnt retires in issue
// order, it also waits for the prefetch and the row stores issued before
// the reload on the common path that did not reload.
__device__ __forceinline__ void wait_vmem()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);
}

// The cell of a target, guessed in f32 and clamped into the grid: used only
// to choose which record to prefetch, never for the result; the exact cell
// test happens when the record is used.  (t - a0) / sp as one FMA with the
// loop-invariant -a0 / sp, clamped by one med3, which returns the smaller of
// its other operands for a NaN (so a NaN guess clamps to 0).
__device__ __forceinline__ int guess_cell(float t, float a0, float inv_sp, float max_cell)
{
    const float f = __builtin_fmaf(t, inv_sp, -(a0 * inv_sp));
    return (int)__builtin_amdgcn_fmed3f(f, 0.0f, max_cell);
}

// ---------------------------------------------------------------------------
// bathymetry: BathymetryData::depth_and_gradient (bathymetry/mod.rs:35-41);
// returns the depth widened to f64, NaN where the reference errs.

__device__ __forceinline__ bool bathy_inside(const BathySlot &s, float px, float py)
{
    return s.valid & (px >= s.ax.z) & (px < s.ax.w) & (py >= s.ay.z) & (py < s.ay.w);
}

// The slot's cell range alone, without its validity: for prefetch tests and
// the branch-free evaluation, where a cell with a corner past the data
// yields NaN all the same (its record holds NaN coefficients and slopes, so
// the derivatives come out all-NaN, as the reference's error makes them).
// A slot never loaded has NaN thresholds and contains no point.
__device__ __forceinline__ bool bathy_in_cell(const BathySlot &s, float px, float py)
{
    return (px >= s.ax.z) & (px < s.ax.w) & (py >= s.ay.z) & (py < s.ay.w);
}

__device__ __forceinline__ void bathy_load(const BathyView &B, const BathyPtrs &P, BathySlot &s,
                                           int i1, int j1)
{
    const uint32_t cell = tiled_cell(i1, j1, B.ntx);
    s.r0 = gld4(P.cells, cell * 32u);
    s.r1 = gld4(P.cells, cell * 32u + 16u);
    s.ax = gld4(P.xa, (uint32_t)i1 * 16u);                  // x[i1], x[i1+1], T
    s.ay = gld4(P.ya, (uint32_t)j1 * 16u);
    s.i1 = i1;
    s.j1 = j1;
    s.valid = (uint32_t)j1 * (uint32_t)B.ncx + (uint32_t)i1 < B.oob_from;
}

// Start loading the record of the cell that (x, y) most likely lies in.
__device__ __forceinline__ void bathy_prefetch(const BathyView &B, const BathyPtrs &P,
                                               BathySlot &s, double x, double y)
{
    const int i1 = guess_cell((float)x, B.ax0, B.pinv_sx, B.pmax_x);
    const int j1 = guess_cell((float)y, B.ay0, B.pinv_sy, B.pmax_y);
    bathy_load(B, P, s, i1, j1);
}

template <int BK>
__device__ __forceinline__ double bathy_query(const BathyView &B, const BathyPtrs &P,
                                              BathySlot &bs, double x, double y, double &hx,
                                              double &hy)
{
    const float px = (float)x, py = (float)y;             // wave_ray_path.rs:122
    if (BK == BK_CONSTANT) {                               // constant_depth.rs:34-40
        const bool nn = __builtin_isnan(px) || __builtin_isnan(py);
        hx = 0.0; hy = 0.0;
        return nn ? MR_NAN : (double)B.h;
    } else if (BK == BK_SLOPE) {                           // constant_slope.rs:63-74
        const bool nn = __builtin_isnan(px) || __builtin_isnan(py);
        const float hv = (B.h0 + B.dhdx * (px - B.x0)) + B.dhdy * (py - B.y0);
        hx = (double)B.dhdx; hy = (double)B.dhdy;
        return nn ? MR_NAN : (double)hv;
    } else if (BK == BK_ARRAY) {                           // array_depth.rs:26-33
        // Rust `as usize`: saturating, NaN -> 0
        const uint64_t xi = px > 0.0f ? (px >= 1.8446744e19f ? UINT64_MAX : (uint64_t)px) : 0;
        const uint64_t yi = py > 0.0f ? (py >= 1.8446744e19f ? UINT64_MAX : (uint64_t)py) : 0;
        hx = 0.0; hy = 0.0;
        if (xi >= (uint64_t)B.an || yi >= (uint64_t)B.an) return MR_NAN;
        return (double)gld(P.arr, (uint32_t)(xi * (uint64_t)B.an + yi));
    } else {                                               // cartesian_netcdf3.rs:98-135
        // in the slot's cell: no index arithmetic (a NaN target fails the
        // comparisons)
        bool ok = bathy_inside(bs, px, py);
        if (__any(!ok)) {
            const bool miss = !ok;
            // nearest (:274-296); a NaN target fails every comparison and
            // yields NaN, as the reference's early return does
            const float ix = index32(px - B.ax0, B.inv_sx);
            const float iy = index32(py - B.ay0, B.inv_sy);
            bool okn = (ix >= 0.0f) & (ix <= B.nxm1) & (iy >= 0.0f) & (iy <= B.nym1);
            // four_corners (:334-390): floor, or n-2 on the last node
            const int i1 = __builtin_elementwise_min((int)ix, B.nx - 2);
            const int j1 = __builtin_elementwise_min((int)iy, B.ny - 2);
            okn = okn & ((uint32_t)j1 * (uint32_t)B.ncx + (uint32_t)i1 < B.oob_from);
            if (miss & okn) bathy_load(B, P, bs, i1, j1);
            wait_vmem();
            ok = ok | (miss & okn);
        }
        const float dxa = px - bs.ax.x, dya = py - bs.ay.x;
        const float4 r0 = bs.r0, r1 = bs.r1;
        float hv = bilin_eval(r0.x, r0.y, r0.z, r0.w, r1.z * dya, r1.w * dxa);
        const float dxd = px - bs.ax.y, dyb = py - bs.ay.y;
        const bool corner = on_corner(dxa, dxd, dya, dyb);
        if (corner & ok)
            hv = (float)gld(P.raw, corner_index(dxa, dya, dyb, B.nx, bs.i1, bs.j1));
        ok = ok & (corner | !__builtin_isnan(r1.z));       // det == 0 errs
        hx = (double)r1.x;
        hy = (double)r1.y;
        return ok ? (double)hv : MR_NAN;
    }
}

// ---------------------------------------------------------------------------
// current: CurrentData::current_and_gradient (current/mod.rs:21-31); returns
// false where the reference errs.

__device__ __forceinline__ bool current_inside(const CurrentSlot &s, double x, double y)
{
    return s.valid & (x >= s.tx.x) & (x < s.tx.y) & (y >= s.ty.x) & (y < s.ty.y);
}

// bathy_in_cell for the current's slots
__device__ __forceinline__ bool current_in_cell(const CurrentSlot &s, double x, double y)
{
    return (x >= s.tx.x) & (x < s.tx.y) & (y >= s.ty.x) & (y < s.ty.y);
}

__device__ __forceinline__ void current_load(const CurrentView &C, const CurrentPtrs &P,
                                             CurrentSlot &s, int i1, int j1)
{
    const uint32_t cell = tiled_cell(i1, j1, C.ntx);
    s.r0 = gld4(P.cells, cell * 64u);
    s.r1 = gld4(P.cells, cell * 64u + 16u);
    s.g0 = gldd2(P.cells, cell * 64u + 32u);
    s.g1 = gldd2(P.cells, cell * 64u + 48u);
    s.xp = gld2(P.xp, (uint32_t)i1);
    s.yp = gld2(P.yp, (uint32_t)j1);
    s.tx = gldd2(P.xt, (uint32_t)i1 * 16u);
    s.ty = gldd2(P.yt, (uint32_t)j1 * 16u);
    s.m = gld2(P.mtab, cell);
    s.i1 = i1;
    s.j1 = j1;
    s.valid = (uint32_t)j1 * (uint32_t)C.ncx + (uint32_t)i1 < C.oob_from;
}

__device__ __forceinline__ void current_prefetch(const CurrentView &C, const CurrentPtrs &P,
                                                 CurrentSlot &s, double x, double y)
{
    const int i1 = guess_cell((float)x, C.px0, C.pinv_sx, C.pmax_x);
    const int j1 = guess_cell((float)y, C.py0, C.pinv_sy, C.pmax_y);
    current_load(C, P, s, i1, j1);
}

template <int CK>
__device__ __forceinline__ bool current_query(const CurrentView &C, const CurrentPtrs &P,
                                              CurrentSlot &cs, double x, double y, double &u,
                                              double &v, double &ux, double &uy, double &vxg,
                                              double &vyg)
{
    if (CK == CK_CONSTANT) {                               // constant_current.rs:66-76
        u = C.u; v = C.v;
        ux = 0.0; uy = 0.0; vxg = 0.0; vyg = 0.0;
        return true;
    } else {                                               // cartesian_current.rs:487-542
        // no NaN guard (:492-495): a NaN target fails the comparisons and
        // then errs, as the reference does
        bool ok = current_inside(cs, x, y);
        if (__any(!ok)) {
            const bool miss = !ok;
            bool okn = true;
            const double fx = index64(x, C.cx0, C.sx, C.inv_sx, C.nxm1, okn);
            const double fy = index64(y, C.cy0, C.sy, C.inv_sy, C.nym1, okn);
            const int i1 = __builtin_elementwise_min((int)fx, C.nx - 2);
            const int j1 = __builtin_elementwise_min((int)fy, C.ny - 2);
            okn = okn & ((uint32_t)j1 * (uint32_t)C.ncx + (uint32_t)i1 < C.oob_from);
            if (miss & okn) current_load(C, P, cs, i1, j1);
            wait_vmem();
            ok = ok | (miss & okn);
        }
        const float tx = (float)x, ty = (float)y;          // cartesian_current.rs:505-512
        const float dxa = tx - cs.xp.x, dya = ty - cs.yp.x;
        const float uw = cs.m.x * dya;
        const float ww = cs.m.y * dxa;
        const float4 r0 = cs.r0, r1 = cs.r1;
        float uf = bilin_eval(r0.x, r0.y, r0.z, r0.w, uw, ww);
        float vf = bilin_eval(r1.x, r1.y, r1.z, r1.w, uw, ww);
        const float dxd = tx - cs.xp.y, dyb = ty - cs.yp.y;
        const bool corner = on_corner(dxa, dxd, dya, dyb);
        if (corner & ok) {
            const uint32_t ic = corner_index(dxa, dya, dyb, C.nx, cs.i1, cs.j1);
            uf = (float)gld(P.raw_u, ic);
            vf = (float)gld(P.raw_v, ic);
        }
        ok = ok & (corner | !__builtin_isnan(cs.m.x));      // det == 0 errs
        u = (double)uf;
        v = (double)vf;
        ux = cs.g0.x; uy = cs.g0.y; vxg = cs.g1.x; vyg = cs.g1.y;
        return ok;
    }
}

// ---------------------------------------------------------------------------
// f64 helpers for the fast dispersion evaluation

// 1/x: hardware reciprocal estimate (about 5e-8 relative) refined by one
// Newton step (about 2e-15).
__device__ __forceinline__ double rcp_nr(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/sqrt(x): hardware estimate (about 5e-8 relative) refined by one Newton
// step (about 4e-15).
__device__ __forceinline__ double rsq_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double t = x * y;
    const double e = __builtin_fma(-t, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

// The f64 constants of the fast dispersion evaluation.  Kernels copy them
// into per-lane registers at start (load_view), so that the loop does not
// rebuild each one in a scalar register pair at every use.
struct FastConsts {
    double inv_ln2, ln2_hi, ln2_lo;   // 1/ln 2; ln 2 split in two parts
    double clamp;                     // |argument| limit of exp_pair
    double expc[10];                  // expm1(f)/f ~ sum c_j f^j, c_9 first (Horner order)
    double g, half_g, m2g;            // g, g/2 and -2g
    double nan;                       // a quiet NaN
};

// e = exp(r) and em = 1 - exp(r) (r <= 0 in use: r = -2kh), each with small
// relative error, from one reduced-argument evaluation: r = m ln2 + f,
// |f| <= ln2/2, expm1(f) = f P(f) with P the degree-9 polynomial
// interpolating expm1(f)/f at the Chebyshev nodes of [-ln2/2, ln2/2] (within
// 1.8e-15 relative, evaluated in f64), e = 2^m (1 + p), em = (1 - 2^m) -
// 2^m p, which has no cancellation for m <= -1 and is exactly -p for m = 0.
// r is clamped to [-1400, 1400], beyond which exp(r) is 0 or overflows in
// f64 either way; fmin first, so that a NaN r gives exp(+1400) = inf, which
// makes the callers' results NaN.
__device__ __forceinline__ void exp_pair(double r, double &e, double &em, const FastConsts &K)
{
    r = __builtin_fmax(__builtin_fmin(r, K.clamp), -K.clamp);
    const double m = __builtin_rint(r * K.inv_ln2);
    double f = __builtin_fma(-m, K.ln2_hi, r);
    f = __builtin_fma(-m, K.ln2_lo, f);
    double P = K.expc[0];
#pragma unroll
    for (int j = 1; j < 10; ++j) P = __builtin_fma(P, f, K.expc[j]);
    const double p = f * P;                             // expm1(f)
    const double s = __builtin_ldexp(1.0, (int)m);
    e = __builtin_fma(s, p, s);
    em = __builtin_fma(-s, p, 1.0 - s);
}

// ---------------------------------------------------------------------------
// WaveRayPath::odes (wave_ray_path.rs:118-150) from the field values at the
// point: depth h (NaN where the bathymetry errs), its gradient, the current
// and its gradients, and okc = false where the current errs.

template <int MATH>
__device__ __forceinline__ void odes(const FastConsts &K, double h, double hx, double hy,
                                     bool okc, double u, double v, double ux, double uy,
                                     double vx, double vy, double kx, double ky, double &dx,
                                     double &dy, double &dkx, double &dky)
{
    if (MATH == MATH_LITERAL) {
        // A NaN depth (a provider error, or a NaN target or datum) makes
        // every derivative NaN through group_velocity and dkdt_bathy, as an
        // error does, except when k <= 0 errs first; both end all-NaN.
        bool ok = okc;
        const double k = sqrt(kx * kx + ky * ky);           // :133-134
        const double theta = atan2(ky, kx);
        // group_velocity :177-188: h <= 0 -> NaN (Ok); k <= 0 -> Err
        if (!(h <= 0.0) && k <= 0.0) ok = false;
        double cg;
        if (h <= 0.0) {
            cg = MR_NAN;
        } else {
            const double ch = cosh(k * h);
            cg = (G / 2.0) * ((tanh(k * h) + (k * h) / (ch * ch)) / sqrt(k * G * tanh(k * h)));
        }
        const double cgx = cg * cos(theta) + u;             // :138-139
        const double cgy = cg * sin(theta) + v;
        // dkdt_bathy :207-216
        const double c = (-0.5) * k / sinh(k * h) / cosh(k * h) * sqrt(G * k * tanh(k * h));
        const double dkxb = c * hx, dkyb = c * hy;
        dx = ok ? cgx : MR_NAN;
        dy = ok ? cgy : MR_NAN;
        dkx = ok ? (dkxb - kx * ux) - ky * vx : MR_NAN;      // :146-147
        dky = ok ? (dkyb - kx * uy) - ky * vy : MR_NAN;
    } else {
        // Every provider error, h <= 0 (or NaN), k <= 0 (or NaN) and a
        // non-finite k*h give all-NaN derivatives in the reference (an
        // error directly; the others through NaN/inf propagation in
        // group_velocity and dkdt_bathy).  Here a provider error poisons h
        // with NaN, which exp_pair's clamp turns into exp(+1400) = inf, so
        // X below is -inf or NaN; h <= 0 makes
        // X = g k (1-e)(1+e) <= 0 and k = 0 makes k = 0 * rsq(0) NaN: each
        // gives a NaN W and so NaN derivatives.  A NaN or infinite k*h
        // propagates through 4kh.
        //
        // With e = exp(-2kh) and W = 1/sqrt(g k (1-e)(1+e)), one reciprocal
        // square root gives every quotient of group_velocity and dkdt_bathy:
        //   tanh kh = (1-e)/(1+e),  sech^2 kh = 4e/(1+e)^2,
        //   1/sigma = (1+e) W,      1/(1+e) = g k (1-e) W^2,
        //   cg = (g/2)(tanh + kh sech^2)/sigma = (g/2) W ((1-e) + 4kh e/(1+e)),
        //   c = (-k/2) sigma/(sinh kh cosh kh) = -2 g k^2 (e/(1+e)) W,
        // the last since sinh kh cosh kh = (1-e)(1+e)/(4e).
        h = okc ? h : MR_NAN;
        const double k2 = __builtin_fma(kx, kx, ky * ky);
        const double rk = rsq_nr(k2);                       // 1/k
        const double k = k2 * rk;
        const double kh = k * h;
        double e, em;
        exp_pair(-2.0 * kh, e, em, K);                      // e = exp(-2kh), em = 1 - e
        const double ope = 1.0 + e;
        const double gke = (K.g * k) * em;
        const double W = rsq_nr(gke * ope);
        const double iope = gke * (W * W);                  // 1/(1+e)
        const double ei = e * iope;
        // cg cos(theta)/kx = cg/k, with (g/2)/k formed beside the chain
        const double cgr = ((K.half_g * rk) * W) * __builtin_fma(4.0 * kh, ei, em);
        const double c = (K.m2g * k2) * (ei * W);
        dx = __builtin_fma(cgr, kx, u);                     // cg cos(atan2(ky, kx)) + u
        dy = __builtin_fma(cgr, ky, v);
        // (c hx - kx ux) - ky vx with one rounding per term
        dkx = __builtin_fma(-ky, vx, __builtin_fma(-kx, ux, c * hx));
        dky = __builtin_fma(-ky, vy, __builtin_fma(-kx, uy, c * hy));
    }
}

// The gridded queries for a target inside the slot's cell and on none of
// its corners: the same operations as bathy_query / current_query take for
// such a target, with no branch and no load.  slow is set where that does
// not hold (a target outside the cell, a NaN target, a slot never loaded, or
// a corner), and those lanes must be answered by the full query.  In a cell
// with a corner past the data the record's NaN values give NaN results, as
// the full query's error does.
__device__ __forceinline__ double bathy_fast(const BathySlot &bs, double x, double y, double &hx,
                                             double &hy, bool &slow)
{
    const float px = (float)x, py = (float)y;
    const bool ok = bathy_in_cell(bs, px, py);
    const float dxa = px - bs.ax.x, dya = py - bs.ay.x;
    const float4 r0 = bs.r0, r1 = bs.r1;
    const float hv = bilin_eval(r0.x, r0.y, r0.z, r0.w, r1.z * dya, r1.w * dxa);
    const float dxd = px - bs.ax.y, dyb = py - bs.ay.y;
    slow = !ok | on_corner(dxa, dxd, dya, dyb);
    hx = (double)r1.x;
    hy = (double)r1.y;
    return __builtin_isnan(r1.z) ? MR_NAN : (double)hv;    // det == 0 errs
}

__device__ __forceinline__ bool current_fast(const CurrentSlot &cs, double x, double y, double &u,
                                             double &v, double &ux, double &uy, double &vxg,
                                             double &vyg, bool &slow)
{
    const bool ok = current_in_cell(cs, x, y);
    const float tx = (float)x, ty = (float)y;
    const float dxa = tx - cs.xp.x, dya = ty - cs.yp.x;
    const float uw = cs.m.x * dya;
    const float ww = cs.m.y * dxa;
    const float4 r0 = cs.r0, r1 = cs.r1;
    const float uf = bilin_eval(r0.x, r0.y, r0.z, r0.w, uw, ww);
    const float vf = bilin_eval(r1.x, r1.y, r1.z, r1.w, uw, ww);
    const float dxd = tx - cs.xp.y, dyb = ty - cs.yp.y;
    slow = !ok | on_corner(dxa, dxd, dya, dyb);
    u = (double)uf;
    v = (double)vf;
    ux = cs.g0.x; uy = cs.g0.y; vxg = cs.g1.x; vyg = cs.g1.y;
    return !__builtin_isnan(cs.m.x);                       // det == 0 errs
}

// ---------------------------------------------------------------------------
// Shared-cell records (SCell) in registers: eight 16-byte pieces, held as
// f64 pairs so that copies between slots move 64 bits per instruction.

struct SharedSlot {
    // a slot never loaded has the empty region (four -inf bounds)
    mr_d2 q[8] = {{MR_EMPTY_REGION, MR_EMPTY_REGION}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0},
                  {0.0, 0.0},                         {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
};

// the four f32 values of a 16-byte record piece.  The whole piece is bit-cast
// at once: hipcc 7.2 (clang 20) miscompiles a bit cast of one element of a
// by-value double vector, d.y, to a float pair, returning d.x's halves.
__device__ __forceinline__ mr_f4 f4of(mr_d2 d)
{
    return __builtin_bit_cast(mr_f4, d);
}

__device__ __forceinline__ void shared_load(const SCell *cells, SharedSlot &s, uint32_t cell)
{
    const MR_GLOBAL mr_d2 *p = (const MR_GLOBAL mr_d2 *)((const MR_GLOBAL char *)cells + cell * 128u);
#pragma unroll
    for (int k = 0; k < 8; ++k) s.q[k] = p[k];
}

// (px, py) in the slot's fast region; a NaN target is in none
__device__ __forceinline__ bool shared_in_region(const SharedSlot &s, float px, float py)
{
    // lo <= p <= hi exactly when med3(p, lo, hi) - p is zero (med3 returns
    // one of its operands, and with f32 denormals kept a difference of two
    // finite floats is zero only when they are equal); the two axes'
    // differences are tested at once, in vector ALU operations with one mask
    // at the end.  A NaN or infinite p makes a difference NaN or infinite, so
    // it is in no region; the empty region, lo = hi = -inf, holds no target.
    const mr_f4 b = f4of(s.q[0]);              // (lo_x, hi_x, lo_y, hi_y)
    const float dx = __builtin_amdgcn_fmed3f(px, b.x, b.y) - px;
    const float dy = __builtin_amdgcn_fmed3f(py, b.z, b.w) - py;
    return __builtin_fabsf(dx) + __builtin_fabsf(dy) == 0.0f;
}

// Both fields at (x, y) from the slot's record, for a target in its fast
// region: the operations bathy_query and current_query take there
// (interpolator.rs:39-84 with u = m01 tty, w = m10 ttx; the u and v pieces
// at the same (u, w)), and slow set where the target is outside the region.
__device__ __forceinline__ void shared_fast(const SharedSlot &s, double x, double y, double &h,
                                            double &hx, double &hy, double &u, double &v,
                                            double &ux, double &uy, double &vx, double &vy,
                                            bool &slow)
{
    const float px = (float)x, py = (float)y;
    slow = !shared_in_region(s, px, py);
    const mr_f4 om = f4of(s.q[1]);             // (x0, y0, m01, m10)
    const float tx = px - om.x, ty = py - om.y;
    const float uu = om.z * ty, ww = om.w * tx;
    const mr_f4 d = f4of(s.q[2]), uc = f4of(s.q[4]), vc = f4of(s.q[5]);
    const float hv = bilin_eval(d.x, d.y, d.z, d.w, uu, ww);
    const float uf = bilin_eval(uc.x, uc.y, uc.z, uc.w, uu, ww);
    const float vf = bilin_eval(vc.x, vc.y, vc.z, vc.w, uu, ww);
    h = (double)hv;
    hx = s.q[3].x;
    hy = s.q[3].y;
    u = (double)uf;
    v = (double)vf;
    ux = s.q[6].x; uy = s.q[6].y; vx = s.q[7].x; vy = s.q[7].y;
}

// Whether any lane of the wave has c, as a wave-uniform value the compiler
// cannot fold back into c (the empty asm).  A rare block guarded as
//   if (__builtin_expect(wave_any(c), 0)) { __asm__ volatile(""); if (c) ... }
// stays behind a scalar branch: when no lane has c, the wave passes it
// without writing EXEC, and the block is laid out after the loop, so the
// common path falls through (the volatile asm keeps the compiler from
// merging the two tests into one EXEC-masked branch over an inline block).
__device__ __forceinline__ bool wave_any(bool c)
{
    uint64_t m = __builtin_amdgcn_ballot_w64(c);
    __asm__("" : "+s"(m));
    return m != 0;
}

// WaveRayPath::system for gridded depth and current on shared cells: every
// lane evaluates from the record in hand; lanes outside its fast region
// evaluate again through the full queries (bathy_query, current_query),
// which find and load their cells' records and handle nodes and errors.
template <int MATH>
__device__ __forceinline__ void rhs_shared(const BathyView &B, const BathyPtrs &BP,
                                           const CurrentView &C, const CurrentPtrs &CP,
                                           const SharedSlot &ss, const FastConsts &K, double x,
                                           double y, double kx, double ky, double &dx,
                                           double &dy, double &dkx, double &dky)
{
    double h, hx, hy, u, v, ux, uy, vx, vy;
    bool slow;
    shared_fast(ss, x, y, h, hx, hy, u, v, ux, uy, vx, vy, slow);
    odes<MATH>(K, h, hx, hy, true, u, v, ux, uy, vx, vy, kx, ky, dx, dy, dkx, dky);
    if (__builtin_expect(wave_any(slow), 0)) {
        __asm__ volatile("");
        if (slow) {
            BathySlot bs;
            CurrentSlot cs;
            const double h2 = bathy_query<BK_GRID>(B, BP, bs, x, y, hx, hy);
            const bool okc = current_query<CK_GRID>(C, CP, cs, x, y, u, v, ux, uy, vx, vy);
            odes<MATH>(K, h2, hx, hy, okc, u, v, ux, uy, vx, vy, kx, ky, dx, dy, dkx, dky);
        }
    }
}

// WaveRayPath::system (wave_ray_path.rs:220-234) around odes (:118-150).
// With a gridded field, every lane first evaluates on the fast path (one
// block of straight-line code in which the compiler can interleave the two
// interpolations with the dispersion arithmetic); the lanes it cannot
// answer then evaluate again with the full queries, which reload records
// and read corner values.
template <int BK, int CK, int MATH>
__device__ __forceinline__ void rhs(const BathyView &B, const BathyPtrs &BP, BathySlot &bs,
                                    const CurrentView &C, const CurrentPtrs &CP,
                                    CurrentSlot &cs, const FastConsts &K, double x, double y,
                                    double kx, double ky, double &dx, double &dy, double &dkx,
                                    double &dky)
{
    double hx, hy, u, v, ux, uy, vx, vy;
    if (BK == BK_GRID || CK == CK_GRID) {
        bool sb = false, sc = false;
        const double h = BK == BK_GRID ? bathy_fast(bs, x, y, hx, hy, sb)
                                       : bathy_query<BK>(B, BP, bs, x, y, hx, hy);
        const bool okc = CK == CK_GRID ? current_fast(cs, x, y, u, v, ux, uy, vx, vy, sc)
                                       : current_query<CK>(C, CP, cs, x, y, u, v, ux, uy, vx, vy);
        odes<MATH>(K, h, hx, hy, okc, u, v, ux, uy, vx, vy, kx, ky, dx, dy, dkx, dky);
        const bool slow = sb | sc;
        if (__builtin_expect(wave_any(slow), 0)) {
            __asm__ volatile("");
            if (slow) {
                const double h2 = bathy_query<BK>(B, BP, bs, x, y, hx, hy);
                const bool okc2 = current_query<CK>(C, CP, cs, x, y, u, v, ux, uy, vx, vy);
                odes<MATH>(K, h2, hx, hy, okc2, u, v, ux, uy, vx, vy, kx, ky, dx, dy, dkx, dky);
            }
        }
        return;
    }
    const double h = bathy_query<BK>(B, BP, bs, x, y, hx, hy);       // odes :120-127
    const bool okc = current_query<CK>(C, CP, cs, x, y, u, v, ux, uy, vx, vy);   // :130
    odes<MATH>(K, h, hx, hy, okc, u, v, ux, uy, vx, vy, kx, ky, dx, dy, dkx, dky);
}

// all four NaN: fmin returns its non-NaN operand, so the minimum of four is
// NaN only when all are (the values tested here are results of arithmetic,
// never signalling NaNs)
__device__ __forceinline__ double min4(double a, double b, double c, double d)
{
    return __builtin_fmin(__builtin_fmin(a, b), __builtin_fmin(c, d));
}

__device__ __forceinline__ bool all_nan4(double a, double b, double c, double d)
{
    return __builtin_isnan(min4(a, b, c, d));
}

}  // namespace mr
