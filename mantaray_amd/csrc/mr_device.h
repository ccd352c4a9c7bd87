// mr_device.h -- device-side field views, field queries and the ray-equation
// right-hand side for gfx950.
//
// Citations are to files of mines-oceanography/mantaray 0.1.0.
//
// Data layout in HBM (built once per field and device by build_bathy_kernel and
// build_current_kernel in mr_kernels.hip):
//   gridded bathymetry: one 32-byte record per cell (i, j), i < nx-1,
//     j < ny-1, row-major with stride nx-1, holding the f32 bilinear
//     coefficients of the cell, the f32 cell gradient and the cell's
//     change-of-basis entries; plus the f32 axes as (a[i], a[i+1]) pairs
//     and the raw f64 depth (read only when the target is exactly a node).
//   gridded current: one 64-byte record per cell with the f32 bilinear
//     coefficients of u and v and the four f64 cell gradients; pairs of the
//     f32 casts of the f64 axes; the change of basis as two constants when the f32 axes
//     are uniform, else a per-cell float2 table; the raw f64 u and v.
// A query loads one record per field (dwordx4 loads) and the two axis pairs
// around the cell instead of 12 scattered corner values, and performs no
// division: everything cell-constant is precomputed with the reference's own
// operations, so the interpolated values are bitwise those of the reference.
// Cells on which the reference errs are recognised without extra loads (see
// the interpolation notes below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mr {

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

struct BathyView {
    int kind;
    float h;                              // ConstantDepth
    float h0, x0, y0, dhdx, dhdy;         // ConstantSlope
    const float *arr; int64_t an;         // ArrayDepth (an x an)
    int nx, ny;                           // CartesianNetcdf3
    float ax0, ay0, nxm1, nym1, nxm2, nym2;   // a[0]; (f32)(n-1); (f32)(n-2)
    double inv_sx, inv_sy;                // RN64(1/(double)|a1-a0|)
    const float2 *xp, *yp;                // axis pairs (a[i], a[i+1]), i < n-1
    const BCell *cells;
    int64_t oob_from;                     // first cell with a corner past the data
    const double *raw;                    // raw depth (exact-corner values)
};

struct CurrentView {
    int kind;
    double u, v;                          // ConstantCurrent
    int nx, ny;                           // CartesianCurrent
    double cx0, cy0, sx, sy, inv_sx, inv_sy;
    double nxm1, nym1, nxm2, nym2;        // n-1, n-2
    const float2 *xp, *yp;                // pairs of the (f32) casts of the f64 axes
    const CCell *cells;
    int64_t oob_from;                     // first cell with a corner past u or v
    const float2 *mtab;                   // per-cell (m01, m10) or null
    float m01, m10;                       // uniform axes
    const double *raw_u, *raw_v;          // raw values (exact-corner values)
};

// Read-only device memory seen through the constant address space, so that
// uniform loads from it are scalar (s_load) loads.
#define MR_CONST __attribute__((address_space(4)))

#define MR_NAN __builtin_nan("")
#define MR_NANF __builtin_nanf("")

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

__device__ __forceinline__ float bilin_eval(float a00, float a10, float a01, float a11,
                                            float u, float w)
{
    return ((a00 + a10 * u) + a01 * w) + (a11 * u) * w;   // interpolator.rs:83
}

// index of the first corner (order A, B, C, D) equal to the target
__device__ __forceinline__ int64_t corner_index(float px, float py, float xa, float xd,
                                                float ya, float yb, int64_t nx, int i1,
                                                int j1)
{
    const int64_t iA = nx * j1 + i1;
    if (px == xa) return py == ya ? iA : iA + nx;          // A, B
    return py == yb ? iA + nx + 1 : iA + 1;                // C, D
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
// when it lies that close to an integer is the exact division taken.
// ok is cleared where the reference returns Err (a NaN index errs later, in
// bilinear, because all four corners collapse onto (0, 0)); i1 is always a
// valid cell index.
__device__ __forceinline__ int index64(double p, double a0, double sp, double inv_sp,
                                       double nm1, double nm2, bool &ok)
{
    const double t = p - a0;
    double q = t * inv_sp;
    double fl = floor(q);
    const double fr = q - fl;
    const double tol = q * 0x1p-50;
    if (fr < tol || fr > 1.0 - tol) {
        q = t / sp;
        fl = floor(q);
    }
    ok = ok && t >= 0.0 && q <= nm1;
    return (int)__builtin_fmin(__builtin_fmax(fl, 0.0), nm2);
}

// ---------------------------------------------------------------------------
// bathymetry: BathymetryData::depth_and_gradient (bathymetry/mod.rs:35-41)

template <int BK>
__device__ __forceinline__ bool bathy_query(const MR_CONST BathyView *B, double x, double y, double &h,
                                            double &hx, double &hy)
{
    const float px = (float)x, py = (float)y;             // wave_ray_path.rs:122
    if (BK == BK_CONSTANT) {                               // constant_depth.rs:34-40
        bool nn = __builtin_isnan(px) || __builtin_isnan(py);
        h = nn ? MR_NAN : (double)B->h;
        hx = nn ? MR_NAN : 0.0;
        hy = nn ? MR_NAN : 0.0;
        return true;
    } else if (BK == BK_SLOPE) {                           // constant_slope.rs:63-74
        bool nn = __builtin_isnan(px) || __builtin_isnan(py);
        float hv = (B->h0 + B->dhdx * (px - B->x0)) + B->dhdy * (py - B->y0);
        h = nn ? MR_NAN : (double)hv;
        hx = nn ? MR_NAN : (double)B->dhdx;
        hy = nn ? MR_NAN : (double)B->dhdy;
        return true;
    } else if (BK == BK_ARRAY) {                           // array_depth.rs:26-33
        // Rust `as usize`: saturating, NaN -> 0
        uint64_t xi = px > 0.0f ? (px >= 1.8446744e19f ? UINT64_MAX : (uint64_t)px) : 0;
        uint64_t yi = py > 0.0f ? (py >= 1.8446744e19f ? UINT64_MAX : (uint64_t)py) : 0;
        if (xi >= (uint64_t)B->an || yi >= (uint64_t)B->an) {
            h = hx = hy = MR_NAN;
            return true;
        }
        h = (double)B->arr[xi * (uint64_t)B->an + yi];
        hx = 0.0; hy = 0.0;
        return true;
    } else {                                               // cartesian_netcdf3.rs:98-135
        // predicated: every load is in bounds, errors only clear ok
        const bool nan_pt = __builtin_isnan(px) || __builtin_isnan(py);
        const float ix = index32(px - B->ax0, B->inv_sx);  // nearest :274-296
        const float iy = index32(py - B->ay0, B->inv_sy);
        bool ok = ix >= 0.0f && ix <= B->nxm1 && iy >= 0.0f && iy <= B->nym1;
        // four_corners :334-390: floor, or n-2 on the last node
        const int i1 = (int)__builtin_fminf(__builtin_fmaxf(ix, 0.0f), B->nxm2);
        const int j1 = (int)__builtin_fminf(__builtin_fmaxf(iy, 0.0f), B->nym2);
        int64_t cell = (int64_t)j1 * (B->nx - 1) + i1;
        ok = ok && cell < B->oob_from;
        cell = ok ? cell : 0;
        const float4 *rec = reinterpret_cast<const float4 *>(B->cells + cell);
        const float4 r0 = rec[0], r1 = rec[1];
        const float2 xp = B->xp[i1];                      // x[i1], x[i1+1]
        const float2 yq = B->yp[j1];
        const float ya = yq.x, yb = yq.y;
        float hv = bilin_eval(r0.x, r0.y, r0.z, r0.w, r1.z * (py - ya), r1.w * (px - xp.x));
        const bool corner = (px == xp.x || px == xp.y) && (py == ya || py == yb);
        if (corner && ok)
            hv = (float)B->raw[corner_index(px, py, xp.x, xp.y, ya, yb, B->nx, i1, j1)];
        ok = ok && (corner || !__builtin_isnan(r1.z));     // det == 0 errs
        h = nan_pt ? MR_NAN : (double)hv;
        hx = nan_pt ? MR_NAN : (double)r1.x;
        hy = nan_pt ? MR_NAN : (double)r1.y;
        return ok || nan_pt;
    }
}

// ---------------------------------------------------------------------------
// current: CurrentData::current_and_gradient (current/mod.rs:21-31)

template <int CK>
__device__ __forceinline__ bool current_query(const MR_CONST CurrentView *C, double x, double y,
                                              double &u, double &v, double &ux, double &uy,
                                              double &vx, double &vy)
{
    if (CK == CK_CONSTANT) {                               // constant_current.rs:66-76
        u = C->u; v = C->v;
        ux = 0.0; uy = 0.0; vx = 0.0; vy = 0.0;
        return true;
    } else {                                               // cartesian_current.rs:487-542
        bool ok = true;
        const int i1 = index64(x, C->cx0, C->sx, C->inv_sx, C->nxm1, C->nxm2, ok);
        const int j1 = index64(y, C->cy0, C->sy, C->inv_sy, C->nym1, C->nym2, ok);
        int64_t cell = (int64_t)j1 * (C->nx - 1) + i1;
        ok = ok && cell < C->oob_from;
        cell = ok ? cell : 0;
        const float4 *rec = reinterpret_cast<const float4 *>(C->cells + cell);
        const float4 r0 = rec[0], r1 = rec[1];
        const double2 g0 = reinterpret_cast<const double2 *>(rec + 2)[0];
        const double2 g1 = reinterpret_cast<const double2 *>(rec + 2)[1];
        float m01 = C->m01, m10 = C->m10;
        if (C->mtab) {
            const float2 m = C->mtab[cell];
            m01 = m.x; m10 = m.y;
        }
        const float tx = (float)x, ty = (float)y;          // cartesian_current.rs:505-512
        const float2 xp = C->xp[i1];
        const float2 yq = C->yp[j1];
        const float ya = yq.x, yb = yq.y;
        const float uw = m01 * (ty - ya);
        const float ww = m10 * (tx - xp.x);
        float uf = bilin_eval(r0.x, r0.y, r0.z, r0.w, uw, ww);
        float vf = bilin_eval(r1.x, r1.y, r1.z, r1.w, uw, ww);
        const bool corner = (tx == xp.x || tx == xp.y) && (ty == ya || ty == yb);
        if (corner && ok) {
            const int64_t ic = corner_index(tx, ty, xp.x, xp.y, ya, yb, C->nx, i1, j1);
            uf = (float)C->raw_u[ic];
            vf = (float)C->raw_v[ic];
        }
        ok = ok && (corner || !__builtin_isnan(m01));       // det == 0 errs
        u = (double)uf;
        v = (double)vf;
        ux = g0.x; uy = g0.y; vx = g1.x; vy = g1.y;
        return ok;
    }
}

// ---------------------------------------------------------------------------
// f64 helpers for the fast dispersion evaluation

// 1/x: hardware reciprocal estimate refined by two Newton steps.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/sqrt(x): hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double t = x * y;
    double e = __builtin_fma(-t, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    t = x * y;
    e = __builtin_fma(-t, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

// e = exp(-r) and em = 1 - exp(-r) for r >= 0, each with small relative
// error, from one reduced-argument evaluation: r = n ln2 + f, |f| <= ln2/2,
// expm1(-f) by its Taylor series through degree 13 (truncation below
// 2^-56 relative on that interval), e = 2^-n (1 + p), em = (1 - 2^-n) -
// 2^-n p, which has no cancellation for n >= 1 and is exactly -p for n = 0.
__device__ __forceinline__ void exp_pair(double r, double &e, double &em)
{
    const double INV_LN2 = 1.4426950408889634;          // 1/ln 2
    const double LN2_HI = 6.93147180369123816490e-01;  // ln 2 split, upper part
    const double LN2_LO = 1.90821492927058770002e-10;  // ln 2 - LN2_HI
    r = r < 1400.0 ? r : 1400.0;                        // exp(-1400) underflows to 0
    const double n = __builtin_rint(r * INV_LN2);
    double f = __builtin_fma(-n, LN2_HI, r);
    f = __builtin_fma(-n, LN2_LO, f);
    const double g = -f;
    // P(g) = sum_{j=0..12} g^j / (j+1)!,  expm1(g) = g * P(g)
    double P = 1.6059043836821614e-10;                  // 1/13!
    P = __builtin_fma(P, g, 2.08767569878680989e-09);   // 1/12!
    P = __builtin_fma(P, g, 2.50521083854417188e-08);   // 1/11!
    P = __builtin_fma(P, g, 2.75573192239858907e-07);   // 1/10!
    P = __builtin_fma(P, g, 2.75573192239858907e-06);   // 1/9!
    P = __builtin_fma(P, g, 2.48015873015873016e-05);   // 1/8!
    P = __builtin_fma(P, g, 1.98412698412698413e-04);   // 1/7!
    P = __builtin_fma(P, g, 1.38888888888888889e-03);   // 1/6!
    P = __builtin_fma(P, g, 8.33333333333333333e-03);   // 1/5!
    P = __builtin_fma(P, g, 4.16666666666666667e-02);   // 1/4!
    P = __builtin_fma(P, g, 1.66666666666666667e-01);   // 1/3!
    P = __builtin_fma(P, g, 0.5);                       // 1/2!
    P = __builtin_fma(P, g, 1.0);                       // 1/1!
    const double p = g * P;
    const double s = __builtin_ldexp(1.0, -(int)n);
    e = __builtin_fma(s, p, s);
    em = __builtin_fma(-s, p, 1.0 - s);
}

// ---------------------------------------------------------------------------
// WaveRayPath::system (wave_ray_path.rs:220-234) around odes (:118-150)

template <int BK, int CK, int MATH>
__device__ __forceinline__ void rhs(const MR_CONST BathyView *B, const MR_CONST CurrentView *C, double x,
                                    double y, double kx, double ky, double &dx, double &dy,
                                    double &dkx, double &dky)
{
    double h, hx, hy, u, v, ux, uy, vx, vy;
    bool ok = bathy_query<BK>(B, x, y, h, hx, hy);         // odes :120-127
    ok = current_query<CK>(C, x, y, u, v, ux, uy, vx, vy) && ok;   // :130

    if (MATH == MATH_LITERAL) {
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
        // non-finite k*h give all-NaN derivatives in the reference (an error
        // directly; the others through NaN/inf propagation in
        // group_velocity and dkdt_bathy), so one mask covers them.
        const double k2 = kx * kx + ky * ky;
        const double rk = rsq_nr(k2);
        const double k = k2 * rk;
        const double kh = k * h;
        const bool bad = !ok || !(h > 0.0) || !(k > 0.0) || !(kh < __builtin_inf());
        const double cth = kx * rk, sth = ky * rk;          // cos, sin of atan2(ky, kx)
        double e, em;
        exp_pair(2.0 * kh, e, em);                          // e = exp(-2kh)
        const double ope = 1.0 + e;
        const double D = rcp_nr(em * ope);                  // 1/(em (1+e))
        const double iope = em * D;                         // 1/(1+e)
        const double T = em * iope;                         // tanh(kh)
        const double sech2 = (4.0 * e) * (iope * iope);     // 1/cosh^2(kh)
        const double gkT = (G * k) * T;
        const double rs = rsq_nr(gkT);
        const double sigma = gkT * rs;                      // sqrt(g k tanh kh)
        const double cg = (0.5 * G) * (T + kh * sech2) * rs;
        // (-k/2)/(sinh cosh) * sigma = -2 k e sigma / (em (1+e))
        const double c = (-2.0 * k) * e * sigma * D;
        dx = bad ? MR_NAN : cg * cth + u;
        dy = bad ? MR_NAN : cg * sth + v;
        dkx = bad ? MR_NAN : (c * hx - kx * ux) - ky * vx;
        dky = bad ? MR_NAN : (c * hy - kx * uy) - ky * vy;
    }
}

__device__ __forceinline__ bool all_nan4(double a, double b, double c, double d)
{
    return __builtin_isnan(a) && __builtin_isnan(b) && __builtin_isnan(c) && __builtin_isnan(d);
}

}  // namespace mr
