// mr_device64.h -- the fp64-interpolation mode (MR_PREC_FP64_INTERP): field
// queries, the right-hand side, and its vector-Jacobian product, for the
// differentiable path.
//
// The reference interpolates in f32 (interpolator.rs:39-84 over f32 corner
// values), so its trajectories are piecewise constant in the field values at
// the scale of an f32 ulp and have no useful derivative with respect to
// them.  This mode keeps the reference's algorithm (the same fractional
// index, cell rule, bilinear formula with its corner and det == 0 exits, and
// cell-difference gradients) but evaluates all of it in f64 on the f64 values,
// with the bathymetry queried at the f64 point instead of its f32 rounding:
//
//   index   (p - a0) / |a1 - a0|         (cartesian_current.rs:244, in f64)
//   cell    floor, or n-2 on the last node (cartesian_current.rs:283-339)
//   value   ((a00 + a10 u) + a01 w) + (a11 u) w, u = m01 (y - y_j),
//           w = m10 (x - x_i)               (interpolator.rs:59-83, in f64)
//   slopes  (D - A) / (x1 - x0), (B - A) / (y1 - y0)
//
// Field values come from kernel arguments (pointers to the raw f64 arrays),
// so a caller can differentiate with respect to them and change them between
// calls without rebuilding anything; the per-cell change of basis (m01, m10)
// depends only on the axes and is tabulated once per field.
//
// Linearisation.  With u, w the bilinear fractions of the query, the value is
//   v = wA A + wB B + wC C + wD D,  wA = (1-u)(1-w), wB = u(1-w), wC = u w,
//   wD = (1-u) w,
// its derivative along x is (a01 + a11 u) m10 and along y (a10 + a11 w) m01;
// the cell-difference slopes are constant inside a cell.  The right-hand
// side (wave_ray_path.rs:118-150) with k = |k|, e = k/k, q = k h,
// T = tanh q, S = sech^2 q, sigma = sqrt(g k T):
//   dx/dt = cg e + U,   cg = (g/2)(T + q S)/sigma
//   dk/dt = c grad h - (grad U)^T k,   c = -k sigma / sinh(2q)
// has the partial derivatives
//   d sigma/dk = cg,  d sigma/dh = (g/2) k^2 S / sigma
//   N = T + q S,  dN/dq = 2 S (1 - q T)
//   d cg/dk = (g/2)(h N_q - N cg/sigma)/sigma
//   d cg/dh = (g/2)(k N_q - N sigma_h/sigma)/sigma
//   d c/dk  = c (1/k + cg/sigma - 2 h coth 2q)
//   d c/dh  = c (sigma_h/sigma - 2 k coth 2q),  coth 2q = (T + 1/T)/2
// and d(cg e)/dk = cg_k e e^T + (cg/k)(I - e e^T).
#pragma once

#include "mr_device.h"

namespace mr {

// Numeric parameters of a field in fp64 mode (per lane, in VGPRs).
struct Grid64View {
    int kind;                             // BK_* for a bathymetry, CK_* for a current
    int nx, ny, ncx;                      // grid; ncx = nx - 1
    int an;                               // ArrayDepth (an x an)
    uint32_t oob_from;                    // first cell (row-major) with a corner past the data
    double a0x, a0y;                      // first coordinate of each axis
    double sx, sy, inv_sx, inv_sy;        // |a1 - a0| and RN(1/|a1 - a0|)
    double nxm1, nym1;                    // n - 1
    double xspace, yspace;                // signed a1 - a0 (the slope denominators)
    double rxs, rys;                      // 1/xspace, 1/yspace (slope weights)
    double h, h0, x0, y0, dhdx, dhdy;     // ConstantDepth / ConstantSlope, widened
    double u, v;                          // ConstantCurrent
};

// Device arrays of a field in fp64 mode (kernel arguments).
struct Grid64Ptrs {
    const double *ax, *ay;                // axes in f64 (a bathymetry's f32 axes widened)
    const double2 *m;                     // per cell, row-major: (m01, m10); NaN m01 marks det == 0
    const float *arr;                     // ArrayDepth values
};

// A gridded value at a point with what its derivatives need.
struct Val64 {
    double val;                           // interpolated value
    double gx, gy;                        // cell-difference slopes
    double dx, dy;                        // derivative of the interpolant along x, y
};

// The cell of a query: flat index of corner A (B = A + nx, C = A + nx + 1,
// D = A + 1) and the bilinear fractions.
struct Cell64 {
    uint32_t iA;
    double u, w;
};

// Cell of (x, y), or ok = false where the reference's index or corner
// lookup errs (a NaN target included); i1 = j1 = 0 then.
__device__ __forceinline__ bool locate64(const Grid64View &G, double x, double y, int &i1,
                                         int &j1)
{
    bool ok = true;
    const double fx = index64(x, G.a0x, G.sx, G.inv_sx, G.nxm1, ok);
    const double fy = index64(y, G.a0y, G.sy, G.inv_sy, G.nym1, ok);
    i1 = ok ? __builtin_elementwise_min((int)fx, G.nx - 2) : 0;
    j1 = ok ? __builtin_elementwise_min((int)fy, G.ny - 2) : 0;
    ok = ok & ((uint32_t)j1 * (uint32_t)G.ncx + (uint32_t)i1 < G.oob_from);
    i1 = ok ? i1 : 0;
    j1 = ok ? j1 : 0;
    return ok;
}

// interpolator::bilinear in f64 on cell (i1, j1) for the value array val,
// with its corner exit: a target on a node returns that corner (first of
// A, B, C, D).  (The det == 0 exit is the caller's: m01 is NaN there.)  The
// corners are read only for a valid cell (ok); the result is NaN otherwise.
__device__ __forceinline__ Val64 interp64(const Grid64View &G, const double *val, bool ok,
                                          double x, double y, double xa, double xd, double ya,
                                          double yb, double m01, double m10, const Cell64 &c)
{
    double vA = MR_NAN, vB = MR_NAN, vC = MR_NAN, vD = MR_NAN;
    if (ok) {                          // the corners exist only in a valid cell
        vA = val[c.iA]; vB = val[c.iA + (uint32_t)G.nx];
        vC = val[c.iA + (uint32_t)G.nx + 1u]; vD = val[c.iA + 1u];
    }
    const double a10 = vB - vA, a01 = vD - vA;
    const double a11 = ((vC - vA) - a10) - a01;
    Val64 r;
    r.val = ((vA + a10 * c.u) + a01 * c.w) + (a11 * c.u) * c.w;
    const bool ex = x == xa, exd = x == xd, eya = y == ya, eyb = y == yb;
    if ((ex | exd) & (eya | eyb))
        r.val = ex ? (eya ? vA : vB) : (eyb ? vC : vD);
    r.gx = (vD - vA) / G.xspace;
    r.gy = (vB - vA) / G.yspace;
    r.dx = (a01 + a11 * c.u) * m10;
    r.dy = (a10 + a11 * c.w) * m01;
    return r;
}

// Corner hit test of interp64 (target exactly on a node of the cell).
__device__ __forceinline__ bool on_node64(double x, double y, double xa, double xd, double ya,
                                          double yb)
{
    return ((x == xa) | (x == xd)) & ((y == ya) | (y == yb));
}

// Bathymetry in fp64 mode.  Returns ok = false where the reference errs;
// h is NaN (not an error) for a NaN point (cartesian_netcdf3.rs:102-105) and
// outside an ArrayDepth.  cell.iA is meaningful for BK_GRID only.
template <int BK>
__device__ __forceinline__ bool bathy64(const Grid64View &B, const Grid64Ptrs &P,
                                        const double *depth, double x, double y, Val64 &h,
                                        Cell64 &cell)
{
    cell.iA = 0; cell.u = 0.0; cell.w = 0.0;
    h.dx = 0.0; h.dy = 0.0;
    const bool nn = __builtin_isnan(x) || __builtin_isnan(y);
    if (BK == BK_CONSTANT) {                               // constant_depth.rs:34-40
        h.val = nn ? MR_NAN : B.h;
        h.gx = 0.0; h.gy = 0.0;
        return true;
    } else if (BK == BK_SLOPE) {                           // constant_slope.rs:63-74
        h.val = nn ? MR_NAN : (B.h0 + B.dhdx * (x - B.x0)) + B.dhdy * (y - B.y0);
        h.gx = B.dhdx; h.gy = B.dhdy;
        h.dx = B.dhdx; h.dy = B.dhdy;
        return true;
    } else if (BK == BK_ARRAY) {                           // array_depth.rs:26-33
        const uint64_t xi = x > 0.0 ? (x >= 1.8446744073709552e19 ? UINT64_MAX : (uint64_t)x) : 0;
        const uint64_t yi = y > 0.0 ? (y >= 1.8446744073709552e19 ? UINT64_MAX : (uint64_t)y) : 0;
        h.gx = 0.0; h.gy = 0.0;
        if (xi >= (uint64_t)B.an || yi >= (uint64_t)B.an) {
            h.val = MR_NAN;
            return true;
        }
        h.val = (double)P.arr[xi * (uint64_t)B.an + yi];
        return true;
    } else {
        int i1, j1;
        bool ok = locate64(B, x, y, i1, j1);
        const double xa = P.ax[i1], xd = P.ax[i1 + 1], ya = P.ay[j1], yb = P.ay[j1 + 1];
        const double2 m = P.m[(uint32_t)j1 * (uint32_t)B.ncx + (uint32_t)i1];
        cell.iA = (uint32_t)B.nx * (uint32_t)j1 + (uint32_t)i1;
        cell.u = m.x * (y - ya);
        cell.w = m.y * (x - xa);
        h = interp64(B, depth, ok, x, y, xa, xd, ya, yb, m.x, m.y, cell);
        ok = ok & (on_node64(x, y, xa, xd, ya, yb) | !__builtin_isnan(m.x));
        if (nn) {
            h.val = MR_NAN; h.gx = MR_NAN; h.gy = MR_NAN;
            ok = true;
        }
        return ok;
    }
}

// Current in fp64 mode; ok = false where the reference errs (no NaN guard:
// a NaN point errs, cartesian_current.rs:492-495).
template <int CK>
__device__ __forceinline__ bool current64(const Grid64View &C, const Grid64Ptrs &P,
                                          const double *cu, const double *cv, double x, double y,
                                          Val64 &u, Val64 &v, Cell64 &cell)
{
    cell.iA = 0; cell.u = 0.0; cell.w = 0.0;
    if (CK == CK_CONSTANT) {                               // constant_current.rs:66-76
        u.val = C.u; v.val = C.v;
        u.gx = u.gy = u.dx = u.dy = 0.0;
        v.gx = v.gy = v.dx = v.dy = 0.0;
        return true;
    } else {
        int i1, j1;
        bool ok = locate64(C, x, y, i1, j1);
        const double xa = P.ax[i1], xd = P.ax[i1 + 1], ya = P.ay[j1], yb = P.ay[j1 + 1];
        const double2 m = P.m[(uint32_t)j1 * (uint32_t)C.ncx + (uint32_t)i1];
        cell.iA = (uint32_t)C.nx * (uint32_t)j1 + (uint32_t)i1;
        cell.u = m.x * (y - ya);
        cell.w = m.y * (x - xa);
        u = interp64(C, cu, ok, x, y, xa, xd, ya, yb, m.x, m.y, cell);
        v = interp64(C, cv, ok, x, y, xa, xd, ya, yb, m.x, m.y, cell);
        ok = ok & (on_node64(x, y, xa, xd, ya, yb) | !__builtin_isnan(m.x));
        return ok;
    }
}

// Field values of one fp64 evaluation.
struct Fields64 {
    const double *depth, *u, *v;
};

// WaveRayPath::system in fp64 mode.
template <int BK, int CK, int MATH>
__device__ __forceinline__ void rhs64(const Grid64View &B, const Grid64Ptrs &BP,
                                      const Grid64View &C, const Grid64Ptrs &CP,
                                      const Fields64 &F, const FastConsts &K, double x, double y,
                                      double kx, double ky, double &dx, double &dy, double &dkx,
                                      double &dky)
{
    Val64 h, u, v;
    Cell64 bc, cc;
    const bool okb = bathy64<BK>(B, BP, F.depth, x, y, h, bc);
    const bool okc = current64<CK>(C, CP, F.u, F.v, x, y, u, v, cc);
    odes<MATH>(K, okb ? h.val : MR_NAN, h.gx, h.gy, okc, u.val, v.val, u.gx, u.gy, v.gx, v.gy,
               kx, ky, dx, dy, dkx, dky);
}

// ---------------------------------------------------------------------------
// Adjoint accumulation into a value grid.  A ray stays in one cell for many
// evaluations, so each lane sums its contributions to the four corners of
// the cell in hand and adds them to the gradient grid (hardware f64 atomics)
// only when it moves to another cell, and once at the end.

struct CornerAcc {
    uint32_t iA = 0xFFFFFFFFu;            // no cell yet
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
};

__device__ __forceinline__ void acc_flush(CornerAcc &s, double *grad, uint32_t nx)
{
    if (s.iA != 0xFFFFFFFFu) {
        unsafeAtomicAdd(grad + s.iA, s.a);
        unsafeAtomicAdd(grad + s.iA + nx, s.b);
        unsafeAtomicAdd(grad + s.iA + nx + 1u, s.c);
        unsafeAtomicAdd(grad + s.iA + 1u, s.d);
    }
    s.a = 0.0; s.b = 0.0; s.c = 0.0; s.d = 0.0;
}

// Adds the cotangent of one query to the corners of its cell: vbar on the
// interpolated value, gxbar and gybar on the two cell-difference slopes.
__device__ __forceinline__ void acc_add(CornerAcc &s, double *grad, uint32_t nx,
                                        const Cell64 &cell, double rxs, double rys, double vbar,
                                        double gxbar, double gybar)
{
    if (cell.iA != s.iA) {
        acc_flush(s, grad, nx);
        s.iA = cell.iA;
    }
    const double u = cell.u, w = cell.w;
    const double gx = gxbar * rxs, gy = gybar * rys;
    s.a += vbar * ((1.0 - u) * (1.0 - w)) - gx - gy;
    s.b += vbar * (u * (1.0 - w)) + gy;
    s.c += vbar * (u * w);
    s.d += vbar * ((1.0 - u) * w) + gx;
}

// Gradient grids of an adjoint sweep (null: not wanted).
struct Grads64 {
    double *depth, *u, *v;
};

struct Accs64 {
    CornerAcc h, u, v;
};

// Vector-Jacobian product of rhs64 at (x, y, kx, ky): given the cotangent
// (g0, g1, g2, g3) of (dx/dt, dy/dt, dkx/dt, dky/dt), adds the state
// cotangent to (bx, by, bkx, bky) and the field-value cotangents to the
// accumulators.  The dispersion terms are evaluated with the closed forms of
// the fast mode whatever the forward's math mode: only the derivative
// depends on them.  The state must be one at which rhs64 is finite.
template <int BK, int CK>
__device__ __forceinline__ void rhs64_vjp(const Grid64View &B, const Grid64Ptrs &BP,
                                          const Grid64View &C, const Grid64Ptrs &CP,
                                          const Fields64 &F, const Grads64 &Gr, Accs64 &acc,
                                          const FastConsts &K, double x, double y, double kx,
                                          double ky, double g0, double g1, double g2, double g3,
                                          double &bx, double &by, double &bkx, double &bky)
{
    Val64 h, u, v;
    Cell64 bc, cc;
    const bool okb = bathy64<BK>(B, BP, F.depth, x, y, h, bc);
    const bool okc = current64<CK>(C, CP, F.u, F.v, x, y, u, v, cc);
    const double hh = h.val;
    const double k2 = __builtin_fma(kx, kx, ky * ky);
    const double rk = rsq_nr(k2);                       // 1/k
    const double k = k2 * rk;
    const double ex = kx * rk, ey = ky * rk;
    const double q = k * hh;
    double e, em;
    exp_pair(2.0 * q, e, em, K);                        // e = exp(-2q), em = 1 - e
    const double ope = 1.0 + e;
    const double D = rcp_nr(em * ope);
    const double iope = em * D;
    const double T = em * iope;                         // tanh q
    const double S = (4.0 * e) * (iope * iope);         // sech^2 q
    const double invT = (ope * ope) * D;                // 1/tanh q
    const double gkT = (K.g * k) * T;
    const double rs = rsq_nr(gkT);                      // 1/sigma
    const double sigma = gkT * rs;
    const double N = __builtin_fma(q, S, T);
    const double cg = K.half_g * N * rs;
    const double c = (-2.0 * k) * e * sigma * D;
    const double sig_h = K.half_g * (k * k) * S * rs;   // d sigma / dh
    const double Nq = 2.0 * S * (1.0 - q * T);
    const double cg_k = K.half_g * rs * (hh * Nq - N * cg * rs);
    const double cg_h = K.half_g * rs * (k * Nq - N * sig_h * rs);
    const double coth2 = 0.5 * (T + invT);
    const double c_k = c * (rk + cg * rs - 2.0 * hh * coth2);
    const double c_h = c * (sig_h * rs - 2.0 * k * coth2);

    const double P = g0 * ex + g1 * ey;                 // cotangent of cg
    const double cbar = g2 * h.gx + g3 * h.gy;          // cotangent of c
    const double hbar = cg_h * P + c_h * cbar;
    const double tk = cg * rk;
    bkx += (cg_k * P) * ex + tk * (g0 - P * ex) + (c_k * cbar) * ex - g2 * u.gx - g3 * u.gy;
    bky += (cg_k * P) * ey + tk * (g1 - P * ey) + (c_k * cbar) * ey - g2 * v.gx - g3 * v.gy;
    bx += hbar * h.dx + g0 * u.dx + g1 * v.dx;
    by += hbar * h.dy + g0 * u.dy + g1 * v.dy;
    // (a state of a finite row never errs; the guards keep a misuse in bounds)
    if (BK == BK_GRID && Gr.depth && okb)
        acc_add(acc.h, Gr.depth, (uint32_t)B.nx, bc, B.rxs, B.rys, hbar, g2 * c, g3 * c);
    if (CK == CK_GRID && okc) {
        if (Gr.u) acc_add(acc.u, Gr.u, (uint32_t)C.nx, cc, C.rxs, C.rys, g0, -g2 * kx, -g3 * kx);
        if (Gr.v) acc_add(acc.v, Gr.v, (uint32_t)C.nx, cc, C.rxs, C.rys, g1, -g2 * ky, -g3 * ky);
    }
}

}  // namespace mr
