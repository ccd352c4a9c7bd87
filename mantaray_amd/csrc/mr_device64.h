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
//   slopes  (D - A) RN(1/(x1 - x0)), (B - A) RN(1/(y1 - y0))
//
//   (the slopes are the cell difference times the rounded reciprocal of the
//   signed spacing, within an ulp of the reference's quotient)
//
// Field values come from kernel arguments (pointers to the raw f64 arrays),
// so a caller can differentiate with respect to them and change them between
// calls without rebuilding anything; what depends only on the axes (cell
// thresholds, node pairs, the change of basis m01, m10) is tabulated once per
// field.  As in the reference mode, each lane holds cells in registers (a
// Slot64 per field) and loads the cell a coming evaluation will most likely
// need ahead of it.
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
    uint32_t last_a, last_b;              // last index of each value array
    double a0x, a0y;                      // first coordinate of each axis
    double sx, sy, inv_sx, inv_sy;        // |a1 - a0| and RN(1/|a1 - a0|)
    double nxm1, nym1;                    // n - 1
    double rxs, rys;                      // RN(1/(a1 - a0)), signed: the slope factors
    double h, h0, x0, y0, dhdx, dhdy;     // ConstantDepth / ConstantSlope, widened
    double u, v;                          // ConstantCurrent
    float gx0, gy0, ginv_x, ginv_y;       // cell guess for prefetch, in f32
    float gmax_x, gmax_y;                 // (f32)(n - 2)
};

// Device arrays of a field in fp64 mode (kernel arguments).
struct Grid64Ptrs {
    const double2 *xt, *yt;               // per axis cell i: fast thresholds, [T[i], T[i+1])
                                          // narrowed to the open node interval
    const double2 *xp, *yp;               // per axis cell i: (a[i], a[i+1]) in f64
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

// One cell of a gridded field held per lane: its fast thresholds, node
// coordinates, change of basis, and the four corners of NV value arrays
// (depth; or u and v).  The fast thresholds of cell i on an axis are the
// f64 index's cell [T[i], T[i+1]) narrowed to the open interval between the
// cell's nodes (a[i], a[i+1]): a target inside them on both axes is in the
// cell and on none of its node lines, so the bilinear value needs neither
// the corner exit nor the node case of the det == 0 exit.  Targets of the
// cell on a node line (rare: the node coordinates themselves) take the full
// query, which locates them by the index.  valid is false for a cell with a
// corner past the data and for a slot never loaded; thresholds start as NaN,
// so such a slot contains no point.
template <int NV>
struct Slot64 {
    double2 tx = {MR_NAN_INIT, MR_NAN_INIT}, ty = {MR_NAN_INIT, MR_NAN_INIT};
    double2 xp = {0.0, 0.0}, yp = {0.0, 0.0}, m = {0.0, 0.0};
    double v[NV][4] = {};                 // A, B, C, D
    uint32_t iA = 0;
    bool valid = false;
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

template <int NV>
__device__ __forceinline__ bool slot64_inside(const Slot64<NV> &s, double x, double y)
{
    return s.valid & (x >= s.tx.x) & (x < s.tx.y) & (y >= s.ty.x) & (y < s.ty.y);
}

// Start loading cell (i1, j1) into s.  Corner indices are clamped into the
// value arrays, so a cell with a corner past the data (never valid) loads
// in-bounds values that are never used.
template <int NV>
__device__ __forceinline__ void slot64_load(const Grid64View &G, const Grid64Ptrs &P,
                                            const double *va, const double *vb, Slot64<NV> &s,
                                            int i1, int j1)
{
    const uint32_t cell = (uint32_t)j1 * (uint32_t)G.ncx + (uint32_t)i1;
    s.tx = gldd2(P.xt, (uint32_t)i1 * 16u);
    s.ty = gldd2(P.yt, (uint32_t)j1 * 16u);
    s.xp = gldd2(P.xp, (uint32_t)i1 * 16u);
    s.yp = gldd2(P.yp, (uint32_t)j1 * 16u);
    s.m = gldd2(P.m, cell * 16u);
    const uint32_t iA = (uint32_t)G.nx * (uint32_t)j1 + (uint32_t)i1;
    const uint32_t idx[4] = {iA, iA + (uint32_t)G.nx, iA + (uint32_t)G.nx + 1u, iA + 1u};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        s.v[0][c] = gld(va, __builtin_elementwise_min(idx[c], G.last_a));
        if (NV > 1) s.v[NV - 1][c] = gld(vb, __builtin_elementwise_min(idx[c], G.last_b));
    }
    s.iA = iA;
    s.valid = cell < G.oob_from;
}

// Start loading the cell that (x, y) most likely lies in (guessed in f32,
// clamped into the grid; used only to choose what to load).
template <int NV>
__device__ __forceinline__ void slot64_prefetch(const Grid64View &G, const Grid64Ptrs &P,
                                                const double *va, const double *vb,
                                                Slot64<NV> &s, double x, double y)
{
    const int i1 = guess_cell((float)x, G.gx0, G.ginv_x, G.gmax_x);
    const int j1 = guess_cell((float)y, G.gy0, G.ginv_y, G.gmax_y);
    slot64_load<NV>(G, P, va, vb, s, i1, j1);
}

// Make s hold the cell of (x, y): lanes whose slot does not contain the
// point locate it (the reference's f64 index) and load it, then wait.
// Returns false where the reference's lookup errs.
template <int NV>
__device__ __forceinline__ bool slot64_query(const Grid64View &G, const Grid64Ptrs &P,
                                             const double *va, const double *vb, Slot64<NV> &s,
                                             double x, double y)
{
    bool ok = slot64_inside(s, x, y);
    if (__any(!ok)) {
        const bool miss = !ok;
        int i1, j1;
        const bool okn = locate64(G, x, y, i1, j1);
        if (miss & okn) slot64_load<NV>(G, P, va, vb, s, i1, j1);
        wait_vmem();
        ok = ok | (miss & okn);
    }
    return ok;
}

// interpolator::bilinear in f64 on the slot's cell for value array q, with
// its corner exit (a target on a node returns that corner, first of A, B,
// C, D; EXITS false leaves it out, for a target inside the slot's fast
// thresholds, which is on no node), and the cell-difference slopes.  The
// det == 0 exit is the caller's (m01 is NaN there).
template <int NV, bool EXITS = true>
__device__ __forceinline__ Val64 slot64_value(const Grid64View &G, const Slot64<NV> &s, int q,
                                              double x, double y, Cell64 &c)
{
    const double vA = s.v[q][0], vB = s.v[q][1], vC = s.v[q][2], vD = s.v[q][3];
    const double a10 = vB - vA, a01 = vD - vA;
    const double a11 = ((vC - vA) - a10) - a01;
    c.iA = s.iA;
    c.u = s.m.x * (y - s.yp.x);
    c.w = s.m.y * (x - s.xp.x);
    Val64 r;
    r.val = ((vA + a10 * c.u) + a01 * c.w) + (a11 * c.u) * c.w;
    if (EXITS) {
        const bool ex = x == s.xp.x, exd = x == s.xp.y, eya = y == s.yp.x, eyb = y == s.yp.y;
        if ((ex | exd) & (eya | eyb))
            r.val = ex ? (eya ? vA : vB) : (eyb ? vC : vD);
    }
    r.gx = (vD - vA) * G.rxs;
    r.gy = (vB - vA) * G.rys;
    r.dx = (a01 + a11 * c.u) * s.m.y;
    r.dy = (a10 + a11 * c.w) * s.m.x;
    return r;
}

// the det == 0 exit: m01 is NaN, unless the target is on a node (EXITS
// false: a target inside the fast thresholds, on no node)
template <int NV, bool EXITS = true>
__device__ __forceinline__ bool slot64_det_ok(const Slot64<NV> &s, double x, double y)
{
    if (!EXITS) return !__builtin_isnan(s.m.x);
    const bool node = ((x == s.xp.x) | (x == s.xp.y)) & ((y == s.yp.x) | (y == s.yp.y));
    return node | !__builtin_isnan(s.m.x);
}

// Bathymetry in fp64 mode.  Returns ok = false where the reference errs;
// h is NaN (not an error) for a NaN point (cartesian_netcdf3.rs:102-105) and
// outside an ArrayDepth.  cell is meaningful for BK_GRID only.
template <int BK>
__device__ __forceinline__ bool bathy64(const Grid64View &B, const Grid64Ptrs &P,
                                        const double *depth, Slot64<1> &bs, double x, double y,
                                        Val64 &h, Cell64 &cell)
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
    } else {                                               // cartesian_netcdf3.rs:98-135
        bool ok = slot64_query<1>(B, P, depth, depth, bs, x, y);
        h = slot64_value<1>(B, bs, 0, x, y, cell);
        ok = ok & slot64_det_ok<1>(bs, x, y);
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
                                          const double *cu, const double *cv, Slot64<2> &cs,
                                          double x, double y, Val64 &u, Val64 &v, Cell64 &cell)
{
    cell.iA = 0; cell.u = 0.0; cell.w = 0.0;
    if (CK == CK_CONSTANT) {                               // constant_current.rs:66-76
        u.val = C.u; v.val = C.v;
        u.gx = u.gy = u.dx = u.dy = 0.0;
        v.gx = v.gy = v.dx = v.dy = 0.0;
        return true;
    } else {                                               // cartesian_current.rs:487-542
        bool ok = slot64_query<2>(C, P, cu, cv, cs, x, y);
        u = slot64_value<2>(C, cs, 0, x, y, cell);
        v = slot64_value<2>(C, cs, 1, x, y, cell);
        return ok & slot64_det_ok<2>(cs, x, y);
    }
}

// Field values of one fp64 evaluation.
struct Fields64 {
    const double *depth, *u, *v;
};

// The cells a lane holds, for the evaluation in hand (q) and the next (n).
struct Slots64 {
    Slot64<1> b;
    Slot64<2> c;
};

// The cells that (x, y) most likely lies in, loaded into s whatever it
// holds (a ray's first cells, or a segment's).
template <int BK, int CK>
__device__ __forceinline__ void load_guess64(const Grid64View &B, const Grid64Ptrs &BP,
                                             const Grid64View &C, const Grid64Ptrs &CP,
                                             const Fields64 &F, Slots64 &s, double x, double y)
{
    if (BK == BK_GRID) slot64_prefetch<1>(B, BP, F.depth, F.depth, s.b, x, y);
    if (CK == CK_GRID) slot64_prefetch<2>(C, CP, F.u, F.v, s.c, x, y);
}

// s_waitcnt on the vector memory count alone, for counts up to 63 (gfx9
// splits vmcnt over bits 3:0 and 15:14; expcnt and lgkmcnt at their no-wait
// maxima).
template <int N>
__device__ __forceinline__ void vm_wait_upto()
{
    static_assert(N >= 0 && N < 64, "vmcnt below 64");
    __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
}

// Start loading, into n, the cells of (x, y) that q does not hold.
template <int BK, int CK>
__device__ __forceinline__ void prefetch64(const Grid64View &B, const Grid64Ptrs &BP,
                                           const Grid64View &C, const Grid64Ptrs &CP,
                                           const Fields64 &F, const Slots64 &q, Slots64 &n,
                                           double x, double y)
{
    if (BK == BK_GRID && !slot64_inside<1>(q.b, x, y))
        slot64_prefetch<1>(B, BP, F.depth, F.depth, n.b, x, y);
    if (CK == CK_GRID && !slot64_inside<2>(q.c, x, y))
        slot64_prefetch<2>(C, CP, F.u, F.v, n.c, x, y);
}

// WaveRayPath::system in fp64 mode, with the cells held in sl.  As rhs()
// does for the reference precision, every lane first evaluates with the
// cells in hand (no branch, no load: the gridded queries' own operations for
// a target inside the slot's cell), and the lanes whose target is not inside
// them evaluate again with the full queries, which locate and load cells.
template <int BK, int CK, int MATH>
__device__ __forceinline__ void rhs64(const Grid64View &B, const Grid64Ptrs &BP,
                                      const Grid64View &C, const Grid64Ptrs &CP,
                                      const Fields64 &F, Slots64 &sl, const FastConsts &K,
                                      double x, double y, double kx, double ky, double &dx,
                                      double &dy, double &dkx, double &dky)
{
    Val64 h, u, v;
    Cell64 bc, cc;
    bool okb, okc;
    bool slow = false;
    if (BK == BK_GRID) {
        slow = !slot64_inside<1>(sl.b, x, y);
        h = slot64_value<1, false>(B, sl.b, 0, x, y, bc);
        okb = slot64_det_ok<1, false>(sl.b, x, y);
    } else {
        okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
    }
    if (CK == CK_GRID) {
        slow = slow | !slot64_inside<2>(sl.c, x, y);
        u = slot64_value<2, false>(C, sl.c, 0, x, y, cc);
        v = slot64_value<2, false>(C, sl.c, 1, x, y, cc);
        okc = slot64_det_ok<2, false>(sl.c, x, y);
    } else {
        okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
    }
    odes<MATH>(K, okb ? h.val : MR_NAN, h.gx, h.gy, okc, u.val, v.val, u.gx, u.gy, v.gx, v.gy,
               kx, ky, dx, dy, dkx, dky);
    if (__builtin_expect(wave_any(slow), 0)) {
        __asm__ volatile("");
        if (slow) {                    // a NaN target lands here too
            okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
            okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
            odes<MATH>(K, okb ? h.val : MR_NAN, h.gx, h.gy, okc, u.val, v.val, u.gx, u.gy,
                       v.gx, v.gy, kx, ky, dx, dy, dkx, dky);
        }
    }
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
    const bool moved = cell.iA != s.iA;
    if (__builtin_expect(wave_any(moved), 0)) {
        __asm__ volatile("");
        if (moved) {
            acc_flush(s, grad, nx);
            s.iA = cell.iA;
        }
    }
    const double u = cell.u, w = cell.w;
    const double gx = gxbar * rxs, gy = gybar * rys;
    s.a += vbar * ((1.0 - u) * (1.0 - w)) - gx - gy;
    s.b += vbar * (u * (1.0 - w)) + gy;
    s.c += vbar * (u * w);
    s.d += vbar * ((1.0 - u) * w) + gx;
}

// acc_add for the current's two grids at once: u and v are gridded on the
// same axes, so one cell test serves both (the test and its scalar branch
// are half of acc_add's cost on the common path).  The sums of a grid not
// wanted (null gu or gv) are formed and never added.  Used by the sweeps
// that apply the vector-Jacobian product in the wave that forms it; the
// four-wave sweep's chain ran slower with it (1.7 %) and keeps acc_add.
__device__ __forceinline__ void acc_add_uv(CornerAcc &su, CornerAcc &sv, double *gu, double *gv,
                                           uint32_t nx, const Cell64 &cell, double rxs,
                                           double rys, double ubar, double uxbar, double uybar,
                                           double vbar, double vxbar, double vybar)
{
    const bool moved = cell.iA != su.iA;
    if (__builtin_expect(wave_any(moved), 0)) {
        __asm__ volatile("");
        if (moved) {
            if (gu) acc_flush(su, gu, nx);
            if (gv) acc_flush(sv, gv, nx);
            su.iA = cell.iA;
            sv.iA = cell.iA;
        }
    }
    const double u = cell.u, w = cell.w;
    const double ux = uxbar * rxs, uy = uybar * rys;
    su.a += ubar * ((1.0 - u) * (1.0 - w)) - ux - uy;
    su.b += ubar * (u * (1.0 - w)) + uy;
    su.c += ubar * (u * w);
    su.d += ubar * ((1.0 - u) * w) + ux;
    const double vx = vxbar * rxs, vy = vybar * rys;
    sv.a += vbar * ((1.0 - u) * (1.0 - w)) - vx - vy;
    sv.b += vbar * (u * (1.0 - w)) + vy;
    sv.c += vbar * (u * w);
    sv.d += vbar * ((1.0 - u) * w) + vx;
}

// Gradient grids of an adjoint sweep (null: not wanted).
struct Grads64 {
    double *depth, *u, *v;
};

struct Accs64 {
    CornerAcc h, u, v;
};

// The linearisation of rhs64 at one stage point, as a record of LIN_N
// doubles, so that the cotangent chain can run apart from the (costlier)
// field queries and dispersion terms:
//   [0, 16)   M, row-major: the state cotangent is z = M g for a cotangent g
//             of (dx/dt, dy/dt, dkx/dt, dky/dt)
//   [16, 20)  hb: the depth-value cotangent is hb . g
//   20, 21, 22  c, kx, ky: the slope cotangents (g2 c, g3 c), (-g2 kx, -g3 kx),
//             (-g2 ky, -g3 ky) of depth, u and v
//   23, 24    bathymetry cell fractions (u, w); 25, 26 the current's
//   27        the two cells' corner-A indices, bit-packed (all ones: none)
constexpr int LIN_N = 28;

__device__ __forceinline__ double pack_cells(uint32_t b, uint32_t c)
{
    return __builtin_bit_cast(double, ((uint64_t)c << 32) | (uint64_t)b);
}
__device__ __forceinline__ void unpack_cells(double v, uint32_t &b, uint32_t &c)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    b = (uint32_t)u;
    c = (uint32_t)(u >> 32);
}

// Whether the depth has no gradient anywhere (ConstantDepth, ArrayDepth:
// their queries return zero slopes and zero derivatives), so that every
// depth-gradient term of the linearisation vanishes.
template <int BK>
__device__ constexpr bool depth_is_flat()
{
    return BK == BK_CONSTANT || BK == BK_ARRAY;
}

// Whether lin_apply reads field f of a linearisation record for the field
// kinds BK, CK (the only fields the ring is given): the depth terms (16-20,
// 23, 24) for gridded depth, the current's (21, 22, 25, 26) for a gridded
// current, and with a flat depth not the four zeros of M's x and y rows.
template <int BK, int CK>
__device__ constexpr bool lin_field_used(int f)
{
    if (f < 16)
        return !depth_is_flat<BK>() || f == 0 || f == 1 || f == 4 || f == 5 || f >= 8;
    if (f == 27) return true;
    if (f <= 20 || f == 23 || f == 24) return BK == BK_GRID;
    return CK == CK_GRID;
}

// The linearisation record of one evaluation from its field values (the
// arithmetic of rhs64_lin below).
template <int BK, int CK>
__device__ __forceinline__ void lin_record(const FastConsts &K, const Val64 &h, const Val64 &u,
                                           const Val64 &v, const Cell64 &bc, const Cell64 &cc,
                                           bool okb, bool okc, double kx, double ky,
                                           double R[LIN_N])
{
    const double hh = h.val;
    const double k2 = __builtin_fma(kx, kx, ky * ky);
    const double rk = rsq_nr(k2);
    const double k = k2 * rk;
    const double ex = kx * rk, ey = ky * rk;
    const double q = k * hh;
    double e, em;
    exp_pair(-2.0 * q, e, em, K);
    const double ope = 1.0 + e;
    // one reciprocal square root, W = 1/sqrt(g k (1-e)(1+e)), gives the
    // reciprocal D = 1/((1-e)(1+e)) = g k W^2 and 1/sigma = W (1+e), as
    // sigma^2 = g k tanh(kh) = g k (1-e)/(1+e)
    const double gk = K.g * k;
    const double W = rsq_nr((gk * em) * ope);
    const double D = gk * (W * W);
    const double iope = em * D;
    const double T = em * iope;
    const double S = (4.0 * e) * (iope * iope);
    if (depth_is_flat<BK>()) {
        // no depth gradient: the rows of M are the current's and the
        // c_g k/|k| terms only
        const double rs = W * ope;
        const double N = __builtin_fma(q, S, T);
        const double cg = K.half_g * N * rs;
        const double Nq = 2.0 * S * (1.0 - q * T);
        const double cg_k = K.half_g * rs * (hh * Nq - N * cg * rs);
        const double tk = cg * rk;
        const double a = cg_k - tk;
        R[0] = u.dx; R[1] = v.dx; R[2] = 0.0; R[3] = 0.0;
        R[4] = u.dy; R[5] = v.dy; R[6] = 0.0; R[7] = 0.0;
        R[8] = a * ex * ex + tk; R[9] = a * ex * ey; R[10] = -u.gx; R[11] = -u.gy;
        R[12] = a * ey * ex; R[13] = a * ey * ey + tk; R[14] = -v.gx; R[15] = -v.gy;
        R[16] = 0.0; R[17] = 0.0; R[18] = 0.0; R[19] = 0.0; R[20] = 0.0;
        R[21] = kx; R[22] = ky;
        R[23] = 0.0; R[24] = 0.0; R[25] = cc.u; R[26] = cc.w;
        R[27] = pack_cells(0xFFFFFFFFu, (CK == CK_GRID && okc) ? cc.iA : 0xFFFFFFFFu);
        return;
    }
    const double invT = (ope * ope) * D;
    const double gkT = gk * T;
    const double rs = W * ope;
    const double sigma = gkT * rs;
    const double N = __builtin_fma(q, S, T);
    const double cg = K.half_g * N * rs;
    const double c = (-2.0 * k) * e * sigma * D;
    const double sig_h = K.half_g * (k * k) * S * rs;
    const double Nq = 2.0 * S * (1.0 - q * T);
    const double cg_k = K.half_g * rs * (hh * Nq - N * cg * rs);
    const double cg_h = K.half_g * rs * (k * Nq - N * sig_h * rs);
    const double coth2 = 0.5 * (T + invT);
    const double c_k = c * (rk + cg * rs - 2.0 * hh * coth2);
    const double c_h = c * (sig_h * rs - 2.0 * k * coth2);
    const double tk = cg * rk;
    const double hb[4] = {cg_h * ex, cg_h * ey, c_h * h.gx, c_h * h.gy};
    const double a = cg_k - tk;
    // rows of M (the partial derivatives of f, transposed)
    R[0] = hb[0] * h.dx + u.dx; R[1] = hb[1] * h.dx + v.dx;
    R[2] = hb[2] * h.dx;        R[3] = hb[3] * h.dx;
    R[4] = hb[0] * h.dy + u.dy; R[5] = hb[1] * h.dy + v.dy;
    R[6] = hb[2] * h.dy;        R[7] = hb[3] * h.dy;
    R[8] = a * ex * ex + tk;    R[9] = a * ex * ey;
    R[10] = c_k * ex * h.gx - u.gx; R[11] = c_k * ex * h.gy - u.gy;
    R[12] = a * ey * ex;        R[13] = a * ey * ey + tk;
    R[14] = c_k * ey * h.gx - v.gx; R[15] = c_k * ey * h.gy - v.gy;
    R[16] = hb[0]; R[17] = hb[1]; R[18] = hb[2]; R[19] = hb[3];
    R[20] = c; R[21] = kx; R[22] = ky;
    R[23] = bc.u; R[24] = bc.w; R[25] = cc.u; R[26] = cc.w;
    // a state of a finite row never errs; the marks keep a misuse in bounds
    R[27] = pack_cells((BK == BK_GRID && okb) ? bc.iA : 0xFFFFFFFFu,
                       (CK == CK_GRID && okc) ? cc.iA : 0xFFFFFFFFu);
}

template <int BK, int CK>
__device__ __forceinline__ void rhs64_lin(const Grid64View &B, const Grid64Ptrs &BP,
                                          const Grid64View &C, const Grid64Ptrs &CP,
                                          const Fields64 &F, Slots64 &sl, const FastConsts &K,
                                          double x, double y, double kx, double ky,
                                          double R[LIN_N])
{
    // as rhs64: first with the cells in hand, then the full queries for the
    // lanes whose target is not inside them
    Val64 h, u, v;
    Cell64 bc, cc;
    bool okb, okc;
    bool slow = false;
    if (BK == BK_GRID) {
        slow = !slot64_inside<1>(sl.b, x, y);
        h = slot64_value<1, false>(B, sl.b, 0, x, y, bc);
        okb = slot64_det_ok<1, false>(sl.b, x, y);
    } else {
        okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
    }
    if (CK == CK_GRID) {
        slow = slow | !slot64_inside<2>(sl.c, x, y);
        u = slot64_value<2, false>(C, sl.c, 0, x, y, cc);
        v = slot64_value<2, false>(C, sl.c, 1, x, y, cc);
        okc = slot64_det_ok<2, false>(sl.c, x, y);
    } else {
        okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
    }
    lin_record<BK, CK>(K, h, u, v, bc, cc, okb, okc, kx, ky, R);
    if (__builtin_expect(wave_any(slow), 0)) {
        __asm__ volatile("");
        if (slow) {
            okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
            okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
            lin_record<BK, CK>(K, h, u, v, bc, cc, okb, okc, kx, ky, R);
        }
    }
}

// Applies a linearisation record R to the cotangent g: z = M g, and the
// field-value cotangents to the accumulators.
template <int BK, int CK>
__device__ __forceinline__ void lin_apply(const double R[LIN_N], const Grid64View &B,
                                          const Grid64View &C, const Grads64 &Gr, Accs64 &acc,
                                          const double g[4], double z[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (depth_is_flat<BK>() && i < 2)      // R[2], R[3], R[6], R[7] are zero
            z[i] = R[4 * i] * g[0] + R[4 * i + 1] * g[1];
        else
            z[i] = ((R[4 * i] * g[0] + R[4 * i + 1] * g[1]) + R[4 * i + 2] * g[2]) +
                   R[4 * i + 3] * g[3];
    }
    uint32_t ib, ic;
    unpack_cells(R[27], ib, ic);
    if (BK == BK_GRID && Gr.depth && ib != 0xFFFFFFFFu) {
        const double hbar = ((R[16] * g[0] + R[17] * g[1]) + R[18] * g[2]) + R[19] * g[3];
        Cell64 cell{ib, R[23], R[24]};
        acc_add(acc.h, Gr.depth, (uint32_t)B.nx, cell, B.rxs, B.rys, hbar, g[2] * R[20],
                g[3] * R[20]);
    }
    if (CK == CK_GRID && ic != 0xFFFFFFFFu) {
        Cell64 cell{ic, R[25], R[26]};
        if (Gr.u) acc_add(acc.u, Gr.u, (uint32_t)C.nx, cell, C.rxs, C.rys, g[0], -g[2] * R[21],
                          -g[3] * R[21]);
        if (Gr.v) acc_add(acc.v, Gr.v, (uint32_t)C.nx, cell, C.rxs, C.rys, g[1], -g[2] * R[22],
                          -g[3] * R[22]);
    }
}

// Vector-Jacobian product of rhs64 at (x, y, kx, ky): given the cotangent
// (g0, g1, g2, g3) of (dx/dt, dy/dt, dkx/dt, dky/dt), adds the state
// cotangent to (bx, by, bkx, bky) and the field-value cotangents to the
// accumulators.  The dispersion terms are evaluated with the closed forms of
// the fast mode whatever the forward's math mode: only the derivative
// depends on them.  The state must be one at which rhs64 is finite.
template <int BK, int CK>
__device__ __forceinline__ void rhs64_vjp(const Grid64View &B, const Grid64Ptrs &BP,
                                          const Grid64View &C, const Grid64Ptrs &CP,
                                          const Fields64 &F, Slots64 &sl, const Grads64 &Gr,
                                          Accs64 &acc, const FastConsts &K, double x, double y,
                                          double kx, double ky, double g0, double g1, double g2,
                                          double g3, double &bx, double &by, double &bkx,
                                          double &bky)
{
    // the field values as in rhs64: from the cells in hand, then the full
    // queries for the lanes whose target is not inside them (the product and
    // the accumulation below run once, after both)
    Val64 h, u, v;
    Cell64 bc, cc;
    bool okb, okc;
    bool slow = false;
    if (BK == BK_GRID) {
        slow = !slot64_inside<1>(sl.b, x, y);
        h = slot64_value<1, false>(B, sl.b, 0, x, y, bc);
        okb = slot64_det_ok<1, false>(sl.b, x, y);
    } else {
        okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
    }
    if (CK == CK_GRID) {
        slow = slow | !slot64_inside<2>(sl.c, x, y);
        u = slot64_value<2, false>(C, sl.c, 0, x, y, cc);
        v = slot64_value<2, false>(C, sl.c, 1, x, y, cc);
        okc = slot64_det_ok<2, false>(sl.c, x, y);
    } else {
        okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
    }
    if (__builtin_expect(wave_any(slow), 0)) {
        __asm__ volatile("");
        if (slow) {
            okb = bathy64<BK>(B, BP, F.depth, sl.b, x, y, h, bc);
            okc = current64<CK>(C, CP, F.u, F.v, sl.c, x, y, u, v, cc);
        }
    }
    const double hh = h.val;
    const double k2 = __builtin_fma(kx, kx, ky * ky);
    const double rk = rsq_nr(k2);                       // 1/k
    const double k = k2 * rk;
    const double ex = kx * rk, ey = ky * rk;
    const double q = k * hh;
    double e, em;
    exp_pair(-2.0 * q, e, em, K);                        // e = exp(-2q), em = 1 - e
    const double ope = 1.0 + e;
    const double D = rcp_nr(em * ope);
    const double iope = em * D;
    const double T = em * iope;                         // tanh q
    const double S = (4.0 * e) * (iope * iope);         // sech^2 q
    const double gkT = (K.g * k) * T;
    const double rs = rsq_nr(gkT);                      // 1/sigma
    const double N = __builtin_fma(q, S, T);
    const double cg = K.half_g * N * rs;
    const double Nq = 2.0 * S * (1.0 - q * T);
    const double cg_k = K.half_g * rs * (hh * Nq - N * cg * rs);
    const double P = g0 * ex + g1 * ey;                 // cotangent of cg
    const double tk = cg * rk;
    if (depth_is_flat<BK>()) {
        // no depth gradient: the terms of c and of the depth cotangent vanish
        bkx += (cg_k * P) * ex + tk * (g0 - P * ex) - g2 * u.gx - g3 * u.gy;
        bky += (cg_k * P) * ey + tk * (g1 - P * ey) - g2 * v.gx - g3 * v.gy;
        bx += g0 * u.dx + g1 * v.dx;
        by += g0 * u.dy + g1 * v.dy;
    } else {
        const double invT = (ope * ope) * D;            // 1/tanh q
        const double sigma = gkT * rs;
        const double c = (-2.0 * k) * e * sigma * D;
        const double sig_h = K.half_g * (k * k) * S * rs;   // d sigma / dh
        const double cg_h = K.half_g * rs * (k * Nq - N * sig_h * rs);
        const double coth2 = 0.5 * (T + invT);
        const double c_k = c * (rk + cg * rs - 2.0 * hh * coth2);
        const double c_h = c * (sig_h * rs - 2.0 * k * coth2);
        const double cbar = g2 * h.gx + g3 * h.gy;      // cotangent of c
        const double hbar = cg_h * P + c_h * cbar;
        bkx += (cg_k * P) * ex + tk * (g0 - P * ex) + (c_k * cbar) * ex - g2 * u.gx - g3 * u.gy;
        bky += (cg_k * P) * ey + tk * (g1 - P * ey) + (c_k * cbar) * ey - g2 * v.gx - g3 * v.gy;
        bx += hbar * h.dx + g0 * u.dx + g1 * v.dx;
        by += hbar * h.dy + g0 * u.dy + g1 * v.dy;
        // (a state of a finite row never errs; the guards keep a misuse in bounds)
        if (BK == BK_GRID && Gr.depth && okb)
            acc_add(acc.h, Gr.depth, (uint32_t)B.nx, bc, B.rxs, B.rys, hbar, g2 * c, g3 * c);
    }
    if (CK == CK_GRID && okc)
        acc_add_uv(acc.u, acc.v, Gr.u, Gr.v, (uint32_t)C.nx, cc, C.rxs, C.rys, g0, -g2 * kx,
                   -g3 * kx, g1, -g2 * ky, -g3 * ky);
}

}  // namespace mr
