/*
 * mr_oracle.c -- CPU oracle for the mantaray many-ray integrator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline, never the
 * product).  A scalar restatement of the reference's hot path, keeping its
 * f32/f64 mixing, branch structure and operation order.  Build with
 * -O2 -ffp-contract=off and no fast-math so that each C operation is one
 * IEEE operation, as in the Rust source (rustc never contracts to FMA).
 *
 * Citations are to files of mines-oceanography/mantaray 0.1.0.
 *
 * Parity status (what pins each function; details in the tests):
 *   or_bilinear32         pinned: interpolator.rs:86-161 known answers,
 *                         brute-force vs an independent numpy restatement.
 *   bathy grid provider   pinned: cartesian_netcdf3.rs:564-834 corner lists,
 *                         linear-gradient and out-of-bounds tests.
 *   current grid provider pinned: cartesian_current.rs:645-991 tests.
 *   group velocity        pinned: wave_ray_path.rs:296-315 known answers,
 *                         deep/shallow closed forms.
 *   dkdt_bathy            pinned: wave_ray_path.rs:574-597 (exact zero) and
 *                         the closed form of d(sigma)/dh.
 *   RK4 stepper           pinned: test_axis, test_shallow (bitwise one-step
 *                         values), test_solout (row count); the rounding of
 *                         the stage states for a non-constant RHS and the
 *                         solout argument are parity unpinned (ode_solvers
 *                         0.4.0 is not vendored); both solout variants are
 *                         implemented.
 *   fp64-interp providers not in the reference (a differentiable variant):
 *                         pinned by closed forms on linear fields, by
 *                         scipy's linear interpolator on random fields, at
 *                         the nodes, and by the reference mode's cell rule
 *                         and errors (tests/test_oracle_fp64.py).
 */
#include "mr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* wave_ray_path.rs:23 */
static const double G = 9.8;

/* Rust `as usize` from a float: saturating, NaN -> 0. */
static uint64_t sat_usize_f32(float v)
{
    if (!(v > 0.0f)) return 0;                       /* NaN, -0, negatives */
    if (v >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)v;
}
static uint64_t sat_usize_f64(double v)
{
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)v;
}

/* interpolator.rs:39-84 -- bilinear interpolation on a 4-corner quad,
 * corners in order A=(x1,y1), B=(x1,y2), C=(x2,y2), D=(x2,y1). */
int or_bilinear32(const float p[4][3], float tx, float ty, float *out)
{
    for (int i = 0; i < 4; i++) {                    /* :47-51 coincidence */
        if (tx == p[i][0] && ty == p[i][1]) { *out = p[i][2]; return OR_OK; }
    }
    float btx = p[1][0] - p[0][0], bty = p[1][1] - p[0][1];   /* :59-61 */
    float dtx = p[3][0] - p[0][0], dty = p[3][1] - p[0][1];
    float ttx = tx - p[0][0], tty = ty - p[0][1];
    float det = (btx * dty) - (dtx * bty);                    /* :64 */
    if (det == 0.0f) return OR_ERR;
    float m00 = dty / det, m01 = -(dtx / det);                /* :69-72 */
    float m10 = -(bty / det), m11 = btx / det;
    float u = m00 * ttx + m01 * tty;                          /* :74-75 */
    float w = m10 * ttx + m11 * tty;
    float a00 = p[0][2];                                      /* :78-81 */
    float a10 = p[1][2] - p[0][2];
    float a01 = p[3][2] - p[0][2];
    float a11 = ((p[2][2] - p[0][2]) - a10) - a01;
    *out = ((a00 + a10 * u) + a01 * w) + (a11 * u) * w;       /* :83 */
    return OR_OK;
}

/* the same algorithm in f64 (fp64-interp mode only) */
static int bilinear64(const double p[4][3], double tx, double ty, double *out)
{
    for (int i = 0; i < 4; i++) {
        if (tx == p[i][0] && ty == p[i][1]) { *out = p[i][2]; return OR_OK; }
    }
    double btx = p[1][0] - p[0][0], bty = p[1][1] - p[0][1];
    double dtx = p[3][0] - p[0][0], dty = p[3][1] - p[0][1];
    double ttx = tx - p[0][0], tty = ty - p[0][1];
    double det = (btx * dty) - (dtx * bty);
    if (det == 0.0) return OR_ERR;
    double m00 = dty / det, m01 = -(dtx / det);
    double m10 = -(bty / det), m11 = btx / det;
    double u = m00 * ttx + m01 * tty;
    double w = m10 * ttx + m11 * tty;
    double a00 = p[0][2];
    double a10 = p[1][2] - p[0][2];
    double a01 = p[3][2] - p[0][2];
    double a11 = ((p[2][2] - p[0][2]) - a10) - a01;
    *out = ((a00 + a10 * u) + a01 * w) + (a11 * u) * w;
    return OR_OK;
}

/* cartesian_netcdf3.rs:274-296 -- fractional index along a f32 axis */
static int nearest32(float target, const float *a, int64_t n, float *idx)
{
    if (n == 0) return OR_ERR;
    if (n == 1) { *idx = 0.0f; return OR_OK; }
    float spacing = fabsf(a[1] - a[0]);
    float index = (target - a[0]) / spacing;
    if (index < 0.0f || index > (float)(n - 1)) return OR_ERR;
    *idx = index;
    return OR_OK;
}

/* cartesian_current.rs:231-253 -- the same along a f64 axis */
static int nearest64(double target, const double *a, int64_t n, double *idx)
{
    if (n == 0) return OR_ERR;
    if (n == 1) { *idx = 0.0; return OR_OK; }
    double spacing = fabs(a[1] - a[0]);
    double index = (target - a[0]) / spacing;
    if (index < 0.0 || index > (double)(n - 1)) return OR_ERR;
    *idx = index;
    return OR_OK;
}

/* cartesian_netcdf3.rs:334-390 -- cell choice along one axis */
static void cell_axis32(float idx, int64_t n, uint64_t *i1, uint64_t *i2)
{
    float low = 0.0f, high = (float)(n - 1);
    if (idx == low) {
        *i1 = sat_usize_f32(idx); *i2 = sat_usize_f32(idx) + 1;
    } else if (idx == high) {
        *i1 = sat_usize_f32(idx) - 1; *i2 = sat_usize_f32(idx);
    } else if (idx - truncf(idx) == 0.0f) {           /* f32::fract */
        *i1 = sat_usize_f32(roundf(idx)); *i2 = *i1 + 1;
    } else {
        *i1 = sat_usize_f32(floorf(idx)); *i2 = sat_usize_f32(ceilf(idx));
    }
}

/* cartesian_current.rs:283-339 */
static void cell_axis64(double idx, int64_t n, uint64_t *i1, uint64_t *i2)
{
    double low = 0.0, high = (double)(n - 1);
    if (idx == low) {
        *i1 = sat_usize_f64(idx); *i2 = sat_usize_f64(idx) + 1;
    } else if (idx == high) {
        *i1 = sat_usize_f64(idx) - 1; *i2 = sat_usize_f64(idx);
    } else if (idx - trunc(idx) == 0.0) {
        *i1 = sat_usize_f64(round(idx)); *i2 = *i1 + 1;
    } else {
        *i1 = sat_usize_f64(floor(idx)); *i2 = sat_usize_f64(ceil(idx));
    }
}

/* cartesian_netcdf3.rs:465-471 */
static int depth_at(const or_bathy *b, uint64_t i, uint64_t j, double *d)
{
    uint64_t index = (uint64_t)b->nx * j + i;
    if (index >= (uint64_t)b->depth_len) return OR_ERR;
    *d = b->depth[index];
    return OR_OK;
}

/* cartesian_current.rs:421-427 */
static int val_at(const or_current *c, uint64_t i, uint64_t j, const double *arr,
                  int64_t len, double *d)
{
    uint64_t index = (uint64_t)c->nx * j + i;
    if (index >= (uint64_t)len) return OR_ERR;
    *d = arr[index];
    return OR_OK;
}

/* corner list [(x1,y1), (x1,y2), (x2,y2), (x2,y1)] (cartesian_netcdf3.rs:389) */
static void corners_of(uint64_t x1, uint64_t x2, uint64_t y1, uint64_t y2, uint64_t ci[4][2])
{
    ci[0][0] = x1; ci[0][1] = y1;
    ci[1][0] = x1; ci[1][1] = y2;
    ci[2][0] = x2; ci[2][1] = y2;
    ci[3][0] = x2; ci[3][1] = y1;
}

/* cartesian_netcdf3.rs:274-296 / cartesian_current.rs:231-253 -- nearest
 * along axis 0 (x) or 1 (y) of a gridded field */
int or_bathy_nearest(const or_bathy *b, float target, int axis, float *idx)
{
    return axis == 0 ? nearest32(target, b->x, b->nx, idx) : nearest32(target, b->y, b->ny, idx);
}

int or_current_nearest(const or_current *c, double target, int axis, double *idx)
{
    return axis == 0 ? nearest64(target, c->x, c->nx, idx) : nearest64(target, c->y, c->ny, idx);
}

/* cartesian_netcdf3.rs:334-390 / cartesian_current.rs:283-339 -- four_corners:
 * [(x1,y1), (x1,y2), (x2,y2), (x2,y1)] as 8 integers */
int or_bathy_corners(const or_bathy *b, float px, float py, uint64_t out[8])
{
    float xi, yi;
    if (nearest32(px, b->x, b->nx, &xi) != OR_OK) return OR_ERR;
    if (nearest32(py, b->y, b->ny, &yi) != OR_OK) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis32(xi, b->nx, &x1, &x2);
    cell_axis32(yi, b->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    for (int q = 0; q < 4; q++) { out[2 * q] = ci[q][0]; out[2 * q + 1] = ci[q][1]; }
    return OR_OK;
}

int or_current_corners(const or_current *c, double x, double y, uint64_t out[8])
{
    double xi, yi;
    if (nearest64(x, c->x, c->nx, &xi) != OR_OK) return OR_ERR;
    if (nearest64(y, c->y, c->ny, &yi) != OR_OK) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis64(xi, c->nx, &x1, &x2);
    cell_axis64(yi, c->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    for (int q = 0; q < 4; q++) { out[2 * q] = ci[q][0]; out[2 * q + 1] = ci[q][1]; }
    return OR_OK;
}

/* cartesian_netcdf3.rs:98-135 (grid) plus the analytic providers
 * constant_depth.rs:26-33, constant_slope.rs:63-74, array_depth.rs:26-33 */
int or_bathy_depth_and_gradient(const or_bathy *b, float px, float py,
                                float *h, float *hx, float *hy)
{
    switch (b->kind) {
    case OR_BATHY_CONSTANT:
        if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else { *h = b->h; *hx = 0.0f; *hy = 0.0f; }
        return OR_OK;
    case OR_BATHY_SLOPE:
        if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else {
            *h = (b->h0 + b->dhdx * (px - b->x0)) + b->dhdy * (py - b->y0);
            *hx = b->dhdx; *hy = b->dhdy;
        }
        return OR_OK;
    case OR_BATHY_ARRAY: {
        uint64_t xi = sat_usize_f32(px), yi = sat_usize_f32(py);
        if (xi >= (uint64_t)b->an || yi >= (uint64_t)b->an) {
            *h = NAN; *hx = NAN; *hy = NAN; return OR_OK;
        }
        *h = b->arr[xi * (uint64_t)b->an + yi]; *hx = 0.0f; *hy = 0.0f;
        return OR_OK;
    }
    default: break;
    }
    if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; return OR_OK; }
    float xi, yi;
    if (nearest32(px, b->x, b->nx, &xi) != OR_OK) return OR_ERR;    /* :313-319 */
    if (nearest32(py, b->y, b->ny, &yi) != OR_OK) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis32(xi, b->nx, &x1, &x2);
    cell_axis32(yi, b->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    /* interpolate :417-445: per corner x[i] (panics past the end), y[j], depth */
    float pts[4][3];
    for (int q = 0; q < 4; q++) {
        if (ci[q][0] >= (uint64_t)b->nx || ci[q][1] >= (uint64_t)b->ny) return OR_PANIC;
        double d;
        if (depth_at(b, ci[q][0], ci[q][1], &d) != OR_OK) return OR_ERR;
        pts[q][0] = b->x[ci[q][0]];
        pts[q][1] = b->y[ci[q][1]];
        pts[q][2] = (float)d;
    }
    float hv;
    if (or_bilinear32((const float (*)[3])pts, px, py, &hv) != OR_OK) return OR_ERR;
    /* gradient :116-134 */
    double x_space = (double)b->x[1] - (double)b->x[0];
    double y_space = (double)b->y[1] - (double)b->y[0];
    double se, sw, nw;
    if (depth_at(b, ci[3][0], ci[3][1], &se) != OR_OK) return OR_ERR;
    if (depth_at(b, ci[0][0], ci[0][1], &sw) != OR_OK) return OR_ERR;
    double xg = (se - sw) / x_space;
    if (depth_at(b, ci[1][0], ci[1][1], &nw) != OR_OK) return OR_ERR;
    double yg = (nw - sw) / y_space;
    *h = hv; *hx = (float)xg; *hy = (float)yg;
    return OR_OK;
}

/* cartesian_current.rs:487-542 (grid) plus constant_current.rs:66-76 */
int or_current_and_gradient(const or_current *c, double x, double y,
                            double *u, double *v, double *ux, double *uy,
                            double *vx, double *vy)
{
    if (c->kind == OR_CURRENT_CONSTANT) {
        *u = c->u; *v = c->v; *ux = 0.0; *uy = 0.0; *vx = 0.0; *vy = 0.0;
        return OR_OK;
    }
    double xi, yi;                                   /* no NaN guard (:492-495) */
    if (nearest64(x, c->x, c->nx, &xi) != OR_OK) return OR_ERR;
    if (nearest64(y, c->y, c->ny, &yi) != OR_OK) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis64(xi, c->nx, &x1, &x2);
    cell_axis64(yi, c->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    float tx = (float)x, ty = (float)y;
    float pu[4][3], pv[4][3];
    /* interpolate(u) then interpolate(v), :365-398 */
    for (int q = 0; q < 4; q++) {
        if (ci[q][0] >= (uint64_t)c->nx || ci[q][1] >= (uint64_t)c->ny) return OR_PANIC;
        double d;
        if (val_at(c, ci[q][0], ci[q][1], c->u_arr, c->u_len, &d) != OR_OK) return OR_ERR;
        pu[q][0] = (float)c->x[ci[q][0]];
        pu[q][1] = (float)c->y[ci[q][1]];
        pu[q][2] = (float)d;
    }
    float uf, vf;
    if (or_bilinear32((const float (*)[3])pu, tx, ty, &uf) != OR_OK) return OR_ERR;
    for (int q = 0; q < 4; q++) {
        double d;
        if (val_at(c, ci[q][0], ci[q][1], c->v_arr, c->v_len, &d) != OR_OK) return OR_ERR;
        pv[q][0] = (float)c->x[ci[q][0]];
        pv[q][1] = (float)c->y[ci[q][1]];
        pv[q][2] = (float)d;
    }
    if (or_bilinear32((const float (*)[3])pv, tx, ty, &vf) != OR_OK) return OR_ERR;
    /* gradients :515-536, signed spacing */
    double x_space = c->x[1] - c->x[0];
    double y_space = c->y[1] - c->y[0];
    double se, sw, nw;
    if (val_at(c, ci[3][0], ci[3][1], c->u_arr, c->u_len, &se) != OR_OK) return OR_ERR;
    if (val_at(c, ci[0][0], ci[0][1], c->u_arr, c->u_len, &sw) != OR_OK) return OR_ERR;
    double dudx = (se - sw) / x_space;
    if (val_at(c, ci[1][0], ci[1][1], c->u_arr, c->u_len, &nw) != OR_OK) return OR_ERR;
    if (val_at(c, ci[0][0], ci[0][1], c->u_arr, c->u_len, &sw) != OR_OK) return OR_ERR;
    double dudy = (nw - sw) / y_space;
    if (val_at(c, ci[3][0], ci[3][1], c->v_arr, c->v_len, &se) != OR_OK) return OR_ERR;
    if (val_at(c, ci[0][0], ci[0][1], c->v_arr, c->v_len, &sw) != OR_OK) return OR_ERR;
    double dvdx = (se - sw) / x_space;
    if (val_at(c, ci[1][0], ci[1][1], c->v_arr, c->v_len, &nw) != OR_OK) return OR_ERR;
    if (val_at(c, ci[0][0], ci[0][1], c->v_arr, c->v_len, &sw) != OR_OK) return OR_ERR;
    double dvdy = (nw - sw) / y_space;
    *u = (double)uf; *v = (double)vf;
    *ux = dudx; *uy = dudy; *vx = dvdx; *vy = dvdy;
    return OR_OK;
}

/* fp64-interp bathymetry: the same index and cell rules in f64 on the widened
 * f32 coordinates, f64 bilinear over f64 depths, f64 slopes (the cell
 * difference times RN(1/spacing), within an ulp of the reference's quotient). */
int or_bathy_depth_and_gradient64(const or_bathy *b, double px, double py,
                                  double *h, double *hx, double *hy)
{
    switch (b->kind) {
    case OR_BATHY_CONSTANT:
        if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else { *h = (double)b->h; *hx = 0.0; *hy = 0.0; }
        return OR_OK;
    case OR_BATHY_SLOPE:
        if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else {
            *h = ((double)b->h0 + (double)b->dhdx * (px - (double)b->x0))
                 + (double)b->dhdy * (py - (double)b->y0);
            *hx = (double)b->dhdx; *hy = (double)b->dhdy;
        }
        return OR_OK;
    case OR_BATHY_ARRAY: {
        uint64_t xi = sat_usize_f64(px), yi = sat_usize_f64(py);
        if (xi >= (uint64_t)b->an || yi >= (uint64_t)b->an) {
            *h = NAN; *hx = NAN; *hy = NAN; return OR_OK;
        }
        *h = (double)b->arr[xi * (uint64_t)b->an + yi]; *hx = 0.0; *hy = 0.0;
        return OR_OK;
    }
    default: break;
    }
    if (isnan(px) || isnan(py)) { *h = NAN; *hx = NAN; *hy = NAN; return OR_OK; }
    if (b->nx < 2 || b->ny < 2) return b->nx == 0 || b->ny == 0 ? OR_ERR : OR_PANIC;
    double sx = fabs((double)b->x[1] - (double)b->x[0]);
    double sy = fabs((double)b->y[1] - (double)b->y[0]);
    double xi = (px - (double)b->x[0]) / sx, yi = (py - (double)b->y[0]) / sy;
    if (xi < 0.0 || xi > (double)(b->nx - 1)) return OR_ERR;
    if (yi < 0.0 || yi > (double)(b->ny - 1)) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis64(xi, b->nx, &x1, &x2);
    cell_axis64(yi, b->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    double pts[4][3];
    for (int q = 0; q < 4; q++) {
        if (ci[q][0] >= (uint64_t)b->nx || ci[q][1] >= (uint64_t)b->ny) return OR_PANIC;
        double d;
        if (depth_at(b, ci[q][0], ci[q][1], &d) != OR_OK) return OR_ERR;
        pts[q][0] = (double)b->x[ci[q][0]];
        pts[q][1] = (double)b->y[ci[q][1]];
        pts[q][2] = d;
    }
    if (bilinear64((const double (*)[3])pts, px, py, h) != OR_OK) return OR_ERR;
    /* slopes: the cell difference times the rounded reciprocal spacing */
    double rx = 1.0 / ((double)b->x[1] - (double)b->x[0]);
    double ry = 1.0 / ((double)b->y[1] - (double)b->y[0]);
    *hx = (pts[3][2] - pts[0][2]) * rx;
    *hy = (pts[1][2] - pts[0][2]) * ry;
    return OR_OK;
}

int or_current_and_gradient64(const or_current *c, double x, double y,
                              double *u, double *v, double *ux, double *uy,
                              double *vx, double *vy)
{
    if (c->kind == OR_CURRENT_CONSTANT) {
        *u = c->u; *v = c->v; *ux = 0.0; *uy = 0.0; *vx = 0.0; *vy = 0.0;
        return OR_OK;
    }
    double xi, yi;
    if (nearest64(x, c->x, c->nx, &xi) != OR_OK) return OR_ERR;
    if (nearest64(y, c->y, c->ny, &yi) != OR_OK) return OR_ERR;
    uint64_t x1, x2, y1, y2, ci[4][2];
    cell_axis64(xi, c->nx, &x1, &x2);
    cell_axis64(yi, c->ny, &y1, &y2);
    corners_of(x1, x2, y1, y2, ci);
    double pu[4][3], pv[4][3];
    for (int q = 0; q < 4; q++) {
        if (ci[q][0] >= (uint64_t)c->nx || ci[q][1] >= (uint64_t)c->ny) return OR_PANIC;
        double du, dv;
        if (val_at(c, ci[q][0], ci[q][1], c->u_arr, c->u_len, &du) != OR_OK) return OR_ERR;
        if (val_at(c, ci[q][0], ci[q][1], c->v_arr, c->v_len, &dv) != OR_OK) return OR_ERR;
        pu[q][0] = pv[q][0] = c->x[ci[q][0]];
        pu[q][1] = pv[q][1] = c->y[ci[q][1]];
        pu[q][2] = du; pv[q][2] = dv;
    }
    if (bilinear64((const double (*)[3])pu, x, y, u) != OR_OK) return OR_ERR;
    if (bilinear64((const double (*)[3])pv, x, y, v) != OR_OK) return OR_ERR;
    double rx = 1.0 / (c->x[1] - c->x[0]);
    double ry = 1.0 / (c->y[1] - c->y[0]);
    *ux = (pu[3][2] - pu[0][2]) * rx;
    *uy = (pu[1][2] - pu[0][2]) * ry;
    *vx = (pv[3][2] - pv[0][2]) * rx;
    *vy = (pv[1][2] - pv[0][2]) * ry;
    return OR_OK;
}

/* wave_ray_path.rs:177-188 */
int or_group_velocity(double k, double h, double *cg)
{
    if (h <= 0.0) { *cg = NAN; return OR_OK; }
    if (k <= 0.0) return OR_ERR;
    double ch = cosh(k * h);
    *cg = (G / 2.0) * ((tanh(k * h) + (k * h) / (ch * ch)) / sqrt(k * G * tanh(k * h)));
    return OR_OK;
}

/* wave_ray_path.rs:207-216; the `* 1.0` factors are exact and kept */
void or_dkdt_bathy(double k, double h, double hx, double hy, double *dkx, double *dky)
{
    *dkx = (-0.5) * k * 1.0 / sinh(k * h) * 1.0 / cosh(k * h)
           * sqrt(G * k * tanh(k * h)) * hx;
    *dky = (-0.5) * k * 1.0 / sinh(k * h) * 1.0 / cosh(k * h)
           * sqrt(G * k * tanh(k * h)) * hy;
}

/* wave_ray_path.rs:118-150 */
int or_odes(const or_bathy *b, const or_current *c, int prec, const double s[4], double ds[4])
{
    double x = s[0], y = s[1], kx = s[2], ky = s[3];
    double h, dhdx, dhdy, u, v, ux, uy, vx, vy;
    int st;
    if (prec == OR_PREC_FP64_INTERP) {
        st = or_bathy_depth_and_gradient64(b, x, y, &h, &dhdx, &dhdy);
        if (st != OR_OK) return st;
        st = or_current_and_gradient64(c, x, y, &u, &v, &ux, &uy, &vx, &vy);
        if (st != OR_OK) return st;
    } else {
        float hf, hxf, hyf;
        st = or_bathy_depth_and_gradient(b, (float)x, (float)y, &hf, &hxf, &hyf);
        if (st != OR_OK) return st;
        h = (double)hf; dhdx = (double)hxf; dhdy = (double)hyf;
        st = or_current_and_gradient(c, x, y, &u, &v, &ux, &uy, &vx, &vy);
        if (st != OR_OK) return st;
    }
    double k = sqrt(kx * kx + ky * ky);
    double theta = atan2(ky, kx);
    double cg;
    st = or_group_velocity(k, h, &cg);
    if (st != OR_OK) return st;
    double cgx = cg * cos(theta) + u;
    double cgy = cg * sin(theta) + v;
    double dkxb, dkyb;
    or_dkdt_bathy(k, h, dhdx, dhdy, &dkxb, &dkyb);
    ds[0] = cgx;
    ds[1] = cgy;
    ds[2] = dkxb - kx * ux - ky * vx;
    ds[3] = dkyb - kx * uy - ky * vy;
    return OR_OK;
}

/* wave_ray_path.rs:220-234: any error -> all-NaN derivatives */
void or_system(const or_bathy *b, const or_current *c, int prec, const double s[4], double ds[4])
{
    if (or_odes(b, c, prec, s, ds) != OR_OK) {
        ds[0] = NAN; ds[1] = NAN; ds[2] = NAN; ds[3] = NAN;
    }
}

static int system_st(const or_bathy *b, const or_current *c, int prec, const double s[4], double ds[4])
{
    int st = or_odes(b, c, prec, s, ds);
    if (st != OR_OK) { ds[0] = NAN; ds[1] = NAN; ds[2] = NAN; ds[3] = NAN; }
    return st;
}

static int all_nan4(const double v[4])
{
    return isnan(v[0]) && isnan(v[1]) && isnan(v[2]) && isnan(v[3]);
}

/* ode_solvers 0.4.0 Rk4::integrate: ceil(...).to_usize().unwrap(); num-traits
 * accepts a float f with -1 < f < 2^64. */
int64_t or_num_steps(double t0, double t_end, double dt)
{
    double q = ceil((t_end - t0) / dt);
    if (!(q > -1.0 && q < 18446744073709551616.0)) return -1;
    if (q >= 9.2e18) return -1;
    return (int64_t)q;
}

/* The stage state y + k h for h = dt/2 or dt, in one of the orders the
 * unvendored ode_solvers 0.4.0 might use (which one it uses is parity
 * unpinned): OR_STAGE_DIRECT y + k*(dt/2) (and y + k*dt), OR_STAGE_SCALE_DT
 * y + (k*dt)/2, OR_STAGE_HALF_FIRST y + (k*0.5)*dt.  For h = dt the three
 * are the same expression. */
static double stage_state(int order, double y, double k, double dt, double half, int is_half)
{
    if (!is_half) return y + k * dt;
    switch (order) {
    case OR_STAGE_SCALE_DT: return y + (k * dt) / 2.0;
    case OR_STAGE_HALF_FIRST: return y + (k * 0.5) * dt;
    default: return y + k * half;
    }
}

/* ray.rs:198-213 + ode_solvers Rk4 (new/integrate/step) + solout
 * (wave_ray_path.rs:236-246).  stage_order: see stage_state; time_order
 * OR_TIME_ACCUMULATE t += dt (Rk4 keeps x = x + step_size) or
 * OR_TIME_PRODUCT t0 + i*dt; combo_order OR_COMBO_SCALE_SIXTH
 * y + (((k1 + k2*2) + k3*2) + k4)*(dt/6) or OR_COMBO_DT_THEN_SIX
 * y + ((((k1 + k2*2) + k3*2) + k4)*dt)/6.  Returns rows written or -1 / -2
 * (panic). */
static int64_t trace_one(const or_bathy *b, const or_current *c, int prec, int solout,
                         int stage_order, int time_order, int combo_order,
                         const double s0[4], double t0, double dt, int64_t n,
                         int64_t max_rows, double *t_rows, double *s_rows)
{
    double half = dt / 2.0, sixth = dt / 6.0;
    double y[4] = {s0[0], s0[1], s0[2], s0[3]};
    double t = t0;
    int64_t rows = 0;
    if (rows < max_rows) { t_rows[rows] = t; memcpy(s_rows + 4 * rows, y, sizeof y); }
    rows++;
    double k1[4], k2[4], k3[4], k4[4], tmp[4], yn[4];
    int have_next = 0;
    for (int64_t i = 0; i < n; i++) {
        int st;
        if (have_next) { memcpy(k1, tmp, sizeof k1); st = OR_OK; }
        else st = system_st(b, c, prec, y, k1);
        if (st == OR_PANIC) return -2;
        for (int q = 0; q < 4; q++) tmp[q] = stage_state(stage_order, y[q], k1[q], dt, half, 1);
        if (system_st(b, c, prec, tmp, k2) == OR_PANIC) return -2;
        for (int q = 0; q < 4; q++) tmp[q] = stage_state(stage_order, y[q], k2[q], dt, half, 1);
        if (system_st(b, c, prec, tmp, k3) == OR_PANIC) return -2;
        for (int q = 0; q < 4; q++) tmp[q] = stage_state(stage_order, y[q], k3[q], dt, half, 0);
        if (system_st(b, c, prec, tmp, k4) == OR_PANIC) return -2;
        for (int q = 0; q < 4; q++) {
            const double sum = ((k1[q] + k2[q] * 2.0) + k3[q] * 2.0) + k4[q];
            yn[q] = y[q] + (combo_order == OR_COMBO_DT_THEN_SIX ? (sum * dt) / 6.0 : sum * sixth);
        }
        t = time_order == OR_TIME_PRODUCT ? t0 + (double)(i + 1) * dt : t + dt;
        if (rows < max_rows) { t_rows[rows] = t; memcpy(s_rows + 4 * rows, yn, sizeof yn); }
        rows++;
        memcpy(y, yn, sizeof y);
        have_next = 0;
        if (all_nan4(yn) || all_nan4(k1)) break;
        if (solout == OR_SOLOUT_NEXT_DERIV) {
            if (system_st(b, c, prec, y, tmp) == OR_PANIC) return -2;
            have_next = 1;
            if (all_nan4(tmp)) break;
        }
    }
    return rows;
}

int64_t or_trace_single(const or_bathy *b, const or_current *c, int prec, int solout,
                        const double s0[4], double t0, double t_end, double dt,
                        int64_t max_rows, double *rows_out)
{
    int64_t n = or_num_steps(t0, t_end, dt);
    if (n < 0) return -1;
    int64_t cap = n + 1;
    double *tr = (double *)malloc(sizeof(double) * (size_t)cap);
    double *sr = (double *)malloc(sizeof(double) * 4 * (size_t)cap);
    if (!tr || !sr) { free(tr); free(sr); return -1; }
    int64_t rows = trace_one(b, c, prec, solout, OR_STAGE_DIRECT, OR_TIME_ACCUMULATE,
                             OR_COMBO_SCALE_SIXTH, s0, t0, dt, n, cap, tr, sr);
    if (rows > 0) {
        for (int64_t r = 0; r < rows && r < max_rows; r++) {
            rows_out[5 * r] = tr[r];
            for (int q = 0; q < 4; q++) rows_out[5 * r + 1 + q] = sr[4 * r + q];
        }
    }
    free(tr); free(sr);
    return rows;
}

/* ray.rs:98-127: one task per ray (rayon par_iter) -> OpenMP dynamic,1 */
int or_trace_many_ordered(const or_bathy *b, const or_current *c, int prec, int solout,
                          int stage_order, int time_order, int combo_order,
                          int64_t n_rays, const double *x0, const double *y0,
                          const double *kx0, const double *ky0,
                          double t0, double t_end, double dt, int nthreads,
                          double *t_out, double *x_out, double *y_out,
                          double *kx_out, double *ky_out, int32_t *len_out,
                          double *end_out)
{
    int64_t n = or_num_steps(t0, t_end, dt);
    if (n < 0) return -1;
    int64_t rows_max = n + 1;
    if (t_out) {
        double t = t0;
        t_out[0] = t;
        for (int64_t s = 1; s <= n; s++) {
            t = time_order == OR_TIME_PRODUCT ? t0 + (double)s * dt : t + dt;
            t_out[s] = t;
        }
    }
    int panicked = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : panicked)
#endif
    {
        double *tr = (double *)malloc(sizeof(double) * (size_t)rows_max);
        double *sr = (double *)malloc(sizeof(double) * 4 * (size_t)rows_max);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t r = 0; r < n_rays; r++) {
            double s0[4] = {x0[r], y0[r], kx0[r], ky0[r]};
            int64_t rows = trace_one(b, c, prec, solout, stage_order, time_order, combo_order,
                                     s0, t0, dt, n, rows_max, tr, sr);
            if (rows < 0) { panicked = 1; rows = 0; }
            if (len_out) len_out[r] = (int32_t)rows;
            for (int64_t s = 0; s < rows_max; s++) {
                int valid = s < rows;
                size_t o = (size_t)s * (size_t)n_rays + (size_t)r;
                if (x_out) x_out[o] = valid ? sr[4 * s + 0] : NAN;
                if (y_out) y_out[o] = valid ? sr[4 * s + 1] : NAN;
                if (kx_out) kx_out[o] = valid ? sr[4 * s + 2] : NAN;
                if (ky_out) ky_out[o] = valid ? sr[4 * s + 3] : NAN;
            }
            if (end_out) {
                for (int q = 0; q < 4; q++)
                    end_out[4 * r + q] = rows > 0 ? sr[4 * (rows - 1) + q] : NAN;
            }
        }
        free(tr); free(sr);
    }
    return panicked ? -2 : 0;
}

int or_trace_many(const or_bathy *b, const or_current *c, int prec, int solout,
                  int64_t n_rays, const double *x0, const double *y0,
                  const double *kx0, const double *ky0,
                  double t0, double t_end, double dt, int nthreads,
                  double *t_out, double *x_out, double *y_out,
                  double *kx_out, double *ky_out, int32_t *len_out,
                  double *end_out)
{
    return or_trace_many_ordered(b, c, prec, solout, OR_STAGE_DIRECT, OR_TIME_ACCUMULATE,
                                 OR_COMBO_SCALE_SIXTH, n_rays,
                                 x0, y0, kx0, ky0, t0, t_end, dt, nthreads, t_out, x_out, y_out,
                                 kx_out, ky_out, len_out, end_out);
}

void or_system_many(const or_bathy *b, const or_current *c, int prec, int64_t n,
                    const double *s, double *ds_out, int32_t *status_out)
{
    for (int64_t i = 0; i < n; i++) {
        int st = system_st(b, c, prec, s + 4 * i, ds_out + 4 * i);
        if (status_out) status_out[i] = st;
    }
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
