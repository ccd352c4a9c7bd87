/*
 * mr_oracle_ld.c -- the fp64-interpolation trace in x87 extended precision.
 *
 * TEST INFRASTRUCTURE ONLY: the finite-difference truth against which the
 * adjoint's gradients are checked.  A central difference of a trace computed
 * in f64 carries the trace's rounding noise divided by the step (about
 * 1e-10 m / 1e-6 m after 2000 steps), which hides errors below 1e-4; the
 * same algorithm carried in long double (64-bit significand) has 2^11 times
 * less noise, and the end-state differences are formed in long double before
 * they are rounded to f64.
 *
 * The algorithm is the oracle's MR_PREC_FP64_INTERP mode (mr_oracle.c:
 * or_bathy_depth_and_gradient64, or_current_and_gradient64, or_odes,
 * trace_one), itself the reference's (cartesian_netcdf3.rs:98-135,
 * cartesian_current.rs:487-542, interpolator.rs:39-84, wave_ray_path.rs
 * :118-216, ode_solvers Rk4) with the interpolation in f64: the same index,
 * cell rule, bilinear formula, cell-difference slopes, right-hand side and
 * RK4 combination, every operation in long double, with g the f64 constant
 * 9.8 widened.  Perturbations are applied in long double: the initial state
 * s0 +- eps ds0 and the values +- eps d.
 *
 * Parity: pinned against the f64 oracle (trajectories agree to the f64
 * rounding level) and, for its derivatives, by the closed forms the tests
 * use; it is not the product and not the reference.
 */
#include "mr_oracle.h"

#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long double ld;

static const ld GL = (ld)9.8;     /* wave_ray_path.rs:23 (the f64 value) */

typedef struct {
    const or_bathy *b;
    const or_current *c;
    const ld *depth, *u, *v;      /* gridded values (perturbed), or NULL */
} fields_ld;

/* The signature of the cells a trace's queries use, in order (FNV-1a over
 * each query's cell indices).  Two traces with the same signature met the
 * same cells at every evaluation, so the cell-difference slopes and the
 * bilinear pieces they used are the same smooth functions: a central
 * difference between them differentiates through no kink. */
static void sig_mix(uint64_t *sig, int64_t v)
{
    if (!sig) return;
    uint64_t h = *sig;
    for (int q = 0; q < 8; q++) {
        h ^= (uint64_t)(v >> (8 * q)) & 0xFFu;
        h *= 0x100000001b3ull;
    }
    *sig = h;
}

/* cartesian_current.rs:231-253 in long double on an axis of n points */
static int nearest_ld(ld target, ld a0, ld a1, int64_t n, ld *idx)
{
    if (n == 0) return OR_ERR;
    if (n == 1) { *idx = 0; return OR_OK; }
    ld sp = fabsl(a1 - a0);
    ld index = (target - a0) / sp;
    if (index < 0.0L || index > (ld)(n - 1)) return OR_ERR;
    *idx = index;
    return OR_OK;
}

/* cartesian_current.rs:283-339: (floor, floor + 1), or (n-2, n-1) on the
 * last node; a NaN index makes every corner (0, 0), where bilinear errs */
static int cell_ld(ld idx, int64_t n, int64_t *i1)
{
    if (isnan(idx)) return OR_ERR;
    int64_t f = (int64_t)floorl(idx);
    *i1 = f > n - 2 ? n - 2 : f;
    return OR_OK;
}

/* interpolator.rs:39-84 in long double; p[q] = (x, y, value) of A, B, C, D */
static int bilinear_ld(const ld p[4][3], ld tx, ld ty, ld *out)
{
    for (int i = 0; i < 4; i++) {
        if (tx == p[i][0] && ty == p[i][1]) { *out = p[i][2]; return OR_OK; }
    }
    ld btx = p[1][0] - p[0][0], bty = p[1][1] - p[0][1];
    ld dtx = p[3][0] - p[0][0], dty = p[3][1] - p[0][1];
    ld ttx = tx - p[0][0], tty = ty - p[0][1];
    ld det = (btx * dty) - (dtx * bty);
    if (det == 0.0L) return OR_ERR;
    ld m00 = dty / det, m01 = -(dtx / det);
    ld m10 = -(bty / det), m11 = btx / det;
    ld u = m00 * ttx + m01 * tty;
    ld w = m10 * ttx + m11 * tty;
    ld a00 = p[0][2];
    ld a10 = p[1][2] - p[0][2];
    ld a01 = p[3][2] - p[0][2];
    ld a11 = ((p[2][2] - p[0][2]) - a10) - a01;
    *out = ((a00 + a10 * u) + a01 * w) + (a11 * u) * w;
    return OR_OK;
}

/* one gridded field at (x, y): the interpolated values of up to two value
 * arrays and their cell-difference slopes */
static int grid_ld(int64_t nx, int64_t ny, ld x0a, ld x1a, ld y0a, ld y1a,
                   const ld *ax, const ld *ay, int64_t len_a, const ld *va, int64_t len_b,
                   const ld *vb, ld x, ld y, ld out[6], uint64_t *sig)
{
    ld xi, yi;
    int64_t i1, j1;
    if (nearest_ld(x, x0a, x1a, nx, &xi) != OR_OK) return OR_ERR;
    if (nearest_ld(y, y0a, y1a, ny, &yi) != OR_OK) return OR_ERR;
    if (cell_ld(xi, nx, &i1) != OR_OK || cell_ld(yi, ny, &j1) != OR_OK) return OR_ERR;
    sig_mix(sig, i1);
    sig_mix(sig, j1);
    const int64_t ci[4][2] = {{i1, j1}, {i1, j1 + 1}, {i1 + 1, j1 + 1}, {i1 + 1, j1}};
    ld pa[4][3], pb[4][3];
    for (int q = 0; q < 4; q++) {
        int64_t k = nx * ci[q][1] + ci[q][0];
        if (k >= len_a || (vb && k >= len_b)) return OR_ERR;
        pa[q][0] = pb[q][0] = ax[ci[q][0]];
        pa[q][1] = pb[q][1] = ay[ci[q][1]];
        pa[q][2] = va[k];
        pb[q][2] = vb ? vb[k] : 0.0L;
    }
    if (bilinear_ld((const ld (*)[3])pa, x, y, &out[0]) != OR_OK) return OR_ERR;
    if (vb && bilinear_ld((const ld (*)[3])pb, x, y, &out[3]) != OR_OK) return OR_ERR;
    ld xs = x1a - x0a, ys = y1a - y0a;
    out[1] = (pa[3][2] - pa[0][2]) / xs;
    out[2] = (pa[1][2] - pa[0][2]) / ys;
    out[4] = (pb[3][2] - pb[0][2]) / xs;
    out[5] = (pb[1][2] - pb[0][2]) / ys;
    return OR_OK;
}

typedef struct {
    ld *bx, *by, *cx, *cy;        /* axes widened */
} axes_ld;

/* Rust's saturating `as usize` of a long double (NaN and negatives to 0) */
static uint64_t sat_usize_ld(ld v)
{
    if (!(v > 0.0L)) return 0;
    if (v >= 18446744073709551616.0L) return UINT64_MAX;
    return (uint64_t)v;
}

static int bathy_ld(const fields_ld *F, const axes_ld *A, ld x, ld y, ld *h, ld *hx, ld *hy,
                    uint64_t *sig)
{
    const or_bathy *b = F->b;
    switch (b->kind) {
    case OR_BATHY_CONSTANT:
        if (isnan(x) || isnan(y)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else { *h = (ld)b->h; *hx = 0.0L; *hy = 0.0L; }
        return OR_OK;
    case OR_BATHY_SLOPE:
        if (isnan(x) || isnan(y)) { *h = NAN; *hx = NAN; *hy = NAN; }
        else {
            *h = ((ld)b->h0 + (ld)b->dhdx * (x - (ld)b->x0)) + (ld)b->dhdy * (y - (ld)b->y0);
            *hx = (ld)b->dhdx; *hy = (ld)b->dhdy;
        }
        return OR_OK;
    case OR_BATHY_ARRAY: {    /* array_depth.rs:26-33: the point truncated to indices */
        const uint64_t xi = sat_usize_ld(x), yi = sat_usize_ld(y);
        if (xi >= (uint64_t)b->an || yi >= (uint64_t)b->an) {
            *h = NAN; *hx = NAN; *hy = NAN;
        } else {
            *h = (ld)b->arr[xi * (uint64_t)b->an + yi]; *hx = 0.0L; *hy = 0.0L;
        }
        return OR_OK;
    }
    default: break;
    }
    if (isnan(x) || isnan(y)) { *h = NAN; *hx = NAN; *hy = NAN; return OR_OK; }
    ld out[6];
    if (grid_ld(b->nx, b->ny, A->bx[0], A->bx[1], A->by[0], A->by[1], A->bx, A->by,
                b->depth_len, F->depth, 0, NULL, x, y, out, sig) != OR_OK)
        return OR_ERR;
    *h = out[0]; *hx = out[1]; *hy = out[2];
    return OR_OK;
}

static int current_ld(const fields_ld *F, const axes_ld *A, ld x, ld y, ld cv[6],
                      uint64_t *sig)
{
    const or_current *c = F->c;
    if (c->kind == OR_CURRENT_CONSTANT) {
        cv[0] = (ld)c->u; cv[3] = (ld)c->v;
        cv[1] = cv[2] = cv[4] = cv[5] = 0.0L;
        return OR_OK;
    }
    return grid_ld(c->nx, c->ny, A->cx[0], A->cx[1], A->cy[0], A->cy[1], A->cx, A->cy,
                   c->u_len, F->u, c->v_len, F->v, x, y, cv, sig);
}

/* wave_ray_path.rs:118-150 with group_velocity :177-188 and dkdt_bathy
 * :207-216; any error -> all-NaN (:222-227) */
static void system_ld(const fields_ld *F, const axes_ld *A, const ld s[4], ld ds[4],
                      uint64_t *sig)
{
    ld h, hx, hy, cv[6];
    ds[0] = ds[1] = ds[2] = ds[3] = NAN;
    if (bathy_ld(F, A, s[0], s[1], &h, &hx, &hy, sig) != OR_OK) return;
    if (current_ld(F, A, s[0], s[1], cv, sig) != OR_OK) return;
    ld kx = s[2], ky = s[3];
    ld k = sqrtl(kx * kx + ky * ky);
    ld theta = atan2l(ky, kx);
    ld cg;
    if (h <= 0.0L) cg = NAN;
    else if (k <= 0.0L) return;
    else {
        ld ch = coshl(k * h);
        cg = (GL / 2.0L) * ((tanhl(k * h) + (k * h) / (ch * ch)) / sqrtl(k * GL * tanhl(k * h)));
    }
    ld c = (-0.5L) * k / sinhl(k * h) / coshl(k * h) * sqrtl(GL * k * tanhl(k * h));
    ds[0] = cg * cosl(theta) + cv[0];
    ds[1] = cg * sinl(theta) + cv[3];
    ds[2] = (c * hx - kx * cv[1]) - ky * cv[4];
    ds[3] = (c * hy - kx * cv[2]) - ky * cv[5];
}

static int all_nan4_ld(const ld v[4])
{
    return isnan(v[0]) && isnan(v[1]) && isnan(v[2]) && isnan(v[3]);
}

/* ode_solvers Rk4 with the MR_SOLOUT_STATE stop; returns the row count and
 * leaves the cell signature of every query in *sig */
static int64_t trace_ld(const fields_ld *F, const axes_ld *A, ld y[4], ld dt, int64_t n,
                        uint64_t *sig)
{
    const ld half = dt / 2.0L, sixth = dt / 6.0L;
    ld k1[4], k2[4], k3[4], k4[4], tmp[4];
    int64_t rows = 1;
    *sig = 0xcbf29ce484222325ull;
    for (int64_t i = 0; i < n; i++) {
        system_ld(F, A, y, k1, sig);
        for (int q = 0; q < 4; q++) tmp[q] = y[q] + k1[q] * half;
        system_ld(F, A, tmp, k2, sig);
        for (int q = 0; q < 4; q++) tmp[q] = y[q] + k2[q] * half;
        system_ld(F, A, tmp, k3, sig);
        for (int q = 0; q < 4; q++) tmp[q] = y[q] + k3[q] * dt;
        system_ld(F, A, tmp, k4, sig);
        for (int q = 0; q < 4; q++)
            y[q] = y[q] + (((k1[q] + k2[q] * 2.0L) + k3[q] * 2.0L) + k4[q]) * sixth;
        rows++;
        if (all_nan4_ld(y)) break;
    }
    return rows;
}

static ld *widen_perturbed(const double *base, const double *dir, int64_t n, ld e)
{
    if (!base || n <= 0) return NULL;
    ld *a = (ld *)malloc(sizeof(ld) * (size_t)n);
    if (!a) return NULL;
    for (int64_t i = 0; i < n; i++) a[i] = (ld)base[i] + (dir ? e * (ld)dir[i] : 0.0L);
    return a;
}

static ld *widen_f32(const float *v, int64_t n)
{
    ld *a = (ld *)malloc(sizeof(ld) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n && a; i++) a[i] = (ld)v[i];
    return a;
}

static ld *widen_f64(const double *v, int64_t n)
{
    ld *a = (ld *)malloc(sizeof(ld) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n && a; i++) a[i] = (ld)v[i];
    return a;
}

int or_fd_end_ld(const or_bathy *b, const or_current *c, int64_t n_rays, const double *s0,
                 const double *ds0, const double *ddepth, const double *du, const double *dv,
                 double eps, double t0, double t_end, double dt, int nthreads, double *diff,
                 double *mid, int32_t *len_plus, int32_t *len_minus, int32_t *same_cells)
{
    int64_t n = or_num_steps(t0, t_end, dt);
    if (n < 0) return -1;
    const int grid_b = b->kind == OR_BATHY_GRID, grid_c = c->kind == OR_CURRENT_GRID;
    axes_ld A = {NULL, NULL, NULL, NULL};
    if (grid_b) { A.bx = widen_f32(b->x, b->nx); A.by = widen_f32(b->y, b->ny); }
    if (grid_c) { A.cx = widen_f64(c->x, c->nx); A.cy = widen_f64(c->y, c->ny); }
    const ld e = (ld)eps;
    fields_ld Fp = {b, c, NULL, NULL, NULL}, Fm = {b, c, NULL, NULL, NULL};
    if (grid_b) {
        Fp.depth = widen_perturbed(b->depth, ddepth, b->depth_len, e);
        Fm.depth = widen_perturbed(b->depth, ddepth, b->depth_len, -e);
    }
    if (grid_c) {
        Fp.u = widen_perturbed(c->u_arr, du, c->u_len, e);
        Fm.u = widen_perturbed(c->u_arr, du, c->u_len, -e);
        Fp.v = widen_perturbed(c->v_arr, dv, c->v_len, e);
        Fm.v = widen_perturbed(c->v_arr, dv, c->v_len, -e);
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
    for (int64_t r = 0; r < n_rays; r++) {
        ld yp[4], ym[4];
        for (int q = 0; q < 4; q++) {
            ld d = ds0 ? e * (ld)ds0[4 * r + q] : 0.0L;
            yp[q] = (ld)s0[4 * r + q] + d;
            ym[q] = (ld)s0[4 * r + q] - d;
        }
        uint64_t sp, sm;
        int64_t lp = trace_ld(&Fp, &A, yp, (ld)dt, n, &sp);
        int64_t lm = trace_ld(&Fm, &A, ym, (ld)dt, n, &sm);
        if (same_cells) same_cells[r] = sp == sm && lp == lm;
        for (int q = 0; q < 4; q++) {
            diff[4 * r + q] = (double)(yp[q] - ym[q]);
            mid[4 * r + q] = (double)((yp[q] + ym[q]) * 0.5L);
        }
        if (len_plus) len_plus[r] = (int32_t)lp;
        if (len_minus) len_minus[r] = (int32_t)lm;
    }
    free(A.bx); free(A.by); free(A.cx); free(A.cy);
    free((void *)Fp.depth); free((void *)Fm.depth);
    free((void *)Fp.u); free((void *)Fm.u); free((void *)Fp.v); free((void *)Fm.v);
    (void)t0;
    return 0;
}
