/*
 * mr_oracle.h -- CPU oracle for the mantaray many-ray integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar C restatement of the reference
 * algorithm (mantaray 0.1.0) used as the checker for the HIP product path and
 * as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls anything in this directory.
 *
 * Every function restates one reference function; the citation is given at
 * the definition in mr_oracle.c.  Parity status of each part is listed in the
 * header comment of mr_oracle.c.
 */
#ifndef MR_ORACLE_H
#define MR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* outcome of a provider / RHS call: the reference's Result<_, Error> plus
 * the cases where the Rust code would panic (Vec index out of range). */
enum { OR_OK = 0, OR_ERR = 1, OR_PANIC = 2 };

/* precision modes */
enum { OR_PREC_REFERENCE = 0, OR_PREC_FP64_INTERP = 1 };

/* solout variants (which derivative ode_solvers hands to solout) */
enum { OR_SOLOUT_STATE = 0, OR_SOLOUT_NEXT_DERIV = 1 };

/* stage-state and time orders of the RK4 step (parity unpinned; the
 * default is the first of each) */
enum { OR_STAGE_DIRECT = 0, OR_STAGE_SCALE_DT = 1, OR_STAGE_HALF_FIRST = 2 };
enum { OR_TIME_ACCUMULATE = 0, OR_TIME_PRODUCT = 1 };
enum { OR_COMBO_SCALE_SIXTH = 0, OR_COMBO_DT_THEN_SIX = 1 };

/* bathymetry kinds */
enum { OR_BATHY_CONSTANT = 0, OR_BATHY_SLOPE = 1, OR_BATHY_ARRAY = 2, OR_BATHY_GRID = 3 };
/* current kinds */
enum { OR_CURRENT_CONSTANT = 0, OR_CURRENT_GRID = 1 };

typedef struct {
    int kind;
    float h;                                   /* ConstantDepth */
    float h0, x0, y0, dhdx, dhdy;              /* ConstantSlope */
    int64_t an; const float *arr;              /* ArrayDepth, square an x an, arr[x*an + y] */
    int64_t nx, ny;                            /* CartesianNetcdf3 */
    const float *x, *y;
    const double *depth; int64_t depth_len;
} or_bathy;

typedef struct {
    int kind;
    double u, v;                               /* ConstantCurrent */
    int64_t nx, ny;                            /* CartesianCurrent */
    const double *x, *y;
    const double *u_arr, *v_arr; int64_t u_len, v_len;
} or_current;

int or_bilinear32(const float pts[4][3], float tx, float ty, float *out);
int or_bathy_nearest(const or_bathy *b, float target, int axis, float *idx);
int or_current_nearest(const or_current *c, double target, int axis, double *idx);
int or_bathy_corners(const or_bathy *b, float px, float py, uint64_t out[8]);
int or_current_corners(const or_current *c, double x, double y, uint64_t out[8]);
int or_bathy_depth_and_gradient(const or_bathy *b, float px, float py,
                                float *h, float *hx, float *hy);
int or_current_and_gradient(const or_current *c, double x, double y,
                            double *u, double *v, double *ux, double *uy,
                            double *vx, double *vy);
int or_group_velocity(double k, double h, double *cg);
void or_dkdt_bathy(double k, double h, double hx, double hy, double *dkx, double *dky);
int or_odes(const or_bathy *b, const or_current *c, int prec, const double s[4], double ds[4]);
void or_system(const or_bathy *b, const or_current *c, int prec, const double s[4], double ds[4]);

/* fp64-interpolation providers (the differentiable variant; not in the reference) */
int or_bathy_depth_and_gradient64(const or_bathy *b, double px, double py,
                                  double *h, double *hx, double *hy);
int or_current_and_gradient64(const or_current *c, double x, double y,
                              double *u, double *v, double *ux, double *uy,
                              double *vx, double *vy);

/* n = ceil((t_end - t0)/dt) as ode_solvers computes it; -1 where it panics */
int64_t or_num_steps(double t0, double t_end, double dt);

/* One ray.  rows_out[r*5 + {0..4}] = (t, x, y, kx, ky); capacity max_rows.
 * Returns the number of rows, or -1 on bad step count. */
int64_t or_trace_single(const or_bathy *b, const or_current *c, int prec, int solout,
                        const double s0[4], double t0, double t_end, double dt,
                        int64_t max_rows, double *rows_out);

/* Many rays, step-major SoA output with NaN padding past len[r].
 * t_out has n+1 entries; x/y/kx/ky_out have (n+1)*n_rays entries
 * (row s, ray r at s*n_rays + r).  Any of the output pointers may be NULL
 * (endpoint-only callers).  end_out (4*n_rays, may be NULL) receives the last
 * valid row of each ray.  nthreads <= 0 uses all OpenMP threads.
 * Returns 0 on success, -1 on bad step count. */
int or_trace_many(const or_bathy *b, const or_current *c, int prec, int solout,
                  int64_t n_rays, const double *x0, const double *y0,
                  const double *kx0, const double *ky0,
                  double t0, double t_end, double dt, int nthreads,
                  double *t_out, double *x_out, double *y_out,
                  double *kx_out, double *ky_out, int32_t *len_out,
                  double *end_out);

/* or_trace_many with another stage-state order (y + k*(dt/2), y + (k*dt)/2
 * or y + (k*0.5)*dt), time column (t += dt or t0 + i*dt) and step
 * combination (sum*(dt/6) or (sum*dt)/6). */
int or_trace_many_ordered(const or_bathy *b, const or_current *c, int prec, int solout,
                          int stage_order, int time_order, int combo_order,
                          int64_t n_rays, const double *x0, const double *y0,
                          const double *kx0, const double *ky0,
                          double t0, double t_end, double dt, int nthreads,
                          double *t_out, double *x_out, double *y_out,
                          double *kx_out, double *ky_out, int32_t *len_out,
                          double *end_out);

/* Batched RHS evaluation for golden-vector checks: ds_out[4*i..] = system(s). */
void or_system_many(const or_bathy *b, const or_current *c, int prec, int64_t n,
                    const double *s, double *ds_out, int32_t *status_out);

int or_max_threads(void);

/* mr_oracle_ld.c: the fp64-interpolation trace in long double, as the
 * finite-difference truth of the adjoint checks.  Each ray is traced from
 * s0 + eps ds0 and from s0 - eps ds0 (s0, ds0: 4 per ray; ds0 may be NULL)
 * with the gridded values + eps d and - eps d (ddepth, du, dv may be NULL);
 * diff = end(+) - end(-) and mid = (end(+) + end(-))/2, each formed in long
 * double, 4 per ray; len_plus / len_minus the row counts (may be NULL);
 * same_cells (may be NULL) 1 where the two traces queried the same grid
 * cells at every evaluation, 0 where one crossed a cell line the other did
 * not (the central difference then spans a kink of the cell slopes).
 * Returns 0, or -1 on a bad step count. */
int or_fd_end_ld(const or_bathy *b, const or_current *c, int64_t n_rays, const double *s0,
                 const double *ds0, const double *ddepth, const double *du, const double *dv,
                 double eps, double t0, double t_end, double dt, int nthreads, double *diff,
                 double *mid, int32_t *len_plus, int32_t *len_minus, int32_t *same_cells);

#ifdef __cplusplus
}
#endif
#endif
