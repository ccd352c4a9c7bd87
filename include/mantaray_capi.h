/*
 * mantaray_capi.h -- C ABI of libmantaray_hip.so, the MI355X (gfx950) many-ray
 * wave integrator.
 *
 * The library replaces the reference's rayon loop ManyRays::trace_many
 * (src/ray.rs:98-127) and the two PyO3 entry points of src/ffi.rs with
 * hand-written HIP kernels: one ray per lane, FP64 state, the reference's RK4
 * step rule and its f32/f64 field interpolation.  Citations are to files of
 * mines-oceanography/mantaray 0.1.0.
 *
 * Conventions
 *   - Every function returns MR_OK (0) or a negative MR_ERR_* code; the
 *     calling thread's last error message is then readable with
 *     mr_last_error().  Where the reference panics (a missing file, a
 *     non-finite step count, a grid with a dimension of length 1) this
 *     library returns an error code instead.
 *   - Field handles are immutable after creation and may be shared between
 *     threads and calls.  Their device copies are built on first use on each
 *     device and kept until mr_field_free.
 *   - Trajectory outputs are step-major SoA: row s of ray r is element
 *     s * n_rays + r.  Rows at or past len[r] are NaN, which is the padded
 *     layout of mantaray.core.ray_tracing (python/mantaray/core.py:114-125).
 *   - Buffers are caller-allocated.  With mr_options.memory == MR_MEM_HOST
 *     the pointers are host memory and the call is synchronous; with
 *     MR_MEM_DEVICE they are device memory of the selected device and the
 *     work is enqueued on mr_options.stream (0 = the null stream) without a
 *     host synchronisation.
 */
#ifndef MANTARAY_CAPI_H
#define MANTARAY_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MR_VERSION_MAJOR 0
#define MR_VERSION_MINOR 1
#define MR_VERSION_PATCH 0

/* status codes */
#define MR_OK                    0
#define MR_ERR_INVALID_ARGUMENT (-1)
#define MR_ERR_IO               (-2)  /* file missing or unreadable */
#define MR_ERR_FORMAT           (-3)  /* not NetCDF-3, or variable missing */
#define MR_ERR_DEVICE           (-4)  /* a HIP runtime call failed */
#define MR_ERR_OUT_OF_MEMORY    (-5)
#define MR_ERR_STEP_COUNT       (-6)  /* ceil((t_end-t0)/dt) not a valid usize */
#define MR_ERR_UNSUPPORTED      (-7)  /* input on which the reference panics */

/* field kinds (mr_field_kind) */
#define MR_BATHY_CONSTANT   0   /* src/bathymetry/constant_depth.rs:15-45 */
#define MR_BATHY_SLOPE      1   /* src/bathymetry/constant_slope.rs:28-76 */
#define MR_BATHY_ARRAY      2   /* src/bathymetry/array_depth.rs:9-34 */
#define MR_BATHY_GRID       3   /* src/bathymetry/cartesian_netcdf3.rs:35-471 */
#define MR_CURRENT_CONSTANT 16  /* src/current/constant_current.rs:19-78 */
#define MR_CURRENT_GRID     17  /* src/current/cartesian_current.rs:18-542 */

/* mr_options.precision */
#define MR_PREC_REFERENCE   0   /* the reference's f32 interpolation, f64 state */
#define MR_PREC_FP64_INTERP 1   /* all interpolation in f64 (differentiable; see below) */

/* mr_options.math */
#define MR_MATH_FAST    0       /* dispersion terms in closed form from one exp */
#define MR_MATH_LITERAL 1       /* atan2/cos/sin/tanh/cosh/sinh as written */

/* mr_options.solout: what stops a ray (src/wave_ray_path.rs:236-246) */
#define MR_SOLOUT_STATE      0  /* the new state is all-NaN */
#define MR_SOLOUT_NEXT_DERIV 1  /* ... or the derivative at the new state is */

/* mr_options.memory */
#define MR_MEM_HOST   0
#define MR_MEM_DEVICE 1

typedef struct mr_field mr_field;

typedef struct mr_options {
    int32_t struct_size;   /* sizeof(mr_options); set by mr_options_init */
    int32_t precision;     /* MR_PREC_* */
    int32_t math;          /* MR_MATH_* */
    int32_t solout;        /* MR_SOLOUT_* */
    int32_t memory;        /* MR_MEM_* */
    int32_t device;        /* HIP device ordinal; -1 = current device */
    int32_t block_size;    /* threads per block, multiple of 64; 0 = tuned */
    int32_t blocks_per_cu; /* mr_batch_run: resident workgroups per CU; 0 = tuned */
    void   *stream;        /* hipStream_t for MR_MEM_DEVICE; NULL = null stream */
    int32_t record_tile;   /* gridded fields, reference precision: cell records
                              are stored in square tiles of 2^record_tile cells
                              a side (0 = row-major ... 5 = 32 x 32); -1 =
                              default (3, 8 x 8).  Results do not depend on it.
                              A field keeps the tables of each tile shape it
                              was used with until it is freed. */
} mr_options;

/* Defaults: reference precision, fast math, MR_SOLOUT_STATE, host memory,
 * current device, tuned block size, 8 x 8 record tiles.  A struct_size of a
 * caller built before record_tile existed (the offset of record_tile) is
 * accepted; record_tile then takes its default. */
void mr_options_init(mr_options *opt);

const char *mr_last_error(void);
int mr_version(int32_t *major, int32_t *minor, int32_t *patch);

/* ---- field handles --------------------------------------------------- */

/* CartesianNetcdf3::open (src/bathymetry/cartesian_netcdf3.rs:167-256):
 * reads the coordinate variables as f32 and the depth as f64 from a
 * NetCDF-3 classic or 64-bit-offset file, in file order, ignoring dimension
 * order and attributes.  Reference names: "x", "y", "depth" (src/ffi.rs:36). */
int mr_bathy_open_nc3(const char *path, const char *xname, const char *yname,
                      const char *depth_name, mr_field **out);

/* CartesianCurrent::open (src/current/cartesian_current.rs:58-213): all four
 * variables promoted to f64.  Reference names "x","y","u","v" (src/ffi.rs:38). */
int mr_current_open_nc3(const char *path, const char *xname, const char *yname,
                        const char *uname, const char *vname, mr_field **out);

/* CartesianNetcdf3 from arrays: depth is the flat row-major array indexed
 * nx * j + i (cartesian_netcdf3.rs:465-471); depth_len may differ from nx*ny
 * exactly as a file's variable may (cells past the end are out of bounds). */
int mr_bathy_from_arrays(int64_t nx, int64_t ny, const float *x, const float *y,
                         int64_t depth_len, const double *depth, mr_field **out);

/* CartesianCurrent from arrays (cartesian_current.rs:18-27, :421-427). */
int mr_current_from_arrays(int64_t nx, int64_t ny, const double *x, const double *y,
                           int64_t u_len, const double *u, int64_t v_len, const double *v,
                           mr_field **out);

/* ConstantDepth::new (src/bathymetry/constant_depth.rs:42-44) */
int mr_bathy_constant(float h, mr_field **out);
/* ConstantSlope (src/bathymetry/constant_slope.rs:28-76) */
int mr_bathy_slope(float h0, float x0, float y0, float dhdx, float dhdy, mr_field **out);
/* ArrayDepth::new (src/bathymetry/array_depth.rs:38-41): square n x n array,
 * a[ix * n + iy] */
int mr_bathy_array(int64_t n, const float *a, mr_field **out);
/* ConstantCurrent::new (src/current/constant_current.rs:30-32) */
int mr_current_constant(double u, double v, mr_field **out);

int mr_field_free(mr_field *f);

/* Introspection (host-only; never touches a device). */
int mr_field_kind(const mr_field *f, int32_t *kind);
int mr_field_grid_shape(const mr_field *f, int64_t *nx, int64_t *ny, int64_t *len_a,
                        int64_t *len_b);
/* Copies the axes widened to f64 and the value arrays (depth, or u then v). */
int mr_field_copy_grid(const mr_field *f, double *x, double *y, double *a, double *b);

/* Builds the field's device tables (8 x 8 record tiles) on `device` (-1 =
 * current) now instead of on first use.  No reference counterpart (the reference has no device);
 * lets callers keep table construction out of a timed region. */
int mr_field_upload(const mr_field *f, int32_t device);

/* ---- integration ----------------------------------------------------- */

/* ode_solvers 0.4.0 Rk4::integrate step count: n = ceil((t_end - t0)/dt).
 * Writes n (the trajectory has n + 1 rows).  MR_ERR_STEP_COUNT where the
 * reference panics in to_usize().unwrap(). */
int mr_num_steps(double t0, double t_end, double dt, int64_t *n);

/* ManyRays::trace_many(t0, t_end, dt) (src/ray.rs:98-127) as called by
 * _mantaray.ray_tracing (src/ffi.rs:51-85).
 *   t_out   [n+1]          time of each row (t accumulates t += dt)
 *   x_out.. [(n+1)*n_rays] step-major trajectories, NaN past len[r]
 *   len_out [n_rays]       rows of each ray, the terminal all-NaN row included
 *   end_out [4*n_rays]     last row of each ray (x, y, kx, ky), may be NULL
 * Any of t_out, x_out, y_out, kx_out, ky_out may be NULL (not written). */
int mr_trace_many(const mr_field *bathy, const mr_field *current, int64_t n_rays,
                  const double *x0, const double *y0, const double *kx0, const double *ky0,
                  double t0, double t_end, double dt, const mr_options *opt,
                  double *t_out, double *x_out, double *y_out, double *kx_out,
                  double *ky_out, int32_t *len_out, double *end_out);

/* SingleRay::trace_individual (src/ray.rs:198-213) as called by
 * _mantaray.single_ray (src/ffi.rs:25-49).  rows_out holds (n+1) rows of
 * (t, x, y, kx, ky); *n_rows receives the rows produced. */
int mr_single_ray(const mr_field *bathy, const mr_field *current,
                  double x0, double y0, double kx0, double ky0,
                  double t0, double t_end, double dt, const mr_options *opt,
                  double *rows_out, int64_t *n_rows);

/* WaveRayPath::system (src/wave_ray_path.rs:220-234) at n states
 * (s[4i..4i+3] = x, y, kx, ky): ds = derivatives, all-NaN where odes errs. */
int mr_rhs_many(const mr_field *bathy, const mr_field *current, int64_t n,
                const double *s, double *ds, const mr_options *opt);

/* ---- differentiable path: fp64 interpolation and the adjoint ----------
 *
 * No counterpart in the reference, which has no derivatives.  These run the
 * reference's algorithm (ManyRays::trace_many, src/ray.rs:98-127, and the
 * providers of src/bathymetry/cartesian_netcdf3.rs and
 * src/current/cartesian_current.rs) with every interpolation in f64
 * (MR_PREC_FP64_INTERP; the bathymetry is queried at the f64 point), and the
 * discrete adjoint of that computation.  Both take device memory only
 * (mr_options.memory == MR_MEM_DEVICE) and enqueue on mr_options.stream.
 *
 * depth, u, v: device arrays that replace the fields' own values (same
 * lengths as the fields' arrays: depth_len, u_len, v_len), or NULL to use
 * them.  They let a caller differentiate with respect to the values and
 * change them between calls without building new fields.
 *
 * Checkpoints: with stride K, checkpoint c (0 <= c < mr_checkpoint_count)
 * is the state at row c*K, stored step-major SoA: ckpt[(4c + q) * n_rays + r]
 * for q = x, y, kx, ky. */

/* Number of checkpoints of a trace: max(1, ceil(n / stride)). */
int mr_checkpoint_count(double t0, double t_end, double dt, int32_t stride, int64_t *count);

/* Bytes of the work buffer mr_adjoint_many needs: 2 * stride * 16 * n_rays
 * doubles (the recomputed states and stage states of two segments). */
int mr_adjoint_workspace_size(int64_t n_rays, int32_t stride, int64_t *bytes);

/* mr_trace_many in fp64-interpolation mode (whatever mr_options.precision
 * says), with value overrides and, when ckpt is not NULL, checkpoints every
 * `stride` steps.  Outputs as mr_trace_many; each may be NULL. */
int mr_trace_fp64(const mr_field *bathy, const mr_field *current,
                  const double *depth, const double *u, const double *v,
                  int64_t n_rays, const double *x0, const double *y0,
                  const double *kx0, const double *ky0,
                  double t0, double t_end, double dt, int32_t stride,
                  const mr_options *opt, double *ckpt,
                  double *t_out, double *x_out, double *y_out, double *kx_out,
                  double *ky_out, int32_t *len_out, double *end_out);

/* Reverse sweep of the discrete adjoint of mr_trace_fp64 (MR_SOLOUT_STATE).
 * Inputs: the checkpoints, len_out and end_out of the forward call (same
 * fields, values, t0, t_end, dt, stride and math mode), and the cotangents
 * of a loss: g_end [4*n_rays] on each ray's last row (end_out's layout),
 * and/or g_x .. g_ky [(n+1)*n_rays] on the rows (step-major); any may be
 * NULL.  A ray's all-NaN last row (edge exit) receives no cotangent.
 * Outputs (each may be NULL): the gradients with respect to the initial
 * states g_x0 .. g_ky0 [n_rays] and to the gridded values g_depth
 * [depth_len], g_u [u_len], g_v [v_len]; the grids are cleared first and
 * summed with f64 atomics, so they vary in the last bits between runs.
 * work: device buffer of mr_adjoint_workspace_size bytes.  opt->block_size
 * selects the sweep kernel: 0 by the launch's size, 64 / 128 / 256 one, two
 * or four waves per 64 rays. */
int mr_adjoint_many(const mr_field *bathy, const mr_field *current,
                    const double *depth, const double *u, const double *v,
                    int64_t n_rays, double t0, double t_end, double dt, int32_t stride,
                    const mr_options *opt, const double *ckpt, const int32_t *len,
                    const double *end, const double *g_end,
                    const double *g_x, const double *g_y, const double *g_kx, const double *g_ky,
                    void *work, int64_t work_bytes,
                    double *g_x0, double *g_y0, double *g_kx0, double *g_ky0,
                    double *g_depth, double *g_u, double *g_v);

/* Size of the device array mr_adjoint_many_noalloc uses as its work buffer:
 * 32 MiB per device, held from the library's first use of it on a device
 * until the process ends.  A launch fits when mr_adjoint_workspace_size
 * (n_rays, stride) is at most this: e.g. up to 2,621 rays at stride 50, or
 * 13,107 rays at stride 10. */
#define MR_ADJOINT_SCRATCH_BYTES (32ll << 20)

/* mr_adjoint_many without a work argument, for callers that cannot
 * allocate device memory (code under HIP-graph capture, serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  The work
 * buffer is a static device array of MR_ADJOINT_SCRATCH_BYTES; sizes that
 * need more are refused with MR_ERR_INVALID_ARGUMENT.  The call allocates
 * nothing: it does not build field tables either, so the fields must already
 * hold their fp64 tables on this device (any mr_trace_fp64 call on the
 * device builds them; the checkpoints this call needs come from one).  Every
 * call on a device shares that device's array: calls that overlap on one
 * device (other streams, other threads) race and give wrong gradients.
 * Arguments, outputs and results are mr_adjoint_many's. */
int mr_adjoint_many_noalloc(const mr_field *bathy, const mr_field *current,
                            const double *depth, const double *u, const double *v,
                            int64_t n_rays, double t0, double t_end, double dt, int32_t stride,
                            const mr_options *opt, const double *ckpt, const int32_t *len,
                            const double *end, const double *g_end,
                            const double *g_x, const double *g_y, const double *g_kx,
                            const double *g_ky,
                            double *g_x0, double *g_y0, double *g_kx0, double *g_ky0,
                            double *g_depth, double *g_u, double *g_v);

/* ---- batched scenes: one persistent launch ----------------------------
 *
 * No counterpart in the reference, which traces one scene per call
 * (ManyRays::trace_many, src/ray.rs:98-127).  Many scenes, each with its own
 * fields, step, step count and rays, are traced by one launch of a fixed
 * number of workgroups that take rays from a shared queue: a lane whose ray
 * has ended takes the next ray of any scene.  Each ray's result is what
 * mr_trace_many gives for its scene (row count and last row; bitwise the
 * same numbers).  All scenes must share the bathymetry kind and the current
 * kind.  Device memory only. */
typedef struct mr_scene {
    const mr_field *bathy, *current;
    int64_t n_rays;
    const double *x0, *y0, *kx0, *ky0;   /* [n_rays] each, device */
    double t0, t_end, dt;
    int32_t *len_out;                    /* [n_rays], device */
    double *end_out;                     /* [4*n_rays], device: last row (x, y, kx, ky) */
} mr_scene;

typedef struct mr_batch mr_batch;

/* Builds the scenes' device tables and the batch's queue and descriptors
 * (the only allocations); scenes[] is copied, but its fields and arrays are
 * referenced: they must outlive the batch. */
int mr_batch_create(int32_t n_scenes, const mr_scene *scenes, const mr_options *opt,
                    mr_batch **out);
/* Enqueues the trace on opt->stream: a queue reset and one persistent
 * launch, nothing allocated and no host synchronisation, so the call can be
 * captured in a HIP graph and replayed.  Calls on one batch must not
 * overlap (they share the queue).  opt: math, solout, block_size and
 * blocks_per_cu (0 = defaults) */
int mr_batch_run(mr_batch *batch, const mr_options *opt);
/* Total ray count and an upper bound of the ray-steps (sum of n_rays * n). */
int mr_batch_info(const mr_batch *batch, int64_t *n_rays, int64_t *max_ray_steps);
int mr_batch_free(mr_batch *batch);

/* ---- result writers (host only) -------------------------------------- */

/* RayResult::from(SolverResult) then as_json (src/ray_result.rs:124-151,
 * :74-76): one ray's rows up to the first row with a NaN component, as
 * {"t_vec":[..],"x_vec":[..],"y_vec":[..],"kx_vec":[..],"ky_vec":[..]} with
 * serde_json's f64 text.  Row r is t[r] and x[r*stride] .. ky[r*stride]
 * (stride = n_rays for one ray of mr_trace_many's step-major output, offset
 * to that ray).  *len receives the text's length; the text and a NUL are
 * written to buf only when cap > *len (call with buf NULL to size it). */
int mr_ray_result_json(const double *t, const double *x, const double *y, const double *kx,
                       const double *ky, int64_t n_rows, int64_t stride, char *buf,
                       int64_t cap, int64_t *len);

/* output_or_append_to_tsv_file (src/ray.rs:218-243): appends to path
 * (created if missing) the header "t x y kx ky", one line per row up to the
 * first with a NaN x, each value followed by a space and the line by " \n",
 * then "END"; values as Rust's Display prints f64.  Rows as above. */
int mr_ray_result_append_tsv(const char *path, const double *t, const double *x,
                             const double *y, const double *kx, const double *ky,
                             int64_t n_rows, int64_t stride);

#ifdef __cplusplus
}
#endif
#endif
