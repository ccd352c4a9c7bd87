// mr_adjoint.hip -- gfx950 kernels of the fp64-interpolation mode and of its
// discrete adjoint.
//
//   trace64_kernel   ManyRays::trace_many (src/ray.rs:98-127) with every
//                    field query in f64 (mr_device64.h); optionally stores the
//                    state every `stride` steps (checkpoints).
//   adjoint_kernel   reverse sweep of the discrete adjoint of the RK4 loop:
//                    gradients of a loss on the rows (or the last row) with
//                    respect to the initial states and to the gridded depth,
//                    u and v values.
//   rhs64_kernel     WaveRayPath::system in fp64 mode at a batch of states.
//   build_m64        per-cell change of basis (m01, m10) of interpolator.rs
//                    :59-72 in f64.
//   fill             constant fill (gradient grids start at zero).
//
// The adjoint is that of the discrete map the forward computes (ode_solvers
// Rk4::step, k1 .. k4 and (((k1 + 2 k2) + 2 k3) + k4) dt/6), so its
// gradients are those of the computed trajectories, not of the ODE.  Each
// lane owns one ray and walks its checkpointed segments backwards: it
// recomputes the segment's states and stage states from the checkpoint
// with the forward's own step code (bitwise the forward's values) into a
// per-lane column of the work buffer, then runs the step adjoints in reverse
// order over them.
#include <mutex>

#include "mantaray_capi.h"
#include "mr_kernels.h"
#include "mr_kernel_common.h"

namespace mr {

// The work buffer of mr_adjoint_many_noalloc: a file-scope device array of
// MR_ADJOINT_SCRATCH_BYTES that each device holds while the library is
// loaded.  Host code never uses the array's name (in host code it is a
// stand-in, not a device address): adjoint_scratch looks its address up on
// each device once, with hipGetSymbolAddress, and every later use goes
// through that pointer.
static __device__ __align__(256) unsigned char scratch[MR_ADJOINT_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t adjoint_scratch(void **ptr)
{
    *ptr = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

// One RK4 step in fp64 mode from s: the first three stage derivatives k, the
// stage states ys (y + k1 dt/2, y + k2 dt/2, y + k3 dt: the states stages 1-3
// evaluate) and the new state o.  The arithmetic is that of trace_kernel's step, and so is
// the cell scheme of its shared-cell step: two slots with fixed roles, sa
// serving st 0 and st 3 (y and y + k3 dt, nearly the next step's start), sb
// st 1 and st 2 (y + k1 dt/2 and y + k2 dt/2).  Each is refilled while the
// other serves, two evaluations before its next use: sa at st 1 for the
// point guessed at y + k1 dt, sb at st 3 for the next step's pair, guessed
// at y + k3 (3 dt / 2); a lane loads only where the guess is not inside the
// cells its slot holds, so no slot is ever copied.  Every evaluation tests
// the cells it uses and queries again where they are not the target's, so
// the slots choose only what is waited for, never a value.
//
// Before each refill the step waits for the vector memory operations issued
// up to then but the last NST (the caller's stores after the step before,
// issued after every load of it): the serving slot's loads, two evaluations
// old, have landed by then.  Without that wait the compiler waits at the
// serving slot's first use, after the refill, and as the refill is
// conditional it cannot count it: it waits for the refill's loads as well.
template <int BK, int CK, int MATH, int NST>
__device__ __forceinline__ void rk4_step64(const Grid64View &B, const Grid64Ptrs &BP,
                                           const Grid64View &C, const Grid64Ptrs &CP,
                                           const Fields64 &F, const FastConsts &K, Slots64 &sa,
                                           Slots64 &sb, double dt, double half, double sixth,
                                           const double s[4], double k[3][4],
                                           double ys[3][4], double o[4])
{
    double sx = s[0], sy = s[1], skx = s[2], sky = s[3];
    double ax = 0.0, ay = 0.0, akx = 0.0, aky = 0.0;
    double px = 0.0, py = 0.0;                 // latest dx/dt, dy/dt, for the guesses
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const double hs = st < 2 ? half : dt;
        if (st == 1 || st == 3) {
            if (st == 1) vm_wait_upto<NST>();
            else vm_wait_upto<0>();
            Slots64 &refill = st == 1 ? sa : sb;
            const double g = st == 1 ? dt : 1.5 * dt;
            prefetch64<BK, CK>(B, BP, C, CP, F, refill, refill, s[0] + px * g, s[1] + py * g);
        }
        double dx, dy, dkx, dky;
        rhs64<BK, CK, MATH>(B, BP, C, CP, F, (st == 0 || st == 3) ? sa : sb, K, sx, sy, skx,
                            sky, dx, dy, dkx, dky);
        px = dx; py = dy;
        if (st < 3) {
            k[st][0] = dx; k[st][1] = dy; k[st][2] = dkx; k[st][3] = dky;
        }
        if (st == 0) {
            ax = dx; ay = dy; akx = dkx; aky = dky;
        } else if (st < 3) {
            ax = __builtin_fma(dx, 2.0, ax); ay = __builtin_fma(dy, 2.0, ay);
            akx = __builtin_fma(dkx, 2.0, akx); aky = __builtin_fma(dky, 2.0, aky);
        } else {
            ax = ax + dx; ay = ay + dy; akx = akx + dkx; aky = aky + dky;
        }
        sx = s[0] + dx * hs; sy = s[1] + dy * hs;
        skx = s[2] + dkx * hs; sky = s[3] + dky * hs;
        if (st < 3) {
            ys[st][0] = sx; ys[st][1] = sy; ys[st][2] = skx; ys[st][3] = sky;
        }
    }
    o[0] = s[0] + ax * sixth; o[1] = s[1] + ay * sixth;
    o[2] = s[2] + akx * sixth; o[3] = s[3] + aky * sixth;
}

// The work buffer: two halves of [stride][8][work_ld] pairs of doubles.
// Step j of a half holds, for ray r, the state and the three stage states as
// eight 16-byte pairs at (j * 8 + p) * work_ld + r, so that one store or load
// of a wave moves 64 consecutive pairs (1 KB) and a step takes eight of them.
__device__ __forceinline__ void work_put(double *half, int64_t ld, int64_t r, int64_t j,
                                         const double s[4], const double ys[3][4])
{
    double2 *w = reinterpret_cast<double2 *>(half) + j * 8 * ld + r;
    w[0] = make_double2(s[0], s[1]);
    w[ld] = make_double2(s[2], s[3]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        w[(2 + 2 * i) * ld] = make_double2(ys[i][0], ys[i][1]);
        w[(3 + 2 * i) * ld] = make_double2(ys[i][2], ys[i][3]);
    }
}

// loads the fields f of step j for which use(f) holds (use(2p) == use(2p+1))
template <class Use>
__device__ __forceinline__ void work_get(const double *half, int64_t ld, int64_t r, int64_t j,
                                         double d[16], Use use)
{
    const double2 *w = reinterpret_cast<const double2 *>(half) + j * 8 * ld + r;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        if (use(2 * p)) {
            const double2 v = w[p * ld];
            d[2 * p] = v.x;
            d[2 * p + 1] = v.y;
        }
    }
}

// stores a recomputing wave issues after each step: its state and three
// stage states, in pairs
constexpr int RECOMPUTE_STORES = 8;

// ---------------------------------------------------------------------------
// fp64 trace: the loop of trace_kernel (mr_kernels.hip) with rhs64 and the
// two fixed-role cell slots of rk4_step64.  Checkpoint c
// holds the state at row c * stride, stored when the step from that row
// starts; a ray's checkpoints up to its last live step are always written.

template <int BK, int CK, int MATH, int OUT>
__global__ void __launch_bounds__(256) trace64_kernel(Trace64Args a)
{
    if ((int)blockIdx.x == a.ray_blocks) {
        if (threadIdx.x == 0) time_column(a.t0, a.dt, a.n_steps, a.t_out);
        return;
    }
    const int64_t nr = a.n_rays;
    const int64_t r = __builtin_elementwise_min((int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                                                nr - 1);
    double s[4] = {a.x0[r], a.y0[r], a.kx0[r], a.ky0[r]};
    if (OUT == 1) {
        a.xo[r] = s[0]; a.yo[r] = s[1]; a.kxo[r] = s[2]; a.kyo[r] = s[3];
    } else if (OUT == 2) {
        if (a.xo) a.xo[r] = s[0];
        if (a.yo) a.yo[r] = s[1];
        if (a.kxo) a.kxo[r] = s[2];
        if (a.kyo) a.kyo[r] = s[3];
    }
    const Grid64View B = load_view<Grid64View>(a.B);
    const Grid64View C = load_view<Grid64View>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    const double half = a.half, dt = a.dt, sixth = a.sixth;
    Slots64 sa, sb;                    // the step's two cell slots
    load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sa, s[0], s[1]);
    load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sb, s[0], s[1]);
    // all of it in registers before the loop, so that the loop head has
    // nothing pending to wait for (as in trace_kernel)
    settle(B);
    settle(C);
    settle(K);
    settle(sa);
    settle(sb);
    bool alive = true;
    int64_t len = 1;
    int64_t s_i = 1;
    int until_ckpt = 0;                // steps to the next checkpoint
    double *ck = a.ckpt ? a.ckpt + r : nullptr;
    for (; s_i <= a.n_steps; ++s_i) {
        if (!__any(alive)) break;
        if (ck) {
            if (until_ckpt == 0) {
                ck[0] = s[0]; ck[nr] = s[1]; ck[2 * nr] = s[2]; ck[3 * nr] = s[3];
                ck += 4 * nr;
                until_ckpt = a.stride;
            }
            --until_ckpt;
        }
        double o[4] = {MR_NAN, MR_NAN, MR_NAN, MR_NAN};
        if (alive) {
            double k[3][4], ys[3][4], n[4];
            // the row stores of the step before are the last operations
            // issued before this step's first refill
            rk4_step64<BK, CK, MATH, OUT == 1 ? 4 : 0>(B, a.bp, C, a.cp, a.f, K, sa, sb, dt,
                                                       half, sixth, s, k, ys, n);
            // the MR_SOLOUT_NEXT_DERIV test applies from the second step on
            // (solout is asked only after a step)
            if (a.solout == 1 && s_i > 1 && all_nan4(k[0][0], k[0][1], k[0][2], k[0][3])) {
                alive = false;         // the ray ended on the row before
            } else {
                o[0] = n[0]; o[1] = n[1]; o[2] = n[2]; o[3] = n[3];
                s[0] = n[0]; s[1] = n[1]; s[2] = n[2]; s[3] = n[3];
                len = s_i + 1;
                if (all_nan4(n[0], n[1], n[2], n[3])) alive = false;
            }
        }
        const int64_t off = s_i * nr + r;
        if (OUT == 1) {
            __builtin_nontemporal_store(o[0], a.xo + off);
            __builtin_nontemporal_store(o[1], a.yo + off);
            __builtin_nontemporal_store(o[2], a.kxo + off);
            __builtin_nontemporal_store(o[3], a.kyo + off);
        } else if (OUT == 2) {
            if (a.xo) a.xo[off] = o[0];
            if (a.yo) a.yo[off] = o[1];
            if (a.kxo) a.kxo[off] = o[2];
            if (a.kyo) a.kyo[off] = o[3];
        }
    }
    if (OUT != 0) {
        for (; s_i <= a.n_steps; ++s_i) {
            const int64_t off = s_i * nr + r;
            if (OUT == 1 || a.xo) __builtin_nontemporal_store(MR_NAN, a.xo + off);
            if (OUT == 1 || a.yo) __builtin_nontemporal_store(MR_NAN, a.yo + off);
            if (OUT == 1 || a.kxo) __builtin_nontemporal_store(MR_NAN, a.kxo + off);
            if (OUT == 1 || a.kyo) __builtin_nontemporal_store(MR_NAN, a.kyo + off);
        }
    }
    if (a.len_out) a.len_out[r] = (int32_t)len;
    if (a.end_out) {
        a.end_out[4 * r + 0] = s[0]; a.end_out[4 * r + 1] = s[1];
        a.end_out[4 * r + 2] = s[2]; a.end_out[4 * r + 3] = s[3];
    }
}

// ---------------------------------------------------------------------------
// Adjoint of one RK4 step.  With lam the cotangent of the new state
// y' = y + (((k1 + 2 k2) + 2 k3) + k4) dt/6 and the stage points
// Y1 = y, Y2 = y + k1 dt/2, Y3 = y + k2 dt/2, Y4 = y + k3 dt:
//   kb4 = kb1 = lam dt/6, kb3 = kb2 = 2 lam dt/6
//   z4 = J(Y4)^T kb4;  kb3 += z4 dt
//   z3 = J(Y3)^T kb3;  kb2 += z3 dt/2
//   z2 = J(Y2)^T kb2;  kb1 += z2 dt/2
//   z1 = J(Y1)^T kb1
//   lam <- lam + z1 + z2 + z3 + z4
// and the field-value cotangents of the four products go to the grids.
template <int BK, int CK>
__device__ __forceinline__ void rk4_step_adjoint(const Grid64View &B, const Grid64Ptrs &BP,
                                                 const Grid64View &C, const Grid64Ptrs &CP,
                                                 const Fields64 &F, Slots64 &sn,
                                                 const Grads64 &Gr, Accs64 &acc,
                                                 const FastConsts &K, double dt, double half,
                                                 double sixth, const double s[4],
                                                 const double ys[3][4], double lam[4])
{
    double kb[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double w = lam[q] * sixth;
        kb[0][q] = w; kb[3][q] = w;
        kb[1][q] = 2.0 * w; kb[2][q] = 2.0 * w;
    }
    double Y[4][2];                    // (x, y) of the four stage points
    Y[0][0] = s[0]; Y[0][1] = s[1];
#pragma unroll
    for (int st = 1; st < 4; ++st) {
        Y[st][0] = ys[st - 1][0];
        Y[st][1] = ys[st - 1][1];
    }
    Slots64 sq = sn;
#pragma unroll
    for (int st = 3; st >= 0; --st) {
        double Yk[2];
        if (st == 0) {
            Yk[0] = s[2]; Yk[1] = s[3];
        } else {
            Yk[0] = ys[st - 1][2];
            Yk[1] = ys[st - 1][3];
        }
        // The stage points pair up: Y2 lies with Y1, and Y0 (this step's
        // start) with the previous step's Y3.  Stages 4 and 2 (st 3, 1) start
        // loading the cells of the pair that follows (known exactly: Y2, Y0),
        // stages 3 and 1 (st 2, 0) take them up.
        if (st == 2 || st == 0)
            sq = sn;
        else
            prefetch64<BK, CK>(B, BP, C, CP, F, sq, sn, Y[st - 1][0], Y[st - 1][1]);
        double z[4] = {0.0, 0.0, 0.0, 0.0};
        rhs64_vjp<BK, CK>(B, BP, C, CP, F, sq, Gr, acc, K, Y[st][0], Y[st][1], Yk[0], Yk[1],
                          kb[st][0], kb[st][1], kb[st][2], kb[st][3], z[0], z[1], z[2], z[3]);
        if (st == 0) sn = sq;          // the cells in hand, for the previous step's Y3
        if (st > 0) {
            const double hs = st < 3 ? half : dt;
#pragma unroll
            for (int q = 0; q < 4; ++q) kb[st - 1][q] += z[q] * hs;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) lam[q] += z[q];
    }
}

// The reverse sweep.  Row m is the ray's last finite row: len-1, or len-2
// when the last row is the all-NaN one (a ray that left the domain, the
// row `end` holds).  Cotangents: g_end on row len-1 when it is finite, g_x ..
// g_ky on rows 0..m; none reaches a NaN row.
//
// A workgroup is two waves over the same 64 rays, pipelined over the
// checkpointed segments, latest first: while wave 1 runs the step adjoints
// of segment c from one half of the work buffer, wave 0 recomputes segment
// c-1 (states and stage states, from its checkpoint) into the other
// half; a barrier separates the rounds.  A segment then costs the longer of
// the two passes instead of their sum, on SIMDs a one-wave-per-CU launch
// leaves idle.  Lanes past n_rays stay in the loop, masked, for the
// barriers.
constexpr int ADJ_BLOCK = 128;

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = __builtin_elementwise_max(v, __shfl_xor(v, off));
    return v;
}

template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(ADJ_BLOCK) adjoint_kernel(AdjointArgs a)
{
    const int64_t nr = a.n_rays;
    const int wave = (int)(threadIdx.x >> 6);          // 0 recomputes, 1 sweeps back
    const int64_t rr = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const bool active = rr < nr;
    const int64_t r = active ? rr : nr - 1;
    const Grid64View B = load_view<Grid64View>(a.B);
    const Grid64View C = load_view<Grid64View>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    const double dt = a.dt, half = a.half, sixth = a.sixth;
    const int stride = a.stride;
    const int64_t ld = a.work_ld;
    const int64_t half_work = (int64_t)stride * 16 * ld;
    const bool dead = all_nan4(a.end[4 * r], a.end[4 * r + 1], a.end[4 * r + 2],
                               a.end[4 * r + 3]);
    const int64_t m = (int64_t)a.len[r] - (dead ? 2 : 1);
    // segments 0 .. cmax_r of this ray; rounds over the workgroup's largest
    const int cmax_r = (active && m >= 1) ? (int)((m - 1) / stride) : -1;
    const int cmax = wave_max(cmax_r);
    if (wave == 0) {
        Slots64 sa, sb;                                // the step's two cell slots
        for (int t = 0; t <= cmax + 1; ++t) {
            const int c = cmax - t;
            if (c >= 0 && c <= cmax_r) {
                const int64_t i0 = (int64_t)c * stride;
                const int nseg = (int)__builtin_elementwise_min((int64_t)stride, m - i0);
                double *wk = a.work + (t & 1) * half_work;
                const double *cp = a.ckpt + ((int64_t)c * 4) * nr + r;
                double s[4] = {cp[0], cp[nr], cp[2 * nr], cp[3 * nr]};
                load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sa, s[0], s[1]);
                load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sb, s[0], s[1]);
                for (int j = 0; j < nseg; ++j) {
                    double k[3][4], ys[3][4], n[4];
                    rk4_step64<BK, CK, MATH, RECOMPUTE_STORES>(B, a.bp, C, a.cp, a.f, K, sa, sb,
                                                               dt, half, sixth, s, k, ys, n);
                    work_put(wk, ld, r, j, s, ys);
#pragma unroll
                    for (int q = 0; q < 4; ++q) s[q] = n[q];
                }
            }
            __syncthreads();
        }
        return;
    }
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    if (active && m >= 0) {
        if (a.g_end && !dead) {
#pragma unroll
            for (int q = 0; q < 4; ++q) lam[q] = a.g_end[4 * r + q];
        }
        const int64_t o = m * nr + r;
        if (a.g_x) lam[0] += a.g_x[o];
        if (a.g_y) lam[1] += a.g_y[o];
        if (a.g_kx) lam[2] += a.g_kx[o];
        if (a.g_ky) lam[3] += a.g_ky[o];
    }
    Slots64 sn;                                        // cells for the next evaluation
    Accs64 acc;
    for (int t = 0; t <= cmax + 1; ++t) {
        const int c = cmax - t + 1;
        if (t >= 1 && c <= cmax_r) {
            const int64_t i0 = (int64_t)c * stride;
            const int nseg = (int)__builtin_elementwise_min((int64_t)stride, m - i0);
            const double *wk = a.work + ((t - 1) & 1) * half_work;
            // each step's state and stage states are loaded one step
            // ahead, while the step before them is swept
            double nxt[16];
            work_get(wk, ld, r, nseg - 1, nxt, [](int) { return true; });
            for (int j = nseg - 1; j >= 0; --j) {
                double ys[3][4], y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    y[q] = nxt[q];
                    ys[0][q] = nxt[4 + q];
                    ys[1][q] = nxt[8 + q];
                    ys[2][q] = nxt[12 + q];
                }
                if (j >= 1) {
                    work_get(wk, ld, r, j - 1, nxt, [](int) { return true; });
                }
                if (j == nseg - 1)     // the sweep starts at the segment's last stage point
                    prefetch64<BK, CK>(B, a.bp, C, a.cp, a.f, sn, sn, ys[2][0], ys[2][1]);
                rk4_step_adjoint<BK, CK>(B, a.bp, C, a.cp, a.f, sn, a.grads, acc, K, dt, half,
                                         sixth, y, ys, lam);
                const int64_t o = (i0 + j) * nr + r;
                if (a.g_x) lam[0] += a.g_x[o];
                if (a.g_y) lam[1] += a.g_y[o];
                if (a.g_kx) lam[2] += a.g_kx[o];
                if (a.g_ky) lam[3] += a.g_ky[o];
            }
        }
        __syncthreads();
    }
    if (!active) return;
    if (a.g_x0) a.g_x0[r] = lam[0];
    if (a.g_y0) a.g_y0[r] = lam[1];
    if (a.g_kx0) a.g_kx0[r] = lam[2];
    if (a.g_ky0) a.g_ky0[r] = lam[3];
    if (BK == BK_GRID && a.grads.depth) acc_flush(acc.h, a.grads.depth, (uint32_t)B.nx);
    if (CK == CK_GRID) {
        if (a.grads.u) acc_flush(acc.u, a.grads.u, (uint32_t)C.nx);
        if (a.grads.v) acc_flush(acc.v, a.grads.v, (uint32_t)C.nx);
    }
}

// The reverse sweep on one wave per 64 rays, for launches with many more
// waves than SIMDs.  The wave recomputes each segment, from the last, from
// its checkpoint into the work buffer and then sweeps it back: no barrier
// and no second buffer.  Two waves that pipeline recomputation and sweep
// (adjoint_kernel) keep the recomputing wave about half idle, since a sweep
// step costs about twice a forward step, and each holds a SIMD of its own;
// once the launch has waves enough to fill every SIMD several times over,
// one wave doing both takes about a quarter less SIMD time.
template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(64) adjoint_seq_kernel(AdjointArgs a)
{
    const int64_t nr = a.n_rays;
    const int64_t rr = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const bool active = rr < nr;
    const int64_t r = active ? rr : nr - 1;
    const Grid64View B = load_view<Grid64View>(a.B);
    const Grid64View C = load_view<Grid64View>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    const double dt = a.dt, half = a.half, sixth = a.sixth;
    const int stride = a.stride;
    const int64_t ld = a.work_ld;
    const bool dead = all_nan4(a.end[4 * r], a.end[4 * r + 1], a.end[4 * r + 2],
                               a.end[4 * r + 3]);
    const int64_t m = (int64_t)a.len[r] - (dead ? 2 : 1);
    const int cmax_r = (active && m >= 1) ? (int)((m - 1) / stride) : -1;
    const int cmax = wave_max(cmax_r);
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    if (active && m >= 0) {
        if (a.g_end && !dead) {
#pragma unroll
            for (int q = 0; q < 4; ++q) lam[q] = a.g_end[4 * r + q];
        }
        const int64_t o = m * nr + r;
        if (a.g_x) lam[0] += a.g_x[o];
        if (a.g_y) lam[1] += a.g_y[o];
        if (a.g_kx) lam[2] += a.g_kx[o];
        if (a.g_ky) lam[3] += a.g_ky[o];
    }
    // the recomputation's two cell slots; the sweep then uses sa (both are
    // loaded whole at each segment's start)
    Slots64 sa, sb;
    Accs64 acc;
    double *wk = a.work;                               // the first half of the buffer
    for (int c = cmax; c >= 0; --c) {
        if (c <= cmax_r) {
            const int64_t i0 = (int64_t)c * stride;
            const int nseg = (int)__builtin_elementwise_min((int64_t)stride, m - i0);
            // recompute the segment from its checkpoint
            const double *cp = a.ckpt + ((int64_t)c * 4) * nr + r;
            double s[4] = {cp[0], cp[nr], cp[2 * nr], cp[3 * nr]};
            load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sa, s[0], s[1]);
            load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sb, s[0], s[1]);
            for (int j = 0; j < nseg; ++j) {
                double k[3][4], ys[3][4], n[4];
                rk4_step64<BK, CK, MATH, RECOMPUTE_STORES>(B, a.bp, C, a.cp, a.f, K, sa, sb, dt,
                                                           half, sixth, s, k, ys, n);
                work_put(wk, ld, r, j, s, ys);
#pragma unroll
                for (int q = 0; q < 4; ++q) s[q] = n[q];
            }
            // and sweep it back, each step's data loaded one step ahead
            double nxt[16];
            work_get(wk, ld, r, nseg - 1, nxt, [](int) { return true; });
            for (int j = nseg - 1; j >= 0; --j) {
                double ys[3][4], y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    y[q] = nxt[q];
                    ys[0][q] = nxt[4 + q];
                    ys[1][q] = nxt[8 + q];
                    ys[2][q] = nxt[12 + q];
                }
                if (j >= 1) {
                    work_get(wk, ld, r, j - 1, nxt, [](int) { return true; });
                }
                if (j == nseg - 1)     // the sweep starts at the segment's last stage point
                    prefetch64<BK, CK>(B, a.bp, C, a.cp, a.f, sa, sa, ys[2][0], ys[2][1]);
                rk4_step_adjoint<BK, CK>(B, a.bp, C, a.cp, a.f, sa, a.grads, acc, K, dt, half,
                                         sixth, y, ys, lam);
                const int64_t o = (i0 + j) * nr + r;
                if (a.g_x) lam[0] += a.g_x[o];
                if (a.g_y) lam[1] += a.g_y[o];
                if (a.g_kx) lam[2] += a.g_kx[o];
                if (a.g_ky) lam[3] += a.g_ky[o];
            }
        }
    }
    if (!active) return;
    if (a.g_x0) a.g_x0[r] = lam[0];
    if (a.g_y0) a.g_y0[r] = lam[1];
    if (a.g_kx0) a.g_kx0[r] = lam[2];
    if (a.g_ky0) a.g_ky0[r] = lam[3];
    if (BK == BK_GRID && a.grads.depth) acc_flush(acc.h, a.grads.depth, (uint32_t)B.nx);
    if (CK == CK_GRID) {
        if (a.grads.u) acc_flush(acc.u, a.grads.u, (uint32_t)C.nx);
        if (a.grads.v) acc_flush(acc.v, a.grads.v, (uint32_t)C.nx);
    }
}

// The reverse sweep with the linearisations on waves of their own.  A
// step's four stage linearisations (field queries, dispersion terms, the
// matrix of rhs64_lin) do not depend on the cotangent chain, only their
// application does.  A workgroup is 2 + NL waves over the same 64 rays, in
// rounds separated by a barrier:
//   wave 0          recomputes one step of segment c-1 (as in adjoint_kernel);
//   waves 1 .. NL   linearise one step of segment c, 4/NL stages each
//                   (latest first), in reverse step order, into a
//                   double-buffered LDS ring; each loads the next round's
//                   step from the work buffer while it works on this one;
//   wave NL + 1     applies the records of the step linearised the round
//                   before (the chain of rk4_step_adjoint), adds the row
//                   cotangents and sums the grid cotangents.
// Segment c's K steps take K rounds; segment cmax is recomputed alone first
// and the chain trails by one round: (cmax + 2) K + 1 rounds in all.  A
// round costs about 4/NL stage linearisations, against eight in the
// two-wave sweep.  Meant for launches of at most one workgroup per CU, where
// the sweep is latency-bound (the LDS ring holds one workgroup per CU).  NL
// = 2 is the one launched: with NL = 4 the six waves share SIMDs and the
// 256 registers a wave then gets spill.
// Whether linearising wave W (NL = 2: 1 or 2; 0: every wave otherwise)
// uses field f of a step's data in the work buffer: f in [0, 4) is the
// state y, [4, 8), [8, 12) and [12, 16) the stage states of stages 1, 2, 3
// (wave 1 evaluates stages 2 and 1 and guesses its next cell from y's
// position; wave 2 evaluates stages 3 and 0).
template <int W>
__device__ constexpr bool lin_uses(int f)
{
    if (W == 1) return f < 2 || (f >= 4 && f < 12);
    if (W == 2) return f < 4 || f >= 12;
    return true;
}

template <int BK, int CK, int MATH, int NL>
__global__ void __launch_bounds__(64 * (NL + 2)) adjoint_lin_kernel(AdjointArgs a)
{
    constexpr int SPW = 4 / NL;                        // stages per linearising wave
    __shared__ double ring[2][4][LIN_N][64];
    const int64_t nr = a.n_rays;
    const int wave = (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int64_t rr = (int64_t)blockIdx.x * 64 + lane;
    const bool active = rr < nr;
    const int64_t r = active ? rr : nr - 1;
    const Grid64View B = load_view<Grid64View>(a.B);
    const Grid64View C = load_view<Grid64View>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    const double dt = a.dt, half = a.half, sixth = a.sixth;
    const int stride = a.stride;
    const int64_t ld = a.work_ld;
    const int64_t half_work = (int64_t)stride * 16 * ld;
    const bool dead = all_nan4(a.end[4 * r], a.end[4 * r + 1], a.end[4 * r + 2],
                               a.end[4 * r + 3]);
    const int64_t m = (int64_t)a.len[r] - (dead ? 2 : 1);
    const int cmax_r = (active && m >= 1) ? (int)((m - 1) / stride) : -1;
    const int cmax = wave_max(cmax_r);
    const int rounds = cmax >= 0 ? (cmax + 2) * stride + 1 : 0;
    const int chain = NL + 1;
    // Each role runs a loop of its own, with the same rounds and one barrier
    // per round: the state of one role is then not live across another's
    // code, and registers go to the largest role, not to all of them.
    if (wave == 0) {
        Slots64 sa, sb;                                // the step's two cell slots
        double s[4] = {0.0, 0.0, 0.0, 0.0};            // the state being advanced
        for (int rho = 0; rho < rounds; ++rho) {
            const int c = cmax - rho / stride, j = rho % stride;
            if (c >= 0 && c <= cmax_r) {
                const int64_t i0 = (int64_t)c * stride;
                if (j < m - i0) {
                    if (j == 0) {
                        const double *cp = a.ckpt + ((int64_t)c * 4) * nr + r;
                        s[0] = cp[0]; s[1] = cp[nr]; s[2] = cp[2 * nr]; s[3] = cp[3 * nr];
                        load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sa, s[0], s[1]);
                        load_guess64<BK, CK>(B, a.bp, C, a.cp, a.f, sb, s[0], s[1]);
                    }
                    double k[3][4], ys[3][4], n[4];
                    rk4_step64<BK, CK, MATH, RECOMPUTE_STORES>(B, a.bp, C, a.cp, a.f, K, sa,
                                                               sb, dt, half, sixth, s, k, ys, n);
                    work_put(a.work + (c & 1) * half_work, ld, r, j, s, ys);
#pragma unroll
                    for (int q = 0; q < 4; ++q) s[q] = n[q];
                }
            }
            __syncthreads();
        }
    } else if (wave < chain) {
        // W: the linearising wave with NL = 2 (1 or 2), as a compile-time
        // value, so that each loads and keeps only the step data its stages
        // use; 0 for the general case
        auto lin_role = [&](auto wc) {
            constexpr int W = decltype(wc)::value;
            Slots64 sn;
            double nxt[16];                                // the next step's data
            bool pre = false;
            for (int rho = 0; rho < rounds; ++rho) {
                // Everything this wave loaded in the round before (the step
                // data and the cells for this round) has landed by now; an
                // explicit wait here says so to the compiler.  Without it the
                // copy of the cell slot below waits for the loads issued
                // before it in this round as well (they are conditional, so
                // the compiler cannot count them): a memory latency a round.
                // For the same reason the next round's step data is loaded
                // after everything this round's stage points need.
                vm_wait_upto<0>();
                const int qq = rho / stride;
                const int c = cmax - qq + 1, j = stride - 1 - rho % stride;
                if (qq >= 1 && c >= 0 && c <= cmax_r && j < m - (int64_t)c * stride) {
                    const double *wb = a.work + (c & 1) * half_work;
                    double d[16];
                    if (pre) {
#pragma unroll
                        for (int f = 0; f < 16; ++f)
                            if (lin_uses<W>(f)) d[f] = nxt[f];
                    } else {                // a segment's first round
                        work_get(wb, ld, r, j, d, [](int f) { return lin_uses<W>(f); });
                        // (waited for here, so that after the join the
                        // compiler sees nothing pending on d)
                        vm_wait_upto<0>();
                    }
                    // the four stage points y, y + k1 dt/2, y + k2 dt/2, y + k3 dt
                    // (the forward's own stage states)
                    double P[4][4];
#pragma unroll
                    for (int st = 0; st < 4; ++st)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (lin_uses<W>(4 * st + q)) P[st][q] = d[4 * st + q];
                    // one step back along k1 from a point p: p - k1 dt, with
                    // k1 dt = 2 (P[1] - y) (only a guess of where to load)
                    const double bx = W == 2 ? 0.0 : 2.0 * (d[0] - P[1][0]);
                    const double by = W == 2 ? 0.0 : 2.0 * (d[1] - P[1][1]);
                    Slots64 sq = sn;
                    if (NL == 2) {
                        // the cells of this wave's next point, into sn (see below)
                        if constexpr (W == 1)
                            prefetch64<BK, CK>(B, a.bp, C, a.cp, a.f, sq, sn, P[1][0] + bx,
                                               P[1][1] + by);
                        else
                            prefetch64<BK, CK>(B, a.bp, C, a.cp, a.f, sq, sn, P[0][0], P[0][1]);
                    }
                    // start loading the next round's step (same segment)
                    pre = j >= 1;
                    if (pre) {
                        work_get(wb, ld, r, j - 1, nxt, [](int f) { return lin_uses<W>(f); });
                    }
                    if (NL == 2) {
                        // The stage points pair up: stages 2 and 1 lie together near
                        // y + k dt/2, and stage 0 here lies with stage 3 of the next
                        // round (one step earlier).  Wave 1 takes stages 2 and 1,
                        // wave 2 stages 3 and 0, so each needs one prefetch per
                        // round: wave 1 for the next round's midpoint (guessed one
                        // step back along k1), wave 2 for stage 0 (known), which
                        // it evaluates in the prefetched slot (the lanes that did
                        // not prefetch hold the same cells in both).
                        double R[LIN_N];
                        if constexpr (W == 1) {
                            rhs64_lin<BK, CK>(B, a.bp, C, a.cp, a.f, sq, K, P[2][0], P[2][1],
                                              P[2][2], P[2][3], R);
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                if (lin_field_used<BK, CK>(f)) ring[rho & 1][2][f][lane] = R[f];
                            rhs64_lin<BK, CK>(B, a.bp, C, a.cp, a.f, sq, K, P[1][0], P[1][1],
                                              P[1][2], P[1][3], R);
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                if (lin_field_used<BK, CK>(f)) ring[rho & 1][1][f][lane] = R[f];
                        } else {
                            rhs64_lin<BK, CK>(B, a.bp, C, a.cp, a.f, sq, K, P[3][0], P[3][1],
                                              P[3][2], P[3][3], R);
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                if (lin_field_used<BK, CK>(f)) ring[rho & 1][3][f][lane] = R[f];
                            rhs64_lin<BK, CK>(B, a.bp, C, a.cp, a.f, sn, K, P[0][0], P[0][1],
                                              P[0][2], P[0][3], R);
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                if (lin_field_used<BK, CK>(f)) ring[rho & 1][0][f][lane] = R[f];
                        }
                    } else {
                        // this wave's stages, latest first (the wave selects values,
                        // the indices stay compile-time)
                        const int first = 3 - (wave - 1) * SPW;
#pragma unroll
                        for (int h = 0; h < SPW; ++h) {
                            const int st = first - h;
                            double Y[4];
#pragma unroll
                            for (int q = 0; q < 4; ++q)
                                Y[q] = st == 3   ? P[3][q]
                                       : st == 2 ? P[2][q]
                                       : st == 1 ? P[1][q]
                                                 : P[0][q];
                            Slots64 sq = sn;
                            // the next point: this wave's next stage, or, after its
                            // last, its first stage one step earlier (guessed one
                            // step back along k1)
                            double gx, gy;
                            if (h + 1 < SPW) {
                                const int sx_ = st - 1;
                                gx = sx_ == 2 ? P[2][0] : sx_ == 1 ? P[1][0] : P[0][0];
                                gy = sx_ == 2 ? P[2][1] : sx_ == 1 ? P[1][1] : P[0][1];
                            } else {
                                const double fx = first == 3 ? P[3][0] : first == 2 ? P[2][0]
                                                : first == 1 ? P[1][0] : P[0][0];
                                const double fy = first == 3 ? P[3][1] : first == 2 ? P[2][1]
                                                : first == 1 ? P[1][1] : P[0][1];
                                gx = fx + bx;
                                gy = fy + by;
                            }
                            prefetch64<BK, CK>(B, a.bp, C, a.cp, a.f, sq, sn, gx, gy);
                            double R[LIN_N];
                            rhs64_lin<BK, CK>(B, a.bp, C, a.cp, a.f, sq, K, Y[0], Y[1], Y[2], Y[3], R);
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                if (lin_field_used<BK, CK>(f)) ring[rho & 1][st][f][lane] = R[f];
                        }
                    }
                } else {
                    pre = false;
                }
                __syncthreads();
            }
        };
        if (NL == 2 && wave == 1) lin_role(std::integral_constant<int, 1>{});
        else if (NL == 2) lin_role(std::integral_constant<int, 2>{});
        else lin_role(std::integral_constant<int, 0>{});
    } else {
        double lam[4] = {0.0, 0.0, 0.0, 0.0};          // the cotangent
        Accs64 acc;
        if (active && m >= 0) {
            if (a.g_end && !dead) {
#pragma unroll
                for (int q = 0; q < 4; ++q) lam[q] = a.g_end[4 * r + q];
            }
            const int64_t o = m * nr + r;
            if (a.g_x) lam[0] += a.g_x[o];
            if (a.g_y) lam[1] += a.g_y[o];
            if (a.g_kx) lam[2] += a.g_kx[o];
            if (a.g_ky) lam[3] += a.g_ky[o];
        }
        for (int rho = 0; rho < rounds; ++rho) {
            if (rho >= 1) {
                const int q1 = (rho - 1) / stride;
                const int c = cmax - q1 + 1, j = stride - 1 - (rho - 1) % stride;
                if (q1 >= 1 && c >= 0 && c <= cmax_r && j < m - (int64_t)c * stride) {
                    double kb[4][4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double w = lam[q] * sixth;
                        kb[0][q] = w; kb[3][q] = w;
                        kb[1][q] = 2.0 * w; kb[2][q] = 2.0 * w;
                    }
                    // each stage's record is read while the stage before it
                    // in the chain is applied, so the LDS latency is hidden
                    double Rn[LIN_N];
#pragma unroll
                    for (int f = 0; f < LIN_N; ++f)
                        Rn[f] = lin_field_used<BK, CK>(f) ? ring[(rho - 1) & 1][3][f][lane] : 0.0;
#pragma unroll
                    for (int st = 3; st >= 0; --st) {
                        double R[LIN_N];
#pragma unroll
                        for (int f = 0; f < LIN_N; ++f) R[f] = Rn[f];
                        if (st > 0) {
#pragma unroll
                            for (int f = 0; f < LIN_N; ++f)
                                Rn[f] = lin_field_used<BK, CK>(f)
                                            ? ring[(rho - 1) & 1][st - 1][f][lane] : 0.0;
                        }
                        double z[4];
                        lin_apply<BK, CK>(R, B, C, a.grads, acc, kb[st], z);
                        if (st > 0) {
                            const double hs = st < 3 ? half : dt;
#pragma unroll
                            for (int q = 0; q < 4; ++q) kb[st - 1][q] += z[q] * hs;
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) lam[q] += z[q];
                    }
                    const int64_t o = ((int64_t)c * stride + j) * nr + r;
                    if (a.g_x) lam[0] += a.g_x[o];
                    if (a.g_y) lam[1] += a.g_y[o];
                    if (a.g_kx) lam[2] += a.g_kx[o];
                    if (a.g_ky) lam[3] += a.g_ky[o];
                }
            }
            __syncthreads();
        }
        if (!active) return;
        if (a.g_x0) a.g_x0[r] = lam[0];
        if (a.g_y0) a.g_y0[r] = lam[1];
        if (a.g_kx0) a.g_kx0[r] = lam[2];
        if (a.g_ky0) a.g_ky0[r] = lam[3];
        if (BK == BK_GRID && a.grads.depth) acc_flush(acc.h, a.grads.depth, (uint32_t)B.nx);
        if (CK == CK_GRID) {
            if (a.grads.u) acc_flush(acc.u, a.grads.u, (uint32_t)C.nx);
            if (a.grads.v) acc_flush(acc.v, a.grads.v, (uint32_t)C.nx);
        }
    }
}

// WaveRayPath::system in fp64 mode at a batch of states
template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(256) rhs64_kernel(Rhs64Args a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Grid64View B = load_view<Grid64View>(a.B);
    const Grid64View C = load_view<Grid64View>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    double dx, dy, dkx, dky;
    Slots64 sl;                        // empty: the query loads the cell
    rhs64<BK, CK, MATH>(B, a.bp, C, a.cp, a.f, sl, K, a.s[4 * i], a.s[4 * i + 1], a.s[4 * i + 2],
                        a.s[4 * i + 3], dx, dy, dkx, dky);
    a.ds[4 * i] = dx; a.ds[4 * i + 1] = dy; a.ds[4 * i + 2] = dkx; a.ds[4 * i + 3] = dky;
}

// per-cell (m01, m10) of interpolator.rs:59-72 in f64, on the corners
// A = (x_i, y_j), B = (x_i, y_j+1), D = (x_i+1, y_j); NaN marks det == 0
__global__ void build_m64_kernel(M64BuildArgs a)
{
    const int64_t ncx = a.nx - 1;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncx * (a.ny - 1)) return;
    const int64_t i = c % ncx, j = c / ncx;
    const double Ax = a.ax[i], Ay = a.ay[j];
    const double btx = Ax - Ax, bty = a.ay[j + 1] - Ay;
    const double dtx = a.ax[i + 1] - Ax, dty = Ay - Ay;
    const double det = (btx * dty) - (dtx * bty);
    a.m[c] = det == 0.0 ? make_double2(MR_NAN, MR_NAN)
                        : make_double2(-(dtx / det), -(bty / det));
}

__global__ void fill_kernel(double *p, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers

hipError_t launch_trace64(const Trace64Args &a, int bk, int ck, int math, int block,
                          hipStream_t stream)
{
    if (a.n_rays <= 0) return hipSuccess;
    if (a.ray_blocks != (int)((a.n_rays + block - 1) / block)) return hipErrorInvalidValue;
    if (a.ckpt && a.stride <= 0) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(a.ray_blocks + (a.t_out ? 1 : 0)));
    const int nout = (a.xo != nullptr) + (a.yo != nullptr) + (a.kxo != nullptr) +
                     (a.kyo != nullptr);
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using KK = decltype(k);
        if (nout == 0)
            hipLaunchKernelGGL((trace64_kernel<KK::bk, KK::ck, KK::math, 0>), grid, dim3(block),
                               0, stream, a);
        else if (nout == 4)
            hipLaunchKernelGGL((trace64_kernel<KK::bk, KK::ck, KK::math, 1>), grid, dim3(block),
                               0, stream, a);
        else
            hipLaunchKernelGGL((trace64_kernel<KK::bk, KK::ck, KK::math, 2>), grid, dim3(block),
                               0, stream, a);
    });
    return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointArgs &a, int bk, int ck, int math, int waves,
                          hipStream_t stream)
{
    if (a.n_rays <= 0) return hipSuccess;
    if (a.stride <= 0 || a.work_ld < a.n_rays) return hipErrorInvalidValue;
    if (waves != 1 && waves != 2 && waves != 4) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((a.n_rays + 63) / 64));    // 64 rays per workgroup
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using KK = decltype(k);
        if (waves == 4)
            hipLaunchKernelGGL((adjoint_lin_kernel<KK::bk, KK::ck, KK::math, 2>), grid,
                               dim3(4 * 64), 0, stream, a);
        else if (waves == 2)
            hipLaunchKernelGGL((adjoint_kernel<KK::bk, KK::ck, KK::math>), grid,
                               dim3(ADJ_BLOCK), 0, stream, a);
        else
            hipLaunchKernelGGL((adjoint_seq_kernel<KK::bk, KK::ck, KK::math>), grid, dim3(64),
                               0, stream, a);
    });
    return hipGetLastError();
}

hipError_t launch_rhs64(const Rhs64Args &a, int bk, int ck, int math, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    const int block = 256;
    const dim3 grid((unsigned)((a.n + block - 1) / block));
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using KK = decltype(k);
        hipLaunchKernelGGL((rhs64_kernel<KK::bk, KK::ck, KK::math>), grid, dim3(block), 0, stream,
                           a);
    });
    return hipGetLastError();
}

hipError_t launch_build_m64(const M64BuildArgs &a, hipStream_t stream)
{
    const int64_t ncell = (a.nx - 1) * (a.ny - 1);
    if (ncell <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(build_m64_kernel, dim3((unsigned)((ncell + block - 1) / block)),
                       dim3(block), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

}  // namespace mr
