// mr_kernels.hip -- gfx950 kernels of the many-ray integrator.
//
//   trace_kernel   ManyRays::trace_many (src/ray.rs:98-127): one ray per
//                  lane, the whole RK4 step loop in registers, rows written
//                  step-major so each wave stores 64 contiguous doubles per
//                  component and step.
//   rhs_kernel     WaveRayPath::system (src/wave_ray_path.rs:220-234) at a
//                  batch of states (parity tests of the right-hand side).
//   build_*        per-cell records of a gridded field (see mr_device.h).
//   fill_time      the time column, t += dt accumulated as ode_solvers does
//                  (one lane: the accumulation is a sequential rounding chain),
//                  for a trace with no rays; otherwise the trace kernel's last
//                  workgroup writes it beside the rays.
#include <cmath>

#include "mr_kernels.h"
#include "mr_kernel_common.h"

namespace mr {

// ---------------------------------------------------------------------------
// ode_solvers 0.4.0 Rk4::step/integrate for one lane:
//   k1 = f(y), k2 = f(y + k1 (dt/2)), k3 = f(y + k2 (dt/2)), k4 = f(y + k3 dt)
//   y <- y + (((k1 + k2 2) + k3 2) + k4) (dt/6)
// with the stop test of WaveRayPath::solout (wave_ray_path.rs:236-246).
//
// Stages 1 and 3 start the loads of the cell records that the following
// stages will most likely use (their points guessed from the latest
// derivative) when those are not the cells already held, so a record is in
// registers one evaluation after it was asked for and a ray crossing into a
// new cell does not wait for memory; a wrong guess only costs a blocking
// reload in the query (rhs, the full-query path).

// One RK4 step of one lane from (x, y, kx, ky): the new state in o, and
// stop_before when (MR_SOLOUT_NEXT_DERIV) the first stage derivative is
// all-NaN, i.e. the ray already ended on the row it starts from; on_stop()
// is then called after the first stage, while (x, y, kx, ky) is still the
// state the step started from.  Every stage state after an all-NaN first
// derivative is NaN, and so is every later derivative and the whole of o.
// bn, cn:
// the records prefetched for the next evaluation; (px, py): the latest
// dx/dt, dy/dt, for the guess.
template <int BK, int CK, int MATH, class OnStop>
__device__ __forceinline__ void rk4_step(const BathyView &B, const BathyPtrs &BP, BathySlot &bn,
                                         const CurrentView &C, const CurrentPtrs &CP,
                                         CurrentSlot &cn, const FastConsts &K, double &px,
                                         double &py, double dt, double half, double sixth,
                                         int solout, double x, double y, double kx, double ky,
                                         double o[4], bool &stop_before, OnStop &&on_stop)
{
    constexpr bool PF_B = BK == BK_GRID, PF_C = CK == CK_GRID;
    double sx = x, sy = y, skx = kx, sky = ky;   // stage state
    double ax = 0.0, ay = 0.0, akx = 0.0, aky = 0.0;
    stop_before = false;
    // The four stage points come in pairs at nearly one place: y + k1 dt/2
    // and y + k2 dt/2, then y + k3 dt and the next step's start.  So the
    // records are prefetched at stages 1 and 3 for the pair that follows
    // (guessed at x + k dt/2 and x + k dt) and taken up at stages 2 and 4;
    // stages 1 and 3 use the records in hand.
    BathySlot bq = bn;
    CurrentSlot cq = cn;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const double hs = st < 2 ? half : dt;
        if (st == 1 || st == 3) {
            bq = bn;
            cq = cn;
        } else {
            const double gx = x + px * hs, gy = y + py * hs;
            if (PF_B && !bathy_in_cell(bq, (float)gx, (float)gy))
                bathy_prefetch(B, BP, bn, gx, gy);
            if (PF_C && !current_in_cell(cq, gx, gy))
                current_prefetch(C, CP, cn, gx, gy);
        }
        double dx, dy, dkx, dky;
        rhs<BK, CK, MATH>(B, BP, bq, C, CP, cq, K, sx, sy, skx, sky, dx, dy, dkx, dky);
        px = dx; py = dy;
        if (st == 0) {
            // With MR_SOLOUT_NEXT_DERIV, solout saw f(y) of the row
            // just pushed: an all-NaN k1 here means the ray already
            // ended on that row.  (Marked unlikely: the test is laid out
            // after the loop, and the default reading falls through.)
            if (__builtin_expect(solout == 1, 0)) {
                stop_before = all_nan4(dx, dy, dkx, dky);
                if (stop_before) on_stop();
            }
            ax = dx; ay = dy; akx = dkx; aky = dky;
        } else if (st < 3) {
            // k * 2 is exact, so fma(k, 2, a) rounds as a + k * 2
            ax = __builtin_fma(dx, 2.0, ax); ay = __builtin_fma(dy, 2.0, ay);
            akx = __builtin_fma(dkx, 2.0, akx); aky = __builtin_fma(dky, 2.0, aky);
        } else {
            ax = ax + dx; ay = ay + dy; akx = akx + dkx; aky = aky + dky;
        }
        sx = x + dx * hs; sy = y + dy * hs;
        skx = kx + dkx * hs; sky = ky + dky * hs;
    }
    o[0] = x + ax * sixth; o[1] = y + ay * sixth;
    o[2] = kx + akx * sixth; o[3] = ky + aky * sixth;
}

// rk4_step for gridded depth and current on shared cells: one combined
// record (SCell) per cell serves both fields, from two slots with fixed
// roles.  In the stage loop below (st = 0..3), sa serves st 0 and st 3 (the
// step's start y and y + k3 dt, nearly the next step's start), sb serves
// st 1 and st 2 (y + k1 dt/2 and y + k2 dt/2).  Each slot is refilled while
// the other one serves, two evaluations before its next use: sa at st 1 for
// the point guessed at x + k1 dt, sb at st 3 for the next step's pair,
// guessed at x + k3 (3 dt / 2).  A lane loads only when the guess lies
// outside the fast region of the record its slot holds, and otherwise keeps
// that record, the right one; so no slot is ever copied.  A guess is only a
// guess: rhs_shared tests the region of the record it uses, and lanes
// outside it take the full queries.  g1, g3: dt and 3 dt / 2 in f32.
//
// Before each refill the step waits for the vector memory operations issued
// up to then but the last NST (the caller's row stores of the step before,
// issued after every load): the serving slot's loads, two evaluations old,
// have landed by then.  Without that wait the compiler waits at the serving
// slot's first use, after the refill, and as the refill is conditional it
// cannot count it: it waits for the refill's loads as well.
//
// DY: dt and dt/2 are powers of two, so k * hs is exact (for every k whose
// product is not subnormal) and the stage state y + k * hs of Rk4::step is
// one rounding: fma(k, hs, y) gives it bitwise in one instruction.
template <int MATH, bool DY, int NST, class OnStop>
__device__ __forceinline__ void rk4_step_shared(const BathyView &B, const BathyPtrs &BP,
                                                const CurrentView &C, const CurrentPtrs &CP,
                                                const SCell *cells, SharedSlot &sa,
                                                SharedSlot &sb, const FastConsts &K, double &px,
                                                double &py, double dt, double half,
                                                double sixth, float g1, float g3, int solout,
                                                double x, double y, double kx, double ky,
                                                double o[4], bool &stop_before, OnStop &&on_stop)
{
    double sx = x, sy = y, skx = kx, sky = ky;   // stage state
    double ax = 0.0, ay = 0.0, akx = 0.0, aky = 0.0;
    stop_before = false;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const double hs = st < 2 ? half : dt;
        if (st == 1 || st == 3) {
            if (st == 1) vm_wait<NST>();
            else vm_wait<0>();
            // the guessed point in f32: it only chooses which record to load
            SharedSlot &refill = st == 1 ? sa : sb;
            const float g = st == 1 ? g1 : g3;
            const float gx = __builtin_fmaf((float)px, g, (float)x);
            const float gy = __builtin_fmaf((float)py, g, (float)y);
            // The record's offset is formed before the test, in every step:
            // some lane of the wave refills in most steps anyway, and these
            // vector operations then stand between the test's compare and
            // the scalar EXEC update that reads it, which otherwise waits for
            // the compare's result (the empty asm keeps them out of the
            // conditional block).
            uint32_t cell = tiled_cell(guess_cell(gx, B.ax0, B.pinv_sx, B.pmax_x),
                                       guess_cell(gy, B.ay0, B.pinv_sy, B.pmax_y), B.ntx);
            __asm__ volatile("" : "+v"(cell));
            if (!shared_in_region(refill, gx, gy)) shared_load(cells, refill, cell);
        }
        double dx, dy, dkx, dky;
        rhs_shared<MATH>(B, BP, C, CP, (st == 0 || st == 3) ? sa : sb, K, sx, sy, skx, sky, dx,
                         dy, dkx, dky);
        px = dx; py = dy;
        if (st == 0) {
            // the MR_SOLOUT_NEXT_DERIV test, as in rk4_step
            if (__builtin_expect(solout == 1, 0)) {
                stop_before = all_nan4(dx, dy, dkx, dky);
                if (stop_before) on_stop();
            }
            ax = dx; ay = dy; akx = dkx; aky = dky;
        } else if (st < 3) {
            ax = __builtin_fma(dx, 2.0, ax); ay = __builtin_fma(dy, 2.0, ay);
            akx = __builtin_fma(dkx, 2.0, akx); aky = __builtin_fma(dky, 2.0, aky);
        } else {
            ax = ax + dx; ay = ay + dy; akx = akx + dkx; aky = aky + dky;
        }
        if (DY) {
            sx = __builtin_fma(dx, hs, x); sy = __builtin_fma(dy, hs, y);
            skx = __builtin_fma(dkx, hs, kx); sky = __builtin_fma(dky, hs, ky);
        } else {
            sx = x + dx * hs; sy = y + dy * hs;
            skx = kx + dkx * hs; sky = ky + dky * hs;
        }
    }
    o[0] = x + ax * sixth; o[1] = y + ay * sixth;
    o[2] = kx + akx * sixth; o[3] = ky + aky * sixth;
}

// the first records of a lane: both slots hold the cell its start most
// likely lies in
__device__ __forceinline__ void shared_prefetch(const BathyView &B, const SCell *cells,
                                                SharedSlot &sa, SharedSlot &sb, double x,
                                                double y)
{
    const int i1 = guess_cell((float)x, B.ax0, B.pinv_sx, B.pmax_x);
    const int j1 = guess_cell((float)y, B.ay0, B.pinv_sy, B.pmax_y);
    shared_load(cells, sa, tiled_cell(i1, j1, B.ntx));
    shared_load(cells, sb, tiled_cell(i1, j1, B.ntx));
}

// A buffer descriptor over the rows from p on (raw: no stride, no swizzle;
// the record count is 2^32 - 1 - k, so no offset below 2^31 is cut off).
// k tells the four descriptors of a trace apart: with equal upper words the
// compiler keeps one copy of those words and re-copies it next to each
// descriptor at every step.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(double *p, int k)
{
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, -1 - k, MR_BUFFER_DWORD3);
}

// one double at byte offset voff + soff of the descriptor's memory, with the
// nontemporal hint
__device__ __forceinline__ void store_row_nt(__amdgpu_buffer_rsrc_t d, uint32_t voff,
                                             uint32_t soff, double v)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), d, (int)voff, (int)soff,
                                          MR_CPOL_NT);
}

// Lanes past n_rays repeat the last ray (same inputs, same arithmetic, so
// their stores write the same bytes to the same addresses): every lane
// stores every row, with no branch around the stores.  OUT: 0 no
// trajectories, 1 all four components, 2 any subset.
// SC: 0 separate records per field; 1 the shared-cell records (SCell) of a
// depth and current on the same cells (rk4_step_shared); 2 the same with dt
// a power of two (its DY).
template <int BK, int CK, int MATH, int OUT, int SC = 0>
__global__ void __launch_bounds__(256) trace_kernel(TraceArgs a)
{
    constexpr bool PF_B = BK == BK_GRID, PF_C = CK == CK_GRID;
    constexpr bool DYAD = SC == 2;
    constexpr int NST = OUT == 1 ? 4 : 0;   // row stores issued per step
    if ((int)blockIdx.x == a.ray_blocks) {
        // the time column: t_s = t_{s-1} + dt (ode_solvers accumulates), a
        // sequential rounding chain, run by one lane beside the rays
        if (threadIdx.x == 0) time_column(a.t0, a.dt, a.n_steps, a.t_out);
        return;
    }
    const int64_t nr = a.n_rays;
    const int64_t r = __builtin_elementwise_min((int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                                                nr - 1);
    double x = a.x0[r], y = a.y0[r], kx = a.kx0[r], ky = a.ky0[r];
    if (OUT == 1) {
        a.xo[r] = x; a.yo[r] = y; a.kxo[r] = kx; a.kyo[r] = ky;
    } else if (OUT == 2) {
        if (a.xo) a.xo[r] = x;
        if (a.yo) a.yo[r] = y;
        if (a.kxo) a.kxo[r] = kx;
        if (a.kyo) a.kyo[r] = ky;
    }
    const BathyView B = load_view<BathyView>(a.B);
    const CurrentView C = load_view<CurrentView>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    const double half = a.half, dt = a.dt, sixth = a.sixth;
    // the views and constants arrive before the loop: otherwise they count
    // as pending at the loop head, and each step waits for every load and
    // store in flight where it first reads them
    settle(B);
    settle(C);
    settle(K);
    BathySlot bn;                      // records prefetched for the next evaluation
    CurrentSlot cn;
    SharedSlot sa, sb;                 // the two record slots, for shared cells
    const float g1 = (float)dt, g3 = (float)(dt + half);
    if (SC) {
        shared_prefetch(B, a.scells, sa, sb, x, y);
        settle(sa);                    // likewise for the first records
        settle(sb);
    } else {
        if (PF_B) bathy_prefetch(B, a.bp, bn, x, y);
        if (PF_C) current_prefetch(C, a.cp, cn, x, y);
        if (PF_B) settle(bn);
        if (PF_C) settle(cn);
    }
    double px = 0.0, py = 0.0;         // latest dx/dt, dy/dt, for the guess
    // A lane's ray is live while its state is not all-NaN.  The step that
    // ends a ray leaves its state all-NaN (an all-NaN new state, or an
    // all-NaN first stage derivative, which makes every component of the new
    // state NaN), so liveness is read off the state at each step, the row
    // stored is the state itself, and no per-lane flag is carried from step
    // to step (a carried flag costs mask merges at every join).  len counts
    // the rows; a ray that the MR_SOLOUT_NEXT_DERIV test stops before a step
    // ends on the state that step started from, written to end_out at once,
    // and is marked by a negative len.
    int32_t len = 1;
    // The integration loop asks solout only after a step, so every ray takes
    // its first step, and the MR_SOLOUT_NEXT_DERIV test (f at the state a
    // step starts from) applies from the second step on: so carries the
    // solout variant into each step, none for the first.
    int so = 0;
    // A ray that starts all-NaN takes that first step in the reference; its
    // row is NaN (every derivative at an all-NaN state is NaN) and counts,
    // and the ray ends there.  Such a lane takes no step here and starts
    // with the row count that step leaves.
    if (a.n_steps > 0 && all_nan4(x, y, kx, ky)) len = 2;
    // min4 of the state, NaN exactly when the state is all-NaN: formed after
    // each step from the step's arithmetic, where the compiler knows no
    // signalling NaN can occur (at the loop head it would quiet all four
    // components first)
    double m = min4(x, y, kx, ky);
    // OUT == 1: the rows go out by buffer stores from one descriptor per
    // component, based at the first row of a block of rows; the row's
    // offset in the block is one scalar register shared by the four stores
    // and the lane's offset one constant vector register.  The launcher
    // keeps a row under 2^30 bytes, and a block's offsets stay below 2^31.
    const uint32_t row_bytes = OUT == 1 ? (uint32_t)nr * 8u : 1u;
    const uint32_t lane_off = (uint32_t)r * 8u;
    const int64_t block_rows =
        OUT == 1 ? __builtin_elementwise_max((int64_t)((0x7fffffffu - row_bytes) / row_bytes),
                                             (int64_t)1)
                 : (int64_t)1 << 30;
    int64_t s = 1;
    while (s <= a.n_steps) {
        const int64_t s_end = __builtin_elementwise_min(a.n_steps, s + block_rows - 1);
        const int64_t row0 = OUT == 1 ? s * nr : 0;
        const __amdgpu_buffer_rsrc_t bx = row_rsrc(a.xo + row0, 0), by = row_rsrc(a.yo + row0, 1),
                                     bkx = row_rsrc(a.kxo + row0, 2),
                                     bky = row_rsrc(a.kyo + row0, 3);
        uint32_t soff = 0;
        // the block's steps counted down in 32 bits: the scalar unit tests a
        // 32-bit count in one instruction, a 64-bit signed bound only through
        // a vector compare.  The loops are left by goto when every lane's ray
        // has ended: a flag tested by the block loop would keep, in the
        // compiler's view, a path from that exit back into the step loop.
        for (uint32_t left = (uint32_t)(s_end - s + 1); left != 0u;
             --left, ++s, soff += row_bytes) {
            const bool live = !__builtin_isnan(m);
            if (__builtin_amdgcn_ballot_w64(live) == 0) goto done;
            if (live) {
                double o[4];
                bool stop_before;
                auto on_stop = [&] {
                    if (a.end_out) {
                        a.end_out[4 * r + 0] = x; a.end_out[4 * r + 1] = y;
                        a.end_out[4 * r + 2] = kx; a.end_out[4 * r + 3] = ky;
                    }
                    len = -(len + 1);
                };
                if constexpr (SC != 0)
                    rk4_step_shared<MATH, DYAD, NST>(B, a.bp, C, a.cp, a.scells, sa, sb, K, px,
                                                     py, dt, half, sixth, g1, g3, so, x, y, kx,
                                                     ky, o, stop_before, on_stop);
                else
                    rk4_step<BK, CK, MATH>(B, a.bp, bn, C, a.cp, cn, K, px, py, dt, half,
                                           sixth, so, x, y, kx, ky, o, stop_before, on_stop);
                x = o[0]; y = o[1]; kx = o[2]; ky = o[3];
                m = min4(x, y, kx, ky);
                ++len;
            }
            if (OUT == 1) {
                // streaming stores: the rows are not read again by the
                // kernel, and keeping them out of L2 leaves it to the cell
                // records
                store_row_nt(bx, lane_off, soff, x);
                store_row_nt(by, lane_off, soff, y);
                store_row_nt(bkx, lane_off, soff, kx);
                store_row_nt(bky, lane_off, soff, ky);
            } else if (OUT == 2) {
                const int64_t o = s * nr + r;
                if (a.xo) a.xo[o] = x;
                if (a.yo) a.yo[o] = y;
                if (a.kxo) a.kxo[o] = kx;
                if (a.kyo) a.kyo[o] = ky;
            }
            so = a.solout;
        }
    }
done:
    // rows past the wave's last live step
    if (OUT != 0) {
        for (; s <= a.n_steps; ++s) {
            const int64_t o = s * nr + r;
            if (OUT == 1 || a.xo) __builtin_nontemporal_store(MR_NAN, a.xo + o);
            if (OUT == 1 || a.yo) __builtin_nontemporal_store(MR_NAN, a.yo + o);
            if (OUT == 1 || a.kxo) __builtin_nontemporal_store(MR_NAN, a.kxo + o);
            if (OUT == 1 || a.kyo) __builtin_nontemporal_store(MR_NAN, a.kyo + o);
        }
    }
    if (a.len_out) a.len_out[r] = len < 0 ? -len : len;
    if (a.end_out && len > 0) {
        a.end_out[4 * r + 0] = x; a.end_out[4 * r + 1] = y;
        a.end_out[4 * r + 2] = kx; a.end_out[4 * r + 3] = ky;
    }
}

// ---------------------------------------------------------------------------
// Batched persistent trace: many scenes (each its own fields, step and step
// count) in one launch of a fixed number of workgroups.  Every lane owns one
// ray at a time; when its ray ends (edge exit, end of its steps), the lane
// takes the next ray of the queue, of whatever scene, so lanes stay busy
// while rays of uneven lifetime come and go (ray compaction at lane level).
// The lanes of a wave that need a ray take consecutive queue entries with one
// atomic add per wave.  Each lane holds its scene's field views, pointers and
// step in registers, and its record slots are emptied when it changes rays.
// Outputs: each ray's row count and last row, as trace_kernel writes them.
template <int BK, int CK, int MATH, bool SC = false>
__global__ void __launch_bounds__(256) batch_kernel(BatchArgs a)
{
    constexpr bool PF_B = BK == BK_GRID, PF_C = CK == CK_GRID;
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    settle(K);                         // as in trace_kernel
    const int lane = (int)(threadIdx.x & 63);
    BathyView B{};
    CurrentView C{};
    BathyPtrs bp{};
    CurrentPtrs cp{};
    BathySlot bn;
    CurrentSlot cn;
    SharedSlot sa, sb;
    const SCell *scells = nullptr;
    double x = 0.0, y = 0.0, kx = 0.0, ky = 0.0, px = 0.0, py = 0.0;
    double dt = 0.0, half = 0.0, sixth = 0.0;
    int64_t local = 0, n_steps = 0, step = 0, len = 1;
    int32_t *len_out = nullptr;
    double *end_out = nullptr;
    bool have = false, done = false;
    for (;;) {
        const bool need = !have && !done;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.next, (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            if (need) {
                const int64_t id =
                    (int64_t)base + __popcll(mask & ((uint64_t(1) << lane) - uint64_t(1)));
                if (id < a.n_rays) {
                    const SceneDesc *S = a.scenes + a.ray_scene[id];
                    B = load_view<BathyView>(S->B);
                    C = load_view<CurrentView>(S->C);
                    settle(B);
                    settle(C);
                    bp = S->bp;
                    cp = S->cp;
                    dt = S->dt; half = S->half; sixth = S->sixth;
                    n_steps = S->n_steps;
                    local = id - S->ray0;
                    len_out = S->len_out;
                    end_out = S->end_out;
                    x = S->x0[local]; y = S->y0[local]; kx = S->kx0[local]; ky = S->ky0[local];
                    step = 0; len = 1; px = 0.0; py = 0.0;
                    if (SC) {
                        scells = S->scells;
                        shared_prefetch(B, scells, sa, sb, x, y);
                    } else {
                        bn = BathySlot{};
                        cn = CurrentSlot{};
                        if (PF_B) bathy_prefetch(B, bp, bn, x, y);
                        if (PF_C) current_prefetch(C, cp, cn, x, y);
                    }
                    have = true;
                } else {
                    done = true;
                }
            }
        }
        if (!__any(have)) break;
        if (have) {
            bool finished = step >= n_steps;
            if (!finished) {
                double o[4];
                bool stop_before;
                if constexpr (SC)
                    rk4_step_shared<MATH, false, 0>(B, bp, C, cp, scells, sa, sb, K, px, py, dt,
                                                 half, sixth, (float)dt, (float)(dt + half),
                                                 a.solout, x, y, kx, ky, o, stop_before, [] {});
                else
                    rk4_step<BK, CK, MATH>(B, bp, bn, C, cp, cn, K, px, py, dt, half, sixth,
                                           a.solout, x, y, kx, ky, o, stop_before, [] {});
                ++step;
                // the MR_SOLOUT_NEXT_DERIV test applies from a ray's second
                // step on (solout is asked only after a step); at the first,
                // an all-NaN k1 makes the new state all-NaN, ending the ray
                // one row later
                if (stop_before && step > 1) {
                    finished = true;
                } else {
                    x = o[0]; y = o[1]; kx = o[2]; ky = o[3];
                    len = step + 1;
                    finished = all_nan4(x, y, kx, ky) || step >= n_steps;
                }
            }
            if (finished) {
                len_out[local] = (int32_t)len;
                end_out[4 * local + 0] = x; end_out[4 * local + 1] = y;
                end_out[4 * local + 2] = kx; end_out[4 * local + 3] = ky;
                have = false;
            }
        }
    }
}

__global__ void clear_counter_kernel(unsigned long long *p)
{
    if (threadIdx.x == 0) *p = 0ull;
}

template <int BK, int CK, int MATH>
__global__ void __launch_bounds__(256) rhs_kernel(RhsArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double dx, dy, dkx, dky;
    const BathyView B = load_view<BathyView>(a.B);
    const CurrentView C = load_view<CurrentView>(a.C);
    const FastConsts K = load_view<FastConsts>(mr_fast_consts);
    BathySlot bs;
    CurrentSlot cs;
    rhs<BK, CK, MATH>(B, a.bp, bs, C, a.cp, cs, K, a.s[4 * i], a.s[4 * i + 1], a.s[4 * i + 2],
                      a.s[4 * i + 3],
                      dx, dy, dkx, dky);
    a.ds[4 * i] = dx; a.ds[4 * i + 1] = dy; a.ds[4 * i + 2] = dkx; a.ds[4 * i + 3] = dky;
}

__global__ void fill_time_kernel(double t0, double dt, int64_t n, double *t)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) time_column(t0, dt, n, t);
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
    a.mtab[t] = cell_basis(a.xp, a.yp, i, j);
}

// Combined records of a depth and a current on shared cells (SCell): the
// depth's coefficients and slopes as build_bathy_kernel makes them, the
// current's as build_current_kernel does, the change of basis of the shared
// f32 axes, and the cell's fast region from the per-axis bounds; empty for a
// cell where either field errs (a corner past its data, det == 0).
__global__ void build_shared_kernel(SharedBuildArgs a)
{
    const int64_t ncx = a.nx - 1;
    const int64_t ncell = ncx * (a.ny - 1);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int64_t i = c % ncx, j = c / ncx;
    SCell r;
    const float2 m = cell_basis(a.xp, a.yp, i, j);   // m01 = NaN marks det == 0
    const float2 xq = a.xp[i], yq = a.yp[j];
    r.x0 = xq.x; r.y0 = yq.x; r.m01 = m.x; r.m10 = m.y;
    const bool err = c >= a.oob_from || __builtin_isnan(m.x);
    if (c >= a.oob_from) {
        r.a00 = r.a10 = r.a01 = r.a11 = MR_NANF;
        r.hx = r.hy = MR_NAN;
        r.ua00 = r.ua10 = r.ua01 = r.ua11 = MR_NANF;
        r.va00 = r.va10 = r.va01 = r.va11 = MR_NANF;
        r.ux = r.uy = r.vx = r.vy = MR_NAN;
    } else {
        const int64_t iA = a.nx * j + i, iB = iA + a.nx, iC = iA + a.nx + 1, iD = iA + 1;
        float vA = (float)a.depth[iA], vB = (float)a.depth[iB];
        float vC = (float)a.depth[iC], vD = (float)a.depth[iD];
        r.a00 = vA;                                           // interpolator.rs:78-81
        r.a10 = vB - vA;
        r.a01 = vD - vA;
        r.a11 = ((vC - vA) - r.a10) - r.a01;
        // cartesian_netcdf3.rs:119-134: f32 slopes, kept widened to f64
        r.hx = (double)(float)((a.depth[iD] - a.depth[iA]) / a.bxspace);
        r.hy = (double)(float)((a.depth[iB] - a.depth[iA]) / a.byspace);
        vA = (float)a.u[iA]; vB = (float)a.u[iB]; vC = (float)a.u[iC]; vD = (float)a.u[iD];
        r.ua00 = vA; r.ua10 = vB - vA; r.ua01 = vD - vA; r.ua11 = ((vC - vA) - r.ua10) - r.ua01;
        vA = (float)a.v[iA]; vB = (float)a.v[iB]; vC = (float)a.v[iC]; vD = (float)a.v[iD];
        r.va00 = vA; r.va10 = vB - vA; r.va01 = vD - vA; r.va11 = ((vC - vA) - r.va10) - r.va01;
        r.ux = (a.u[iD] - a.u[iA]) / a.cxspace;               // cartesian_current.rs:515-536
        r.uy = (a.u[iB] - a.u[iA]) / a.cyspace;
        r.vx = (a.v[iD] - a.v[iA]) / a.cxspace;
        r.vy = (a.v[iB] - a.v[iA]) / a.cyspace;
    }
    // An empty region is lo = hi = -inf, which holds no target under the
    // region test (shared_in_region): a target in a cell where the reference
    // errs always takes the full queries.
    const float2 bx = a.fx[i], by = a.fy[j];
    const float ninf = -__builtin_inff();
    r.lo_x = err ? ninf : bx.x;
    r.hi_x = err ? ninf : bx.y;
    r.lo_y = err ? ninf : by.x;
    r.hi_y = err ? ninf : by.y;
    a.cells[tiled_cell((int)i, (int)j, a.ntx)] = r;
}

// ---------------------------------------------------------------------------
// launchers

hipError_t launch_trace(const TraceArgs &a, int bk, int ck, int math, int block,
                        hipStream_t stream)
{
    if (a.n_rays <= 0) return hipSuccess;
    if (a.ray_blocks != (int)((a.n_rays + block - 1) / block)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(a.ray_blocks + (a.t_out ? 1 : 0)));
    // all four components go out by buffer stores (trace_kernel's OUT 1)
    // while one row stays under 2^30 bytes; otherwise by the kernel that
    // takes any subset (OUT 2)
    int nout = (a.xo != nullptr) + (a.yo != nullptr) + (a.kxo != nullptr) + (a.kyo != nullptr);
    if (nout == 4 && a.n_rays >= (int64_t(1) << 27)) nout = 5;
    if (a.shared_cells && bk == BK_GRID && ck == CK_GRID) {
        // a power-of-two step: dt/2 and dt scale exactly
        int e2 = 0;
        const bool dy = std::fabs(std::frexp(a.dt, &e2)) == 0.5 && std::isnormal(a.half);
        auto go = [&](auto m, auto sc) {
            constexpr int M = decltype(m)::value, S = decltype(sc)::value;
            if (nout == 0)
                hipLaunchKernelGGL((trace_kernel<BK_GRID, CK_GRID, M, 0, S>), grid, dim3(block),
                                   0, stream, a);
            else if (nout == 4)
                hipLaunchKernelGGL((trace_kernel<BK_GRID, CK_GRID, M, 1, S>), grid, dim3(block),
                                   0, stream, a);
            else
                hipLaunchKernelGGL((trace_kernel<BK_GRID, CK_GRID, M, 2, S>), grid, dim3(block),
                                   0, stream, a);
        };
        auto by_sc = [&](auto m) {
            if (dy) go(m, std::integral_constant<int, 2>{});
            else go(m, std::integral_constant<int, 1>{});
        };
        if (math == MATH_FAST) by_sc(std::integral_constant<int, MATH_FAST>{});
        else by_sc(std::integral_constant<int, MATH_LITERAL>{});
        return hipGetLastError();
    }
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using K = decltype(k);
        if (nout == 0)
            hipLaunchKernelGGL((trace_kernel<K::bk, K::ck, K::math, 0>), grid, dim3(block), 0,
                               stream, a);
        else if (nout == 4)
            hipLaunchKernelGGL((trace_kernel<K::bk, K::ck, K::math, 1>), grid, dim3(block), 0,
                               stream, a);
        else
            hipLaunchKernelGGL((trace_kernel<K::bk, K::ck, K::math, 2>), grid, dim3(block), 0,
                               stream, a);
    });
    return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &a, int bk, int ck, int math, int block, int blocks,
                        hipStream_t stream)
{
    if (a.n_rays <= 0) return hipSuccess;
    if (block < 64 || block > 256 || block % 64 || blocks < 1) return hipErrorInvalidValue;
    if (a.shared_cells && bk == BK_GRID && ck == CK_GRID) {
        if (math == MATH_FAST)
            hipLaunchKernelGGL((batch_kernel<BK_GRID, CK_GRID, MATH_FAST, true>),
                               dim3((unsigned)blocks), dim3(block), 0, stream, a);
        else
            hipLaunchKernelGGL((batch_kernel<BK_GRID, CK_GRID, MATH_LITERAL, true>),
                               dim3((unsigned)blocks), dim3(block), 0, stream, a);
        return hipGetLastError();
    }
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using KK = decltype(k);
        hipLaunchKernelGGL((batch_kernel<KK::bk, KK::ck, KK::math>), dim3((unsigned)blocks),
                           dim3(block), 0, stream, a);
    });
    return hipGetLastError();
}

hipError_t launch_clear_counter(unsigned long long *p, hipStream_t stream)
{
    hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(64), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_rhs(const RhsArgs &a, int bk, int ck, int math, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    const int block = 256;
    const dim3 grid((unsigned)((a.n + block - 1) / block));
    dispatch_kinds(bk, ck, math, [&](auto k) {
        using K = decltype(k);
        hipLaunchKernelGGL((rhs_kernel<K::bk, K::ck, K::math>), grid, dim3(block), 0, stream, a);
    });
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

hipError_t launch_build_shared(const SharedBuildArgs &a, hipStream_t stream)
{
    const int64_t ncell = (a.nx - 1) * (a.ny - 1);
    if (ncell <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(build_shared_kernel, dim3((unsigned)((ncell + block - 1) / block)),
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
