"""Benchmark: FP64 RK4 ray-steps/s of the many-ray integrator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (a "step" = one pass of the hot path over one batch): the
seamount-jet case, a 1024 x 1024 grid at 200 m with a Gaussian seamount and a
zonal jet, 65,536 rays, 5000 RK4 steps of 1 s each, full step-major
trajectories written to HBM.  Initial conditions, field tables and output
buffers are resident on the GPU before the timed region; the timed region is
K back-to-back launches of the trace kernel.  Before it, the autotuner picks
the record tile shape (4 x 4 to 32 x 32 cells; --tile fixes it) from two
timed launches of each.

The second half of the metric, the adjoint's cost, is measured on the
eddy-adjoint case (512 x 512 grid at 250 m, constant depth 50 m, 20 Gaussian
eddies, 16,384 rays, 2000 steps, checkpoint stride 50, loss on the ray
endpoints): the fp64-interpolation forward alone, and that forward with its
checkpoints followed by the reverse sweep (gradients with respect to the
initial states and the u, v grids), each timed with events on its stream.

The inversion case (2048 x 2048 grid at 100 m, depth 30 m, 262,144 rays,
3000 steps) runs 20 descent iterations of forward + adjoint from a zero
current against endpoints traced through the true current, and reports the
time per iteration, memory and the recovered current's RMS error.

The batched case (64 scenes of 256^2 to 2048^2 cells, 2^20 rays, up to
10,000 steps each) times one persistent launch of all scenes (launch shape
and record tile shape from the autotuner, and the same launch replayed from
a HIP graph) against one launch per scene.

Printed: one JSON line with the throughput, the dominant kernel's roofline
(achieved FP64 rate from its algorithmic work, 745 flop per ray-step, over
its own event-timed duration, against the FP64 vector peak: the kernel is
bound by FP64 vector issue; HBM bytes per launch and record-fetch bytes per
ray-step from the rocprofv3 counter summary in profiles/, reported only
while that summary's source hash matches the kernel's sources), the same
workload with literal math, and the CPU oracle (OpenMP over rays, all host
cores) timed on the first 16,384 rays of the same workload (median of three
runs), with a single-thread figure beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FLOP_PER_RAY_STEP = 745          # 265 basic FP64 ops + 24 transcendentals x 20
PEAK_FP64_TFLOPS = 78.6          # MI355X FP64 (vector and matrix)
# HBM bytes per launch of the trace kernel on this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel of known size
# (mantaray_amd.profiling); used only while its source hash matches the
# kernel's sources
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_kernel_seamount_jet_traffic.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--rays", type=int, default=65536)
    p.add_argument("--n-steps", type=int, default=5000, help="RK4 steps per ray")
    p.add_argument("--math", default="fast", choices=["fast", "literal"])
    p.add_argument("--block", type=int, default=0)
    p.add_argument("--tile", type=int, default=None,
                   help="log2 of the record tile edge (default: the autotuner picks it)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    p.add_argument("--no-inversion", action="store_true", help="skip the inversion run")
    p.add_argument("--no-batch", action="store_true", help="skip the batched-scenes run")
    p.add_argument("--no-literal", action="store_true",
                   help="skip the literal-math launches of the same workload")
    p.add_argument("--cpu-rays", type=int, default=16384,
                   help="rays of the workload the CPU oracle traces (from the first)")
    return p.parse_args()


def _host():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(cfg, n_rays, runs=3, single_rays=256):
    """The CPU oracle (OpenMP over rays, dynamic schedule, all host threads)
    on the first n_rays rays: median of `runs` timed runs after a warm-up,
    plus a single-thread figure on the first single_rays rays."""
    import numpy as np
    import oracle
    oracle.build()
    b = oracle.Bathymetry.grid(cfg["bathy"]["x"], cfg["bathy"]["y"], cfg["bathy"]["depth"])
    c = oracle.Current.grid(cfg["current"]["x"], cfg["current"]["y"], cfg["current"]["u"],
                            cfg["current"]["v"])

    def run(n, threads):
        sl = slice(0, n)
        t0 = time.perf_counter()
        r = oracle.trace_many(b, c, cfg["x0"][sl], cfg["y0"][sl], cfg["kx0"][sl], cfg["ky0"][sl],
                              cfg["t0"], cfg["t_end"], cfg["dt"], trajectories=False,
                              nthreads=threads)
        el = time.perf_counter() - t0
        return int((r["len"].astype(np.int64) - 1).sum()), el

    run(min(n_rays, 1024), 0)                                   # warm-up
    timed = [run(n_rays, 0) for _ in range(runs)]
    steps = timed[0][0]
    el = float(np.median([t for _, t in timed]))
    s1, e1 = run(single_rays, 1)
    model, ncpu = _host()
    return {"value": steps / el, "unit": "ray-steps/s", "cores": oracle.max_threads(),
            "kind": "port",
            "sample": f"first {n_rays} rays of the same workload, all steps ({steps} ray-steps), "
                      f"median of {runs} runs ({el:.2f} s each)",
            "single_thread": {"value": s1 / e1, "sample": f"first {single_rays} rays, 1 thread"},
            "host": {"cpu": model, "nproc": ncpu}}


def _event_ms(fn, stream, reps):
    """Median duration of fn() between events on `stream` over reps runs."""
    import numpy as np
    import torch
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def adjoint_cost(dev, reps=5):
    """fwd+bwd / fwd on the eddy-adjoint workload (see the module docstring)."""
    import torch

    from mantaray_amd import adjoint, core, workloads
    cfg = workloads.eddy_adjoint()
    b, c = workloads.make_fields(cfg)
    b.upload(0)
    c.upload(0)
    T, dt, stride = cfg["t_end"], cfg["dt"], cfg["stride"]
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    u = torch.tensor(cfg["current"]["u"].ravel(), device=dev)
    v = torch.tensor(cfg["current"]["v"].ravel(), device=dev)
    tx = torch.tensor(cfg["target_x"], device=dev)
    ty = torch.tensor(cfg["target_y"], device=dev)
    nr = tx.numel()
    stream = torch.cuda.current_stream(dev)
    work = torch.empty(adjoint.workspace_bytes(nr, stride), dtype=torch.uint8, device=dev)
    state = {}

    def fwd_plain():
        adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, u=u, v=v, stream=stream)

    def fwd_bwd():
        f = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, u=u, v=v, stride=stride, stream=stream)
        e = f["end"]
        g = torch.zeros_like(e)
        g[:, 0] = 2.0 * (e[:, 0] - tx)
        g[:, 1] = 2.0 * (e[:, 1] - ty)
        state["g"] = adjoint.adjoint(b, c, f, 0.0, T, dt, stride, g_end=g, u=u, v=v,
                                     stream=stream, work=work)
        state["f"] = f

    def fwd_ref():
        core.trace_many_device(b, c, *ic, 0.0, T, dt, trajectories=False, stream=stream)

    for fn in (fwd_plain, fwd_bwd, fwd_ref):
        fn()
    torch.cuda.synchronize()
    t_f = _event_ms(fwd_plain, stream, reps)
    t_fb = _event_ms(fwd_bwd, stream, reps)
    t_ref = _event_ms(fwd_ref, stream, reps)
    f = state["f"]
    ray_steps = int((f["len"].to(torch.int64) - 1).sum().item())
    from mantaray_amd.autotune import Autotuner
    best, table = adjoint.tune_stride(b, c, *ic, 0.0, T, dt, u=u, v=v, force=True,
                                      tuner=Autotuner())
    return {"workload": "eddy_adjoint", "grid": "512x512 @ 250 m", "rays": nr,
            "rk4_steps": int(T / dt), "checkpoint_stride": stride,
            "loss": "sum_r |(x, y)_r(T) - target_r|^2",
            "gradients": "initial (x, y, kx, ky) of every ray; every u and v grid value",
            "fwd_ms": t_f, "fwd_bwd_ms": t_fb, "ratio": t_fb / t_f,
            "fwd_reference_precision_ms": t_ref, "ratio_vs_reference_precision_fwd": t_fb / t_ref,
            "ray_steps": ray_steps,
            "checkpoint_bytes": int(f["ckpt"].numel() * 8),
            "workspace_bytes": int(work.numel()),
            "autotuned_stride": best,
            "stride_table_ms": table,
            "note": "fwd = the fp64-interpolation forward the adjoint differentiates, endpoints "
                    "only; fwd+bwd = that forward with checkpoints, then the reverse sweep"}


def inversion_cost(dev, iterations=20):
    """The inversion workload: 20 descent iterations of forward + adjoint on
    the device, from a zero current, against endpoints traced through the
    true current (see mantaray_amd.inversion)."""
    import numpy as np
    import torch

    from mantaray_amd import adjoint, inversion, workloads
    cfg = workloads.eddy_inversion()
    b, c = workloads.make_fields(cfg)
    b.upload(0)
    c.upload(0)
    T, dt, stride = cfg["t_end"], cfg["dt"], cfg["stride"]
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    ut = torch.tensor(cfg["current"]["u"], device=dev)
    vt = torch.tensor(cfg["current"]["v"], device=dev)
    obs = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, u=ut.view(-1), v=vt.view(-1))["end"][:, :2]
    obs = obs.clone()
    z = torch.zeros_like(ut)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    free0, total = torch.cuda.mem_get_info(dev)
    sigma, first = 100.0, 0.5
    r = inversion.invert(b, c, ic, obs, T, dt, u0=z, v0=z, iterations=iterations,
                         stride=stride, sigma_cells=sigma, first_step=first, method="bb",
                         u_true=ut, v_true=vt)
    h = r["history"]
    per_iter = [x["seconds"] for x in h[:-1]]
    nr = obs.shape[0]
    return {"workload": "eddy_inversion", "grid": "2048x2048 @ 100 m", "rays": nr,
            "rk4_steps": int(T / dt), "checkpoint_stride": stride, "iterations": iterations,
            "method": f"gradient descent, Barzilai-Borwein steps, gradient smoothed by a "
                      f"Gaussian of {sigma:g} cells",
            "seconds_per_iteration": float(np.median(per_iter)),
            "seconds_total": float(sum(per_iter)),
            "ray_steps_per_iteration": nr * int(T / dt),
            "loss_initial": h[0]["loss"], "loss_final": h[-1]["loss"],
            "rms_error_initial": h[0]["rms_error"], "rms_error_final": h[-1]["rms_error"],
            "rms_true_current": float(torch.sqrt((ut ** 2 + vt ** 2).mean())),
            "peak_torch_bytes": int(torch.cuda.max_memory_allocated(dev)),
            "checkpoint_bytes": int(adjoint.checkpoint_count(0.0, T, dt, stride) * 4 * nr * 8),
            "workspace_bytes": int(adjoint.workspace_bytes(nr, stride)),
            "device_bytes_in_use_at_start": int(total - free0)}


def batch_cost(dev):
    """The batched workload: 64 scenes (grid sides 256-2048 at 250 m, random
    depth and eddies, 2^20 rays of uneven lifetime, steps 0.5-2 s, up to
    10,000 steps) in one persistent launch, against one launch per scene;
    launch shape, then record tile shape, from the autotuner (the per-scene
    launches use the same tiles); the launch replayed from a HIP graph."""
    import numpy as np
    import torch

    from mantaray_amd import core, workloads
    from mantaray_amd.autotune import Autotuner
    from mantaray_amd.batch import Batch
    t0 = time.perf_counter()
    cfgs = workloads.batched_scenes()
    gen_s = time.perf_counter() - t0
    scenes = []
    for cfg in cfgs:
        b, c = workloads.make_fields(cfg)
        scenes.append({"bathymetry": b, "current": c,
                       **{k: cfg[k] for k in ("x0", "y0", "kx0", "ky0", "t0", "t_end", "dt")}})
    t0 = time.perf_counter()
    batch = Batch(scenes)
    build_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream(dev)
    best, table, tile_table = batch.autotune("batched_scenes_seed2", tuner=Autotuner(),
                                             force=True)
    batch.run()
    torch.cuda.synchronize()
    ray_steps = int(sum(int((o["len"].to(torch.int64) - 1).sum()) for o in batch.results()))
    t_batch = _event_ms(batch.run, stream, 3)
    ics = [[torch.tensor(sc[k], dtype=torch.float64, device=dev) for k in
            ("x0", "y0", "kx0", "ky0")] for sc in scenes]

    def per_scene():
        for sc, ic in zip(scenes, ics):
            core.trace_many_device(sc["bathymetry"], sc["current"], *ic, sc["t0"], sc["t_end"],
                                   sc["dt"], trajectories=False, stream=stream,
                                   record_tile=batch.record_tile)

    per_scene()
    t_scenes = _event_ms(per_scene, stream, 3)
    s = torch.cuda.Stream(dev)
    s.wait_stream(stream)
    with torch.cuda.stream(s):
        batch.run(stream=s)
    stream.wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        batch.run(stream=s)
    t_graph = _event_ms(g.replay, stream, 3)
    n_rays, bound = batch.info()
    batch.close()
    return {"workload": "batched_scenes", "scenes": len(scenes),
            "grid_sides": sorted({int(sc["bathy"]["x"].size) for sc in cfgs}), "rays": n_rays,
            "ray_steps": ray_steps, "ray_steps_bound": bound,
            "persistent_ms": t_batch, "persistent_ray_steps_per_s": ray_steps / (t_batch * 1e-3),
            "frac_at_745_flop": ray_steps * FLOP_PER_RAY_STEP / (t_batch * 1e-3) / 1e12
            / PEAK_FP64_TFLOPS,
            "graph_replay_ms": t_graph,
            "per_scene_launches_ms": t_scenes,
            "per_scene_ray_steps_per_s": ray_steps / (t_scenes * 1e-3),
            "speedup_vs_per_scene": t_scenes / t_batch,
            "autotuned": {**best, "record_tile": batch.record_tile},
            "autotune_table_ms": [[c["block_size"], c["blocks_per_cu"], t] for c, t in table],
            "record_tile_table_ms": [[c["record_tile"], t] for c, t in tile_table],
            "scene_generation_s": gen_s, "tables_build_s": build_s,
            "outputs": "row count and last row of every ray"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(rays shard across GPUs with no exchange step)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mantaray_amd import core, workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    cfg = workloads.seamount_jet(n_rays=args.rays, n_steps=args.n_steps)
    b, c = workloads.make_fields(cfg)
    b.upload(0)
    c.upload(0)
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    n = core.num_steps(cfg["t0"], cfg["t_end"], cfg["dt"])
    nr = args.rays
    out = {"t": torch.empty(n + 1, dtype=torch.float64, device=dev),
           "len": torch.empty(nr, dtype=torch.int32, device=dev),
           "end": torch.empty((nr, 4), dtype=torch.float64, device=dev)}
    for k in ("x", "y", "kx", "ky"):
        out[k] = torch.empty((n + 1, nr), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def launch(record_tile=-1):
        core.trace_many_device(b, c, *ic, cfg["t0"], cfg["t_end"], cfg["dt"], math=args.math,
                               block_size=args.block, out=out, stream=stream,
                               record_tile=record_tile)

    # record tile shape: the autotuner's pick among 4 x 4 .. 32 x 32 (each
    # candidate's tables are built in its untimed warm-up launch)
    tile_table = []
    if args.tile is None:
        from mantaray_amd.autotune import Autotuner
        best_tile, tile_table = Autotuner().tune(
            f"seamount_jet|{nr}|record_tile", [{"record_tile": t} for t in (2, 3, 4, 5)],
            launch, reps=2, warmup=1, force=True)
        tile = best_tile["record_tile"]
    else:
        tile = args.tile
    run = lambda: launch(tile)                                   # noqa: E731

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    ray_steps = int((out["len"].to(torch.int64) - 1).sum().item())

    # timed region: K launches bracketed by synchronisation; each launch also
    # bracketed by events on the launch stream for the kernel's own duration
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))

    value = ray_steps * args.steps / wall
    achieved = ray_steps * FLOP_PER_RAY_STEP / (kern_ms * 1e-3) / 1e12
    from mantaray_amd.profiling import load_traffic, source_hash
    tfile = load_traffic(TRAFFIC_FILE) if args.math == "fast" else None
    traffic = None if tfile is None else tfile["hbm_bytes_per_launch"]

    # the same launches with literal math (libm-style exp, sqrt and divide
    # in place of the fast dispersion form), beside the headline
    literal = None
    if not args.no_literal and args.math == "fast":
        def run_lit():
            core.trace_many_device(b, c, *ic, cfg["t0"], cfg["t_end"], cfg["dt"],
                                   math="literal", block_size=args.block, out=out,
                                   stream=stream, record_tile=tile)
        run_lit()
        lit_steps = int((out["len"].to(torch.int64) - 1).sum().item())
        reps = max(2, args.steps // 4)
        lev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        torch.cuda.synchronize()
        for e0, e1 in lev:
            e0.record(stream)
            run_lit()
            e1.record(stream)
        torch.cuda.synchronize()
        lit_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in lev]))
        literal = {"ray_steps_per_s": lit_steps / (lit_ms * 1e-3), "kernel_ms": lit_ms,
                   "launches": reps, "ray_steps_per_launch": lit_steps,
                   "frac_at_745_flop": lit_steps * FLOP_PER_RAY_STEP / (lit_ms * 1e-3) / 1e12
                   / PEAK_FP64_TFLOPS}
    res = {
        "metric": "FP64 RK ray-steps/s on 1 MI355X vs host CPU; adjoint fwd+bwd / fwd cost",
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "seamount_jet", "grid": "1024x1024 @ 200 m", "rays": nr,
                   "rk4_steps": n, "dt_s": cfg["dt"], "ray_steps_per_launch": ray_steps,
                   "outputs": "full step-major trajectories (x, y, kx, ky) in HBM",
                   "math": args.math, "parallelism": "rays, one per lane",
                   "record_tile": tile,
                   "record_tile_table_ms": [[ct["record_tile"], t] for ct, t in tile_table]},
        "roofline": {"bound": "fp64-valu",
                     "roof": "FP64 vector issue (no MFMA: one ray per lane, 78.6 TFLOP/s is "
                             "the MI355X FP64 vector peak)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel_ms": kern_ms,
                     "work": f"{FLOP_PER_RAY_STEP} flop/ray-step x {ray_steps} ray-steps",
                     "hbm_GBps": None if traffic is None else traffic / (kern_ms * 1e-3) / 1e9,
                     "hbm_peak_GBps": 8000.0,
                     "record_fetch_B_per_ray_step": None if tfile is None
                     else tfile["fetch_bytes_per_ray_step"],
                     "traffic_source_hash": source_hash(),
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE "
                                     "passes scaled by a 1 GiB copy; null when they were not "
                                     "measured on this kernel's sources"},
        "literal_math": literal,
    }
    # the trajectories of the forward measurement (10.5 GB) are not needed by
    # the adjoint and inversion measurements, whose memory use is reported
    del out, ic
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_adjoint:
        res["adjoint"] = adjoint_cost(dev)
    if not args.no_inversion:
        res["inversion"] = inversion_cost(dev)
    if not args.no_batch:
        res["batch"] = batch_cost(dev)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(cfg, min(args.cpu_rays, nr))
        res["speedup_vs_cpu"] = value / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
