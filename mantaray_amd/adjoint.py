"""Differentiable ray tracing on the GPU: the fp64-interpolation trace, its
discrete adjoint, and a torch.autograd.Function over both.

  trace_fp64(...)        mr_trace_fp64: the trace with every field query in
                         f64, optional value overrides and checkpoints
  adjoint(...)           mr_adjoint_many: the reverse sweep
  RayTrace               torch.autograd.Function: (x0, y0, kx0, ky0, depth,
                         u, v) -> last rows (and optionally every row)
  trace_differentiable   the convenience call of RayTrace

The reference (mantaray 0.1.0) has no derivatives; this is the path that
ManyRays::trace_many (src/ray.rs:98-127) takes when a loss on the traced rays
must be differentiated with respect to the initial states or the gridded
fields.  The gradients are those of the discrete RK4 map the forward
computes (checked against central finite differences of the same forward).
Everything runs on the HIP kernels of libmantaray_hip.so on torch device
tensors; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _capi as C
from .core import _opts, num_steps
from .fields import Bathymetry, Current


def _ptr(t):
    return None if t is None else t.data_ptr()


def value_lengths(field) -> tuple[int, int]:
    """Lengths of a field's value arrays: (depth_len, 0) or (u_len, v_len)."""
    nx, ny, la, lb = (ctypes.c_int64() for _ in range(4))
    C.check(C.lib().mr_field_grid_shape(field.handle, *[ctypes.byref(v) for v in (nx, ny, la, lb)]))
    return la.value, lb.value


def checkpoint_count(t0: float, t_end: float, dt: float, stride: int) -> int:
    n = ctypes.c_int64()
    C.check(C.lib().mr_checkpoint_count(float(t0), float(t_end), float(dt), int(stride),
                                        ctypes.byref(n)))
    return n.value


def workspace_bytes(n_rays: int, stride: int) -> int:
    n = ctypes.c_int64()
    C.check(C.lib().mr_adjoint_workspace_size(int(n_rays), int(stride), ctypes.byref(n)))
    return n.value


def _check_values(field, arr, which, length, device=None):
    import torch
    if arr is None:
        return None
    if not isinstance(arr, torch.Tensor) or arr.dtype != torch.float64 or not arr.is_cuda:
        raise TypeError(f"{which} must be a float64 CUDA tensor")
    if device is not None and arr.device != device:
        raise ValueError(f"{which} is on {arr.device}; the trace runs on {device}")
    if arr.numel() != length:
        raise ValueError(f"{which} has {arr.numel()} values; the field has {length}")
    return arr.contiguous()


def _device_opts(dev, stream, math, block_size):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(dev)
    return _opts(math, "state", "fp64", block_size=block_size, memory=C.MR_MEM_DEVICE,
                 device=dev.index if dev.index is not None else -1,
                 stream=ctypes.c_void_p(stream.cuda_stream))


def trace_fp64(bathymetry: Bathymetry, current: Current, x0, y0, kx0, ky0, t0: float,
               t_end: float, dt: float, *, depth=None, u=None, v=None, stride: int = 0,
               trajectories: bool = False, math: str = "fast", block_size: int = 0,
               stream=None) -> dict:
    """Trace float64 device tensors in fp64-interpolation mode.

    depth / u / v (float64 device tensors of the field's value-array
    lengths) replace the fields' own values.  Returns a dict of tensors:
    len (n_rays,) int32, end (n_rays, 4), and with trajectories t (n+1,) and
    x, y, kx, ky (n+1, n_rays); with stride > 0 also ckpt
    (n_ckpt, 4, n_rays), the state every stride steps.
    """
    import torch

    from ._tensors import check_initial_conditions
    dev = check_initial_conditions((x0, y0, kx0, ky0))
    ics = [a.contiguous() for a in (x0, y0, kx0, ky0)]
    nr = min(a.numel() for a in ics)
    la, _ = value_lengths(bathymetry)
    lu, lv = value_lengths(current)
    depth = _check_values(bathymetry, depth, "depth", la, dev)
    u = _check_values(current, u, "u", lu, dev)
    v = _check_values(current, v, "v", lv, dev)
    n = num_steps(t0, t_end, dt)
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"len": torch.empty(nr, dtype=torch.int32, device=dev),
           "end": torch.empty((nr, 4), **f64)}
    if trajectories:
        out["t"] = torch.empty(n + 1, **f64)
        for k in ("x", "y", "kx", "ky"):
            out[k] = torch.empty((n + 1, nr), **f64)
    if stride > 0:
        out["ckpt"] = torch.empty((checkpoint_count(t0, t_end, dt, stride), 4, nr), **f64)
    o = _device_opts(dev, stream, math, block_size)
    C.check(C.lib().mr_trace_fp64(
        bathymetry.handle, current.handle, _ptr(depth), _ptr(u), _ptr(v), nr,
        *[a.data_ptr() for a in ics], float(t0), float(t_end), float(dt), int(stride),
        ctypes.byref(o), _ptr(out.get("ckpt")), _ptr(out.get("t")), _ptr(out.get("x")),
        _ptr(out.get("y")), _ptr(out.get("kx")), _ptr(out.get("ky")), out["len"].data_ptr(),
        out["end"].data_ptr()))
    return out


def adjoint(bathymetry: Bathymetry, current: Current, fwd: dict, t0: float, t_end: float,
            dt: float, stride: int, *, depth=None, u=None, v=None, g_end=None, g_rows=None,
            want=("x0", "y0", "kx0", "ky0", "depth", "u", "v"), math: str = "fast",
            block_size: int = 0, stream=None, work=None) -> dict:
    """Reverse sweep over a trace_fp64 result (made with the same stride,
    fields, values and math mode).

    g_end: cotangent of fwd["end"] (n_rays, 4); g_rows: dict with any of
    "x", "y", "kx", "ky" -> cotangent of that trajectory (n+1, n_rays).
    Returns the gradients named in `want`: x0 .. ky0 (n_rays,), and depth,
    u, v (the value arrays; only for gridded fields).
    """
    import torch

    from ._tensors import check_device_tensor
    if fwd.get("ckpt") is None:
        raise ValueError("fwd holds no checkpoints: trace it with stride > 0")
    ck = fwd["ckpt"]
    dev = ck.device
    nr = fwd["len"].numel()
    n = num_steps(t0, t_end, dt)
    check_device_tensor(ck, "fwd['ckpt']", (checkpoint_count(t0, t_end, dt, stride), 4, nr),
                        torch.float64, dev)
    check_device_tensor(fwd["len"], "fwd['len']", (nr,), torch.int32, dev)
    check_device_tensor(fwd["end"], "fwd['end']", (nr, 4), torch.float64, dev)
    if g_end is not None:
        check_device_tensor(g_end.contiguous(), "g_end", (nr, 4), torch.float64, dev)
    for k, t in (g_rows or {}).items():
        if k not in ("x", "y", "kx", "ky"):
            raise ValueError(f"unknown row cotangent {k!r}")
        if t is not None:
            check_device_tensor(t.contiguous(), f"g_rows[{k!r}]", (n + 1, nr), torch.float64,
                                dev)
    la, _ = value_lengths(bathymetry)
    lu, lv = value_lengths(current)
    depth = _check_values(bathymetry, depth, "depth", la, dev)
    u = _check_values(current, u, "u", lu, dev)
    v = _check_values(current, v, "v", lv, dev)
    f64 = dict(dtype=torch.float64, device=dev)
    res = {}
    for k in ("x0", "y0", "kx0", "ky0"):
        res[k] = torch.empty(nr, **f64) if k in want else None
    res["depth"] = (torch.empty(la, **f64)
                    if "depth" in want and bathymetry.kind == C.MR_BATHY_GRID else None)
    grid_c = current.kind == C.MR_CURRENT_GRID
    res["u"] = torch.empty(lu, **f64) if "u" in want and grid_c else None
    res["v"] = torch.empty(lv, **f64) if "v" in want and grid_c else None
    nbytes = workspace_bytes(nr, stride)
    if work is not None and (not isinstance(work, torch.Tensor) or work.device != dev or
                             not work.is_contiguous()):
        raise ValueError("work must be a contiguous tensor on the trace's device")
    if work is None or work.numel() * work.element_size() < nbytes:
        work = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=dev)
    rows = {k: (g_rows or {}).get(k) for k in ("x", "y", "kx", "ky")}
    rows = {k: (None if t is None else t.contiguous()) for k, t in rows.items()}
    ge = None if g_end is None else g_end.contiguous()
    o = _device_opts(dev, stream, math, block_size)
    C.check(C.lib().mr_adjoint_many(
        bathymetry.handle, current.handle, _ptr(depth), _ptr(u), _ptr(v), nr, float(t0),
        float(t_end), float(dt), int(stride), ctypes.byref(o), ck.data_ptr(),
        fwd["len"].data_ptr(), fwd["end"].data_ptr(), _ptr(ge), _ptr(rows["x"]),
        _ptr(rows["y"]), _ptr(rows["kx"]), _ptr(rows["ky"]), work.data_ptr(), nbytes,
        _ptr(res["x0"]), _ptr(res["y0"]), _ptr(res["kx0"]), _ptr(res["ky0"]),
        _ptr(res["depth"]), _ptr(res["u"]), _ptr(res["v"])))
    return res


def tune_stride(bathymetry: Bathymetry, current: Current, x0, y0, kx0, ky0, t0: float,
                t_end: float, dt: float, *, depth=None, u=None, v=None,
                candidates=(10, 25, 50, 100, 200), memory_budget: int | None = None,
                key: str | None = None, math: str = "fast", force: bool = False, tuner=None):
    """Pick the checkpoint stride by timing forward-with-checkpoints plus the
    reverse sweep (endpoint loss, initial-state and grid gradients) for each
    candidate whose checkpoints and work buffer fit memory_budget bytes
    (default: no limit).  A short stride stores more checkpoints; a long one
    a larger work buffer.  Returns (stride, [(candidate, ms), ...])."""
    import torch

    from .autotune import Autotuner
    nr = min(t.numel() for t in (x0, y0, kx0, ky0))

    def mem(k):
        return checkpoint_count(t0, t_end, dt, k) * 4 * nr * 8 + workspace_bytes(nr, k)

    cands = [{"stride": int(k)} for k in candidates
             if memory_budget is None or mem(k) <= memory_budget]
    if not cands:
        raise ValueError("no checkpoint stride fits the memory budget")
    want = tuple(n for n, ok in (("x0", True), ("depth", bathymetry.kind == C.MR_BATHY_GRID),
                                 ("u", current.kind == C.MR_CURRENT_GRID),
                                 ("v", current.kind == C.MR_CURRENT_GRID)) if ok)

    def run(stride):
        f = trace_fp64(bathymetry, current, x0, y0, kx0, ky0, t0, t_end, dt, depth=depth, u=u,
                       v=v, stride=stride, math=math)
        g = torch.ones_like(f["end"])
        adjoint(bathymetry, current, f, t0, t_end, dt, stride, depth=depth, u=u, v=v, g_end=g,
                want=want, math=math)

    k = key or f"stride|{nr}|{t0}|{t_end}|{dt}|{math}|{memory_budget}"
    best, table = (tuner or Autotuner()).tune(k, cands, run, force=force)
    return best["stride"], [(c["stride"], t) for c, t in table]


@dataclass
class TraceSpec:
    """Non-tensor arguments of RayTrace."""
    bathymetry: Bathymetry
    current: Current
    t0: float
    t_end: float
    dt: float
    stride: int = 50
    trajectories: bool = False
    math: str = "fast"
    block_size: int = 0


def _ray_trace_function():
    import torch

    class RayTrace(torch.autograd.Function):
        """(spec, x0, y0, kx0, ky0, depth, u, v) -> (end, len[, x, y, kx, ky]).

        end (n_rays, 4) is each ray's last row; len (n_rays,) int32 is not
        differentiable; with spec.trajectories the rows x, y, kx, ky
        (n+1, n_rays) follow.  depth, u, v may be None (the fields' own
        values, no gradient).  The forward stores the state every
        spec.stride steps; the backward recomputes each segment from its
        checkpoint and runs the discrete adjoint over it.  A ray whose last
        row is all-NaN (it left the domain) passes no gradient through that
        row.
        """

        @staticmethod
        def forward(ctx, spec, x0, y0, kx0, ky0, depth, u, v):
            ctx.set_materialize_grads(False)
            fwd = trace_fp64(spec.bathymetry, spec.current, x0, y0, kx0, ky0, spec.t0,
                             spec.t_end, spec.dt, depth=depth, u=u, v=v, stride=spec.stride,
                             trajectories=spec.trajectories, math=spec.math,
                             block_size=spec.block_size)
            ctx.spec = spec
            ctx.save_for_backward(fwd["ckpt"], fwd["len"], fwd["end"], depth, u, v)
            ctx.mark_non_differentiable(fwd["len"])
            outs = (fwd["end"], fwd["len"])
            if spec.trajectories:
                outs = outs + (fwd["x"], fwd["y"], fwd["kx"], fwd["ky"])
            return outs

        @staticmethod
        def backward(ctx, g_end, _g_len, *g_rows):
            spec = ctx.spec
            ck, ln, end, depth, u, v = ctx.saved_tensors
            need = ctx.needs_input_grad
            want = [name for name, i in (("x0", 1), ("y0", 2), ("kx0", 3), ("ky0", 4),
                                         ("depth", 5), ("u", 6), ("v", 7)) if need[i]]
            rows = None
            if g_rows:
                rows = dict(zip(("x", "y", "kx", "ky"), g_rows))
            if g_end is None and not (rows and any(r is not None for r in rows.values())):
                return (None,) * 8
            g = adjoint(spec.bathymetry, spec.current, {"ckpt": ck, "len": ln, "end": end},
                        spec.t0, spec.t_end, spec.dt, spec.stride, depth=depth, u=u, v=v,
                        g_end=g_end, g_rows=rows, want=want, math=spec.math,
                        block_size=spec.block_size)
            return (None, g["x0"], g["y0"], g["kx0"], g["ky0"], g["depth"], g["u"], g["v"])

    return RayTrace


_RAY_TRACE = None


def ray_trace_function():
    """The torch.autograd.Function class (built on first use, so that this
    module imports without torch)."""
    global _RAY_TRACE
    if _RAY_TRACE is None:
        _RAY_TRACE = _ray_trace_function()
    return _RAY_TRACE


def trace_differentiable(bathymetry: Bathymetry, current: Current, x0, y0, kx0, ky0,
                         t0: float, t_end: float, dt: float, *, depth=None, u=None, v=None,
                         stride: int = 50, trajectories: bool = False, math: str = "fast",
                         block_size: int = 0):
    """Differentiable trace: returns (end, len) or (end, len, x, y, kx, ky)
    as RayTrace does; gradients flow to whichever of x0 .. ky0, depth, u, v
    require them."""
    spec = TraceSpec(bathymetry, current, t0, t_end, dt, stride, trajectories, math, block_size)
    return ray_trace_function().apply(spec, x0, y0, kx0, ky0, depth, u, v)
