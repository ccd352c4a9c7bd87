"""Ray tracing API: the drop-in for mantaray.core and _mantaray.

  single_ray(...)   mantaray.core.single_ray   (python/mantaray/core.py:9-65)
                    over _mantaray.single_ray  (src/ffi.rs:25-49)
  ray_tracing(...)  mantaray.core.ray_tracing  (python/mantaray/core.py:68-132)
                    over _mantaray.ray_tracing (src/ffi.rs:51-85)
  trace_many(...)   ManyRays::trace_many (src/ray.rs:98-127), numpy in/out
  trace_many_device(...)  the same on torch device tensors, on a stream
  rhs(...)          WaveRayPath::system (src/wave_ray_path.rs:220-234)

Every call runs the HIP kernels of libmantaray_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import datetime
from collections.abc import Sequence

import numpy as np

from . import _capi as C
from .fields import Bathymetry, Current

try:
    import xarray as _xr
except ImportError:  # the image has no xarray; use the small stand-in
    _xr = None
from . import _dataset

_MATH = {"fast": C.MR_MATH_FAST, "literal": C.MR_MATH_LITERAL}
_SOLOUT = {"state": C.MR_SOLOUT_STATE, "next_deriv": C.MR_SOLOUT_NEXT_DERIV}
_PREC = {"reference": C.MR_PREC_REFERENCE, "fp64": C.MR_PREC_FP64_INTERP}


def _Dataset(data_vars, attrs=None):
    if _xr is not None:
        return _xr.Dataset(data_vars=data_vars, attrs=attrs or {})
    return _dataset.Dataset(data_vars=data_vars, attrs=attrs)


def num_steps(t0: float, t_end: float, dt: float) -> int:
    """ode_solvers Rk4 step count ceil((t_end - t0)/dt); ValueError where it panics."""
    n = ctypes.c_int64()
    rc = C.lib().mr_num_steps(float(t0), float(t_end), float(dt), ctypes.byref(n))
    if rc == C.MR_ERR_STEP_COUNT:
        raise ValueError(C.lib().mr_last_error().decode())
    C.check(rc)
    return n.value


def _opts(math="fast", solout="state", precision="reference", **kw):
    return C.options(math=_MATH[math], solout=_SOLOUT[solout], precision=_PREC[precision], **kw)


def trace_many(bathymetry: Bathymetry, current: Current, x0, y0, kx0, ky0,
               t0: float = 0.0, t_end: float = 1.0, dt: float = 1.0, *,
               math: str = "fast", solout: str = "state", trajectories: bool = True,
               block_size: int = 0, device: int = -1, precision: str = "reference",
               record_tile: int = -1) -> dict:
    """Trace rays from host arrays; returns a dict of numpy arrays.

    precision "reference" interpolates as the reference does (f32); "fp64"
    interpolates in f64 (the differentiable mode of mantaray_amd.adjoint).

    t (n+1,), x/y/kx/ky (n+1, n_rays) step-major with NaN past len[r] (only
    with trajectories=True), len (n_rays,) rows per ray including the
    terminal all-NaN row, end (n_rays, 4) the last row of each ray.
    Mismatched initial-condition lengths are truncated to the shortest, as
    the reference's zip does (src/ffi.rs:65-69).
    """
    ics = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()) for a in (x0, y0, kx0, ky0)]
    nr = min(a.size for a in ics)
    ics = [np.ascontiguousarray(a[:nr]) for a in ics]
    n = num_steps(t0, t_end, dt)
    t = np.empty(n + 1)
    outs = [np.empty((n + 1, nr)) if trajectories else None for _ in range(4)]
    ln = np.empty(nr, dtype=np.int32)
    end = np.empty((nr, 4))
    o = _opts(math, solout, precision, block_size=block_size, device=device,
              record_tile=record_tile)
    C.check(C.lib().mr_trace_many(
        bathymetry.handle, current.handle, nr, *[a.ctypes.data for a in ics],
        float(t0), float(t_end), float(dt), ctypes.byref(o), t.ctypes.data,
        *[(a.ctypes.data if a is not None else None) for a in outs],
        ln.ctypes.data, end.ctypes.data))
    res = {"t": t, "len": ln, "end": end}
    for name, a in zip(("x", "y", "kx", "ky"), outs):
        res[name] = a
    return res


def trace_many_device(bathymetry: Bathymetry, current: Current, x0, y0, kx0, ky0,
                      t0: float = 0.0, t_end: float = 1.0, dt: float = 1.0, *,
                      math: str = "fast", solout: str = "state", trajectories: bool = True,
                      block_size: int = 0, out: dict | None = None, stream=None,
                      precision: str = "reference", record_tile: int = -1) -> dict:
    """Trace rays whose initial conditions are float64 torch tensors on a GPU.

    Outputs are torch tensors on the same device (or the preallocated ones in
    `out`), written by kernels enqueued on `stream` (default: torch's current
    stream); the call does not synchronise.  record_tile: log2 of the edge of
    the square tiles the gridded fields' cell records are stored in (-1 = 8 x
    8; results do not depend on it).
    """
    import torch

    from ._tensors import check_device_tensor, check_initial_conditions
    dev = check_initial_conditions((x0, y0, kx0, ky0))
    ics = [a.contiguous() for a in (x0, y0, kx0, ky0)]
    nr = min(a.numel() for a in ics)
    n = num_steps(t0, t_end, dt)
    if out is None:
        out = {"t": torch.empty(n + 1, dtype=torch.float64, device=dev),
               "len": torch.empty(nr, dtype=torch.int32, device=dev),
               "end": torch.empty((nr, 4), dtype=torch.float64, device=dev)}
        for k in ("x", "y", "kx", "ky"):
            out[k] = (torch.empty((n + 1, nr), dtype=torch.float64, device=dev)
                      if trajectories else None)
    else:
        shapes = {"t": (n + 1,), "len": (nr,), "end": (nr, 4), "x": (n + 1, nr),
                  "y": (n + 1, nr), "kx": (n + 1, nr), "ky": (n + 1, nr)}
        unknown = set(out) - set(shapes)
        if unknown:
            raise ValueError(f"unknown output names {sorted(unknown)}")
        for k, shape in shapes.items():
            if out.get(k) is not None:
                check_device_tensor(out[k], f"out[{k!r}]", shape,
                                    torch.int32 if k == "len" else torch.float64, dev)
    if stream is None:
        stream = torch.cuda.current_stream(dev)
    o = _opts(math, solout, precision, block_size=block_size, memory=C.MR_MEM_DEVICE,
              device=dev.index if dev.index is not None else -1,
              stream=ctypes.c_void_p(stream.cuda_stream), record_tile=record_tile)

    def p(k):
        v = out.get(k)
        return None if v is None else v.data_ptr()

    C.check(C.lib().mr_trace_many(
        bathymetry.handle, current.handle, nr, *[a.data_ptr() for a in ics],
        float(t0), float(t_end), float(dt), ctypes.byref(o), p("t"), p("x"), p("y"),
        p("kx"), p("ky"), p("len"), p("end")))
    return out


def rhs(bathymetry: Bathymetry, current: Current, states, *, math: str = "fast",
        precision: str = "reference") -> np.ndarray:
    """WaveRayPath::system at states (N, 4) -> derivatives (N, 4), all-NaN on error."""
    s = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(-1, 4))
    ds = np.empty_like(s)
    o = _opts(math, precision=precision)
    C.check(C.lib().mr_rhs_many(bathymetry.handle, current.handle, s.shape[0], s.ctypes.data,
                                ds.ctypes.data, ctypes.byref(o)))
    return ds


def _fields(bathymetry, current):
    b = bathymetry if isinstance(bathymetry, Bathymetry) else Bathymetry.from_netcdf(str(bathymetry))
    c = current if isinstance(current, Current) else Current.from_netcdf(str(current))
    return b, c


def single_ray(x0: float, y0: float, kx0: float, ky0: float, duration: float,
               step_size: float, bathymetry, current, **kw):
    """Propagate a single ray; see mantaray.core.single_ray (core.py:9-65).

    `bathymetry` and `current` are NetCDF-3 paths (variables x, y, depth and
    x, y, u, v, as src/ffi.rs:36-38 opens them) or field objects.  Returns a
    Dataset with variables time, x, y, kx, ky over dimension time_step, every
    row including the terminal all-NaN one; time, x, y are coordinates.
    """
    b, c = _fields(bathymetry, current)
    return single_dataset(trace_many(b, c, [x0], [y0], [kx0], [ky0], 0.0, duration, step_size,
                                     **kw))


def single_dataset(r: dict):
    """The single_ray Dataset (core.py:56-63) of a one-ray trace_many result:
    every row of the ray, the terminal all-NaN one included."""
    ln = int(r["len"][0])
    cols = [r["t"][:ln], r["x"][:ln, 0], r["y"][:ln, 0], r["kx"][:ln, 0], r["ky"][:ln, 0]]
    varnames = ["time", "x", "y", "kx", "ky"]
    ds = _Dataset({v: (["time_step"], col.copy()) for v, col in zip(varnames, cols)},
                  attrs={"date_created": str(datetime.datetime.now())})
    return ds.set_coords(["time", "x", "y"])


def ray_tracing(x0: Sequence[float], y0: Sequence[float], kx0: Sequence[float],
                ky0: Sequence[float], duration: float, step_size: float, bathymetry,
                current, **kw):
    """Ray tracing for many initial conditions; see mantaray.core.ray_tracing
    (core.py:68-132).

    Returns a Dataset with variables time, x, y, kx, ky over dimensions
    (time_step, ray), each ray NaN-padded (time included) to the longest
    ray, as core.py:114-125 pads the per-ray lists.
    """
    b, c = _fields(bathymetry, current)
    return rays_dataset(trace_many(b, c, x0, y0, kx0, ky0, 0.0, duration, step_size, **kw))


def rays_dataset(r: dict):
    """The ray_tracing Dataset (core.py:104-132) of a trace_many result: each
    ray's rows, then all-NaN rows (time too) up to the longest ray."""
    ln = r["len"]
    nr = ln.size
    if nr == 0:
        raise ValueError("max() arg is an empty sequence")   # core.py:111 on no rays
    longest = int(ln.max())
    valid = np.arange(longest)[:, None] < ln[None, :]
    time = np.where(valid, r["t"][:longest, None], np.nan)
    data = {"time": time}
    for k in ("x", "y", "kx", "ky"):
        data[k] = r[k][:longest].copy()
    ds = _Dataset({name: (["time_step", "ray"], v) for name, v in data.items()})
    ds = ds.set_coords(["time", "x", "y"])
    ds["time_step"] = range(longest)
    ds["ray"] = np.arange(nr)
    ds.attrs["date_created"] = str(datetime.datetime.now())
    return ds
