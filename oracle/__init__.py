"""ctypes wrapper for the CPU oracle (test infrastructure only).

The oracle is a scalar C restatement of the reference's many-ray integrator
(see mr_oracle.c for citations and the parity status of each function).  It is
the checker for the HIP product path and the CPU baseline of bench.py; only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmr_oracle.so")

OK, ERR, PANIC = 0, 1, 2
PREC_REFERENCE, PREC_FP64_INTERP = 0, 1
SOLOUT_STATE, SOLOUT_NEXT_DERIV = 0, 1
STAGE_DIRECT, STAGE_SCALE_DT, STAGE_HALF_FIRST = 0, 1, 2
TIME_ACCUMULATE, TIME_PRODUCT = 0, 1
COMBO_SCALE_SIXTH, COMBO_DT_THEN_SIX = 0, 1
BATHY_CONSTANT, BATHY_SLOPE, BATHY_ARRAY, BATHY_GRID = 0, 1, 2, 3
CURRENT_CONSTANT, CURRENT_GRID = 0, 1

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


class _Bathy(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("h", ctypes.c_float),
        ("h0", ctypes.c_float), ("x0", ctypes.c_float), ("y0", ctypes.c_float),
        ("dhdx", ctypes.c_float), ("dhdy", ctypes.c_float),
        ("an", ctypes.c_int64), ("arr", _f32p),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("x", _f32p), ("y", _f32p),
        ("depth", _f64p), ("depth_len", ctypes.c_int64),
    ]


class _Current(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("u", ctypes.c_double), ("v", ctypes.c_double),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("x", _f64p), ("y", _f64p),
        ("u_arr", _f64p), ("v_arr", _f64p),
        ("u_len", ctypes.c_int64), ("v_len", ctypes.c_int64),
    ]


def build(force: bool = False) -> str:
    """Compile libmr_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_bilinear32.argtypes = [ctypes.POINTER(ctypes.c_float * 3 * 4), ctypes.c_float,
                                    ctypes.c_float, _f32p]
        L.or_bathy_depth_and_gradient.argtypes = [ctypes.POINTER(_Bathy), ctypes.c_float,
                                                  ctypes.c_float, _f32p, _f32p, _f32p]
        L.or_current_and_gradient.argtypes = [ctypes.POINTER(_Current), ctypes.c_double,
                                              ctypes.c_double] + [_f64p] * 6
        L.or_bathy_depth_and_gradient64.argtypes = [ctypes.POINTER(_Bathy), ctypes.c_double,
                                                    ctypes.c_double, _f64p, _f64p, _f64p]
        L.or_current_and_gradient64.argtypes = [ctypes.POINTER(_Current), ctypes.c_double,
                                                ctypes.c_double] + [_f64p] * 6
        L.or_group_velocity.argtypes = [ctypes.c_double, ctypes.c_double, _f64p]
        L.or_dkdt_bathy.argtypes = [ctypes.c_double] * 4 + [_f64p, _f64p]
        L.or_dkdt_bathy.restype = None
        L.or_num_steps.argtypes = [ctypes.c_double] * 3
        L.or_num_steps.restype = ctypes.c_int64
        L.or_trace_single.argtypes = [ctypes.POINTER(_Bathy), ctypes.POINTER(_Current),
                                      ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int64, _f64p]
        L.or_trace_single.restype = ctypes.c_int64
        L.or_trace_many.argtypes = [ctypes.POINTER(_Bathy), ctypes.POINTER(_Current),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                    _f64p, _f64p, _f64p, _f64p,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p,
                                    _i32p, _f64p]
        L.or_trace_many_ordered.argtypes = [ctypes.POINTER(_Bathy), ctypes.POINTER(_Current),
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                            _f64p, _f64p, _f64p, _f64p,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p,
                                            _i32p, _f64p]
        L.or_system_many.argtypes = [ctypes.POINTER(_Bathy), ctypes.POINTER(_Current),
                                     ctypes.c_int, ctypes.c_int64, _f64p, _f64p, _i32p]
        L.or_system_many.restype = None
        L.or_max_threads.restype = ctypes.c_int
        L.or_fd_end_ld.argtypes = [ctypes.POINTER(_Bathy), ctypes.POINTER(_Current),
                                   ctypes.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, _f64p, _f64p, _i32p, _i32p,
                                   _i32p]
        L.or_bathy_nearest.argtypes = [ctypes.POINTER(_Bathy), ctypes.c_float, ctypes.c_int, _f32p]
        L.or_current_nearest.argtypes = [ctypes.POINTER(_Current), ctypes.c_double, ctypes.c_int,
                                         _f64p]
        L.or_bathy_corners.argtypes = [ctypes.POINTER(_Bathy), ctypes.c_float, ctypes.c_float,
                                       ctypes.POINTER(ctypes.c_uint64 * 8)]
        L.or_current_corners.argtypes = [ctypes.POINTER(_Current), ctypes.c_double,
                                         ctypes.c_double, ctypes.POINTER(ctypes.c_uint64 * 8)]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Bathymetry:
    """Oracle-side bathymetry source; keeps its numpy buffers alive."""

    def __init__(self, kind, **kw):
        self.kind = kind
        self._keep = []
        s = _Bathy()
        s.kind = kind
        if kind == BATHY_CONSTANT:
            s.h = kw["h"]
        elif kind == BATHY_SLOPE:
            s.h0, s.x0, s.y0 = kw.get("h0", 50.0), kw.get("x0", 0.0), kw.get("y0", 0.0)
            s.dhdx, s.dhdy = kw.get("dhdx", -5e-2), kw.get("dhdy", 0.0)
        elif kind == BATHY_ARRAY:
            arr = np.ascontiguousarray(kw["array"], dtype=np.float32)
            assert arr.ndim == 2 and arr.shape[0] == arr.shape[1]
            self._keep.append(arr)
            s.an, s.arr = arr.shape[0], _p(arr, _f32p)
        elif kind == BATHY_GRID:
            x = np.ascontiguousarray(kw["x"], dtype=np.float32)
            y = np.ascontiguousarray(kw["y"], dtype=np.float32)
            d = np.ascontiguousarray(kw["depth"], dtype=np.float64).ravel()
            self._keep += [x, y, d]
            s.nx, s.ny = x.size, y.size
            s.x, s.y = _p(x, _f32p), _p(y, _f32p)
            s.depth, s.depth_len = _p(d, _f64p), d.size
        self.s = s

    @classmethod
    def constant(cls, h):
        return cls(BATHY_CONSTANT, h=h)

    @classmethod
    def slope(cls, h0=50.0, x0=0.0, y0=0.0, dhdx=-5e-2, dhdy=0.0):
        return cls(BATHY_SLOPE, h0=h0, x0=x0, y0=y0, dhdx=dhdx, dhdy=dhdy)

    @classmethod
    def array(cls, a):
        return cls(BATHY_ARRAY, array=a)

    @classmethod
    def grid(cls, x, y, depth):
        return cls(BATHY_GRID, x=x, y=y, depth=depth)


class Current:
    def __init__(self, kind, **kw):
        self.kind = kind
        self._keep = []
        s = _Current()
        s.kind = kind
        if kind == CURRENT_CONSTANT:
            s.u, s.v = kw.get("u", 0.0), kw.get("v", 0.0)
        else:
            x = np.ascontiguousarray(kw["x"], dtype=np.float64)
            y = np.ascontiguousarray(kw["y"], dtype=np.float64)
            u = np.ascontiguousarray(kw["u"], dtype=np.float64).ravel()
            v = np.ascontiguousarray(kw["v"], dtype=np.float64).ravel()
            self._keep += [x, y, u, v]
            s.nx, s.ny = x.size, y.size
            s.x, s.y = _p(x, _f64p), _p(y, _f64p)
            s.u_arr, s.v_arr = _p(u, _f64p), _p(v, _f64p)
            s.u_len, s.v_len = u.size, v.size
        self.s = s

    @classmethod
    def constant(cls, u=0.0, v=0.0):
        return cls(CURRENT_CONSTANT, u=u, v=v)

    @classmethod
    def grid(cls, x, y, u, v):
        return cls(CURRENT_GRID, x=x, y=y, u=u, v=v)


def bilinear32(points, target):
    """points: 4 x (x, y, value) in corner order A, B, C, D."""
    arr = (ctypes.c_float * 3 * 4)()
    for i in range(4):
        for j in range(3):
            arr[i][j] = points[i][j]
    out = ctypes.c_float()
    st = lib().or_bilinear32(ctypes.byref(arr), target[0], target[1], ctypes.byref(out))
    return st, out.value


def bathy_nearest(b: Bathymetry, target, axis):
    idx = ctypes.c_float()
    st = lib().or_bathy_nearest(ctypes.byref(b.s), np.float32(target), axis, ctypes.byref(idx))
    return st, np.float32(idx.value)


def current_nearest(c: Current, target, axis):
    idx = ctypes.c_double()
    st = lib().or_current_nearest(ctypes.byref(c.s), target, axis, ctypes.byref(idx))
    return st, idx.value


def bathy_corners(b: Bathymetry, px, py):
    out = (ctypes.c_uint64 * 8)()
    st = lib().or_bathy_corners(ctypes.byref(b.s), np.float32(px), np.float32(py), out)
    return st, [(int(out[2 * q]), int(out[2 * q + 1])) for q in range(4)]


def current_corners(c: Current, x, y):
    out = (ctypes.c_uint64 * 8)()
    st = lib().or_current_corners(ctypes.byref(c.s), x, y, out)
    return st, [(int(out[2 * q]), int(out[2 * q + 1])) for q in range(4)]


def bathy_query(b: Bathymetry, px, py):
    h, hx, hy = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    st = lib().or_bathy_depth_and_gradient(ctypes.byref(b.s), np.float32(px), np.float32(py),
                                           ctypes.byref(h), ctypes.byref(hx), ctypes.byref(hy))
    return st, (np.float32(h.value), np.float32(hx.value), np.float32(hy.value))


def current_query(c: Current, x, y):
    out = [ctypes.c_double() for _ in range(6)]
    st = lib().or_current_and_gradient(ctypes.byref(c.s), x, y, *[ctypes.byref(o) for o in out])
    return st, tuple(o.value for o in out)


def bathy_query64(b: Bathymetry, x, y):
    out = [ctypes.c_double() for _ in range(3)]
    st = lib().or_bathy_depth_and_gradient64(ctypes.byref(b.s), x, y,
                                             *[ctypes.byref(o) for o in out])
    return st, tuple(o.value for o in out)


def current_query64(c: Current, x, y):
    out = [ctypes.c_double() for _ in range(6)]
    st = lib().or_current_and_gradient64(ctypes.byref(c.s), x, y,
                                         *[ctypes.byref(o) for o in out])
    return st, tuple(o.value for o in out)


def group_velocity(k, h):
    out = ctypes.c_double()
    st = lib().or_group_velocity(k, h, ctypes.byref(out))
    return st, out.value


def dkdt_bathy(k, h, hx, hy):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().or_dkdt_bathy(k, h, hx, hy, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def num_steps(t0, t_end, dt):
    return int(lib().or_num_steps(t0, t_end, dt))


def system_many(b, c, states, prec=PREC_REFERENCE):
    s = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 4)
    ds = np.empty_like(s)
    st = np.empty(s.shape[0], dtype=np.int32)
    lib().or_system_many(ctypes.byref(b.s), ctypes.byref(c.s), prec, s.shape[0],
                         _p(s, _f64p), _p(ds, _f64p), _p(st, _i32p))
    return ds, st


def trace_single(b, c, s0, t0, t_end, dt, prec=PREC_REFERENCE, solout=SOLOUT_STATE):
    """Rows (t, x, y, kx, ky) of one ray, as SingleRay::trace_individual returns."""
    n = num_steps(t0, t_end, dt)
    if n < 0:
        raise ValueError("invalid step count (the reference panics)")
    out = np.empty((n + 1, 5))
    s0a = np.ascontiguousarray(s0, dtype=np.float64)
    rows = lib().or_trace_single(ctypes.byref(b.s), ctypes.byref(c.s), prec, solout,
                                 _p(s0a, _f64p), t0, t_end, dt, n + 1, _p(out, _f64p))
    if rows == -2:
        raise RuntimeError("reference would panic (grid dimension < 2)")
    return out[:rows].copy()


def trace_many(b, c, x0, y0, kx0, ky0, t0, t_end, dt, prec=PREC_REFERENCE,
               solout=SOLOUT_STATE, nthreads=0, trajectories=True, stage_order=STAGE_DIRECT,
               time_order=TIME_ACCUMULATE, combo_order=COMBO_SCALE_SIXTH):
    """Step-major SoA: dict with t (n+1), x/y/kx/ky (n+1, n_rays), len (n_rays), end (n_rays, 4).

    stage_order / time_order select the rounding order of the RK4 stage
    states (STAGE_DIRECT y + k*(dt/2), STAGE_SCALE_DT y + (k*dt)/2,
    STAGE_HALF_FIRST y + (k*0.5)*dt), of the time column
    (TIME_ACCUMULATE t += dt, TIME_PRODUCT t0 + i*dt) and of the step's
    combination (COMBO_SCALE_SIXTH sum*(dt/6), COMBO_DT_THEN_SIX (sum*dt)/6)."""
    x0, y0, kx0, ky0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (x0, y0, kx0, ky0))
    nr = x0.size
    n = num_steps(t0, t_end, dt)
    if n < 0:
        raise ValueError("invalid step count (the reference panics)")
    t = np.empty(n + 1)
    outs = [np.empty((n + 1, nr)) if trajectories else None for _ in range(4)]
    ln = np.empty(nr, dtype=np.int32)
    end = np.empty((nr, 4))
    rc = lib().or_trace_many_ordered(ctypes.byref(b.s), ctypes.byref(c.s), prec, solout,
                                     stage_order, time_order, combo_order, nr,
                             _p(x0, _f64p), _p(y0, _f64p), _p(kx0, _f64p), _p(ky0, _f64p),
                             t0, t_end, dt, nthreads, _p(t, _f64p),
                             *[_p(o, _f64p) for o in outs], _p(ln, _i32p), _p(end, _f64p))
    if rc == -2:
        raise RuntimeError("reference would panic (grid dimension < 2)")
    res = {"t": t, "len": ln, "end": end}
    for name, o in zip(("x", "y", "kx", "ky"), outs):
        res[name] = o
    return res


def fd_end_ld(b, c, s0, t0, t_end, dt, eps=0.0, ds0=None, ddepth=None, du=None, dv=None,
              nthreads=0, with_cells=False):
    """Long-double traces from s0 +- eps ds0 with values +- eps d (fp64-interp
    algorithm): returns diff = end(+) - end(-), mid = (end(+) + end(-))/2
    (n_rays, 4) and the two row counts; with_cells adds a boolean per ray,
    True where both traces queried the same grid cells at every evaluation
    (their central difference spans no cell line)."""
    s0 = np.ascontiguousarray(s0, dtype=np.float64).reshape(-1, 4)
    nr = s0.shape[0]
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
            for a in (ds0, ddepth, du, dv)]
    diff = np.empty((nr, 4))
    mid = np.empty((nr, 4))
    lp = np.empty(nr, dtype=np.int32)
    lm = np.empty(nr, dtype=np.int32)
    same = np.empty(nr, dtype=np.int32)
    rc = lib().or_fd_end_ld(ctypes.byref(b.s), ctypes.byref(c.s), nr, _p(s0, _f64p),
                            *[_p(a, _f64p) for a in arrs], eps, t0, t_end, dt, nthreads,
                            _p(diff, _f64p), _p(mid, _f64p), _p(lp, _i32p), _p(lm, _i32p),
                            _p(same, _i32p))
    if rc != 0:
        raise ValueError("invalid step count (the reference panics)")
    if with_cells:
        return diff, mid, lp, lm, same.astype(bool)
    return diff, mid, lp, lm


def max_threads():
    return int(lib().or_max_threads())
