"""Field handles: bathymetry and current sources of the ray equations.

Python mirror of the reference's field providers:
  Bathymetry.from_netcdf   CartesianNetcdf3::open (src/bathymetry/cartesian_netcdf3.rs:167-256)
  Bathymetry.from_arrays   CartesianNetcdf3 { x, y, depth } (cartesian_netcdf3.rs:35-43)
  Bathymetry.constant      ConstantDepth::new (src/bathymetry/constant_depth.rs:42-44)
  Bathymetry.slope         ConstantSlope builder (src/bathymetry/constant_slope.rs:28-76)
  Bathymetry.array         ArrayDepth::new (src/bathymetry/array_depth.rs:38-41)
  Current.from_netcdf      CartesianCurrent::open (src/current/cartesian_current.rs:58-213)
  Current.from_arrays      CartesianCurrent { x_vec, y_vec, u_vec, v_vec } (:18-27)
  Current.constant         ConstantCurrent::new (src/current/constant_current.rs:30-32)
Each wraps an immutable mr_field handle of libmantaray_hip.so.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi as C

# builder defaults of the reference
DEFAULT_BATHYMETRY = 2000.0          # constant_depth.rs:9
CONSTANT_DEPTH_DEFAULT = 1000.0      # constant_depth.rs:16
SLOPE_DEFAULTS = dict(h0=50.0, x0=0.0, y0=0.0, dhdx=-5e-2, dhdy=0.0)   # constant_slope.rs:30-42


class _Field:
    role = ""

    def __init__(self, handle: int, keep=None):
        self._h = ctypes.c_void_p(handle)
        self._keep = keep

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def kind(self) -> int:
        k = ctypes.c_int32()
        C.check(C.lib().mr_field_kind(self._h, ctypes.byref(k)))
        return k.value

    def grid(self):
        """(x, y, values...) of a gridded field as numpy arrays (f64)."""
        nx, ny, la, lb = (ctypes.c_int64() for _ in range(4))
        C.check(C.lib().mr_field_grid_shape(self._h, *[ctypes.byref(v) for v in (nx, ny, la, lb)]))
        x, y = np.empty(nx.value), np.empty(ny.value)
        a, b = np.empty(la.value), np.empty(lb.value)
        C.check(C.lib().mr_field_copy_grid(self._h, x.ctypes.data, y.ctypes.data,
                                           a.ctypes.data, b.ctypes.data))
        return (x, y, a) if self.role == "bathymetry" else (x, y, a, b)

    def upload(self, device: int = -1) -> None:
        """Build the device tables now (otherwise built on first use)."""
        C.check(C.lib().mr_field_upload(self._h, device))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            C.lib().mr_field_free(self._h)
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _new(fn, *args) -> int:
    out = ctypes.c_void_p()
    C.check(fn(*args, ctypes.byref(out)))
    return out.value


class Bathymetry(_Field):
    role = "bathymetry"

    @classmethod
    def from_netcdf(cls, path, xname="x", yname="y", depth_name="depth") -> "Bathymetry":
        return cls(_new(C.lib().mr_bathy_open_nc3, os.fsencode(str(path)), xname.encode(),
                        yname.encode(), depth_name.encode()))

    @classmethod
    def from_arrays(cls, x, y, depth) -> "Bathymetry":
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        d = np.ascontiguousarray(depth, dtype=np.float64).ravel()
        return cls(_new(C.lib().mr_bathy_from_arrays, x.size, y.size, x.ctypes.data,
                        y.ctypes.data, d.size, d.ctypes.data))

    @classmethod
    def constant(cls, h=CONSTANT_DEPTH_DEFAULT) -> "Bathymetry":
        return cls(_new(C.lib().mr_bathy_constant, float(h)))

    @classmethod
    def slope(cls, h0=SLOPE_DEFAULTS["h0"], x0=SLOPE_DEFAULTS["x0"], y0=SLOPE_DEFAULTS["y0"],
              dhdx=SLOPE_DEFAULTS["dhdx"], dhdy=SLOPE_DEFAULTS["dhdy"]) -> "Bathymetry":
        return cls(_new(C.lib().mr_bathy_slope, float(h0), float(x0), float(y0), float(dhdx),
                        float(dhdy)))

    @classmethod
    def array(cls, a) -> "Bathymetry":
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise ValueError("ArrayDepth takes a square array")
        return cls(_new(C.lib().mr_bathy_array, a.shape[0], a.ctypes.data))


class Current(_Field):
    role = "current"

    @classmethod
    def from_netcdf(cls, path, xname="x", yname="y", uname="u", vname="v") -> "Current":
        return cls(_new(C.lib().mr_current_open_nc3, os.fsencode(str(path)), xname.encode(),
                        yname.encode(), uname.encode(), vname.encode()))

    @classmethod
    def from_arrays(cls, x, y, u, v) -> "Current":
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        return cls(_new(C.lib().mr_current_from_arrays, x.size, y.size, x.ctypes.data,
                        y.ctypes.data, u.size, u.ctypes.data, v.size, v.ctypes.data))

    @classmethod
    def constant(cls, u=0.0, v=0.0) -> "Current":
        return cls(_new(C.lib().mr_current_constant, float(u), float(v)))
