"""NetCDF-3 fixture files in the layout the reference's tests build them
(src/io/utility.rs:29-89 and :105-181): fixed dimensions y then x, f32
coordinate variables "y" and "x" (x[i] = i * x_step as f32), and the field
variables over (y, x), so value (i, j) sits at flat index nx * j + i.
Written with scipy.io.netcdf_file (classic format)."""
from __future__ import annotations

import numpy as np


def _axes(x_num, y_num, x_step, y_step):
    x = (np.arange(x_num, dtype=np.float32) * np.float32(x_step)).astype(np.float32)
    y = (np.arange(y_num, dtype=np.float32) * np.float32(y_step)).astype(np.float32)
    return x, y


def create_bathymetry(path, x_num, y_num, x_step, y_step, depth_fn, version=1):
    """depth_fn(x f32 array, y f32 array) -> f64 array, evaluated on the
    (y, x) mesh; the reference's create_netcdf3_bathymetry."""
    from scipy.io import netcdf_file
    x, y = _axes(x_num, y_num, x_step, y_step)
    X, Y = np.meshgrid(x, y)                       # (y_num, x_num)
    depth = np.asarray(depth_fn(X, Y), dtype=np.float64).reshape(y_num, x_num)
    with netcdf_file(str(path), "w", version=version) as f:
        f.createDimension("y", y_num)
        f.createDimension("x", x_num)
        f.createVariable("y", "f4", ("y",))[:] = y
        f.createVariable("x", "f4", ("x",))[:] = x
        f.createVariable("depth", "f8", ("y", "x"))[:] = depth
    return x, y, depth


def create_current(path, x_num, y_num, x_step, y_step, current_fn, version=1):
    """current_fn(x, y) -> (u, v) f64 arrays; create_netcdf3_current."""
    from scipy.io import netcdf_file
    x, y = _axes(x_num, y_num, x_step, y_step)
    X, Y = np.meshgrid(x, y)
    u, v = current_fn(X, Y)
    u = np.broadcast_to(np.asarray(u, dtype=np.float64), X.shape).copy()
    v = np.broadcast_to(np.asarray(v, dtype=np.float64), X.shape).copy()
    with netcdf_file(str(path), "w", version=version) as f:
        f.createDimension("y", y_num)
        f.createDimension("x", x_num)
        f.createVariable("y", "f4", ("y",))[:] = y
        f.createVariable("x", "f4", ("x",))[:] = x
        f.createVariable("u", "f8", ("y", "x"))[:] = u
        f.createVariable("v", "f8", ("y", "x"))[:] = v
    return x, y, u, v


def four_depth_fn(x, y):
    """cartesian_netcdf3.rs:487-501: four quadrants of depth 20, 10, 5, 15."""
    return np.where(x < 25000.0, np.where(y < 12500.0, 20.0, 10.0),
                    np.where(y < 12500.0, 5.0, 15.0))


def simple_current(x, y):
    """cartesian_current.rs:558-560: (5, 0) everywhere."""
    return np.full(x.shape, 5.0), np.zeros(x.shape)
