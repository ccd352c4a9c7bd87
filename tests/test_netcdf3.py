"""The product's NetCDF-3 reader (mr_bathy_open_nc3 / mr_current_open_nc3)
against files written by scipy.io.netcdf_file.

Reference behaviour (cartesian_netcdf3.rs:167-256, cartesian_current.rs:58-213):
variables are read whole, in file order, whatever their dimensions; every
numeric type is converted with Rust `as` (bathymetry axes to f32 and depth
to f64, all current variables to f64); attributes are ignored.  The netcdf3
crate maps NC_BYTE to i8 and NC_CHAR to u8 (cartesian_current.rs:579-657
reads an i8 y, an i16 x, a u8 u and an i32 v)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import mantaray_amd as mr


def write(path, dims, vars_, version=1, unlimited=None, gattrs=None):
    with netcdf_file(str(path), "w", version=version) as f:
        for name, n in dims.items():
            f.createDimension(name, None if name == unlimited else n)
        for k, v in (gattrs or {}).items():
            setattr(f, k, v)
        for name, (typ, vdims, data, attrs) in vars_.items():
            var = f.createVariable(name, typ, vdims)
            for k, v in (attrs or {}).items():
                setattr(var, k, v)
            var[:] = data


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("xt,dt", [("f4", "f8"), ("f8", "f4"), ("i2", "i4"), ("b", "f8"),
                                   ("i4", "i2")])
def test_bathymetry_types(tmp_path, version, xt, dt):
    x = np.array([0, 2, 4, 6], dtype=xt)
    y = np.array([-3, 0, 3], dtype=xt)
    d = (np.arange(12).reshape(3, 4) * 3 - 7).astype(dt)
    p = tmp_path / "b.nc"
    write(p, {"y": 3, "x": 4}, {"y": (xt, ("y",), y, None), "x": (xt, ("x",), x, None),
                                "depth": (dt, ("y", "x"), d, {"units": "m"})},
          version=version, gattrs={"title": "fixture"})
    gx, gy, gd = mr.Bathymetry.from_netcdf(p).grid()
    assert np.array_equal(gx, x.astype(np.float32).astype(np.float64))
    assert np.array_equal(gy, y.astype(np.float32).astype(np.float64))
    assert np.array_equal(gd, d.astype(np.float64).ravel())


def test_current_reference_dtypes(tmp_path):
    """cartesian_current.rs:579-657: y i8, x i16, u u8 (NC_CHAR), v i32."""
    p = tmp_path / "c.nc"
    with netcdf_file(str(p), "w") as f:
        f.createDimension("y", 2)
        f.createDimension("x", 3)
        f.createVariable("y", "b", ("y",))[:] = np.array([-1, 1], dtype=np.int8)
        f.createVariable("x", "i2", ("x",))[:] = np.array([0, 5, 10], dtype=np.int16)
        f.createVariable("u", "c", ("y", "x"))[:] = np.full((2, 3), b"\x05", dtype="S1")
        f.createVariable("v", "i4", ("y", "x"))[:] = np.zeros((2, 3), dtype=np.int32)
    gx, gy, gu, gv = mr.Current.from_netcdf(p).grid()
    assert np.array_equal(gx, [0.0, 5.0, 10.0]) and np.array_equal(gy, [-1.0, 1.0])
    assert np.array_equal(gu, np.full(6, 5.0)) and np.array_equal(gv, np.zeros(6))


def test_dimension_order_ignored(tmp_path):
    """Values are taken in file order whatever the dimension order."""
    d = np.arange(12.0).reshape(4, 3)              # stored over (x, y)
    p = tmp_path / "t.nc"
    write(p, {"x": 4, "y": 3}, {"x": ("f4", ("x",), np.arange(4.0), None),
                                "y": ("f4", ("y",), np.arange(3.0), None),
                                "depth": ("f8", ("x", "y"), d, None)})
    _, _, gd = mr.Bathymetry.from_netcdf(p).grid()
    assert np.array_equal(gd, d.ravel())


def test_record_variables(tmp_path):
    """A record (unlimited) dimension: records are interleaved on disk."""
    x = np.arange(5.0)
    y = np.arange(4.0) * 2.0
    u = np.arange(20.0).reshape(4, 5)
    v = -u
    p = tmp_path / "r.nc"
    write(p, {"y": 4, "x": 5}, {"x": ("f8", ("x",), x, None), "y": ("f8", ("y",), y, None),
                                "u": ("f8", ("y", "x"), u, None),
                                "v": ("f4", ("y", "x"), v, None)}, unlimited="y")
    gx, gy, gu, gv = mr.Current.from_netcdf(p).grid()
    assert np.array_equal(gx, x) and np.array_equal(gy, y)
    assert np.array_equal(gu, u.ravel()) and np.array_equal(gv, v.ravel())


def test_custom_names(tmp_path):
    p = tmp_path / "n.nc"
    write(p, {"lat": 2, "lon": 3}, {"lon": ("f4", ("lon",), np.arange(3.0), None),
                                    "lat": ("f4", ("lat",), np.arange(2.0), None),
                                    "h": ("f8", ("lat", "lon"), np.ones((2, 3)), None)})
    gx, gy, gd = mr.Bathymetry.from_netcdf(p, "lon", "lat", "h").grid()
    assert gx.size == 3 and gy.size == 2 and np.array_equal(gd, np.ones(6))


def test_large_values_cast_like_rust_as(tmp_path):
    """f64 -> f32 axes round to nearest; i32 beyond 2^24 rounds too."""
    x = np.array([0.1, 16777217.0, -1e30], dtype=np.float64)
    p = tmp_path / "v.nc"
    write(p, {"y": 2, "x": 3}, {"x": ("f8", ("x",), x, None),
                                "y": ("i4", ("y",), np.array([16777217, 3], dtype=np.int32), None),
                                "depth": ("f8", ("y", "x"), np.zeros((2, 3)), None)})
    gx, gy, _ = mr.Bathymetry.from_netcdf(p).grid()
    assert np.array_equal(gx, x.astype(np.float32).astype(np.float64))
    assert np.array_equal(gy, np.array([16777217, 3]).astype(np.float32).astype(np.float64))
