"""Single-ray behaviour cases of the reference, as data.

Each case restates one test of src/ray.rs (test_single_wave, :268-1315) or
of src/tests/test_constant_depth.rs: a field pair, one initial state, the
time span, and the reference's checks on the rows (x, y, kx, ky):
  const   column stays bitwise equal to a value on every row;
  nondec / noninc   column is monotone over the rows;
  gt / lt   last row's value is above / below the first row's;
  near    |column - value| < 10 f64::EPSILON on every row.
The same table is run against the CPU oracle and against the HIP path.
"""
from __future__ import annotations

import numpy as np

X, Y, KX, KY = 0, 1, 2, 3


def _grid(n=100):
    return np.arange(n, dtype=np.float32)


def two_depth(x, y):
    """ray.rs:258-264: 50 m for x < 50, 20 m from x = 50 on."""
    return np.where(x >= np.float32(50.0), 20.0, 50.0)


# current fields of ray.rs:879-1259: f32 (x, y) in, (f32 expression) as f64 out
CURRENTS = {
    "dudx": lambda x, y: ((x / np.float32(100.0)).astype(np.float64), np.zeros(x.shape)),
    "dudy": lambda x, y: ((y / np.float32(100.0)).astype(np.float64), np.zeros(x.shape)),
    "dvdy": lambda x, y: (np.zeros(x.shape), (y / np.float32(100.0)).astype(np.float64)),
    "dvdx": lambda x, y: (np.zeros(x.shape), (x / np.float32(100.0)).astype(np.float64)),
    "all": lambda x, y: (((x + y) / np.float32(100.0)).astype(np.float64),
                         ((x + y) / np.float32(100.0)).astype(np.float64)),
}


def bathy_spec_arrays(spec):
    """(x, y, depth[ny, nx]) of a gridded bathymetry spec."""
    kind, fn = spec
    assert kind == "grid"
    x = y = _grid()
    Xm, Ym = np.meshgrid(x, y)
    return x, y, np.asarray(fn(Xm, Ym), dtype=np.float64)


def current_spec_arrays(spec):
    kind, name = spec
    assert kind == "grid"
    x = y = _grid()
    Xm, Ym = np.meshgrid(x, y)
    u, v = CURRENTS[name](Xm, Ym)
    return x.astype(np.float64), y.astype(np.float64), u, v


# (name, bathymetry, current, (x0, y0, kx0, ky0), (t0, t_end, dt), checks, citation)
CASES = [
    ("constant_wave_shallow_x", ("constant", 10.0), ("constant", 0.0, 0.0),
     (10.0, 50.0, 0.01, 0.0), (0.0, 8.0, 1.0),
     {"const": {Y: 50.0, KX: 0.01, KY: 0.0}, "nondec": [X]}, "ray.rs:271-293"),
    ("constant_wave_shallow_xy", ("constant", 10.0), ("constant", 0.0, 0.0),
     (10.0, 10.0, 0.007, 0.007), (0.0, 8.0, 1.0),
     {"const": {KX: 0.007, KY: 0.007}, "nondec": [X, Y]}, "ray.rs:298-331"),
    ("constant_wave_deep_x", ("constant", 10.0), ("constant", 0.0, 0.0),
     (10.0, 50.0, 1.0, 0.0), (0.0, 18.0, 1.0),
     {"const": {Y: 50.0, KX: 1.0, KY: 0.0}, "nondec": [X]}, "ray.rs:336-360"),
    ("constant_wave_deep_xy", ("constant", 10.0), ("constant", 0.0, 0.0),
     (10.0, 10.0, 0.7, 0.7), (0.0, 18.0, 1.0),
     {"const": {KX: 0.7, KY: 0.7}, "nondec": [X, Y]}, "ray.rs:365-397"),
    ("two_depth_wave_shallow_x", ("grid", two_depth), ("constant", 0.0, 0.0),
     (10.0, 50.0, 0.01, 0.0), (0.0, 5.0, 1.0),
     {"const": {Y: 50.0, KY: 0.0}, "nondec": [X, KX], "gt": [KX]}, "ray.rs:402-436"),
    ("two_depth_wave_shallow_xy", ("grid", two_depth), ("constant", 0.0, 0.0),
     (10.0, 10.0, 0.007, 0.007), (0.0, 6.8, 0.1),
     {"const": {KY: 0.007}, "nondec": [X, Y, KX], "x_ge_y": True, "gt": [KX]},
     "ray.rs:441-477"),
    ("two_depth_wave_deep_x", ("grid", two_depth), ("constant", 0.0, 0.0),
     (10.0, 50.0, 1.0, 0.0), (0.0, 30.0, 1.0),
     {"const": {KY: 0.0, Y: 50.0, KX: 1.0}, "nondec": [X]}, "ray.rs:482-510"),
    ("two_depth_wave_deep_xy", ("grid", two_depth), ("constant", 0.0, 0.0),
     (10.0, 10.0, 0.7, 0.7), (0.0, 40.0, 1.0),
     {"const": {KX: 0.7, KY: 0.7}, "nondec": [X]}, "ray.rs:515-548"),
    ("slope_depth_wave_x", ("slope",), ("constant", 0.0, 0.0),
     (10.0, 1000.0, 0.1, 0.0), (0.0, 100.0, 1.0),
     {"nondec_finite": [X], "gt_finite": [KX]}, "ray.rs:553-586"),
    ("constant_depth_zero_current", ("constant", 10.0), ("constant", 0.0, 0.0),
     (0.0, 0.0, 0.1, 0.0), (100.0, 102.0, 1.0),
     {"const": {Y: 0.0, KX: 0.1, KY: 0.0}, "nondec": [X]}, "ray.rs:592-614"),
    ("constant_depth_and_current", ("constant", 10.0), ("constant", 0.5, 0.0),
     (0.0, 0.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {Y: 0.0, KX: 0.1, KY: 0.0}, "nondec": [X]}, "ray.rs:621-643"),
    ("positive_v", ("constant", 1000.0), ("constant", 0.0, 0.5),
     (0.0, 0.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1, KY: 0.0}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:648-679"),
    ("negative_v", ("constant", 1000.0), ("constant", 0.0, -0.5),
     (0.0, 0.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1, KY: 0.0}, "nondec": [X], "noninc": [Y], "gt": [X], "lt": [Y]},
     "ray.rs:686-717"),
    ("positive_u", ("constant", 1000.0), ("constant", 0.5, 0.0),
     (0.0, 0.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KX: 0.0, KY: 0.1}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:723-755"),
    ("negative_u", ("constant", 1000.0), ("constant", -0.5, 0.0),
     (0.0, 0.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KX: 0.0, KY: 0.1}, "noninc": [X], "nondec": [Y], "lt": [X], "gt": [Y]},
     "ray.rs:761-792"),
    ("positive_u_and_v", ("constant", 1000.0), ("constant", 0.5, 0.5),
     (0.0, 0.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1, KY: 0.0}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:798-830"),
    ("negative_u_and_v", ("constant", 1000.0), ("constant", -0.5, -0.5),
     (0.0, 0.0, -0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: -0.1, KY: 0.0}, "noninc": [X, Y], "lt": [X, Y]}, "ray.rs:836-867"),
    ("dudx_along_x", ("constant", 1000.0), ("grid", "dudx"),
     (1.0, 1.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KY: 0.0, Y: 1.0}, "nondec": [X], "noninc": [KX], "lt": [KX], "gt": [X]},
     "ray.rs:879-930"),
    ("dudx_along_y", ("constant", 1000.0), ("grid", "dudx"),
     (1.0, 1.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KY: 0.1, KX: 0.0}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:931-958"),
    ("dudy_along_x", ("constant", 1000.0), ("grid", "dudy"),
     (1.0, 50.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1}, "nondec": [X], "gt": [X], "noninc": [Y, KY], "lt": [Y, KY]},
     "ray.rs:968-1022"),
    ("dudy_along_y", ("constant", 1000.0), ("grid", "dudy"),
     (1.0, 1.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KX: 0.0, KY: 0.1}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:1024-1051"),
    ("dvdy_along_x", ("constant", 1000.0), ("grid", "dvdy"),
     (1.0, 1.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1, KY: 0.0}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:1063-1107"),
    ("dvdy_along_y", ("constant", 1000.0), ("grid", "dvdy"),
     (1.0, 1.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KX: 0.0}, "near": {X: 1.0}, "nondec": [Y], "gt": [Y], "noninc": [KY],
      "lt": [KY]}, "ray.rs:1109-1150"),
    ("dvdx_along_x", ("constant", 1000.0), ("grid", "dvdx"),
     (1.0, 1.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"const": {KX: 0.1, KY: 0.0}, "nondec": [X, Y], "gt": [X, Y]}, "ray.rs:1163-1207"),
    ("dvdx_along_y", ("constant", 1000.0), ("grid", "dvdx"),
     (50.0, 1.0, 0.0, 0.1), (1.0, 10.0, 1.0),
     {"const": {KY: 0.1}, "nondec": [Y], "gt": [Y], "noninc": [X, KX], "lt": [X, KX]},
     "ray.rs:1209-1250"),
    ("all_gradients", ("constant", 1000.0), ("grid", "all"),
     (1.0, 1.0, 0.1, 0.0), (1.0, 10.0, 1.0),
     {"nondec": [X, Y], "gt": [X, Y], "noninc": [KX, KY], "lt": [KX, KY]},
     "ray.rs:1255-1314"),
]

# Many-ray integration cases (src/tests/), per ray the strict helper checks of
# src/tests/helper/mod.rs:13-47 over the rows whose x is not NaN:
# "inc" / "dec" strictly monotone, "same" bitwise constant.
_PI = np.pi
_K = 0.05


def _dir_rays():
    # test_constant_depth.rs:15-24: 12 directions from (50 km, 25 km), |k| 0.05
    rays, checks = [], []
    for i in range(12):
        kx = 0.05 * np.cos(_PI * i / 6.0)
        ky = 0.05 * np.sin(_PI * i / 6.0)
        rays.append((50_000.0, 25_000.0, kx, ky))
        c = {"same": [KX, KY], "inc": [], "dec": []}
        for col, k in ((X, kx), (Y, ky)):
            if abs(k) < np.finfo(np.float64).eps:
                c["same"].append(col)
            elif k > 0:
                c["inc"].append(col)
            else:
                c["dec"].append(col)
        checks.append(c)
    return rays, checks


def _beach(up, down, straight, cu, cd, cs):
    return [up, down, straight], [cu, cd, cs]


_c = lambda inc=(), dec=(), same=(): {"inc": list(inc), "dec": list(dec), "same": list(same)}

MANY_CASES = [
    ("constant_depth_deep", ("constant", 2000.0), ("constant", 0.0, 0.0), *_dir_rays(),
     (0.0, 5000.0, 1.0), "test_constant_depth.rs:10-60"),
    ("constant_depth_shallow", ("constant", 10.0), ("constant", 0.0, 0.0), *_dir_rays(),
     (0.0, 5000.0, 1.0), "test_constant_depth.rs:62-112"),
    ("linear_beach_right", ("slope", 100.0, 0.0, 0.0, -0.05, 0.0), ("constant", 0.0, 0.0),
     *_beach((0.0, 0.0, _K * np.cos(_PI / 6.0), _K * np.sin(_PI / 6.0)),
             (0.0, 0.0, _K * np.cos(-_PI / 6.0), _K * np.sin(-_PI / 6.0)),
             (0.0, 0.0, _K, 0.0),
             _c(inc=[X, Y, KX], same=[KY]), _c(inc=[X, KX], dec=[Y], same=[KY]),
             _c(inc=[X, KX], same=[Y, KY])),
     (0.0, 1000.0, 1.0), "linear_beach.rs:12-71"),
    ("linear_beach_left", ("slope", 0.0, 0.0, 0.0, 0.05, 0.0), ("constant", 0.0, 0.0),
     *_beach((2000.0, 0.0, -_K * np.cos(_PI / 6.0), _K * np.sin(_PI / 6.0)),
             (2000.0, 0.0, -_K * np.cos(-_PI / 6.0), _K * np.sin(-_PI / 6.0)),
             (2000.0, 100.0, -_K, 0.0),
             _c(dec=[X, KX], inc=[Y], same=[KY]), _c(dec=[X, Y, KX], same=[KY]),
             _c(dec=[X, KX], same=[Y, KY])),
     (0.0, 1000.0, 1.0), "linear_beach.rs:73-131"),
    ("linear_beach_top", ("slope", 100.0, 0.0, 0.0, 0.0, -0.05), ("constant", 0.0, 0.0),
     *_beach((0.0, 0.0, _K * np.cos(4.0 * _PI / 6.0), _K * np.sin(4.0 * _PI / 6.0)),
             (0.0, 0.0, _K * np.cos(2.0 * _PI / 6.0), _K * np.sin(2.0 * _PI / 6.0)),
             (100.0, 0.0, 0.0, _K),
             _c(dec=[X], inc=[Y, KY], same=[KX]), _c(inc=[X, Y, KY], same=[KX]),
             _c(same=[X, KX], inc=[Y, KY])),
     (0.0, 1000.0, 1.0), "linear_beach.rs:133-192"),
    ("linear_beach_bottom", ("slope", 0.0, 0.0, 0.0, 0.0, 0.05), ("constant", 0.0, 0.0),
     *_beach((0.0, 2000.0, _K * np.cos(4.0 * _PI / 6.0), -_K * np.sin(4.0 * _PI / 6.0)),
             (0.0, 2000.0, _K * np.cos(2.0 * _PI / 6.0), -_K * np.sin(2.0 * _PI / 6.0)),
             (100.0, 2000.0, 0.0, -_K),
             _c(dec=[X, Y, KY], same=[KX]), _c(inc=[X], dec=[Y, KY], same=[KX]),
             _c(same=[X, KX], dec=[Y, KY])),
     (0.0, 1000.0, 1.0), "linear_beach.rs:194-253"),
]


def check_helper(rows, checks):
    """helper/mod.rs increase / decrease / same over rows with finite x."""
    fin = rows[~np.isnan(rows[:, X])]
    for c in checks.get("inc", []):
        assert np.all(np.diff(fin[:, c]) > 0), ("inc", c)
    for c in checks.get("dec", []):
        assert np.all(np.diff(fin[:, c]) < 0), ("dec", c)
    for c in checks.get("same", []):
        assert np.all(fin[:, c] == fin[0, c]), ("same", c)


def check_rows(rows, checks):
    """Assert the reference's checks on rows[:, 0:4] = (x, y, kx, ky)."""
    eps = np.finfo(np.float64).eps
    for c, v in checks.get("const", {}).items():
        assert np.all(rows[:, c] == v), (c, v, rows[:, c])
    for c, v in checks.get("near", {}).items():
        assert np.all(np.abs(rows[:, c] - v) < 10 * eps), (c, rows[:, c])
    for c in checks.get("nondec", []):
        assert np.all(np.diff(rows[:, c]) >= 0), (c, rows[:, c])
    for c in checks.get("noninc", []):
        assert np.all(np.diff(rows[:, c]) <= 0), (c, rows[:, c])
    for c in checks.get("gt", []):
        assert rows[-1, c] > rows[0, c], c
    for c in checks.get("lt", []):
        assert rows[-1, c] < rows[0, c], c
    if checks.get("x_ge_y"):
        assert np.all(rows[:, X] >= rows[:, Y])
    fin = rows[~np.isnan(rows[:, X])]
    for c in checks.get("nondec_finite", []):
        assert np.all(np.diff(fin[:, c]) >= 0), c
    for c in checks.get("gt_finite", []):
        assert fin[-1, c] > fin[0, c], c
