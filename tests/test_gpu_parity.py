"""Parity of the HIP path (libmantaray_hip.so, through its C ABI) with the
CPU oracle, on the GPU.

Tolerances:
  - trajectories on the linear-beach workload: |dx|, |dy| <= 1e-6 m, the
    bound the reference comparison is held to; identical row counts;
  - right-hand side at arbitrary states: identical NaN pattern (every error
    and termination case), finite values within 1e-12 relative (literal
    transcendentals) / 1e-11 (fast dispersion form) of the column scale;
  - the interpolated depth/current themselves are f32 values computed with
    the reference's operations, so states whose derivatives depend only on
    them (constant-depth rays) are compared exactly.
"""
import numpy as np
import pytest

from _helpers import compare_rhs, oracle_fields, random_grid_case, random_states

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    from mantaray_amd import _capi
    _capi.lib()
    return mantaray_amd


@pytest.fixture(scope="module")
def jet_small():
    from mantaray_amd import workloads
    cfg = workloads.seamount_jet(n_rays=1024, n_steps=400)
    b, c = workloads.make_fields(cfg)
    return cfg, b, c


@pytest.mark.parametrize("tile", [0, 2, 4, 5])
def test_record_tiles_give_identical_traces(mr, jet_small, tile):
    """The record tile shape (row-major, 4 x 4, 16 x 16, 32 x 32 cells)
    moves cell records in memory and changes nothing else: trajectories,
    row counts and right-hand sides are bitwise those of the 8 x 8 default
    (seamount-jet fields, 1024 rays, 400 steps)."""
    cfg, b, c = jet_small
    args = (b, c, cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"], cfg["t_end"],
            cfg["dt"])
    ref = mr.trace_many(*args)
    got = mr.trace_many(*args, record_tile=tile)
    assert np.array_equal(got["len"], ref["len"])
    for k in ("t", "x", "y", "kx", "ky", "end"):
        assert np.array_equal(got[k], ref[k], equal_nan=True), k
    s = np.stack([ref["x"][::37].ravel(), ref["y"][::37].ravel(), ref["kx"][::37].ravel(),
                  ref["ky"][::37].ravel()], axis=1)
    from mantaray_amd import _capi as C
    import ctypes
    ds = {}
    for t in (-1, tile):
        o = C.options(record_tile=t)
        out = np.empty_like(s)
        C.check(C.lib().mr_rhs_many(b.handle, c.handle, s.shape[0], s.ctypes.data,
                                    out.ctypes.data, ctypes.byref(o)))
        ds[t] = out
    assert np.isfinite(ds[-1]).sum() > s.size // 2
    assert np.array_equal(ds[tile], ds[-1], equal_nan=True)


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_linear_beach_trajectories(mr, oracle_mod, math):
    from mantaray_amd import workloads
    cfg = workloads.linear_beach()
    b, c = workloads.make_fields(cfg)
    got = mr.trace_many(b, c, cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"],
                        cfg["t_end"], cfg["dt"], math=math)
    ob, oc = oracle_fields(oracle_mod, cfg)
    ref = oracle_mod.trace_many(ob, oc, cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"],
                                cfg["t0"], cfg["t_end"], cfg["dt"])
    assert np.array_equal(got["len"], ref["len"])
    assert np.array_equal(got["t"], ref["t"])
    for k in ("x", "y"):
        assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k]))
        d = np.nanmax(np.abs(got[k] - ref[k]))
        print(f"{math} max |d{k}| = {d:.3e} m")
        assert d <= 1e-6
    for k in ("kx", "ky"):
        d = np.nanmax(np.abs(got[k] - ref[k]) / np.abs(ref[k]).max())
        assert d <= 1e-9


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("math", ["fast", "literal"])
def test_rhs_random_states(mr, oracle_mod, seed, math):
    cfg = random_grid_case(seed)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 4096)
    ds = mr.rhs(b, c, s, math=math)
    ref, st = oracle_mod.system_many(ob, oc, s)
    assert np.array_equal(np.isnan(ds), np.isnan(ref))
    scale = np.nanmax(np.abs(ref), axis=0)
    err = np.nanmax(np.abs(ds - ref) / scale)
    print(f"seed {seed} {math}: max scaled error {err:.3e}, {int((st != 0).sum())} error rows")
    assert err <= (1e-12 if math == "literal" else 1e-11)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("math", ["fast", "literal"])
def test_short_traces_from_special_states(mr, oracle_mod, seed, math):
    """One and three RK4 steps from the states of test_rhs_random_states
    (exact nodes, cell lines, the last node row and column, points off the
    grid, NaN and zero k).  The trace kernel answers an evaluation from the
    cell records in hand and passes nodes, cell changes and errors to the
    full queries; its rows match the oracle's row for row: identical row
    counts and NaN patterns, positions within 1e-8 m, wavenumbers within
    1e-10 of their scale."""
    cfg = random_grid_case(seed)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 4096)
    for n in (1, 3):
        got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, float(n), 1.0,
                            math=math)
        ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0,
                                    float(n), 1.0)
        assert np.array_equal(got["len"], ref["len"])
        live = int((ref["len"] == n + 1).sum())
        assert live > 1000, live                # most rays take every step
        for k in ("x", "y", "kx", "ky"):
            assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), (n, k)
        dxy = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("x", "y"))
        kscale = np.nanmax(np.abs(ref["kx"])) + np.nanmax(np.abs(ref["ky"]))
        dk = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("kx", "ky")) / kscale
        print(f"seed {seed} {math} {n} steps: {live} rays live, max |d(x,y)| {dxy:.2e} m, "
              f"max scaled |dk| {dk:.2e}")
        assert dxy <= 1e-8 and dk <= 1e-10


@pytest.mark.parametrize("shift", [0.0, 0.37, 1.0])
def test_depth_and_current_cells(mr, oracle_mod, shift):
    """The trace predicts the next cells with one test when the depth and
    current grids share their cells (shift 0: current axes equal to the
    depth's) and with a test per field otherwise (the current's axes moved
    by 0.37 and by 1 cell).  Either way a prediction only chooses what to
    load ahead: trajectories against the oracle, identical NaN patterns and
    row counts, |dx|, |dy| <= 1e-6 m."""
    cfg = random_grid_case(11)
    bx, by = cfg["bathy"]["x"], cfg["bathy"]["y"]
    cx = cfg["current"]["x"] + shift * float(bx[1] - bx[0])
    cy = cfg["current"]["y"] + shift * float(by[1] - by[0])
    b = mr.Bathymetry.from_arrays(bx, by, cfg["bathy"]["depth"])
    c = mr.Current.from_arrays(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
    ob = oracle_mod.Bathymetry.grid(bx, by, cfg["bathy"]["depth"])
    oc = oracle_mod.Current.grid(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
    rng = np.random.default_rng(11)
    n = 2048
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.02, 0.2, n)
    x, y = bx.astype(np.float64), by.astype(np.float64)
    s = np.stack([rng.uniform(x[2], x[-3], n), rng.uniform(y[2], y[-3], n),
                  k * np.cos(th), k * np.sin(th)], 1)
    got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, 80.0, 1.0)
    tr = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, 80.0, 1.0)
    assert np.array_equal(got["len"], tr["len"])
    assert (tr["len"] > 20).sum() > n // 2
    for kk in ("x", "y", "kx", "ky"):
        assert np.array_equal(np.isnan(got[kk]), np.isnan(tr[kk])), kk
    assert max(np.nanmax(np.abs(got[kk] - tr[kk])) for kk in ("x", "y")) <= 1e-6


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_truncated_value_arrays(mr, oracle_mod, math):
    """Value arrays shorter than nx * ny, as a file's variable may be
    (cartesian_netcdf3.rs:465-471, cartesian_current.rs:365-398): a cell
    with a corner past the data errs, so a ray that reaches one ends with an
    all-NaN row.  Depth, u and v are cut at different points; the states
    start near that end of the grid and move into it.  Right-hand side and
    trajectories against the oracle: identical NaN patterns and row counts,
    values within the tolerances of the other tests."""
    cfg = random_grid_case(7)
    nx, ny = cfg["bathy"]["x"].size, cfg["bathy"]["y"].size
    depth = cfg["bathy"]["depth"].ravel()[: nx * ny - nx - 3]
    u = cfg["current"]["u"].ravel()[: nx * ny - 5]
    v = cfg["current"]["v"].ravel()[: nx * ny - 2 * nx + 4]
    b = mr.Bathymetry.from_arrays(cfg["bathy"]["x"], cfg["bathy"]["y"], depth)
    c = mr.Current.from_arrays(cfg["current"]["x"], cfg["current"]["y"], u, v)
    ob = oracle_mod.Bathymetry.grid(cfg["bathy"]["x"], cfg["bathy"]["y"], depth)
    oc = oracle_mod.Current.grid(cfg["current"]["x"], cfg["current"]["y"], u, v)
    x, y = cfg["bathy"]["x"].astype(np.float64), cfg["bathy"]["y"].astype(np.float64)
    rng = np.random.default_rng(7)
    n = 2048
    th = rng.uniform(0.2, np.pi - 0.2, n)                  # heading towards +y
    k = rng.uniform(0.02, 0.2, n)
    s = np.stack([rng.uniform(x[0], x[-1], n), rng.uniform(y[-5], y[-1], n),
                  k * np.cos(th), k * np.sin(th)], 1)
    s[:64, 0] = rng.choice(x, 64)                          # on grid lines and nodes
    s[32:96, 1] = rng.choice(y[-5:], 64)
    ds = mr.rhs(b, c, s, math=math)
    ref, st = oracle_mod.system_many(ob, oc, s)
    assert np.array_equal(np.isnan(ds), np.isnan(ref))
    assert (st != 0).sum() > 100                           # states in error cells
    scale = np.nanmax(np.abs(ref), axis=0)
    assert np.nanmax(np.abs(ds - ref) / scale) <= (1e-12 if math == "literal" else 1e-11)
    got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, 60.0, 1.0, math=math)
    tr = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, 60.0, 1.0)
    assert np.array_equal(got["len"], tr["len"])
    ended = int((tr["len"] < 61).sum())
    print(f"{math}: {int((st != 0).sum())} error states, {ended} of {n} rays end early")
    assert ended > 200
    for kk in ("x", "y", "kx", "ky"):
        assert np.array_equal(np.isnan(got[kk]), np.isnan(tr[kk])), kk
    assert max(np.nanmax(np.abs(got[kk] - tr[kk])) for kk in ("x", "y")) <= 1e-6


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_reference_one_step_values(mr, math):
    """wave_ray_path.rs test_axis (:405-416) and test_shallow (:521-548):
    one RK4 step from (0, 0) with t0 = 0, tf = 1, dt = 1."""
    c = mr.Current.constant(0.0, 0.0)
    g2 = np.sqrt(9.8) / 2.0
    for h, cases in (
        (1000.0, [(0.0, 1.0, 0.0, g2), (1.0, 0.0, g2, 0.0), (0.0, -1.0, 0.0, -g2),
                  (-1.0, 0.0, -g2, 0.0)]),
        (0.1, [(0.0, 1.0, 6.031543168844801e-17, 0.9850257515953494),
               (1.0, 0.0, 0.9850257515953494, 0.0),
               (0.0, -1.0, 6.031543168844801e-17, -0.9850257515953494),
               (-1.0, 0.0, -0.9850257515953494, 1.2063086337689602e-16)]),
    ):
        b = mr.Bathymetry.constant(h)
        for kx, ky, xf, yf in cases:
            r = mr.trace_many(b, c, [0.0], [0.0], [kx], [ky], 0.0, 1.0, 1.0, math=math)
            x, y = r["x"][-1, 0], r["y"][-1, 0]
            # the reference asserts |x - xf| < f64::EPSILON
            assert abs(x - xf) < 4 * np.finfo(float).eps and abs(y - yf) < 4 * np.finfo(float).eps


def test_reference_termination_cases(mr):
    """wave_ray_path.rs test_solout (:601-626), test_zero_k (:362-381),
    test_zero_h (:383-401), out_of_range_give_nan (:552-570)."""
    c = mr.Current.constant(0.0, 0.0)
    a = mr.Bathymetry.array(np.full((2, 2), 1000.0))
    r = mr.trace_many(a, c, [0.0], [0.0], [0.0], [1.0], 0.0, 10.0, 1.0)
    assert r["len"][0] == 3 and np.isnan(r["end"][0]).all()
    r = mr.trace_many(mr.Bathymetry.constant(1000.0), c, [0.0], [0.0], [0.0], [0.0], 0.0, 10.0, 1.0)
    assert np.isnan(r["end"][0]).all()
    r = mr.trace_many(mr.Bathymetry.constant(0.0), c, [0.0], [0.0], [1.0], [1.0], 0.0, 10.0, 1.0)
    assert np.isnan(r["end"][0]).all()
    r = mr.trace_many(a, c, [0.0], [0.0], [0.0], [1.0], 0.0, 10.0, 1.0)
    assert np.isnan(r["end"][0][:2]).all()


def test_core_api_netcdf(mr, tmp_path):
    """python/tests/test_core.py:40-78 on the HIP path: 3x3 NetCDF-3 files."""
    from scipy.io import netcdf_file
    with netcdf_file(tmp_path / "island.nc", "w") as f:
        f.createDimension("x", 3); f.createDimension("y", 3)
        f.createVariable("x", "f8", ("x",))[:] = [-1e4, 0, 1e4]
        f.createVariable("y", "f8", ("y",))[:] = [-1e4, 0, 1e4]
        f.createVariable("depth", "f8", ("x", "y"))[:] = 10_000 * np.ones((3, 3))
    with netcdf_file(tmp_path / "current.nc", "w") as f:
        f.createDimension("x", 3); f.createDimension("y", 3)
        f.createVariable("x", "f8", ("x",))[:] = [-1e8, 0, 1e8]
        f.createVariable("y", "f8", ("y",))[:] = [-1e8, 0, 1e8]
        f.createVariable("u", "f8", ("x", "y"))[:] = 0.01 * np.ones((3, 3))
        f.createVariable("v", "f8", ("x", "y"))[:] = 0.01 * np.ones((3, 3))
    ds = mr.single_ray(-1000, 0, 0.01, 0, 10, 2, tmp_path / "island.nc", tmp_path / "current.nc")
    assert ds.sizes["time_step"] == 6
    assert (ds.kx == 0.01).all()
    assert (ds.ky == 0.0).all()
    ds = mr.ray_tracing(3 * [-1000], 3 * [0], 3 * [0.01], 3 * [0], 10, 2,
                        str(tmp_path / "island.nc"), str(tmp_path / "current.nc"))
    assert ds.sizes["time_step"] == 6
    assert ds.sizes["ray"] == 3
    assert (ds.kx == 0.01).all()
    assert (ds.ky == 0.0).all()
