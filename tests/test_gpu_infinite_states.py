"""Rays that start at infinite positions, on grids with error cells.

The shared-cell trace tests a target against a record's fast region
through med3(p, lo, hi) - p per axis, and a cell where the reference errs (a
corner past a value array) has the empty region lo = hi = -inf.  An
infinite target makes a difference infinite or NaN and must take the full
queries, which err, giving the reference's all-NaN row.  A ray at x = -inf
guesses cell column 0 and one at x = +inf the last, so with y in the last
cell rows of a grid whose value arrays are cut short (every cell of those
rows errs) the record in hand has the empty region, whose bounds equal that
target; x = +inf and y = +-inf meet ordinary records too.
Row counts, NaN patterns and values against the oracle, for both math modes,
with trajectories (buffer-store kernel) and endpoints only.
"""
import numpy as np
import pytest

from _helpers import random_grid_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    from mantaray_amd import _capi
    _capi.lib()
    return mantaray_amd


def _states(cfg, n, rng):
    x = cfg["bathy"]["x"].astype(np.float64)
    y = cfg["bathy"]["y"].astype(np.float64)
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    k = rng.uniform(0.02, 0.2, n)
    s = np.stack([rng.uniform(x[0], x[-1], n), rng.uniform(y[0], y[-1], n),
                  k * np.cos(th), k * np.sin(th)], 1)
    q = n // 8
    s[:q, 0] = -np.inf                                     # x = -inf, y anywhere ...
    s[q:2 * q, 0] = -np.inf                                # ... and in the error rows
    s[q:2 * q, 1] = rng.uniform(y[-3], y[-1], q)
    s[2 * q:3 * q, 0] = np.inf
    s[3 * q:4 * q, 1] = -np.inf
    s[4 * q:5 * q, 1] = np.inf
    s[5 * q:5 * q + 8, :2] = -np.inf                       # both coordinates
    s[5 * q + 8:5 * q + 16, 0] = -np.inf                   # x = -inf, y NaN
    s[5 * q + 8:5 * q + 16, 1] = np.nan
    return s


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_infinite_positions_with_error_cells(mr, oracle_mod, math):
    import torch

    cfg = random_grid_case(17)
    nx, ny = cfg["bathy"]["x"].size, cfg["bathy"]["y"].size
    depth = cfg["bathy"]["depth"].ravel()[: nx * ny - nx - 3]
    u = cfg["current"]["u"].ravel()[: nx * ny - 5]
    v = cfg["current"]["v"].ravel()[: nx * ny - 2 * nx + 4]
    b = mr.Bathymetry.from_arrays(cfg["bathy"]["x"], cfg["bathy"]["y"], depth)
    c = mr.Current.from_arrays(cfg["current"]["x"], cfg["current"]["y"], u, v)
    ob = oracle_mod.Bathymetry.grid(cfg["bathy"]["x"], cfg["bathy"]["y"], depth)
    oc = oracle_mod.Current.grid(cfg["current"]["x"], cfg["current"]["y"], u, v)
    s = _states(cfg, 1024, np.random.default_rng(17))
    for n in (1, 5):
        got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, float(n), 1.0,
                            math=math)
        ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0,
                                    float(n), 1.0)
        assert np.array_equal(got["len"], ref["len"]), n
        inf_rows = ~np.isfinite(s[:, :2]).all(axis=1)
        assert (ref["len"][inf_rows] == 2).all()           # one all-NaN row, then the end
        for k in ("x", "y", "kx", "ky"):
            a, r = got[k], ref[k]
            assert np.array_equal(np.isnan(a), np.isnan(r)), (n, k)
            fin = np.isfinite(r)
            assert np.array_equal(a[~np.isnan(r) & ~fin], r[~np.isnan(r) & ~fin]), (n, k)
            scale = 1.0 if k in ("x", "y") else np.nanmax(np.abs(r[fin]))
            tol = 1e-8 if k in ("x", "y") else 1e-10
            assert np.max(np.abs(a[fin] - r[fin])) <= tol * scale, (n, k)
        ic = [torch.tensor(np.ascontiguousarray(s[:, q]), device="cuda:0") for q in range(4)]
        dev = mr.trace_many_device(b, c, *ic, 0.0, float(n), 1.0, math=math,
                                   trajectories=False)
        assert np.array_equal(dev["len"].cpu().numpy(), ref["len"]), n
        end = dev["end"].cpu().numpy()
        assert np.array_equal(np.isnan(end), np.isnan(ref["end"])), n
    print(f"{math}: {int(inf_rows.sum())} rays start at an infinite position")
