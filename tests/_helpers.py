"""Shared helpers for the tests: field pairs built identically for the HIP
product (mantaray_amd) and for the CPU oracle."""
from __future__ import annotations

import numpy as np


def oracle_fields(O, cfg):
    if "constant" in cfg["bathy"]:
        b = O.Bathymetry.constant(cfg["bathy"]["constant"])
    else:
        b = O.Bathymetry.grid(cfg["bathy"]["x"], cfg["bathy"]["y"], cfg["bathy"]["depth"])
    c = O.Current.grid(cfg["current"]["x"], cfg["current"]["y"], cfg["current"]["u"],
                       cfg["current"]["v"])
    return b, c


def random_grid_case(seed: int, nx: int = 21, ny: int = 17, dx: float = 37.5,
                     dy: float = 52.0, x0: float = -400.0, y0: float = 120.0):
    """Small gridded fields with smooth random depth and current."""
    rng = np.random.default_rng(seed)
    x = (x0 + dx * np.arange(nx)).astype(np.float32)
    y = (y0 + dy * np.arange(ny)).astype(np.float32)
    X, Y = np.meshgrid(x.astype(np.float64), y.astype(np.float64))
    depth = 30.0 + 20.0 * np.sin(X / 300.0 + rng.uniform(0, 6)) * np.cos(Y / 250.0) \
        + rng.uniform(0, 5, size=X.shape)
    u = 0.5 * np.sin(Y / 200.0 + rng.uniform(0, 6)) + rng.normal(0, 0.05, size=X.shape)
    v = 0.4 * np.cos(X / 180.0) + rng.normal(0, 0.05, size=X.shape)
    cfg = {
        "bathy": {"x": x, "y": y, "depth": depth},
        "current": {"x": x.astype(np.float64), "y": y.astype(np.float64), "u": u, "v": v},
    }
    return cfg


def random_states(seed: int, cfg, n: int, k_lo=0.005, k_hi=0.5, margin=0.1):
    """States spread over (and slightly beyond) the grid, plus exact nodes,
    cell edges and NaN / zero-k cases."""
    rng = np.random.default_rng(seed)
    x = cfg["bathy"]["x"].astype(np.float64)
    y = cfg["bathy"]["y"].astype(np.float64)
    lx, ly = x[-1] - x[0], y[-1] - y[0]
    xs = rng.uniform(x[0] - margin * lx, x[-1] + margin * lx, n)
    ys = rng.uniform(y[0] - margin * ly, y[-1] + margin * ly, n)
    k = np.exp(rng.uniform(np.log(k_lo), np.log(k_hi), n))
    th = rng.uniform(0, 2 * np.pi, n)
    s = np.stack([xs, ys, k * np.cos(th), k * np.sin(th)], axis=1)
    m = n // 8
    # exact nodes (corner coincidence) and exact cell lines
    s[:m, 0] = rng.choice(x, m)
    s[:m, 1] = rng.choice(y, m)
    s[m:2 * m, 0] = rng.choice(x, m)
    s[2 * m:3 * m, 1] = rng.choice(y, m)
    # last row / column (index == n-1)
    s[3 * m:3 * m + 4, 0] = x[-1]
    s[3 * m + 4:3 * m + 8, 1] = y[-1]
    # special values
    s[-1] = [np.nan, y[3], 0.1, 0.0]
    s[-2] = [x[3], np.nan, 0.1, 0.0]
    s[-3] = [x[3] + 1.0, y[3] + 1.0, 0.0, 0.0]
    s[-4] = [x[3] + 1.0, y[3] + 1.0, np.nan, 0.0]
    s[-5] = [x[3] + 1.0, y[3] + 1.0, 0.0, np.nan]
    return s


def compare_rhs(ds, ref, rtol):
    """NaN pattern identical, finite entries within rtol (relative to the
    larger of |ref| and the row's largest finite magnitude for that column)."""
    nan_a, nan_b = np.isnan(ds), np.isnan(ref)
    assert np.array_equal(nan_a, nan_b), f"NaN pattern differs at rows {np.nonzero((nan_a != nan_b).any(1))[0][:10]}"
    fin = ~nan_b
    scale = np.maximum(np.abs(ref), 1e-300)
    err = np.where(fin, np.abs(ds - ref) / scale, 0.0)
    # entries that are zero in the reference must be (near) zero here too
    return float(err.max()) if err.size else 0.0
