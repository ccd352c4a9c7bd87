"""The reference's single-ray and many-ray behaviour tests (src/ray.rs,
src/tests/) on the CPU oracle; tests/ray_cases.py holds the cases."""
import numpy as np
import pytest

import ray_cases as RC


def oracle_fields(O, bspec, cspec):
    if bspec[0] == "constant":
        b = O.Bathymetry.constant(bspec[1])
    elif bspec[0] == "slope":
        b = O.Bathymetry.slope(*bspec[1:]) if len(bspec) > 1 else O.Bathymetry.slope()
    else:
        b = O.Bathymetry.grid(*RC.bathy_spec_arrays(bspec))
    if cspec[0] == "constant":
        c = O.Current.constant(cspec[1], cspec[2])
    else:
        c = O.Current.grid(*RC.current_spec_arrays(cspec))
    return b, c


@pytest.mark.parametrize("case", RC.CASES, ids=[c[0] for c in RC.CASES])
def test_single_ray_case(oracle_mod, case):
    name, bspec, cspec, s0, (t0, t1, dt), checks, _ = case
    b, c = oracle_fields(oracle_mod, bspec, cspec)
    rows = oracle_mod.trace_single(b, c, list(s0), t0, t1, dt)
    RC.check_rows(rows[:, 1:], checks)


def test_slope_depths(oracle_mod):
    """ray.rs:562-563: the default ConstantSlope is 49.5 m at x = 10 and 35 m
    at x = 300 (y = 1000)."""
    b = oracle_mod.Bathymetry.slope()
    assert oracle_mod.bathy_query(b, 10.0, 1000.0)[1][0] == np.float32(49.5)
    assert oracle_mod.bathy_query(b, 300.0, 1000.0)[1][0] == np.float32(35.0)


@pytest.mark.parametrize("case", RC.MANY_CASES, ids=[c[0] for c in RC.MANY_CASES])
def test_many_ray_case(oracle_mod, case):
    name, bspec, cspec, rays, checks, (t0, t1, dt), _ = case
    b, c = oracle_fields(oracle_mod, bspec, cspec)
    r = np.array(rays)
    res = oracle_mod.trace_many(b, c, r[:, 0], r[:, 1], r[:, 2], r[:, 3], t0, t1, dt)
    for i, ch in enumerate(checks):
        rows = np.stack([res[k][: res["len"][i], i] for k in ("x", "y", "kx", "ky")], axis=1)
        RC.check_helper(rows, ch)


def test_many_waves_ok(oracle_mod):
    """ray.rs:1331-1361: nine rays on the default slope for 1e5 s all return."""
    O = oracle_mod
    y0 = np.arange(10.0, 100.0, 10.0)
    res = O.trace_many(O.Bathymetry.slope(), O.Current.constant(), np.full(9, 10.0), y0,
                       np.ones(9), np.zeros(9), 0.0, 100000.0, 1.0, trajectories=False)
    assert res["len"].shape == (9,) and (res["len"] >= 1).all()
