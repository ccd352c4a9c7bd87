"""The reference's single-ray and many-ray behaviour tests (src/ray.rs,
src/tests/) on the HIP path, with the same cases and checks as the oracle
runs them (tests/ray_cases.py), and each trajectory compared with the
oracle's: rows within 1e-9 relative per component, identical row counts."""
import numpy as np
import pytest

import ray_cases as RC
from test_ray_cases import oracle_fields

pytestmark = pytest.mark.gpu


def hip_fields(bspec, cspec):
    import mantaray_amd as mr
    if bspec[0] == "constant":
        b = mr.Bathymetry.constant(bspec[1])
    elif bspec[0] == "slope":
        b = mr.Bathymetry.slope(*bspec[1:]) if len(bspec) > 1 else mr.Bathymetry.slope()
    else:
        b = mr.Bathymetry.from_arrays(*RC.bathy_spec_arrays(bspec))
    if cspec[0] == "constant":
        c = mr.Current.constant(cspec[1], cspec[2])
    else:
        c = mr.Current.from_arrays(*RC.current_spec_arrays(cspec))
    return b, c


def close_rows(got, ref):
    assert got.shape == ref.shape
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    scale = np.maximum(np.nanmax(np.abs(ref), axis=0), 1e-300)
    err = np.nanmax(np.abs(got - ref) / scale) if np.isfinite(ref).any() else 0.0
    assert err <= 1e-9, err


@pytest.mark.parametrize("math", ["fast", "literal"])
@pytest.mark.parametrize("case", RC.CASES, ids=[c[0] for c in RC.CASES])
def test_single_ray_case_gpu(oracle_mod, case, math):
    import mantaray_amd as mr
    name, bspec, cspec, s0, (t0, t1, dt), checks, _ = case
    b, c = hip_fields(bspec, cspec)
    r = mr.trace_many(b, c, [s0[0]], [s0[1]], [s0[2]], [s0[3]], t0, t1, dt, math=math)
    n = int(r["len"][0])
    rows = np.stack([r[k][:n, 0] for k in ("x", "y", "kx", "ky")], axis=1)
    RC.check_rows(rows, checks)
    ob, oc = oracle_fields(oracle_mod, bspec, cspec)
    ref = oracle_mod.trace_single(ob, oc, list(s0), t0, t1, dt)
    close_rows(rows, ref[:, 1:])


@pytest.mark.parametrize("math", ["fast", "literal"])
@pytest.mark.parametrize("case", RC.MANY_CASES, ids=[c[0] for c in RC.MANY_CASES])
def test_many_ray_case_gpu(oracle_mod, case, math):
    import mantaray_amd as mr
    name, bspec, cspec, rays, checks, (t0, t1, dt), _ = case
    b, c = hip_fields(bspec, cspec)
    a = np.array(rays)
    res = mr.trace_many(b, c, a[:, 0], a[:, 1], a[:, 2], a[:, 3], t0, t1, dt, math=math)
    ob, oc = oracle_fields(oracle_mod, bspec, cspec)
    ref = oracle_mod.trace_many(ob, oc, a[:, 0], a[:, 1], a[:, 2], a[:, 3], t0, t1, dt)
    assert np.array_equal(res["len"], ref["len"])
    for i, ch in enumerate(checks):
        n = int(res["len"][i])
        rows = np.stack([res[k][:n, i] for k in ("x", "y", "kx", "ky")], axis=1)
        RC.check_helper(rows, ch)
        close_rows(rows, np.stack([ref[k][:n, i] for k in ("x", "y", "kx", "ky")], axis=1))
