"""HIP path at full size and across its options, against the CPU oracle.

  - the seamount-jet workload at its full size (1024 x 1024 grid, 65,536
    rays, 5000 steps) on the GPU; every 64th ray traced by the oracle:
    identical row counts, end states within 1e-6 m / 1e-9 relative in k;
  - the golden fixtures (independent restatement) on the GPU;
  - MR_SOLOUT_NEXT_DERIV against the oracle's same variant;
  - host-memory and device-memory entry points, and every block size, give
    bit-identical results;
  - the analytic sources (ConstantDepth, ConstantSlope, ArrayDepth) in the
    right-hand side at random states.
"""
import os

import numpy as np
import pytest

from _helpers import oracle_fields

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    mantaray_amd._capi.lib()
    return mantaray_amd


@pytest.fixture(scope="module")
def seamount():
    from mantaray_amd import workloads
    return workloads.seamount_jet()


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_seamount_jet_full_size(mr, oracle_mod, seamount, math):
    from mantaray_amd import workloads
    cfg = seamount
    b, c = workloads.make_fields(cfg)
    got = mr.trace_many(b, c, cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"],
                        cfg["t_end"], cfg["dt"], math=math, trajectories=False)
    sub = slice(0, None, 64)
    ob, oc = oracle_fields(oracle_mod, cfg)
    ref = oracle_mod.trace_many(ob, oc, cfg["x0"][sub], cfg["y0"][sub], cfg["kx0"][sub],
                                cfg["ky0"][sub], cfg["t0"], cfg["t_end"], cfg["dt"],
                                trajectories=False)
    gl, rl = got["len"][sub], ref["len"]
    print(f"{math}: {int((gl != rl).sum())} length mismatches of {rl.size}; "
          f"ray-steps {int((got['len'] - 1).sum())}")
    assert np.array_equal(gl, rl)
    ge, re_ = got["end"][sub], ref["end"]
    assert np.array_equal(np.isnan(ge), np.isnan(re_))
    fin = ~np.isnan(re_[:, 0])
    dxy = np.abs(ge[fin, :2] - re_[fin, :2]).max()
    dk = (np.abs(ge[fin, 2:] - re_[fin, 2:]) / np.abs(re_[fin, 2:]).max()).max()
    print(f"{math}: max |d(x,y)| {dxy:.3e} m, max relative |dk| {dk:.3e}")
    assert dxy <= 1e-6 and dk <= 1e-9


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_golden_rhs_on_gpu(mr, math):
    r = np.load(os.path.join(GOLDEN, "rhs_cases.npz"))
    for ci in range(int(r["n_cases"])):
        g = {k[len(f"g{ci}_"):]: r[k] for k in r.files if k.startswith(f"g{ci}_")}
        b = mr.Bathymetry.from_arrays(g["bx"], g["by"], g["depth"])
        c = mr.Current.from_arrays(g["cx"], g["cy"], g["u"], g["v"])
        ds = mr.rhs(b, c, g["states"], math=math)
        ref = g["ds"]
        assert np.array_equal(np.isnan(ds), np.isnan(ref)), ci
        scale = np.nanmax(np.abs(ref), axis=0)
        err = np.nanmax(np.abs(ds - ref) / scale)
        assert err <= (1e-12 if math == "literal" else 1e-11), (ci, err)


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_golden_trajectories_on_gpu(mr, math):
    t = np.load(os.path.join(GOLDEN, "trajectories.npz"))
    b = mr.Bathymetry.from_arrays(t["g_bx"], t["g_by"], t["g_depth"])
    c = mr.Current.from_arrays(t["g_cx"], t["g_cy"], t["g_u"], t["g_v"])
    s0 = t["s0"]
    n, dt = int(t["n"]), float(t["dt"])
    r = mr.trace_many(b, c, s0[:, 0], s0[:, 1], s0[:, 2], s0[:, 3], 0.0, n * dt, dt, math=math)
    assert np.array_equal(r["len"], t["len"])
    got = np.stack([r[k] for k in ("x", "y", "kx", "ky")], axis=2)
    ref = t["rows"]
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.nanmax(np.abs(got[..., :2] - ref[..., :2])) <= 1e-6


def test_solout_next_deriv(mr, oracle_mod):
    """The other reading of ode_solvers' solout argument (f at the new
    state): same rows as the oracle's variant, on rays leaving the grid."""
    from _helpers import random_grid_case
    cfg = random_grid_case(5)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    rng = np.random.default_rng(5)
    n = 256
    x0 = rng.uniform(-300.0, 300.0, n)
    y0 = rng.uniform(200.0, 900.0, n)
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.01, 0.2, n)
    args = (x0, y0, k * np.cos(th), k * np.sin(th), 0.0, 400.0, 2.0)
    got = mr.trace_many(b, c, *args, solout="next_deriv")
    ref = oracle_mod.trace_many(ob, oc, *args, solout=oracle_mod.SOLOUT_NEXT_DERIV)
    assert np.array_equal(got["len"], ref["len"])
    for key in ("x", "y"):
        assert np.array_equal(np.isnan(got[key]), np.isnan(ref[key]))
        assert np.nanmax(np.abs(got[key] - ref[key])) <= 1e-6


def test_memory_modes_and_block_sizes_agree(mr):
    import torch
    from mantaray_amd import core, workloads
    cfg = workloads.seamount_jet(n_rays=3000, n_steps=300)
    b, c = workloads.make_fields(cfg)
    args = (cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"], cfg["t_end"], cfg["dt"])
    base = mr.trace_many(b, c, *args)
    for block in (64, 128, 192, 256):
        r = mr.trace_many(b, c, *args, block_size=block)
        for k in ("t", "len", "x", "y", "kx", "ky", "end"):
            assert np.array_equal(r[k], base[k], equal_nan=True), (block, k)
    dev = torch.device("cuda:0")
    ic = [torch.tensor(a, dtype=torch.float64, device=dev) for a in args[:4]]
    out = core.trace_many_device(b, c, *ic, *args[4:])
    torch.cuda.synchronize()
    for k in ("t", "len", "x", "y", "kx", "ky", "end"):
        assert np.array_equal(out[k].cpu().numpy(), base[k], equal_nan=True), k


@pytest.mark.parametrize("kind", ["constant", "slope", "array"])
@pytest.mark.parametrize("math", ["fast", "literal"])
def test_analytic_sources_rhs(mr, oracle_mod, kind, math):
    rng = np.random.default_rng(11)
    n = 2048
    s = np.stack([rng.uniform(-5, 60, n), rng.uniform(-5, 60, n),
                  rng.normal(0, 0.1, n), rng.normal(0, 0.1, n)], axis=1)
    s[:8, 2:] = 0.0
    s[8] = [np.nan, 1.0, 0.1, 0.1]
    if kind == "constant":
        b, ob = mr.Bathymetry.constant(35.0), oracle_mod.Bathymetry.constant(35.0)
    elif kind == "slope":
        b, ob = mr.Bathymetry.slope(), oracle_mod.Bathymetry.slope()
    else:
        a = rng.uniform(1.0, 80.0, (50, 50))
        b, ob = mr.Bathymetry.array(a), oracle_mod.Bathymetry.array(a)
    c, oc = mr.Current.constant(0.3, -0.2), oracle_mod.Current.constant(0.3, -0.2)
    ds = mr.rhs(b, c, s, math=math)
    ref, _ = oracle_mod.system_many(ob, oc, s)
    assert np.array_equal(np.isnan(ds), np.isnan(ref))
    scale = np.nanmax(np.abs(ref), axis=0)
    assert np.nanmax(np.abs(ds - ref) / scale) <= (1e-12 if math == "literal" else 1e-11)
