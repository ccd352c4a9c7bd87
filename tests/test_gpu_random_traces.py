"""Randomized traces against the oracle.

Twelve small scenes drawn from a seed: grid sizes 6-40 nodes a side, random
spacing and origin, axes uniform or jittered by up to a fifth of a cell (the
reference's index rule uses the first spacing, (p - a0)/|a1 - a0|, and the
interpolation the true corner coordinates; oracle and product both follow
it), rough depth that crosses zero in places (h <= 0 ends a ray), rough
current, and a step of 0.3, 0.5, 0.75, 1 or 1.7 s.  Depth and current share
their cells in most scenes (the shared-cell kernel, with dt a power of two
or not) and are offset by a fraction of a cell in the rest (separate
records).  The states include exact nodes, cell lines, the last row and
column, points off the grid, NaN and zero k.  Each scene runs in both math
modes and both solout readings, with trajectories and endpoints only:
identical row counts and NaN patterns, positions within 1e-6 m,
wavenumbers within 1e-8 of their scale, and endpoints equal to the last
rows.  The same scenes also run together in one batched launch, and in the
fp64-interpolation mode the adjoint differentiates.
"""
import numpy as np
import pytest

from _helpers import oracle_fields, random_states

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    from mantaray_amd import _capi
    _capi.lib()
    return mantaray_amd


def _scene(seed):
    rng = np.random.default_rng(7000 + seed)
    nx, ny = (int(v) for v in rng.integers(6, 41, size=2))
    dx, dy = rng.uniform(20.0, 80.0, size=2)
    x0, y0 = rng.uniform(-1000.0, 1000.0, size=2)
    x = x0 + dx * np.arange(nx)
    y = y0 + dy * np.arange(ny)
    if seed % 2:
        x[1:] += rng.uniform(-0.2, 0.2, nx - 1) * dx
        y[1:] += rng.uniform(-0.2, 0.2, ny - 1) * dy
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    X, Y = np.meshgrid(x32.astype(np.float64), y32.astype(np.float64))
    depth = (rng.uniform(5.0, 40.0) + rng.uniform(5.0, 30.0)
             * np.sin(X / rng.uniform(60.0, 300.0) + rng.uniform(0, 6))
             * np.cos(Y / rng.uniform(60.0, 300.0)) + rng.normal(0.0, 2.0, X.shape))
    u = rng.uniform(-0.8, 0.8) * np.sin(Y / rng.uniform(80.0, 300.0)) + rng.normal(0, 0.05, X.shape)
    v = rng.uniform(-0.8, 0.8) * np.cos(X / rng.uniform(80.0, 300.0)) + rng.normal(0, 0.05, X.shape)
    cx, cy = x32.astype(np.float64), y32.astype(np.float64)
    if seed % 4 == 3:                                    # separate cells
        cx = cx + 0.31 * dx
        cy = cy - 0.17 * dy
    cfg = {"bathy": {"x": x32, "y": y32, "depth": depth},
           "current": {"x": cx, "y": cy, "u": u, "v": v}}
    dt = float(rng.choice([0.3, 0.5, 0.75, 1.0, 1.7]))
    steps = int(rng.integers(20, 61))
    return cfg, dt, steps


@pytest.mark.parametrize("seed", range(12))
def test_random_scene(mr, oracle_mod, seed):
    import torch

    cfg, dt, steps = _scene(seed)
    b = mr.Bathymetry.from_arrays(**cfg["bathy"])
    c = mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 512)
    t_end = steps * dt
    ended = 0
    for math in ("fast", "literal"):
        for solout in ("state", "next_deriv"):
            osol = oracle_mod.SOLOUT_STATE if solout == "state" else oracle_mod.SOLOUT_NEXT_DERIV
            ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, t_end,
                                        dt, solout=osol)
            got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, t_end, dt,
                                math=math, solout=solout)
            tag = (seed, math, solout)
            assert np.array_equal(got["len"], ref["len"]), tag
            for k in ("x", "y", "kx", "ky"):
                assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), (tag, k)
            dxy = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("x", "y"))
            kscale = np.nanmax(np.abs(ref["kx"])) + np.nanmax(np.abs(ref["ky"]))
            dk = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("kx", "ky")) / kscale
            assert dxy <= 1e-6 and dk <= 1e-8, (tag, dxy, dk)
            ic = [torch.tensor(np.ascontiguousarray(s[:, q]), device="cuda:0") for q in range(4)]
            dev = mr.trace_many_device(b, c, *ic, 0.0, t_end, dt, math=math, solout=solout,
                                       trajectories=False)
            assert np.array_equal(dev["len"].cpu().numpy(), got["len"]), tag
            end = dev["end"].cpu().numpy()
            assert np.array_equal(np.isnan(end), np.isnan(got["end"])), tag
            fin = ~np.isnan(end)
            assert np.array_equal(end[fin], got["end"][fin]), tag
            ended = max(ended, int((ref["len"] < steps + 1).sum()))
    print(f"seed {seed}: {cfg['bathy']['x'].size} x {cfg['bathy']['y'].size} nodes, dt {dt}, "
          f"{steps} steps, {ended} of 512 rays end early")


def test_random_scenes_batched(mr):
    """All twelve scenes in one persistent launch: every ray's row count and
    last row equal bitwise those of its scene's own trace (some scenes share
    cells and some do not, so the batch runs on separate records while most
    single traces use the shared-cell records)."""
    import torch
    from mantaray_amd.batch import Batch
    scenes = []
    for seed in range(12):
        cfg, dt, steps = _scene(seed)
        s = random_states(seed, cfg, 512)
        scenes.append({"bathymetry": mr.Bathymetry.from_arrays(**cfg["bathy"]),
                       "current": mr.Current.from_arrays(**cfg["current"]),
                       "x0": s[:, 0], "y0": s[:, 1], "kx0": s[:, 2], "ky0": s[:, 3],
                       "t0": 0.0, "t_end": steps * dt, "dt": dt})
    batch = Batch(scenes)
    batch.run()
    torch.cuda.synchronize()
    for i, (sc, out) in enumerate(zip(scenes, batch.results())):
        ref = mr.trace_many(sc["bathymetry"], sc["current"], sc["x0"], sc["y0"], sc["kx0"],
                            sc["ky0"], sc["t0"], sc["t_end"], sc["dt"], trajectories=False)
        assert np.array_equal(out["len"].cpu().numpy(), ref["len"]), i
        end = out["end"].cpu().numpy()
        assert np.array_equal(np.isnan(end), np.isnan(ref["end"])), i
        fin = ~np.isnan(end)
        assert np.array_equal(end[fin], ref["end"][fin]), i
    batch.close()


@pytest.mark.parametrize("seed", range(0, 12, 3))
def test_random_scene_fp64_interpolation(mr, oracle_mod, seed):
    """The fp64-interpolation trace (the path the adjoint differentiates) on
    the same scenes against the oracle in that mode: identical row counts and
    NaN patterns, positions within 1e-7 m."""
    cfg, dt, steps = _scene(seed)
    b = mr.Bathymetry.from_arrays(**cfg["bathy"])
    c = mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 512)
    args = (s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, steps * dt, dt)
    got = mr.trace_many(b, c, *args, precision="fp64")
    ref = oracle_mod.trace_many(ob, oc, *args, prec=oracle_mod.PREC_FP64_INTERP)
    assert np.array_equal(got["len"], ref["len"])
    for k in ("x", "y", "kx", "ky"):
        assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), k
    d = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("x", "y"))
    print(f"seed {seed} fp64: max |d(x, y)| {d:.2e} m")
    assert d <= 1e-7


@pytest.mark.parametrize("kind", ["constant", "slope", "array"])
@pytest.mark.parametrize("current", ["grid", "constant"])
def test_random_analytic_sources(mr, oracle_mod, kind, current):
    """Analytic depths (ConstantDepth, ConstantSlope with random parameters,
    a random ArrayDepth) with a scene's gridded current (jittered axes) or a
    constant one, traced in both math modes against the oracle: identical row
    counts and NaN patterns, positions within 1e-6 m."""
    seed = {"constant": 1, "slope": 5, "array": 9}[kind]
    cfg, dt, steps = _scene(seed)
    rng = np.random.default_rng(300 + seed)
    cx, cy = cfg["current"]["x"], cfg["current"]["y"]
    if kind == "constant":
        h = float(rng.uniform(3.0, 60.0))
        b, ob = mr.Bathymetry.constant(h), oracle_mod.Bathymetry.constant(h)
    elif kind == "slope":
        p = dict(h0=float(rng.uniform(10.0, 60.0)), x0=float(cx[len(cx) // 2]),
                 y0=float(cy[len(cy) // 2]), dhdx=float(rng.uniform(-0.1, 0.1)),
                 dhdy=float(rng.uniform(-0.1, 0.1)))
        b, ob = mr.Bathymetry.slope(**p), oracle_mod.Bathymetry.slope(**p)
    else:
        a = rng.uniform(-2.0, 60.0, (60, 60))
        b, ob = mr.Bathymetry.array(a), oracle_mod.Bathymetry.array(a)
    if current == "grid":
        if kind == "array":                  # ArrayDepth's domain is [0, 59]^2
            cx = (cx - cx[0]) * (59.0 / (cx[-1] - cx[0]))
            cy = (cy - cy[0]) * (59.0 / (cy[-1] - cy[0]))
        c = mr.Current.from_arrays(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
        oc = oracle_mod.Current.grid(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
        lo_x, hi_x, lo_y, hi_y = cx[0], cx[-1], cy[0], cy[-1]
    else:
        u, v = (float(w) for w in rng.uniform(-0.5, 0.5, 2))
        c, oc = mr.Current.constant(u, v), oracle_mod.Current.constant(u, v)
        lo_x, hi_x, lo_y, hi_y = (0.0, 59.0, 0.0, 59.0) if kind == "array" else \
            (cx[0], cx[-1], cy[0], cy[-1])
    n = 512
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    k = np.exp(rng.uniform(np.log(0.01), np.log(0.5), n))
    s = np.stack([rng.uniform(lo_x, hi_x, n), rng.uniform(lo_y, hi_y, n),
                  k * np.cos(th), k * np.sin(th)], 1)
    s[:8, 2:] = 0.0
    s[8] = [np.nan, lo_y, 0.1, 0.1]
    for math in ("fast", "literal"):
        got = mr.trace_many(b, c, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, steps * dt, dt,
                            math=math)
        ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0,
                                    steps * dt, dt)
        assert np.array_equal(got["len"], ref["len"]), math
        for key in ("x", "y", "kx", "ky"):
            assert np.array_equal(np.isnan(got[key]), np.isnan(ref[key])), (math, key)
        d = max(np.nanmax(np.abs(got[key] - ref[key])) for key in ("x", "y"))
        assert d <= 1e-6, (math, d)


@pytest.mark.parametrize("seed", [1, 3, 5, 7])
@pytest.mark.parametrize("math", ["fast", "literal"])
def test_random_scene_rhs(mr, oracle_mod, seed, math):
    """The right-hand side (the full field queries, no cell records carried
    between evaluations) at the scenes' special states on jittered axes,
    against the oracle: identical NaN patterns, values within the tolerance
    of the uniform-grid RHS tests."""
    cfg, dt, steps = _scene(seed)
    b = mr.Bathymetry.from_arrays(**cfg["bathy"])
    c = mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 4096)
    ds = mr.rhs(b, c, s, math=math)
    ref, st = oracle_mod.system_many(ob, oc, s)
    assert np.array_equal(np.isnan(ds), np.isnan(ref))
    scale = np.nanmax(np.abs(ref), axis=0)
    assert np.nanmax(np.abs(ds - ref) / scale) <= (1e-12 if math == "literal" else 1e-11)


@pytest.mark.parametrize("seed", [1, 3])
def test_random_scene_gradients(mr, oracle_mod, seed):
    """Gradients on jittered axes (shared cells for seed 1, separate for 3):
    the 4 x 4 Jacobian of each of 64 rays' end state with respect to its
    initial state, by four reverse sweeps with unit cotangents, and the
    derivative of sum_r x_r(T) along a random direction of u, against
    central differences of the long-double trace (eps 1e-6 m for positions,
    1e-12 for wavenumbers, 1e-6 for u).  Rays whose two perturbed traces met
    different cells (where the cell slopes jump) or that end on an all-NaN
    row (which passes no gradient) are left out; per ray and
    end component the Jacobian row agrees within 1e-6 of its largest entry,
    and the directional derivative within 1e-6 of its terms."""
    import torch

    from mantaray_amd import adjoint
    cfg, dt, steps = _scene(seed)
    b = mr.Bathymetry.from_arrays(**cfg["bathy"])
    c = mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    rng = np.random.default_rng(500 + seed)
    x, y = cfg["bathy"]["x"].astype(np.float64), cfg["bathy"]["y"].astype(np.float64)
    n = 64
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    k = rng.uniform(0.03, 0.2, n)
    s0 = np.stack([rng.uniform(x[1], x[-2], n), rng.uniform(y[1], y[-2], n),
                   k * np.cos(th), k * np.sin(th)], 1)
    T = steps * dt
    dev = lambda a: torch.tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda:0")  # noqa: E731
    ic = [dev(s0[:, q]) for q in range(4)]
    u, v = dev(cfg["current"]["u"].ravel()), dev(cfg["current"]["v"].ravel())
    stride = 10
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, stride=stride, u=u, v=v)
    J = np.zeros((n, 4, 4))
    gu = None
    for q in range(4):
        ge = torch.zeros((n, 4), dtype=torch.float64, device="cuda:0")
        ge[:, q] = 1.0
        g = adjoint.adjoint(b, c, fwd, 0.0, T, dt, stride, g_end=ge, u=u, v=v,
                            want=("x0", "y0", "kx0", "ky0", "u"))
        for j, name in enumerate(("x0", "y0", "kx0", "ky0")):
            J[:, q, j] = g[name].cpu().numpy()
        if q == 0:
            gu = g["u"].cpu().numpy()
    scale = np.array([1e-6, 1e-6, 1e-12, 1e-12])
    Jfd = np.zeros_like(J)
    same = np.ones(n, dtype=bool)
    for j in range(4):
        ds0 = np.zeros_like(s0)
        ds0[:, j] = 1.0
        diff, _, lp, lm, sc = oracle_mod.fd_end_ld(ob, oc, s0, 0.0, T, dt, eps=scale[j],
                                                   ds0=ds0, with_cells=True)
        Jfd[:, :, j] = diff / (2 * scale[j])
        # both perturbed traces in the same cells, and both with the full
        # row count and a finite last row (a ray whose last row is all-NaN
        # passes no gradient through it)
        same &= sc & (lp == lm) & (lp == steps + 1) & np.isfinite(diff).all(axis=1)
    assert same.sum() >= 16, int(same.sum())
    A, B = J * scale, Jfd * scale
    err = np.abs(A - B).max(axis=2) / np.abs(B).max(axis=2)
    print(f"seed {seed}: {int(same.sum())} of {n} rays usable; per-ray Jacobian rows "
          f"max error {err[same].max():.2e}")
    assert err[same].max() <= 1e-6
    # d(sum of x_r(T) over the usable rays)/du along a random direction
    d = rng.normal(0.0, 1.0, cfg["current"]["u"].size)
    diff, _, lp, lm, sc = oracle_mod.fd_end_ld(ob, oc, s0[same], 0.0, T, dt, eps=1e-6, du=d,
                                               with_cells=True)
    ok = sc & (lp == lm)
    fd = float(np.sum(diff[ok, 0])) / 2e-6
    ge = torch.zeros((n, 4), dtype=torch.float64, device="cuda:0")
    idx = np.nonzero(same)[0][ok]
    ge[torch.tensor(idx, device="cuda:0"), 0] = 1.0
    g = adjoint.adjoint(b, c, fwd, 0.0, T, dt, stride, g_end=ge, u=u, v=v, want=("u",))
    ad = float(np.dot(g["u"].cpu().numpy(), d))
    terms = float(np.sum(np.abs(g["u"].cpu().numpy() * d)))
    print(f"seed {seed}: d(sum x)/du along a random direction: adjoint {ad:.9e}, "
          f"long-double FD {fd:.9e} over {int(ok.sum())} rays")
    assert abs(ad - fd) <= 1e-6 * terms


@pytest.mark.parametrize("kind", ["constant", "slope", "array"])
def test_analytic_depth_gradients(mr, oracle_mod, kind):
    """Gradients over analytic depths (ConstantDepth and ArrayDepth, whose
    slopes are zero, take the sweeps' flat-depth arithmetic; ConstantSlope the
    general one) with a scene's gridded current on axes scaled into
    ArrayDepth's [0, 59] m domain: each of 32 rays' 4 x 4 Jacobian of its end
    state with respect to its initial state, against central differences of
    the long-double trace, as in the scene test above, on every sweep kernel
    (block sizes 64, 128 and 256)."""
    import torch

    from mantaray_amd import adjoint
    cfg, _, _ = _scene(5)
    rng = np.random.default_rng(900 + ["constant", "slope", "array"].index(kind))
    cx, cy = cfg["current"]["x"], cfg["current"]["y"]
    cx = (cx - cx[0]) * (59.0 / (cx[-1] - cx[0]))
    cy = (cy - cy[0]) * (59.0 / (cy[-1] - cy[0]))
    c = mr.Current.from_arrays(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
    oc = oracle_mod.Current.grid(cx, cy, cfg["current"]["u"], cfg["current"]["v"])
    if kind == "constant":
        b, ob = mr.Bathymetry.constant(12.0), oracle_mod.Bathymetry.constant(12.0)
    elif kind == "slope":
        p = dict(h0=15.0, x0=30.0, y0=30.0, dhdx=0.08, dhdy=-0.05)
        b, ob = mr.Bathymetry.slope(**p), oracle_mod.Bathymetry.slope(**p)
    else:
        a = rng.uniform(10.0, 20.0, (60, 60))
        b, ob = mr.Bathymetry.array(a), oracle_mod.Bathymetry.array(a)
    n, dt, steps, stride = 32, 0.1, 60, 10
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    k = rng.uniform(0.2, 1.0, n)
    s0 = np.stack([rng.uniform(20.0, 40.0, n), rng.uniform(20.0, 40.0, n),
                   k * np.cos(th), k * np.sin(th)], 1)
    T = steps * dt
    dev = lambda x: torch.tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda:0")  # noqa: E731
    ic = [dev(s0[:, q]) for q in range(4)]
    u, v = dev(cfg["current"]["u"].ravel()), dev(cfg["current"]["v"].ravel())
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, stride=stride, u=u, v=v)
    scale = np.array([1e-6, 1e-6, 1e-12, 1e-12])
    Jfd = np.zeros((n, 4, 4))
    same = np.ones(n, dtype=bool)
    for j in range(4):
        ds0 = np.zeros_like(s0)
        ds0[:, j] = 1.0
        diff, _, lp, lm, sc = oracle_mod.fd_end_ld(ob, oc, s0, 0.0, T, dt, eps=scale[j],
                                                   ds0=ds0, with_cells=True)
        Jfd[:, :, j] = diff / (2 * scale[j])
        same &= sc & (lp == lm) & (lp == steps + 1) & np.isfinite(diff).all(axis=1)
    assert same.sum() >= 24, int(same.sum())
    for block in (64, 128, 256):
        J = np.zeros((n, 4, 4))
        for q in range(4):
            ge = torch.zeros((n, 4), dtype=torch.float64, device="cuda:0")
            ge[:, q] = 1.0
            g = adjoint.adjoint(b, c, fwd, 0.0, T, dt, stride, g_end=ge, u=u, v=v,
                                want=("x0", "y0", "kx0", "ky0"), block_size=block)
            for j, name in enumerate(("x0", "y0", "kx0", "ky0")):
                J[:, q, j] = g[name].cpu().numpy()
        A, B = J * scale, Jfd * scale
        err = np.abs(A - B).max(axis=2) / np.abs(B).max(axis=2)
        print(f"{kind}, block {block}: {int(same.sum())} of {n} rays usable; per-ray "
              f"Jacobian rows max error {err[same].max():.2e}")
        assert err[same].max() <= 1e-6, (kind, block)

