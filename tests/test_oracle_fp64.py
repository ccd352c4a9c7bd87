"""The oracle's fp64-interpolation mode (the algorithm of the differentiable
path) and its long-double finite-difference truth, pinned on the CPU.

fp64 mode queries:
  - on linear fields the bilinear value is the linear function and the
    cell-difference slopes are its slopes (closed form);
  - on random fields the value is scipy's RegularGridInterpolator (linear),
    an independent implementation of the same interpolant, and a node
    returns its value exactly;
  - the cell rule and the errors are those of the reference's f64 index
    (the current's statuses in both modes agree), the values differ from
    the reference's f32 interpolation by f32 rounding only, and the slopes
    (cell difference times RN(1/spacing)) are within an ulp of its quotients.
Long-double traces (or_fd_end_ld):
  - unperturbed, they end where the f64 oracle's traces end, to the f64
    trace's rounding (below 1e-8 m after 2000 steps);
  - their central differences reproduce closed-form derivatives of straight
    deep-water rays: d x(T)/d x0 = 1, d x(T)/d kx0 and d y(T)/d kx0 from
    cg = sqrt(g/k)/2, and d x(T)/d U = T for a uniform shift of a gridded
    current.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from _helpers import oracle_fields, random_grid_case

G = 9.8


def _bathy(O, x, y, d):
    return O.Bathymetry.grid(np.asarray(x, np.float32), np.asarray(y, np.float32), d)


def test_fp64_linear_fields(oracle_mod):
    O = oracle_mod
    x = (-300.0 + 25.0 * np.arange(30)).astype(np.float32)
    y = (100.0 + 40.0 * np.arange(20)).astype(np.float32)
    X, Y = np.meshgrid(x.astype(np.float64), y.astype(np.float64))
    a, bx, by = 37.25, 0.0123, -0.00731
    b = _bathy(O, x, y, a + bx * X + by * Y)
    cu = (0.3 - 0.001 * X + 0.002 * Y, -0.1 + 0.0005 * X - 0.0015 * Y)
    c = O.Current.grid(x.astype(np.float64), y.astype(np.float64), *cu)
    rng = np.random.default_rng(0)
    for _ in range(500):
        px = rng.uniform(x[0], x[-1])
        py = rng.uniform(y[0], y[-1])
        st, (h, hx, hy) = O.bathy_query64(b, px, py)
        assert st == 0
        assert abs(h - (a + bx * px + by * py)) <= 1e-12 * abs(h)
        assert abs(hx - bx) <= 1e-13 and abs(hy - by) <= 1e-13
        st, (u, v, ux, uy, vx, vy) = O.current_query64(c, px, py)
        assert st == 0
        assert abs(u - (0.3 - 0.001 * px + 0.002 * py)) <= 1e-13
        assert abs(v - (-0.1 + 0.0005 * px - 0.0015 * py)) <= 1e-13
        assert abs(ux + 0.001) <= 1e-15 and abs(uy - 0.002) <= 1e-15
        assert abs(vx - 0.0005) <= 1e-15 and abs(vy + 0.0015) <= 1e-15


@pytest.mark.parametrize("seed", [1, 2])
def test_fp64_matches_scipy_and_nodes(oracle_mod, seed):
    O = oracle_mod
    cfg = random_grid_case(seed)
    ob, oc = oracle_fields(O, cfg)
    x = cfg["bathy"]["x"].astype(np.float64)
    y = cfg["bathy"]["y"].astype(np.float64)
    depth = cfg["bathy"]["depth"]
    ip = RegularGridInterpolator((y, x), depth, method="linear")
    ipu = RegularGridInterpolator((y, x), cfg["current"]["u"], method="linear")
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(x[0], x[-1], 400), rng.uniform(y[0], y[-1], 400)], 1)
    for px, py in pts:
        st, (h, _, _) = O.bathy_query64(ob, px, py)
        assert st == 0 and abs(h - ip([[py, px]])[0]) <= 1e-12 * np.abs(depth).max()
        st, (u, *_r) = O.current_query64(oc, px, py)
        assert st == 0
        assert abs(u - ipu([[py, px]])[0]) <= 1e-12 * np.abs(cfg["current"]["u"]).max()
    for i in (0, 3, x.size - 1):
        for j in (0, 5, y.size - 1):
            st, (h, _, _) = O.bathy_query64(ob, x[i], y[j])
            assert st == 0 and h == depth[j, i]
            st, (u, v, *_r) = O.current_query64(oc, x[i], y[j])
            assert u == cfg["current"]["u"][j, i] and v == cfg["current"]["v"][j, i]


def test_fp64_cell_rule_and_errors_match_reference_mode(oracle_mod):
    O = oracle_mod
    cfg = random_grid_case(4)
    ob, oc = oracle_fields(O, cfg)
    x = cfg["current"]["x"]
    y = cfg["current"]["y"]
    rng = np.random.default_rng(4)
    lx, ly = x[-1] - x[0], y[-1] - y[0]
    pts = np.stack([rng.uniform(x[0] - 0.2 * lx, x[-1] + 0.2 * lx, 3000),
                    rng.uniform(y[0] - 0.2 * ly, y[-1] + 0.2 * ly, 3000)], 1)
    pts[:50, 0] = x[-1]
    pts[50:100, 1] = y[0]
    pts[100:150] = np.stack([rng.choice(x, 50), rng.choice(y, 50)], 1)
    pts[150] = [np.nan, y[2]]
    scale = np.abs(cfg["current"]["u"]).max()
    for px, py in pts:
        st64, v64 = O.current_query64(oc, px, py)
        st32, v32 = O.current_query(oc, px, py)
        assert st64 == st32, (px, py)
        if st64 == 0:
            assert abs(v64[0] - v32[0]) <= 1e-6 * scale       # f32 rounding of the point and arithmetic
            # slopes: the same f64 cell difference, times RN(1/spacing) here and
            # divided by the spacing in the reference mode (within an ulp)
            for a64, a32 in zip(v64[2:], v32[2:]):
                assert abs(a64 - a32) <= 2.3e-16 * abs(a32)


def _deep_case(O, h=2000.0):
    return O.Bathymetry.constant(h), O.Current.constant(0.0, 0.0)


def test_ld_traces_end_where_f64_traces_end(oracle_mod):
    O = oracle_mod
    cfg = random_grid_case(5)
    ob, oc = oracle_fields(O, cfg)
    rng = np.random.default_rng(5)
    n = 64
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.03, 0.2, n)
    s0 = np.stack([rng.uniform(-300, 300, n), rng.uniform(200, 900, n), k * np.cos(th),
                   k * np.sin(th)], 1)
    r = O.trace_many(ob, oc, *s0.T, 0.0, 80.0, 1.0, prec=O.PREC_FP64_INTERP,
                     trajectories=False)
    diff, mid, lp, lm = O.fd_end_ld(ob, oc, s0, 0.0, 80.0, 1.0)
    assert np.array_equal(lp, r["len"]) and np.array_equal(lm, r["len"])
    assert np.all((diff == 0) | np.isnan(diff))
    fin = ~np.isnan(r["end"][:, 0])
    assert fin.sum() > 10 and (~fin).sum() > 10
    assert np.abs(mid[fin] - r["end"][fin]).max(axis=0)[:2].max() <= 1e-8
    from mantaray_amd import workloads
    cfg = workloads.eddy_adjoint(n_rays=64)
    ob, oc = oracle_fields(O, cfg)
    s0 = np.stack([cfg[k] for k in ("x0", "y0", "kx0", "ky0")], 1)
    r = O.trace_many(ob, oc, *s0.T, 0.0, 2000.0, 1.0, prec=O.PREC_FP64_INTERP,
                     trajectories=False)
    _, mid, lp, _ = O.fd_end_ld(ob, oc, s0, 0.0, 2000.0, 1.0)
    assert np.array_equal(lp, r["len"])
    assert np.abs(mid[:, :2] - r["end"][:, :2]).max() <= 1e-8


@pytest.mark.parametrize("kind", ["slope", "array"])
def test_ld_traces_analytic_depths(oracle_mod, kind):
    """The long-double trace over ConstantSlope and ArrayDepth (the array
    truncates the point to indices, array_depth.rs:26-33, and a point off
    the array has NaN depth) ends where the f64 oracle's fp64 mode ends: the
    same row counts, and the end states within the f64 trace's rounding."""
    O = oracle_mod
    rng = np.random.default_rng(41)
    if kind == "slope":
        ob = O.Bathymetry.slope(h0=15.0, x0=30.0, y0=30.0, dhdx=0.08, dhdy=-0.05)
    else:
        ob = O.Bathymetry.array(rng.uniform(10.0, 20.0, (60, 60)))
    ax = np.linspace(0.0, 59.0, 12)
    X, Y = np.meshgrid(ax, ax)
    oc = O.Current.grid(ax, ax, 0.3 * np.sin(Y / 9.0), 0.2 * np.cos(X / 7.0))
    n = 64
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    k = rng.uniform(0.2, 1.0, n)
    s0 = np.stack([rng.uniform(2.0, 57.0, n), rng.uniform(2.0, 57.0, n), k * np.cos(th),
                   k * np.sin(th)], 1)
    r = O.trace_many(ob, oc, *s0.T, 0.0, 12.0, 0.1, prec=O.PREC_FP64_INTERP,
                     trajectories=False)
    _, mid, lp, _ = O.fd_end_ld(ob, oc, s0, 0.0, 12.0, 0.1)
    assert np.array_equal(lp, r["len"])
    fin = ~np.isnan(r["end"][:, 0])
    assert fin.sum() >= 32 and (kind == "slope" or (~fin).sum() >= 1)
    assert np.array_equal(np.isnan(mid), np.isnan(r["end"]))
    assert np.abs(mid[fin, :2] - r["end"][fin, :2]).max() <= 1e-10


def test_ld_central_differences_closed_form(oracle_mod):
    """Straight deep-water rays (h = 2000 m, k >= 0.03: tanh(kh) = 1 in
    long double): x(T) = x0 + T cg kx/|k| with cg = sqrt(g/k)/2."""
    O = oracle_mod
    b, c = _deep_case(O)
    rng = np.random.default_rng(8)
    n = 32
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.03, 0.2, n)
    s0 = np.stack([rng.uniform(-1e4, 1e4, n), rng.uniform(-1e4, 1e4, n), k * np.cos(th),
                   k * np.sin(th)], 1)
    T = 500.0
    ex, ey = np.cos(th), np.sin(th)
    cg = 0.5 * np.sqrt(G / k)
    dcg = -0.25 * np.sqrt(G) * k ** -1.5
    for comp, want in ((0, (np.ones(n), np.zeros(n))),
                       (2, (T * (dcg * ex * ex + cg * ey * ey / k),
                            T * (dcg - cg / k) * ex * ey))):
        ds0 = np.zeros_like(s0)
        ds0[:, comp] = 1.0
        eps = 1e-6 if comp == 0 else 1e-9
        diff, _, _, _ = O.fd_end_ld(b, c, s0, 0.0, T, 1.0, eps=eps, ds0=ds0)
        got = diff[:, :2] / (2 * eps)
        for q in range(2):
            err = np.abs(got[:, q] - want[q]).max() / max(np.abs(want[q]).max(), 1.0)
            assert err <= 1e-7, (comp, q, err)


def test_ld_value_perturbation_closed_form(oracle_mod):
    """A uniform shift of a gridded current's u moves x(T) by T per unit."""
    O = oracle_mod
    x = np.arange(0.0, 50001.0, 500.0)
    X, Y = np.meshgrid(x, x)
    u = np.full(X.shape, 0.4)
    v = np.full(X.shape, -0.2)
    b = O.Bathymetry.constant(2000.0)
    c = O.Current.grid(x, x, u, v)
    s0 = np.array([[20000.0, 25000.0, 0.05, 0.02], [30000.0, 20000.0, -0.04, 0.06]])
    T = 300.0
    for name, q in (("du", 0), ("dv", 1)):
        diff, _, _, _ = O.fd_end_ld(b, c, s0, 0.0, T, 1.0, eps=1e-6,
                                    **{name: np.ones(u.size)})
        d = diff[:, q] / 2e-6
        assert np.abs(d - T).max() <= 1e-6 * T, (name, d)
        assert np.abs(diff[:, 1 - q]).max() <= 1e-15 * T
