"""The fp64-interpolation mode and its discrete adjoint on the GPU.

  - mr_rhs_many and mr_trace_many in MR_PREC_FP64_INTERP against the oracle's
    fp64-interp mode (same algorithm, f64 throughout): right-hand sides at
    random states (grid nodes, cell lines, out of range, NaN) within 1e-12
    relative, trajectories within 1e-9 m with identical row counts;
  - checkpoints are bitwise the trajectory rows they stand for;
  - mr_adjoint_many against central finite differences of the same forward:
    one step (tight), many steps on small gridded depth and current (initial
    states per ray, directional derivatives of the depth, u and v grids,
    single grid values), a trajectory (all-rows) loss, rays that leave the
    grid, and the eddy workload (512 x 512 grid, 16,384 rays, 2000 steps;
    finite differences on 256 rays) with the loss sum |x(T) - target|^2;
  - gradients with respect to the initial states do not depend on the
    checkpoint stride (bitwise), and the torch.autograd.Function returns the
    same gradients as the direct call.

Finite differences of a 2000-step f64 trace carry its rounding noise (about
1e-10 m in the endpoints), so with eps = 1e-6 m they are good to about 1e-4
relative only; the eddy check therefore takes its truth from the same
algorithm traced in long double (the oracle's or_fd_end_ld, 2^11 times less
noise).  Steps stay small enough that no perturbed stage point crosses a
cell line, where the cell-difference slopes jump.
"""
import numpy as np
import pytest

from _helpers import oracle_fields, random_grid_case, random_states

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    mantaray_amd._capi.lib()
    return mantaray_amd


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    return _t


def _dev(torch, a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device="cuda:0")


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("math", ["fast", "literal"])
def test_fp64_rhs_matches_oracle(mr, oracle_mod, seed, math):
    cfg = random_grid_case(seed)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = random_states(seed, cfg, 4096)
    ds = mr.rhs(b, c, s, math=math, precision="fp64")
    ref, _ = oracle_mod.system_many(ob, oc, s, prec=oracle_mod.PREC_FP64_INTERP)
    assert np.array_equal(np.isnan(ds), np.isnan(ref))
    scale = np.nanmax(np.abs(ref), axis=0)
    err = np.nanmax(np.abs(ds - ref) / scale)
    assert err <= (1e-13 if math == "literal" else 1e-12), err


@pytest.mark.parametrize("math", ["fast", "literal"])
def test_fp64_trace_matches_oracle(mr, oracle_mod, math):
    cfg = random_grid_case(5)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    rng = np.random.default_rng(5)
    n = 512
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.01, 0.2, n)
    args = (rng.uniform(-300.0, 300.0, n), rng.uniform(200.0, 900.0, n), k * np.cos(th),
            k * np.sin(th), 0.0, 400.0, 2.0)
    got = mr.trace_many(b, c, *args, math=math, precision="fp64")
    ref = oracle_mod.trace_many(ob, oc, *args, prec=oracle_mod.PREC_FP64_INTERP)
    assert np.array_equal(got["len"], ref["len"])
    assert (ref["len"] < 201).sum() > 50          # many rays leave the grid
    for key in ("x", "y", "kx", "ky"):
        assert np.array_equal(np.isnan(got[key]), np.isnan(ref[key]))
    assert np.nanmax(np.abs(got["x"] - ref["x"])) <= 1e-9
    assert np.nanmax(np.abs(got["y"] - ref["y"])) <= 1e-9


def test_fp64_trace_eddy_subset(mr, oracle_mod):
    from mantaray_amd import workloads
    cfg = workloads.eddy_adjoint(n_rays=512)
    b, c = workloads.make_fields(cfg)
    ob, oc = oracle_fields(oracle_mod, cfg)
    args = (cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"], cfg["t_end"], cfg["dt"])
    got = mr.trace_many(b, c, *args, precision="fp64", trajectories=False)
    ref = oracle_mod.trace_many(ob, oc, *args, prec=oracle_mod.PREC_FP64_INTERP,
                                trajectories=False)
    assert np.array_equal(got["len"], ref["len"])
    d = np.abs(got["end"][:, :2] - ref["end"][:, :2]).max()
    print(f"eddy fp64 endpoints: max |d(x, y)| {d:.3e} m")
    assert d <= 1e-8


def test_checkpoints_are_rows(mr, torch):
    from mantaray_amd import adjoint, workloads
    cfg = workloads.eddy_adjoint(n_rays=300, n_steps=230)
    b, c = workloads.make_fields(cfg)
    ic = [_dev(torch, cfg[k]) for k in ("x0", "y0", "kx0", "ky0")]
    r = adjoint.trace_fp64(b, c, *ic, 0.0, 230.0, 1.0, stride=50, trajectories=True)
    torch.cuda.synchronize()
    ck = r["ckpt"].cpu().numpy()
    assert ck.shape == (5, 4, 300)
    rows = np.stack([r[k].cpu().numpy() for k in ("x", "y", "kx", "ky")], axis=1)
    for i in range(5):
        assert np.array_equal(ck[i], rows[50 * i])
    full = mr.trace_many(b, c, *cfg_args(cfg, 230.0), precision="fp64")
    assert np.array_equal(full["x"], r["x"].cpu().numpy())
    assert np.array_equal(full["end"], r["end"].cpu().numpy())


def cfg_args(cfg, t_end):
    return (cfg["x0"], cfg["y0"], cfg["kx0"], cfg["ky0"], cfg["t0"], t_end, cfg["dt"])


# ---------------------------------------------------------------------------
# finite differences


def _fd_ic(adjoint, b, c, ic, comp, eps, t_end, dt, tx, ty, **kw):
    """Per-ray central difference of L_r = (x_r - tx)^2 + (y_r - ty)^2."""
    plus = [t.clone() for t in ic]
    minus = [t.clone() for t in ic]
    plus[comp] += eps
    minus[comp] -= eps
    ep = adjoint.trace_fp64(b, c, *plus, 0.0, t_end, dt, **kw)["end"].cpu().numpy()
    em = adjoint.trace_fp64(b, c, *minus, 0.0, t_end, dt, **kw)["end"].cpu().numpy()
    dl = ((ep[:, 0] - em[:, 0]) * (ep[:, 0] + em[:, 0] - 2 * tx)
          + (ep[:, 1] - em[:, 1]) * (ep[:, 1] + em[:, 1] - 2 * ty))
    step = (plus[comp] - minus[comp]).cpu().numpy()      # the step as represented
    return dl / step


def _fd_values(adjoint, b, c, ic, t_end, dt, tx, ty, eps, vals, pert):
    """Central difference of sum_r L_r along a direction in the value arrays
    (vals: name -> values; pert: name -> direction tensor)."""
    vals = {k: vals.get(k) for k in ("depth", "u", "v")}
    plus = {k: (None if a is None else a.clone()) for k, a in vals.items()}
    minus = {k: (None if a is None else a.clone()) for k, a in vals.items()}
    for k, d in pert.items():
        plus[k] = plus[k] + eps * d
        minus[k] = minus[k] - eps * d
    ep = adjoint.trace_fp64(b, c, *ic, 0.0, t_end, dt, **plus)["end"].cpu().numpy()
    em = adjoint.trace_fp64(b, c, *ic, 0.0, t_end, dt, **minus)["end"].cpu().numpy()
    ok = ~np.isnan(ep[:, 0]) & ~np.isnan(em[:, 0])
    dl = ((ep[:, 0] - em[:, 0]) * (ep[:, 0] + em[:, 0] - 2 * tx)
          + (ep[:, 1] - em[:, 1]) * (ep[:, 1] + em[:, 1] - 2 * ty))
    return float(np.sum(dl[ok])) / (2 * eps)


def _g_end(end, tx, ty, torch):
    e = end.cpu().numpy()
    g = np.zeros_like(e)
    g[:, 0] = 2 * (e[:, 0] - tx)
    g[:, 1] = 2 * (e[:, 1] - ty)
    g[np.isnan(g)] = 0.0
    return torch.tensor(g, device=end.device)


def _report(tag, g, fd):
    """Print the distribution of per-ray relative errors."""
    r = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)
    q = np.quantile(r, [0.5, 0.9, 0.99, 1.0])
    print(f"{tag}: per-ray rel err median {q[0]:.1e} p90 {q[1]:.1e} p99 {q[2]:.1e} "
          f"max {q[3]:.1e}; normwise {_rel(g, fd):.1e}")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _small_case(mr, torch, n=256, seed=7):
    cfg = random_grid_case(seed, nx=41, ny=37, dx=40.0, dy=45.0, x0=-800.0, y0=-700.0)
    b, c = mr.Bathymetry.from_arrays(**cfg["bathy"]), mr.Current.from_arrays(**cfg["current"])
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    k = rng.uniform(0.03, 0.15, n)
    x0 = rng.uniform(-500.0, 500.0, n)
    y0 = rng.uniform(-400.0, 500.0, n)
    ic = [_dev(torch, a) for a in (x0, y0, k * np.cos(th), k * np.sin(th))]
    tx = x0 + rng.uniform(-200, 200, n)
    ty = y0 + rng.uniform(-200, 200, n)
    vals = {"depth": _dev(torch, cfg["bathy"]["depth"].ravel()),
            "u": _dev(torch, cfg["current"]["u"].ravel()),
            "v": _dev(torch, cfg["current"]["v"].ravel())}
    return b, c, ic, tx, ty, vals


def test_adjoint_one_step(mr, torch):
    """One RK4 step: the adjoint of the four stage products against central
    differences (no accumulated rounding, so the check is tight)."""
    from mantaray_amd import adjoint
    b, c, ic, tx, ty, vals = _small_case(mr, torch)
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, 1.0, 1.0, stride=1, **vals)
    ge = _g_end(fwd["end"], tx, ty, torch)
    g = adjoint.adjoint(b, c, fwd, 0.0, 1.0, 1.0, 1, g_end=ge, **vals)
    torch.cuda.synchronize()
    for comp, name, eps in ((0, "x0", 1e-4), (1, "y0", 1e-4), (2, "kx0", 1e-8), (3, "ky0", 1e-8)):
        fd = _fd_ic(adjoint, b, c, ic, comp, eps, 1.0, 1.0, tx, ty, **vals)
        err = _rel(g[name].cpu().numpy(), fd)
        print(f"one step d/d{name}: rel err {err:.2e}")
        assert err <= 1e-6, (name, err)


@pytest.mark.parametrize("stride", [1, 16])
def test_adjoint_small_grid(mr, torch, stride):
    """120 steps over rough gridded depth and current (rays leave the grid):
    initial-state gradients per ray, and depth / u / v directional
    derivatives, against central differences."""
    from mantaray_amd import adjoint
    b, c, ic, tx, ty, vals = _small_case(mr, torch)
    T = 120.0
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=stride, **vals)
    ln = fwd["len"].cpu().numpy()
    assert (ln < 121).sum() > 10                     # some rays leave the grid
    ge = _g_end(fwd["end"], tx, ty, torch)
    g = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, stride, g_end=ge, **vals)
    torch.cuda.synchronize()
    dead = ln < 121
    for comp, name, eps in ((0, "x0", 1e-5), (1, "y0", 1e-5), (2, "kx0", 1e-9), (3, "ky0", 1e-9)):
        fd = _fd_ic(adjoint, b, c, ic, comp, eps, T, 1.0, tx, ty, **vals)
        gg = g[name].cpu().numpy()
        assert np.all(gg[dead] == 0.0)               # a NaN last row passes nothing
        live = ~dead & np.isfinite(fd)
        assert live.sum() > 150
        _report(f"stride {stride} d/d{name}", gg[live], fd[live])
        err = _rel(gg[live], fd[live])
        print(f"stride {stride} d/d{name}: rel err {err:.2e}")
        assert err <= 1e-5, (name, err)
    rng = np.random.default_rng(3)
    for name in ("depth", "u", "v"):
        d = rng.normal(size=vals[name].numel())
        d /= np.linalg.norm(d)
        dt_ = _dev(torch, d)
        fd = _fd_values(adjoint, b, c, ic, T, 1.0, tx, ty, 1e-6, vals, {name: dt_})
        an = float(np.dot(g[name].cpu().numpy(), d))
        err = abs(an - fd) / abs(fd)
        print(f"stride {stride} directional d/d{name}: adjoint {an:.9e} fd {fd:.9e} rel {err:.2e}")
        assert err <= 1e-5, (name, err)


def test_adjoint_single_values(mr, torch):
    """Gradient entries of single u and depth values near many rays."""
    from mantaray_amd import adjoint
    b, c, ic, tx, ty, vals = _small_case(mr, torch, n=1024)
    T = 60.0
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=10, **vals)
    ge = _g_end(fwd["end"], tx, ty, torch)
    g = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, 10, g_end=ge, **vals)
    torch.cuda.synchronize()
    for name in ("depth", "u", "v"):
        ga = g[name].cpu().numpy()
        for idx in np.argsort(-np.abs(ga))[:3]:
            d = torch.zeros_like(vals[name])
            d[idx] = 1.0
            fd = _fd_values(adjoint, b, c, ic, T, 1.0, tx, ty, 1e-6, vals, {name: d})
            err = abs(ga[idx] - fd) / abs(fd)
            print(f"d/d{name}[{idx}]: adjoint {ga[idx]:.9e} fd {fd:.9e} rel {err:.2e}")
            assert err <= 1e-5, (name, idx, err)


def test_adjoint_rows_loss(mr, torch):
    """A loss on every row: L = sum_rows w_s . (x_s, y_s, kx_s, ky_s)."""
    from mantaray_amd import adjoint
    b, c, ic, tx, ty, vals = _small_case(mr, torch)
    T = 80.0
    n = 80
    rng = np.random.default_rng(11)
    w = {k: rng.normal(size=(n + 1, 256)) * s for k, s in
         (("x", 1.0), ("y", 1.0), ("kx", 1e3), ("ky", 1e3))}
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=7, trajectories=True, **vals)
    rows = {k: _dev(torch, w[k]) for k in w}
    g = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, 7, g_rows=rows, **vals)
    torch.cuda.synchronize()

    def loss_rows(icx):
        r = adjoint.trace_fp64(b, c, *icx, 0.0, T, 1.0, trajectories=True, **vals)
        tot = np.zeros(256)
        for k in w:
            a = r[k].cpu().numpy()
            tot += np.nansum(np.where(np.isnan(a), 0.0, a * w[k]), axis=0)
        return tot

    for comp, name, eps in ((0, "x0", 1e-5), (3, "ky0", 1e-9)):
        plus = [t.clone() for t in ic]
        minus = [t.clone() for t in ic]
        plus[comp] += eps
        minus[comp] -= eps
        fd = (loss_rows(plus) - loss_rows(minus)) / (2 * eps)
        err = _rel(g[name].cpu().numpy(), fd)
        print(f"rows loss d/d{name}: rel err {err:.2e}")
        assert err <= 1e-5, (name, err)


def test_stride_independence_and_autograd(mr, torch):
    from mantaray_amd import adjoint, workloads
    cfg = workloads.eddy_adjoint(n_rays=700, n_steps=300)
    b, c = workloads.make_fields(cfg)
    ic = [_dev(torch, cfg[k]) for k in ("x0", "y0", "kx0", "ky0")]
    u = _dev(torch, cfg["current"]["u"].ravel())
    v = _dev(torch, cfg["current"]["v"].ravel())
    tx, ty = cfg["target_x"], cfg["target_y"]
    res = {}
    for stride in (1, 7, 50, 300, 1000):
        fwd = adjoint.trace_fp64(b, c, *ic, 0.0, 300.0, 1.0, stride=stride, u=u, v=v)
        ge = _g_end(fwd["end"], tx, ty, torch)
        res[stride] = adjoint.adjoint(b, c, fwd, 0.0, 300.0, 1.0, stride, g_end=ge, u=u, v=v)
    for stride in (7, 50, 300, 1000):
        for k in ("x0", "y0", "kx0", "ky0"):
            assert torch.equal(res[stride][k], res[1][k]), (stride, k)
        for k in ("u", "v"):
            d = (res[stride][k] - res[1][k]).abs().max().item()
            assert d <= 1e-9 * res[1][k].abs().max().item(), (stride, k, d)
    # the autograd.Function
    leaves = [t.clone().requires_grad_(True) for t in ic]
    ul, vl = u.clone().requires_grad_(True), v.clone().requires_grad_(True)
    end, ln = adjoint.trace_differentiable(b, c, *leaves, 0.0, 300.0, 1.0, u=ul, v=vl,
                                           stride=50)
    txt, tyt = _dev(torch, tx), _dev(torch, ty)
    loss = ((end[:, 0] - txt) ** 2 + (end[:, 1] - tyt) ** 2).sum()
    loss.backward()
    for t, k in zip(leaves, ("x0", "y0", "kx0", "ky0")):
        assert torch.equal(t.grad, res[50][k]), k
    for t, k in ((ul, "u"), (vl, "v")):
        d = (t.grad - res[50][k]).abs().max().item()
        assert d <= 1e-9 * res[50][k].abs().max().item(), k


def _ld_loss_fd(oracle_mod, ob, oc, s0, T, tx, ty, eps, **dirs):
    """Central difference of sum_r L_r from the long-double traces."""
    diff, mid, _, _ = oracle_mod.fd_end_ld(ob, oc, s0, 0.0, T, 1.0, eps=eps, **dirs)
    return (diff[:, 0] * 2 * (mid[:, 0] - tx) + diff[:, 1] * 2 * (mid[:, 1] - ty)) / (2 * eps)


def test_adjoint_eddy_workload(mr, oracle_mod, torch):
    """The eddy workload (512 x 512, 16,384 rays, 2000 steps, checkpoint
    stride 50): gradients of sum_r |x_r(T) - target_r|^2 on a 256-ray subset
    against central differences with eps = 1e-6 (wavenumbers: 1e-12, which
    moves the endpoints as far).  Truth: the same algorithm traced in long
    double (or_fd_end_ld), within 1e-6 normwise; the f64 forward's own
    central differences agree to their rounding noise (1e-3)."""
    from mantaray_amd import adjoint, workloads
    cfg = workloads.eddy_adjoint()
    b, c = workloads.make_fields(cfg)
    ob, oc = oracle_fields(oracle_mod, cfg)
    ic = [_dev(torch, cfg[k]) for k in ("x0", "y0", "kx0", "ky0")]
    u = _dev(torch, cfg["current"]["u"].ravel())
    v = _dev(torch, cfg["current"]["v"].ravel())
    tx, ty = cfg["target_x"], cfg["target_y"]
    T = cfg["t_end"]
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=50, u=u, v=v)
    ge = _g_end(fwd["end"], tx, ty, torch)
    g = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, 50, g_end=ge, u=u, v=v)
    torch.cuda.synchronize()
    assert int((fwd["len"] < 2001).sum()) == 0
    sub = slice(0, 256)
    s0 = np.stack([cfg[k][sub] for k in ("x0", "y0", "kx0", "ky0")], 1)
    ics = [t[sub].clone() for t in ic]
    tsx, tsy = tx[sub], ty[sub]
    for comp, name, scale in ((0, "x0", 1.0), (1, "y0", 1.0), (2, "kx0", 1e-6),
                              (3, "ky0", 1e-6)):
        gg = g[name][sub].cpu().numpy()
        ds0 = np.zeros_like(s0)
        ds0[:, comp] = 1.0
        fd = _ld_loss_fd(oracle_mod, ob, oc, s0, T, tsx, tsy, 1e-6 * scale, ds0=ds0)
        _report(f"eddy d/d{name} (long-double FD)", gg, fd)
        err = _rel(gg, fd)
        assert err <= 1e-6, (name, err)
        fd64 = _fd_ic(adjoint, b, c, ics, comp, 1e-6 * scale, T, 1.0, tsx, tsy, u=u, v=v)
        _report(f"eddy d/d{name} (f64 FD)", gg, fd64)
        assert _rel(gg, fd64) <= 1e-3, name
    # grid values: directional derivatives of the subset's loss
    fws = adjoint.trace_fp64(b, c, *ics, 0.0, T, 1.0, stride=50, u=u, v=v)
    gs = adjoint.adjoint(b, c, fws, 0.0, T, 1.0, 50, g_end=_g_end(fws["end"], tsx, tsy, torch),
                         u=u, v=v)
    rng = np.random.default_rng(4)
    for name in ("u", "v"):
        d = rng.normal(size=u.numel())
        d /= np.linalg.norm(d)
        an = float(np.dot(gs[name].cpu().numpy(), d))
        fd = float(np.sum(_ld_loss_fd(oracle_mod, ob, oc, s0, T, tsx, tsy, 1e-6,
                                      **{"d" + name: d})))
        err = abs(an - fd) / abs(fd)
        print(f"eddy directional d/d{name}: adjoint {an:.12e} long-double fd {fd:.12e} "
              f"rel {err:.2e}")
        assert err <= 1e-6, (name, err)
        fd64 = _fd_values(adjoint, b, c, ics, T, 1.0, tsx, tsy, 1e-6, {"u": u, "v": v},
                          {name: _dev(torch, d)})
        print(f"eddy directional d/d{name}: f64 fd {fd64:.12e} rel {abs(an - fd64) / abs(fd64):.2e}")
        assert abs(an - fd64) <= 1e-3 * abs(fd64)
    # single values: the largest entries of the u gradient
    gu = gs["u"].cpu().numpy()
    for idx in np.argsort(-np.abs(gu))[:2]:
        d = np.zeros(u.numel())
        d[idx] = 1.0
        fd = float(np.sum(_ld_loss_fd(oracle_mod, ob, oc, s0, T, tsx, tsy, 1e-6, du=d)))
        err = abs(gu[idx] - fd) / abs(fd)
        print(f"eddy d/du[{idx}]: adjoint {gu[idx]:.12e} long-double fd {fd:.12e} rel {err:.2e}")
        assert err <= 1e-6, (idx, err)


def test_adjoint_edge_cases(mr, torch):
    """Zero steps; a stride longer than the trace; NaN and out-of-grid initial
    states (a NaN row at once: only row 0's cotangent reaches them); no rays."""
    from mantaray_amd import adjoint
    b, c, ic, tx, ty, vals = _small_case(mr, torch, n=64)
    ic = [t.clone() for t in ic]
    ic[0][3] = float("nan")                     # NaN initial x
    ic[1][5] = 1e6                              # outside the grid
    rng = np.random.default_rng(2)
    # zero steps: the only row is the initial state
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, 0.0, 1.0, stride=4, trajectories=True, **vals)
    assert fwd["ckpt"].shape[0] == 1 and int(fwd["len"].max()) == 1
    ge = _dev(torch, rng.normal(size=(64, 4)))
    g = adjoint.adjoint(b, c, fwd, 0.0, 0.0, 1.0, 4, g_end=ge, **vals)
    torch.cuda.synchronize()
    want = ge.cpu().numpy()
    end = fwd["end"].cpu().numpy()
    dead = np.isnan(end).all(axis=1)
    for q, k in enumerate(("x0", "y0", "kx0", "ky0")):
        got = g[k].cpu().numpy()
        assert np.array_equal(got[~dead], want[~dead, q]), k
    # 40 steps, stride 1000 (one segment), rows cotangent on every row
    T = 40.0
    rows = {k: _dev(torch, rng.normal(size=(41, 64))) for k in ("x", "y")}
    f1 = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=1000, trajectories=True, **vals)
    f2 = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=3, trajectories=True, **vals)
    g1 = adjoint.adjoint(b, c, f1, 0.0, T, 1.0, 1000, g_rows=rows, **vals)
    g2 = adjoint.adjoint(b, c, f2, 0.0, T, 1.0, 3, g_rows=rows, **vals)
    torch.cuda.synchronize()
    ln = f1["len"].cpu().numpy()
    assert ln[3] == 2 and ln[5] == 2            # a NaN row straight away
    for k in ("x0", "y0", "kx0", "ky0"):
        assert torch.equal(g1[k], g2[k]), k
    # the dead-at-once rays get row 0's cotangent only
    for r_ in (3, 5):
        assert g1["x0"][r_].item() == rows["x"][0, r_].item()
        assert g1["y0"][r_].item() == rows["y"][0, r_].item()
        assert g1["kx0"][r_].item() == 0.0 and g1["ky0"][r_].item() == 0.0
    # no rays: nothing traced, the grid gradients are cleared
    empty = [t[:0].clone() for t in ic]
    f0 = adjoint.trace_fp64(b, c, *empty, 0.0, T, 1.0, stride=5, **vals)
    g0 = adjoint.adjoint(b, c, f0, 0.0, T, 1.0, 5, **vals)
    torch.cuda.synchronize()
    assert g0["x0"].numel() == 0 and float(g0["u"].abs().sum()) == 0.0


@pytest.mark.parametrize("case", ["small", "eddy"])
def test_sweep_kernels_agree(mr, torch, case):
    """The three reverse-sweep kernels (two waves: recompute | sweep; four:
    recompute | two linearise | chain, the default at one workgroup per CU;
    one wave recomputing and sweeping each segment in turn) give the same
    gradients: they differ only in the order of a few roundings (the
    four-wave chain applies each stage as an explicit 4 x 4 matrix; the
    order of concurrent atomic additions to the grids)."""
    from mantaray_amd import adjoint, workloads
    if case == "small":
        b, c, ic, tx, ty, vals = _small_case(mr, torch)
        T, stride = 90.0, 7
    else:
        cfg = workloads.eddy_adjoint(n_rays=3000, n_steps=400)
        b, c = workloads.make_fields(cfg)
        ic = [_dev(torch, cfg[k]) for k in ("x0", "y0", "kx0", "ky0")]
        vals = {"u": _dev(torch, cfg["current"]["u"].ravel()),
                "v": _dev(torch, cfg["current"]["v"].ravel())}
        tx, ty = cfg["target_x"], cfg["target_y"]
        T, stride = 400.0, 50
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=stride, **vals)
    ge = _g_end(fwd["end"], tx, ty, torch)
    rows = {"kx": _dev(torch, np.random.default_rng(1).normal(size=(int(T) + 1, ic[0].numel())))}
    g2 = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, stride, g_end=ge, g_rows=rows, block_size=128,
                         **vals)
    for block in (256, 64):
        g4 = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, stride, g_end=ge, g_rows=rows,
                             block_size=block, **vals)
        torch.cuda.synchronize()
        for k in ("x0", "y0", "kx0", "ky0", "depth", "u", "v"):
            if g2.get(k) is None:
                continue
            a2, a4 = g2[k].cpu().numpy(), g4[k].cpu().numpy()
            d = np.abs(a2 - a4).max() / max(np.abs(a2).max(), 1e-300)
            print(f"{case} {k}: two vs {block // 64} waves {d:.2e}")
            assert d <= 1e-12, (k, block, d)


def test_adjoint_eddy_per_ray(mr, oracle_mod, torch):
    """The eddy workload's gradients ray by ray.  Four reverse sweeps over
    all 16,384 rays (the four-wave kernel this launch size selects), each
    with a unit cotangent on one end-state component of the 256 subset rays
    and zero elsewhere, give each subset ray's 4 x 4 Jacobian of its end
    state with respect to its initial state; the same Jacobian comes from
    central differences of the long-double trace (eps 1e-6 m for positions,
    1e-12 for wavenumbers).  Per ray and end-state component, the row of
    the Jacobian (columns scaled by those eps, so that positions and
    wavenumbers weigh alike) must agree within 1e-6 of its largest entry.
    Rays whose two perturbed traces met different grid cells (a jump of the
    cell slopes lies between them) are left out and counted; on this
    workload there are none.  The loss gradient of each ray, normwise over
    its four components scaled alike, must agree within 1e-6 of the size of
    its terms 2 dx . x' + 2 dy . y' (which can cancel): a component that is a
    small entry of its Jacobian row carries that row's error, so a relative
    test component by component would measure the row's conditioning."""
    from mantaray_amd import adjoint, workloads
    cfg = workloads.eddy_adjoint()
    b, c = workloads.make_fields(cfg)
    ob, oc = oracle_fields(oracle_mod, cfg)
    ic = [_dev(torch, cfg[k]) for k in ("x0", "y0", "kx0", "ky0")]
    u = _dev(torch, cfg["current"]["u"].ravel())
    v = _dev(torch, cfg["current"]["v"].ravel())
    T = cfg["t_end"]
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=50, u=u, v=v)
    nr = ic[0].numel()
    sub = np.arange(256)
    J = np.zeros((256, 4, 4))                     # J[r, q, j] = d end_q / d s0_j
    for q in range(4):
        ge = torch.zeros((nr, 4), dtype=torch.float64, device="cuda:0")
        ge[:256, q] = 1.0
        g = adjoint.adjoint(b, c, fwd, 0.0, T, 1.0, 50, g_end=ge, u=u, v=v,
                            want=("x0", "y0", "kx0", "ky0"))
        for j, k in enumerate(("x0", "y0", "kx0", "ky0")):
            J[:, q, j] = g[k][:256].cpu().numpy()
    torch.cuda.synchronize()
    s0 = np.stack([cfg[k][sub] for k in ("x0", "y0", "kx0", "ky0")], 1)
    scale = np.array([1e-6, 1e-6, 1e-12, 1e-12])
    Jfd = np.zeros_like(J)
    same = np.ones(256, dtype=bool)
    mids = None
    for j in range(4):
        ds0 = np.zeros_like(s0)
        ds0[:, j] = 1.0
        diff, mid, lp, lm, sc = oracle_mod.fd_end_ld(ob, oc, s0, 0.0, T, 1.0, eps=scale[j],
                                                     ds0=ds0, with_cells=True)
        Jfd[:, :, j] = diff / (2 * scale[j])
        same &= sc & (lp == lm)
        mids = mid
    print(f"eddy per ray: {int((~same).sum())} of 256 rays crossed a cell line between "
          f"their perturbed traces")
    assert same.sum() >= 250
    A, B = J * scale, Jfd * scale                 # columns in units of the perturbations
    err = np.abs(A - B).max(axis=2) / np.abs(B).max(axis=2)     # per ray and end component
    for q, name in enumerate(("x", "y", "kx", "ky")):
        e = err[same, q]
        print(f"eddy per-ray Jacobian row d{name}(T)/ds0: max {e.max():.2e} "
              f"p99 {np.quantile(e, 0.99):.2e} median {np.median(e):.2e}")
    assert err[same].max() <= 1e-6, float(err[same].max())
    # the loss gradient per ray: normwise over the ray's four components
    # (scaled as the Jacobian's columns), against the size of its terms
    # 2 dx . x' and 2 dy . y' (which can cancel)
    tx, ty = cfg["target_x"][sub], cfg["target_y"][sub]
    dx, dy = 2 * (mids[:, 0] - tx), 2 * (mids[:, 1] - ty)
    g_loss = (dx[:, None] * J[:, 0, :] + dy[:, None] * J[:, 1, :]) * scale
    f_loss = (dx[:, None] * Jfd[:, 0, :] + dy[:, None] * Jfd[:, 1, :]) * scale
    terms = (np.abs(dx[:, None] * Jfd[:, 0, :]) + np.abs(dy[:, None] * Jfd[:, 1, :])) * scale
    rel = np.abs(g_loss - f_loss).max(axis=1) / terms.max(axis=1)
    comp = np.abs(g_loss - f_loss) / terms
    print(f"eddy per-ray loss gradient: max normwise error / terms {rel[same].max():.2e} "
          f"(p99 {np.quantile(rel[same], 0.99):.2e}); one component at a time "
          f"max {comp[same].max():.2e}")
    assert rel[same].max() <= 1e-6
