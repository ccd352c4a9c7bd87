"""The batched-scenes and inversion workloads at their own sizes.

Batched scenes (64 scenes of 256^2 to 2048^2 cells at 250 m, 2^20 rays, steps
0.5-2 s, up to 10,000 of them; the multi-scene form of the loop of
src/ray.rs:98-127): one persistent launch, at the default launch shape and
at the shape bench.py's autotuner picks (64 threads, 4 workgroups per CU).
Every ray's row count and last row equal bitwise those of a per-scene
trace_many launch, and every 1024th ray (1,024 rays, in scenes of every
size) matches the oracle: equal row counts, end positions within 1e-6 m,
wavenumbers within 1e-9 relative.

Inversion (2048 x 2048 grid at 100 m, depth 30 m, 262,144 rays, 3000 steps):
two descent iterations from a zero current lower the endpoint misfit.  At
the second iterate the gradient of the misfit of 128 rays is taken by the
full 262,144-ray reverse sweep (the one-wave kernel this launch size
selects: 4,096 workgroups of 64 rays, 16 per CU), with a zero cotangent on
every other ray, and checked against central differences of the same
algorithm traced in long double (the oracle's or_fd_end_ld, eps 1e-6 m for
positions, 1e-12 for wavenumbers, 1e-6 m/s along a unit direction of u or
v): per ray for the initial states, normwise over the ray's four components
scaled by the perturbations, within 1e-6 of the size of the terms of its
misfit derivative (the misfit 2 dx . x' + 2 dy . y' can cancel, and no
derivative is more accurate than its terms), and within 1e-6 relative for
one u and one v directional derivative.  Rays whose two
perturbed traces met different grid cells (their difference spans a jump of
the cell slopes) are left out of the per-ray check and counted.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    return _t


def _bits(a):
    a = np.asarray(a, dtype=np.float64)
    return np.where(np.isnan(a), np.uint64(0x7FF8000000000000), a.view(np.uint64))


def test_batched_scenes_full_size(oracle_mod, torch):
    from mantaray_amd import core, workloads
    from mantaray_amd.batch import Batch
    cfgs = workloads.batched_scenes()
    print(f"batched scenes: {len(cfgs)} scenes generated", flush=True)
    scenes = []
    for cfg in cfgs:
        b, c = workloads.make_fields(cfg)
        scenes.append({"bathymetry": b, "current": c,
                       **{k: cfg[k] for k in ("x0", "y0", "kx0", "ky0", "t0", "t_end", "dt")}})
    batch = Batch(scenes)
    dev = torch.device("cuda:0")
    per_scene = []
    for sc in scenes:
        ic = [torch.tensor(sc[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
        o = core.trace_many_device(sc["bathymetry"], sc["current"], *ic, sc["t0"], sc["t_end"],
                                   sc["dt"], trajectories=False)
        per_scene.append((o["len"].cpu().numpy(), o["end"].cpu().numpy()))
    for shape in ({}, {"block_size": 64, "blocks_per_cu": 4}):
        batch.run(**shape)
        torch.cuda.synchronize()
        res = [(o["len"].cpu().numpy(), o["end"].cpu().numpy()) for o in batch.results()]
        for i, ((bl, be), (sl, se)) in enumerate(zip(res, per_scene)):
            assert np.array_equal(bl, sl), (shape, i)
            assert np.array_equal(_bits(be), _bits(se)), (shape, i)
    n_rays, bound = batch.info()
    steps = sum(int((bl - 1).sum()) for bl, _ in res)
    print(f"batched scenes: {n_rays} rays, {steps} ray-steps of at most {bound}", flush=True)
    assert n_rays == 1 << 20 and steps < bound
    # every 1024th ray of the batch against the oracle
    checked, sides = 0, set()
    first = 0
    for i, (cfg, (bl, be)) in enumerate(zip(cfgs, res)):
        nr = bl.size
        idx = np.arange((-first) % 1024, nr, 1024)
        first += nr
        if idx.size == 0:
            continue
        ob = oracle_mod.Bathymetry.grid(cfg["bathy"]["x"], cfg["bathy"]["y"], cfg["bathy"]["depth"])
        oc = oracle_mod.Current.grid(cfg["current"]["x"], cfg["current"]["y"], cfg["current"]["u"],
                                     cfg["current"]["v"])
        ref = oracle_mod.trace_many(ob, oc, cfg["x0"][idx], cfg["y0"][idx], cfg["kx0"][idx],
                                    cfg["ky0"][idx], cfg["t0"], cfg["t_end"], cfg["dt"],
                                    trajectories=False)
        assert np.array_equal(bl[idx], ref["len"]), i
        ge, re_ = be[idx], ref["end"]
        assert np.array_equal(np.isnan(ge), np.isnan(re_)), i
        fin = ~np.isnan(re_[:, 0])
        if fin.any():
            assert np.abs(ge[fin, :2] - re_[fin, :2]).max() <= 1e-6, i
            ks = np.abs(re_[fin, 2:]).max()
            assert np.abs(ge[fin, 2:] - re_[fin, 2:]).max() <= 1e-9 * ks, i
        checked += idx.size
        sides.add(int(cfg["bathy"]["x"].size))
    print(f"batched scenes: {checked} rays against the oracle, grid sides {sorted(sides)}")
    assert checked == 1024 and sides == {256, 512, 1024, 2048}
    batch.close()


def _misfit_fd(oracle_mod, ob, oc, s0, T, obs, eps, **dirs):
    """Per ray: central difference of |(x, y)(T) - obs|^2, the sum of the
    absolute values of its two terms, and whether both perturbed traces
    met the same cells."""
    diff, mid, lp, lm, same = oracle_mod.fd_end_ld(ob, oc, s0, 0.0, T, 1.0, eps=eps,
                                                   with_cells=True, **dirs)
    tx = diff[:, 0] * 2 * (mid[:, 0] - obs[:, 0]) / (2 * eps)
    ty = diff[:, 1] * 2 * (mid[:, 1] - obs[:, 1]) / (2 * eps)
    return tx + ty, np.abs(tx) + np.abs(ty), same & (lp == lm)


def test_inversion_full_size(oracle_mod, torch):
    from mantaray_amd import adjoint, inversion, workloads
    cfg = workloads.eddy_inversion()
    b, c = workloads.make_fields(cfg)
    dev = torch.device("cuda:0")
    T, dt, stride = cfg["t_end"], cfg["dt"], cfg["stride"]
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    ut = torch.tensor(cfg["current"]["u"], device=dev)
    vt = torch.tensor(cfg["current"]["v"], device=dev)
    obs = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, u=ut.view(-1), v=vt.view(-1))["end"][:, :2]
    obs = obs.clone()
    z = torch.zeros_like(ut)
    r = inversion.invert(b, c, ic, obs, T, dt, u0=z, v0=z, iterations=2, stride=stride,
                         sigma_cells=100.0, first_step=0.5, method="bb", u_true=ut, v_true=vt)
    h = r["history"]
    print("inversion losses", [f"{x['loss']:.4e}" for x in h],
          "rms", [f"{x['rms_error']:.3f}" for x in h], flush=True)
    assert h[1]["loss"] < h[0]["loss"] and h[2]["loss"] < h[1]["loss"]
    assert h[2]["rms_error"] < h[0]["rms_error"]
    # the gradient at the second iterate, by the full-size sweep
    u2, v2 = r["u"].reshape(-1).contiguous(), r["v"].reshape(-1).contiguous()
    fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, dt, u=u2, v=v2, stride=stride)
    nr = obs.shape[0]
    end = fwd["end"].cpu().numpy()
    cand = np.arange(0, nr, 1024)
    sub = cand[np.isfinite(end[cand]).all(axis=1)][:128]   # rays still on the grid
    assert sub.size == 128
    g_end = torch.zeros_like(fwd["end"])
    sub_t = torch.tensor(sub, device=dev)
    g_end[sub_t, :2] = 2.0 * (fwd["end"][sub_t, :2] - obs[sub_t])
    g = adjoint.adjoint(b, c, fwd, 0.0, T, dt, stride, g_end=g_end, u=u2, v=v2)
    torch.cuda.synchronize()
    # rays outside the subset carry no cotangent
    for k in ("x0", "y0", "kx0", "ky0"):
        gk = g[k].cpu().numpy()
        assert np.all(gk[np.setdiff1d(np.arange(nr), sub)] == 0.0), k
    ob = oracle_mod.Bathymetry.constant(cfg["bathy"]["constant"])
    oc = oracle_mod.Current.grid(cfg["current"]["x"], cfg["current"]["y"],
                                 u2.cpu().numpy().reshape(cfg["current"]["u"].shape),
                                 v2.cpu().numpy().reshape(cfg["current"]["v"].shape))
    s0 = np.stack([cfg[k][sub] for k in ("x0", "y0", "kx0", "ky0")], 1)
    ob_sub = obs.cpu().numpy()[sub]
    scales = np.array([1.0, 1.0, 1e-6, 1e-6]) * 1e-6
    an, fd, terms = (np.zeros((sub.size, 4)) for _ in range(3))
    same = np.ones(sub.size, dtype=bool)
    for comp, name in enumerate(("x0", "y0", "kx0", "ky0")):
        ds0 = np.zeros_like(s0)
        ds0[:, comp] = 1.0
        f, t, sm = _misfit_fd(oracle_mod, ob, oc, s0, T, ob_sub, scales[comp], ds0=ds0)
        an[:, comp], fd[:, comp], terms[:, comp] = g[name].cpu().numpy()[sub], f, t
        same &= sm
    # per ray, normwise over its four components (scaled by the
    # perturbations), against the size of the misfit's terms
    err = (np.abs(an - fd) * scales).max(axis=1) / (terms * scales).max(axis=1)
    comp_err = np.abs(an - fd) / terms
    print(f"inversion per-ray initial-state gradient: max normwise error / terms "
          f"{err[same].max():.2e}, median {np.median(err[same]):.2e}; one component at a time "
          f"max {comp_err[same].max():.2e}; {int((~same).sum())} of {sub.size} rays crossed a "
          f"cell line between their perturbed traces", flush=True)
    assert same.sum() >= 120
    assert err[same].max() <= 1e-6, float(err[same].max())
    rng = np.random.default_rng(7)
    for name in ("u", "v"):
        d = rng.normal(size=u2.numel())
        d /= np.linalg.norm(d)
        an = float(np.dot(g[name].cpu().numpy(), d))
        fd, terms, same = _misfit_fd(oracle_mod, ob, oc, s0, T, ob_sub, 1e-6, **{"d" + name: d})
        fds = float(fd.sum())
        err = abs(an - fds) / abs(fds)
        print(f"inversion directional d/d{name}: adjoint {an:.12e} long-double fd {fds:.12e} "
              f"rel {err:.2e} ({int((~same).sum())} rays crossed a cell line)", flush=True)
        assert err <= 1e-6, (name, err)
