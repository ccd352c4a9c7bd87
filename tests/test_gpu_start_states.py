"""Rays that start all-NaN, partly NaN or off the grid, beside ordinary rays.

In the reference a ray takes its first step whatever its start (ode_solvers'
Rk4::integrate asks solout only after a step, src/wave_ray_path.rs:236-246),
and the "next_deriv" reading's test (f at the state a step starts from)
applies from the second step on: an all-NaN start, or a start off the grid
(an all-NaN first derivative), gives a NaN first row that counts under both
readings.  The trace kernel reads liveness off the state and settles an
all-NaN start without stepping; these traces check that against the oracle
for both solout readings, for 0, 1, 5 and 40 steps, with all four trajectory
components, a subset of them and none, and through the batched kernel:
identical row counts and NaN patterns, positions within 1e-8 m, wavenumbers
within 1e-10 of their scale, and the same last rows.  Waves holding only
all-NaN starts and waves mixing them with live rays are both present.
"""
import numpy as np
import pytest

from _helpers import oracle_fields, random_grid_case, random_states

pytestmark = pytest.mark.gpu


def _starts(cfg):
    s = random_states(3, cfg, 256)
    x, y = cfg["bathy"]["x"], cfg["bathy"]["y"]
    s[0:64] = np.nan                                   # a wave of all-NaN starts
    s[64:256:7] = np.nan                               # all-NaN starts among live rays
    s[65:256:11, 0] = np.nan                           # x alone NaN
    s[66:256:13, 2:] = np.nan                          # both wavenumbers NaN
    s[67:256:17, 0] = x[-1] + 500.0                    # off the grid
    s[68:256:19, 1] = y[0] - 500.0
    return s


def _close(got, ref, tag):
    assert np.array_equal(np.isnan(got), np.isnan(ref)), tag
    if np.isfinite(ref).any():
        d = np.nanmax(np.abs(got - ref))
        return d
    return 0.0


@pytest.mark.parametrize("solout", ["state", "next_deriv"])
def test_start_states_against_the_oracle(oracle_mod, solout):
    import torch

    from mantaray_amd import core
    cfg = random_grid_case(3)
    b = core.Bathymetry.from_arrays(**cfg["bathy"])
    c = core.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = _starts(cfg)
    dev = torch.device("cuda:0")
    ic = [torch.tensor(np.ascontiguousarray(s[:, k]), device=dev) for k in range(4)]
    so = oracle_mod.SOLOUT_STATE if solout == "state" else oracle_mod.SOLOUT_NEXT_DERIV
    for n in (0, 1, 5, 40):
        ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, float(n),
                                    1.0, solout=so)
        if n > 0:
            assert ref["len"].min() == 2, "every ray takes its first step"
        kscale = max(np.nanmax(np.abs(ref["kx"])), np.nanmax(np.abs(ref["ky"])))
        for mode in ("all", "subset", "none"):
            out = None
            if mode == "subset":
                out = {"t": torch.empty(n + 1, dtype=torch.float64, device=dev),
                       "len": torch.empty(256, dtype=torch.int32, device=dev),
                       "end": torch.empty((256, 4), dtype=torch.float64, device=dev),
                       "x": torch.empty((n + 1, 256), dtype=torch.float64, device=dev),
                       "ky": torch.empty((n + 1, 256), dtype=torch.float64, device=dev)}
            got = core.trace_many_device(b, c, *ic, 0.0, float(n), 1.0, solout=solout,
                                         trajectories=mode == "all", out=out)
            torch.cuda.synchronize()
            tag = (solout, n, mode)
            assert np.array_equal(got["len"].cpu().numpy(), ref["len"]), tag
            end = got["end"].cpu().numpy()
            assert _close(end[:, :2], ref["end"][:, :2], tag) <= 1e-8, tag
            assert _close(end[:, 2:], ref["end"][:, 2:], tag) <= 1e-10 * kscale, tag
            keys = {"all": ("x", "y", "kx", "ky"), "subset": ("x", "ky"), "none": ()}[mode]
            for k in keys:
                tol = 1e-8 if k in ("x", "y") else 1e-10 * kscale
                assert _close(got[k].cpu().numpy(), ref[k], tag + (k,)) <= tol, tag + (k,)
        print(f"{solout} {n} steps: row counts {np.bincount(ref['len'])[1:6]} ... "
              f"{int((ref['len'] == n + 1).sum())} rays take every step")


@pytest.mark.parametrize("solout", ["state", "next_deriv"])
def test_start_states_batched(oracle_mod, solout):
    from mantaray_amd import core
    from mantaray_amd.batch import Batch
    cfg = random_grid_case(3)
    b = core.Bathymetry.from_arrays(**cfg["bathy"])
    c = core.Current.from_arrays(**cfg["current"])
    ob, oc = oracle_fields(oracle_mod, cfg)
    s = _starts(cfg)
    so = oracle_mod.SOLOUT_STATE if solout == "state" else oracle_mod.SOLOUT_NEXT_DERIV
    scenes = [{"bathymetry": b, "current": c, "x0": s[:, 0], "y0": s[:, 1], "kx0": s[:, 2],
               "ky0": s[:, 3], "t0": 0.0, "t_end": float(n), "dt": 1.0} for n in (1, 5, 40)]
    batch = Batch(scenes, solout=solout)
    batch.run()
    import torch
    torch.cuda.synchronize()
    for sc, o in zip(scenes, batch.results()):
        ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0,
                                    sc["t_end"], 1.0, solout=so, trajectories=False)
        tag = (solout, sc["t_end"])
        assert np.array_equal(o["len"].cpu().numpy(), ref["len"]), tag
        end = o["end"].cpu().numpy()
        kscale = np.nanmax(np.abs(ref["end"][:, 2:]))
        assert _close(end[:, :2], ref["end"][:, :2], tag) <= 1e-8, tag
        assert _close(end[:, 2:], ref["end"][:, 2:], tag) <= 1e-10 * kscale, tag
    batch.close()
