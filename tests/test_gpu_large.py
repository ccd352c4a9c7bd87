"""Launches at the trace kernel's size limits.

All four trajectory components go out by buffer stores (one descriptor per
component, one scalar row offset) while a row is under 2^30 bytes, i.e. up to
2^27 - 1 rays; at 2^27 rays and beyond the launcher takes the kernel that
stores any subset with plain stores.  These traces run 2^27 - 64 rays (the
buffer-store path with blocks of one row each, the most row blocks a launch
can have) and 2^27 rays (the fallback), three steps each, every ray a copy of
one of 4,096 states on a small random grid (nodes, cell edges, points off the
grid, NaN and zero k among them).  Rays sampled across the whole launch, the
last ones included, must equal bitwise the same states traced in a launch of
4,096 rays, whose rows match the oracle.
"""
import numpy as np
import pytest

from _helpers import oracle_fields, random_grid_case, random_states

pytestmark = pytest.mark.gpu

BASE = 4096


def _bits(a):
    a = np.asarray(a, dtype=np.float64)
    return np.where(np.isnan(a), np.uint64(0x7FF8000000000000), a.view(np.uint64))


@pytest.mark.parametrize("n_rays", [(1 << 27) - 64, 1 << 27])
def test_largest_launches(oracle_mod, n_rays):
    import torch

    import mantaray_amd as core
    cfg = random_grid_case(5)
    b = core.Bathymetry.from_arrays(**cfg["bathy"])
    c = core.Current.from_arrays(**cfg["current"])
    s = random_states(5, cfg, BASE)
    dev = torch.device("cuda:0")
    base = [torch.tensor(np.ascontiguousarray(s[:, k]), device=dev) for k in range(4)]
    small = core.trace_many_device(b, c, *base, 0.0, 3.0, 1.0)
    ob, oc = oracle_fields(oracle_mod, cfg)
    ref = oracle_mod.trace_many(ob, oc, s[:, 0], s[:, 1], s[:, 2], s[:, 3], 0.0, 3.0, 1.0)
    assert np.array_equal(small["len"].cpu().numpy(), ref["len"])
    for k in ("x", "y"):
        got = small[k].cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(ref[k])), k
        assert np.nanmax(np.abs(got - ref[k])) <= 1e-8, k
    reps = -(-n_rays // BASE)
    ic = [t.repeat(reps)[:n_rays].contiguous() for t in base]
    print(f"{n_rays} rays: initial conditions built", flush=True)
    big = core.trace_many_device(b, c, *ic, 0.0, 3.0, 1.0)
    torch.cuda.synchronize()
    del ic
    idx = np.unique(np.concatenate([np.arange(0, n_rays, 1_000_003), np.arange(n_rays - 300, n_rays)]))
    it = torch.tensor(idx, device=dev)
    src = idx % BASE
    st = torch.tensor(src, device=dev)
    assert np.array_equal(big["len"][it].cpu().numpy(), small["len"][st].cpu().numpy())
    assert np.array_equal(_bits(big["end"][it].cpu().numpy()), _bits(small["end"][st].cpu().numpy()))
    for k in ("x", "y", "kx", "ky"):
        assert np.array_equal(_bits(big[k][:, it].cpu().numpy()),
                              _bits(small[k][:, st].cpu().numpy())), k
    print(f"{n_rays} rays: {idx.size} sampled rays equal the 4,096-ray launch bitwise", flush=True)
    del big
    torch.cuda.empty_cache()
