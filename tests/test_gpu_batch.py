"""Batched persistent trace (mr_batch_*) on the GPU.

Scenes of different grid sizes, steps, step counts and ray counts (one with
no rays), most rays leaving their grid early:
  - every ray's row count and last row are bitwise those of mr_trace_many
    on its own scene, for both solout variants and several launch shapes;
  - a batch captured in a HIP graph (torch.cuda.CUDAGraph) and replayed,
    with eager work between the replays, gives the eager results each time;
  - the oracle agrees on a sample of rays (row counts equal, endpoints
    within 1e-6 m);
  - scenes of different field kinds are refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(mr, seed=5, shift=0.0):
    """shift: the current's axes moved by that many cells from the depth's
    (0: the two fields share their cells)."""
    from mantaray_amd import workloads as W
    rs = W.Stream(seed)
    specs = [(64, 300, 700, 1.0), (128, 0, 200, 0.5), (96, 1500, 900, 2.0),
             (200, 4000, 400, 1.25), (64, 77, 1200, 0.75), (160, 2500, 600, 1.5)]
    scenes, cfgs = [], []
    for n, nr, steps, dt in specs:
        dx = 250.0
        xd = np.arange(n) * dx
        depth = np.maximum(W.gaussian_random_field(n, dx, rs, 2000.0, 800.0, 8000.0), 1.0)
        u, v = W.eddies(xd, xd, rs, n_eddies=5, radius=(3000.0, 8000.0))
        x0 = rs.uniform(nr, dx, xd[-1] - dx)
        y0 = rs.uniform(nr, dx, xd[-1] - dx)
        th = rs.uniform(nr, 0.0, 2 * np.pi)
        k = W.wavenumber(2 * np.pi / rs.uniform(nr, 6.0, 16.0), np.full(nr, 2000.0))
        cfg = {"bathy": {"x": xd.astype(np.float32), "y": xd.astype(np.float32), "depth": depth},
               "current": {"x": xd + shift * dx, "y": xd + shift * dx, "u": u, "v": v},
               "x0": x0, "y0": y0, "kx0": k * np.cos(th), "ky0": k * np.sin(th),
               "t0": 0.0, "t_end": steps * dt, "dt": dt}
        b, c = W.make_fields(cfg)
        cfgs.append(cfg)
        scenes.append({"bathymetry": b, "current": c, **{k: cfg[k] for k in
                       ("x0", "y0", "kx0", "ky0", "t0", "t_end", "dt")}})
    return scenes, cfgs


@pytest.fixture(scope="module")
def mr():
    import mantaray_amd
    mantaray_amd._capi.lib()
    return mantaray_amd


@pytest.mark.parametrize("solout", ["state", "next_deriv"])
def test_batch_matches_per_scene_traces(mr, solout):
    import torch
    from mantaray_amd.batch import Batch
    scenes, cfgs = _scenes(mr)
    batch = Batch(scenes, solout=solout)
    total, bound = batch.info()
    assert total == sum(s["x0"].size for s in scenes)
    refs = []
    for sc in scenes:
        refs.append(mr.trace_many(sc["bathymetry"], sc["current"], sc["x0"], sc["y0"], sc["kx0"],
                                  sc["ky0"], sc["t0"], sc["t_end"], sc["dt"], solout=solout,
                                  trajectories=False))
    early = sum(int((r["len"] < round(sc["t_end"] / sc["dt"]) + 1).sum())
                for r, sc in zip(refs, scenes))
    assert early > total // 4                    # lifetimes are uneven
    for block, per_cu in ((0, 0), (64, 1), (128, 2), (256, 4)):
        for out in batch.results():
            out["len"].fill_(-1)
            out["end"].fill_(-7.0)
        batch.run(block_size=block, blocks_per_cu=per_cu)
        torch.cuda.synchronize()
        for i, (out, ref) in enumerate(zip(batch.results(), refs)):
            assert np.array_equal(out["len"].cpu().numpy(), ref["len"]), (block, per_cu, i)
            assert np.array_equal(out["end"].cpu().numpy(), ref["end"], equal_nan=True), (block, i)
    batch.close()


def test_batch_graph_replay(mr):
    import torch
    from mantaray_amd.batch import Batch
    scenes, _ = _scenes(mr, seed=6)
    batch = Batch(scenes)
    batch.run()
    torch.cuda.synchronize()
    eager = [{k: v.clone() for k, v in o.items()} for o in batch.results()]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        batch.run(stream=s)                      # warm the stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        batch.run(stream=s)
    for rep in range(3):
        for o in batch.results():
            o["len"].fill_(-1)
            o["end"].fill_(-7.0)
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for o, e in zip(batch.results(), eager):
            assert torch.equal(o["len"], e["len"]), rep
            assert torch.equal(torch.nan_to_num(o["end"], nan=1e300),
                               torch.nan_to_num(e["end"], nan=1e300)), rep
    batch.close()


def test_batch_record_tiles_and_tile_tuning(mr, tmp_path):
    """The record tile shape moves cell records, not results: batches with
    row-major, 4 x 4 and 32 x 32 tiles give the 8 x 8 default's results
    bitwise, and after the autotuner's tile search the batch keeps one
    handle, of the chosen shape, with the same results."""
    import torch
    from mantaray_amd.autotune import Autotuner
    from mantaray_amd.batch import Batch
    scenes, _ = _scenes(mr, seed=7)
    ref = Batch(scenes)
    ref.run()
    torch.cuda.synchronize()
    want = [{k: v.cpu().numpy() for k, v in o.items()} for o in ref.results()]
    ref.close()

    def check(batch, what):
        for o in batch.results():
            o["len"].fill_(-1)
            o["end"].fill_(-7.0)
        batch.run()
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(batch.results(), want)):
            assert np.array_equal(o["len"].cpu().numpy(), w["len"]), (what, i)
            assert np.array_equal(o["end"].cpu().numpy(), w["end"], equal_nan=True), (what, i)

    for tile in (0, 2, 5):
        b = Batch(scenes, record_tile=tile)
        assert b.record_tile == tile
        check(b, tile)
        b.close()
    b = Batch(scenes)
    best, table, tile_table = b.autotune("tiles", tuner=Autotuner(str(tmp_path / "at.json")),
                                         force=True, tiles=(2, 3, 4))
    assert len(table) == 9 and [c["record_tile"] for c, _ in tile_table] == [2, 3, 4]
    assert b.record_tile == min(tile_table, key=lambda ct: ct[1])[0]["record_tile"]
    check(b, "tuned")
    b.close()


@pytest.mark.parametrize("shift", [0.0, 0.5])
def test_batch_against_oracle(mr, oracle_mod, shift):
    """Endpoints against the oracle, with the depth and current on the same
    cells (one prefetch test for both) and on cells half a cell apart."""
    import torch
    from _helpers import oracle_fields
    from mantaray_amd.batch import Batch
    scenes, cfgs = _scenes(mr, seed=7, shift=shift)
    batch = Batch(scenes)
    batch.run()
    torch.cuda.synchronize()
    for out, cfg in zip(batch.results(), cfgs):
        sub = slice(0, None, 17)
        if cfg["x0"][sub].size == 0:
            continue
        ob, oc = oracle_fields(oracle_mod, cfg)
        ref = oracle_mod.trace_many(ob, oc, cfg["x0"][sub], cfg["y0"][sub], cfg["kx0"][sub],
                                    cfg["ky0"][sub], cfg["t0"], cfg["t_end"], cfg["dt"],
                                    trajectories=False)
        assert np.array_equal(out["len"].cpu().numpy()[sub], ref["len"])
        e = out["end"].cpu().numpy()[sub]
        fin = ~np.isnan(ref["end"][:, 0])
        assert np.array_equal(np.isnan(e[:, 0]), ~fin)
        if fin.any():
            assert np.abs(e[fin, :2] - ref["end"][fin, :2]).max() <= 1e-6


def test_batch_refuses_mixed_kinds(mr):
    from mantaray_amd.batch import Batch
    scenes, _ = _scenes(mr, seed=8)
    scenes[1] = dict(scenes[1], bathymetry=mr.Bathymetry.constant(100.0))
    with pytest.raises(mr._capi.MantarayError):
        Batch(scenes[:3])
