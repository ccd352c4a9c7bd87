"""Many scenes in one persistent launch (mr_batch_create / mr_batch_run).

    b = Batch([{"bathymetry": B1, "current": C1, "x0": .., "y0": .., "kx0": ..,
                "ky0": .., "t0": 0.0, "t_end": 5000.0, "dt": 1.0}, ...])
    b.run()                      # enqueued on torch's current stream
    res = b.results()            # per scene: {"len": (n,), "end": (n, 4)}

record_tile (keyword of Batch) sets the record tile shape of the gridded
fields; autotune() picks the launch shape and then the tile shape.  Each
scene has its own fields, step, step count and rays; every ray's row
count and last row are what trace_many gives for its scene.  The launch takes
rays from one queue, so lanes freed by rays that end early (edge exits,
short scenes) take new rays at once.  run() allocates nothing and does not
synchronise, so it can be captured in a torch.cuda.CUDAGraph and replayed.
The reference traces one scene per call (ManyRays::trace_many,
src/ray.rs:98-127); batches have no counterpart there.
"""
from __future__ import annotations

import ctypes

from . import _capi as C
from .core import _MATH, _SOLOUT


class Batch:
    def __init__(self, scenes: list, *, math: str = "fast", solout: str = "state",
                 device=None, record_tile: int = -1):
        import torch
        dev = (torch.device("cuda", torch.cuda.current_device()) if device is None
               else torch.device(device))
        self.device = dev
        self._keep = []
        self.outputs = []
        arr = (C.Scene * max(1, len(scenes)))()
        for i, sc in enumerate(scenes):
            ics = [torch.as_tensor(sc[k], dtype=torch.float64, device=dev).contiguous()
                   for k in ("x0", "y0", "kx0", "ky0")]
            nr = min(t.numel() for t in ics)
            out = {"len": torch.empty(nr, dtype=torch.int32, device=dev),
                   "end": torch.empty((nr, 4), dtype=torch.float64, device=dev)}
            self._keep.append((sc["bathymetry"], sc["current"], ics))
            self.outputs.append(out)
            s = arr[i]
            s.bathy = sc["bathymetry"].handle.value
            s.current = sc["current"].handle.value
            s.n_rays = nr
            s.x0, s.y0, s.kx0, s.ky0 = (t.data_ptr() for t in ics)
            s.t0, s.t_end, s.dt = float(sc["t0"]), float(sc["t_end"]), float(sc["dt"])
            s.len_out = out["len"].data_ptr()
            s.end_out = out["end"].data_ptr()
        self._math, self._solout = _MATH[math], _SOLOUT[solout]
        self._arr, self._n = arr, len(scenes)
        self._h = None
        self._h = self._create(record_tile)
        self.record_tile = 3 if record_tile == -1 else record_tile   # -1: 8 x 8

    def _dev_index(self) -> int:
        return self.device.index if self.device.index is not None else -1

    def _create(self, record_tile: int):
        """A batch handle over the same scenes with cell records in tiles of
        2^record_tile cells a side (-1 = 8 x 8)."""
        o = C.options(math=self._math, solout=self._solout, memory=C.MR_MEM_DEVICE,
                      device=self._dev_index(), record_tile=record_tile)
        h = ctypes.c_void_p()
        C.check(C.lib().mr_batch_create(self._n, self._arr, ctypes.byref(o), ctypes.byref(h)))
        return h

    def info(self) -> tuple[int, int]:
        """(total rays, upper bound of the ray-steps)."""
        n, m = ctypes.c_int64(), ctypes.c_int64()
        C.check(C.lib().mr_batch_info(self._h, ctypes.byref(n), ctypes.byref(m)))
        return n.value, m.value

    tuned = None

    def run(self, stream=None, block_size: int = 0, blocks_per_cu: int = 0,
            _handle=None) -> None:
        import torch
        if not block_size and not blocks_per_cu and self.tuned:
            block_size, blocks_per_cu = self.tuned["block_size"], self.tuned["blocks_per_cu"]
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        o = C.options(math=self._math, solout=self._solout, memory=C.MR_MEM_DEVICE,
                      device=self._dev_index(), block_size=block_size,
                      blocks_per_cu=blocks_per_cu, stream=ctypes.c_void_p(stream.cuda_stream))
        C.check(C.lib().mr_batch_run(_handle or self._h, ctypes.byref(o)))

    def autotune(self, key: str, force: bool = False, tuner=None, tiles=(2, 3, 4, 5)):
        """Pick block size and workgroups per CU for this batch by timing the
        candidates (mantaray_amd.autotune), then, with that launch shape, the
        record tile shape among `tiles` (log2 of the tile edge); later run()
        calls use the choice.  Returns (launch shape, its search table, tile
        search table)."""
        from .autotune import Autotuner
        tuner = tuner or Autotuner()
        cands = [{"block_size": b, "blocks_per_cu": p}
                 for b in (64, 128, 256) for p in (1, 2, 4) if b * p <= 1024]
        best, table = tuner.tune(f"batch|{key}", cands, lambda **c: self.run(**c), force=force)
        self.tuned = best
        tile_table = []
        if tiles:
            handles = {}

            def run_tile(record_tile):
                h = handles.get(record_tile)
                if h is None:
                    h = handles[record_tile] = self._create(record_tile)
                self.run(_handle=h)

            tbest, tile_table = tuner.tune(f"batch-tile|{key}",
                                           [{"record_tile": t} for t in tiles], run_tile,
                                           force=force)
            chosen = tbest["record_tile"]
            if chosen != self.record_tile:
                h = handles.pop(chosen, None) or self._create(chosen)
                C.lib().mr_batch_free(self._h)
                self._h, self.record_tile = h, chosen
            for h in handles.values():
                if h.value != self._h.value:
                    C.lib().mr_batch_free(h)
        return best, table, tile_table

    def results(self) -> list:
        return self.outputs

    def close(self) -> None:
        if self._h:
            C.lib().mr_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
