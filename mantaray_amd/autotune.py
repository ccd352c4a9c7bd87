"""Autotuner: times candidate launch settings of a GPU operation on its own
inputs and keeps the fastest, cached per device and workload key.

    tuner = Autotuner()
    best, table = tuner.tune("batch:seed2", [{"block_size": 64, "blocks_per_cu": 1}, ...],
                             lambda **c: batch.run(**c))

Each candidate runs `warmup` times untimed and `reps` times between HIP
events on the current stream; the median decides.  Results are kept in a
JSON file (default ~/.cache/mantaray_amd/autotune.json, or the path in
MANTARAY_AUTOTUNE_CACHE) under "<device name>|<key>", so a later process
with the same key skips the search; `force=True` searches again.
"""
from __future__ import annotations

import json
import os


def _default_path() -> str:
    return os.environ.get("MANTARAY_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mantaray_amd",
                                       "autotune.json"))


class Autotuner:
    def __init__(self, cache_path: str | None = None):
        self.path = cache_path or _default_path()
        self._cache = None

    def _load(self) -> dict:
        if self._cache is None:
            try:
                with open(self.path) as f:
                    self._cache = json.load(f)
            except (OSError, ValueError):
                self._cache = {}
        return self._cache

    def _save(self) -> None:
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self._cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass                       # a read-only home only loses the cache

    @staticmethod
    def _time(run, cand, reps, warmup) -> float:
        import numpy as np
        import torch
        for _ in range(warmup):
            run(**cand)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            run(**cand)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    def tune(self, key: str, candidates: list, run, reps: int = 2, warmup: int = 1,
             force: bool = False):
        """Returns (best candidate, [(candidate, ms), ...] of this search, or
        the cached entry's table)."""
        import torch
        full = f"{torch.cuda.get_device_name()}|{key}"
        cache = self._load()
        if not force and full in cache:
            ent = cache[full]
            return ent["best"], [(c, t) for c, t in ent["table"]]
        table = [(dict(c), self._time(run, c, reps, warmup)) for c in candidates]
        best = min(table, key=lambda ct: ct[1])[0]
        cache[full] = {"best": best, "table": table}
        self._save()
        return best, table
