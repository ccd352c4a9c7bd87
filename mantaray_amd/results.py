"""The reference's result writers over trace results.

  ray_result_json(r, ray)        RayResult::from + as_json (src/ray_result.rs:74-151)
  save_ray_result(path, r, ray)  RayResult::save_file (src/ray_result.rs:116-120)
  append_tsv(path, r, ray)       output_or_append_to_tsv_file (src/ray.rs:218-243)

`r` is a trace_many result (t, x, y, kx, ky step-major, with trajectories),
or a single_ray-style table of rows (t, x, y, kx, ky) with ray=None.  The
text is produced by libmantaray_hip.so (host code; no GPU needed).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


def _columns(r, ray):
    if ray is None:
        rows = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1, 5).T)
        return [np.ascontiguousarray(c) for c in rows], 1, rows.shape[1]
    t = np.ascontiguousarray(r["t"], dtype=np.float64)
    cols = [np.ascontiguousarray(r[k], dtype=np.float64) for k in ("x", "y", "kx", "ky")]
    nr = cols[0].shape[1]
    if not 0 <= ray < nr:
        raise IndexError(f"ray {ray} of {nr}")
    # pointers to the ray's column of each step-major array, stride n_rays
    return [t] + [c.ravel()[ray:] for c in cols], nr, t.size


def ray_result_json(r, ray: int | None = None) -> str:
    """The JSON text of one ray: rows up to the first with a NaN component."""
    cols, stride, n = _columns(r, ray)
    ln = ctypes.c_int64()
    ptrs = [c.ctypes.data for c in cols]
    C.check(C.lib().mr_ray_result_json(*ptrs, n, stride, None, 0, ctypes.byref(ln)))
    buf = ctypes.create_string_buffer(ln.value + 1)
    C.check(C.lib().mr_ray_result_json(*ptrs, n, stride, buf, ln.value + 1, ctypes.byref(ln)))
    return buf.value.decode()


def save_ray_result(path, r, ray: int | None = None) -> int:
    """Write the JSON text to path; returns the bytes written."""
    s = ray_result_json(r, ray)
    with open(path, "w") as f:
        f.write(s)
    return len(s.encode())


def append_tsv(path, r, ray: int | None = None) -> None:
    """Append one ray's rows (up to the first NaN x) to a TSV file."""
    cols, stride, n = _columns(r, ray)
    C.check(C.lib().mr_ray_result_append_tsv(str(path).encode(), *[c.ctypes.data for c in cols],
                                             n, stride))
