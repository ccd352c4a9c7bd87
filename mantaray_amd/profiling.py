"""HBM traffic of the trace kernel from rocprofv3 counter passes.

    python -m mantaray_amd.profiling traffic --fetch DIR --write DIR \\
        --copy-fetch DIR --copy-write DIR --ray-steps N --out FILE.json

Each DIR is the output directory of one `rocprofv3 --pmc <counter> -f csv`
run (one counter per run: FETCH_SIZE, then WRITE_SIZE, each for the traced
program and for a 1 GiB device-to-device copy).  The counters read KiB per
dispatch; each is scaled by what it reads for the copy, whose bytes moved are
known (FETCH_SIZE reads half of the copy's bytes on MI355X, WRITE_SIZE all of
them).  The summary is stamped with a hash of the sources the trace kernel
is compiled from, and `load_traffic` returns it only while that hash still
matches: a figure measured on another kernel is never reported for this one.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import statistics

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
# the translation unit of the trace kernel and the headers it includes
TRACE_SOURCES = ("mr_kernels.hip", "mr_kernels.h", "mr_kernel_common.h", "mr_device.h")
COPY_BYTES = 1 << 30                   # the calibration copy: 2^27 doubles


def source_hash(files=TRACE_SOURCES, root: str = CSRC) -> str:
    """sha256 (first 16 hex digits) over the named source files, in order."""
    h = hashlib.sha256()
    for name in files:
        with open(os.path.join(root, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read() + b"\0")
    return h.hexdigest()[:16]


def _counter_csv(d: str) -> str:
    found = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    if not found:
        raise FileNotFoundError(f"no rocprofv3 counter_collection.csv under {d}")
    return found[0]


def per_dispatch(d: str, counter: str, kernel: str) -> list[float]:
    """The counter's value (summed over its instances) for every dispatch of
    the kernels whose name contains `kernel`, in dispatch order."""
    sums: dict[int, float] = {}
    with open(_counter_csv(d), newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter or kernel not in row["Kernel_Name"]:
                continue
            i = int(row["Dispatch_Id"])
            sums[i] = sums.get(i, 0.0) + float(row["Counter_Value"])
    if not sums:
        raise ValueError(f"no {counter} rows for a kernel named *{kernel}* under {d}")
    return [sums[i] for i in sorted(sums)]


def traffic_summary(fetch: str, write: str, copy_fetch: str, copy_write: str, ray_steps: int,
                    kernel: str = "trace_kernel", copy_kernel: str = "copyBuffer",
                    commands: dict | None = None) -> dict:
    f_k = per_dispatch(fetch, "FETCH_SIZE", kernel)
    w_k = per_dispatch(write, "WRITE_SIZE", kernel)
    cf = statistics.median(per_dispatch(copy_fetch, "FETCH_SIZE", copy_kernel))
    cw = statistics.median(per_dispatch(copy_write, "WRITE_SIZE", copy_kernel))
    # bytes the copy moved / bytes the counter reported for it
    f_scale = COPY_BYTES / (cf * 1024.0)
    w_scale = COPY_BYTES / (cw * 1024.0)
    fetch_b = statistics.median(f_k) * 1024.0 * f_scale
    write_b = statistics.median(w_k) * 1024.0 * w_scale
    return {
        "kernel_match": kernel,
        "source_hash": source_hash(),
        "sources": list(TRACE_SOURCES),
        "fetch_kib_per_dispatch": f_k,
        "write_kib_per_dispatch": w_k,
        "copy_calibration": {"bytes_moved_per_dispatch": COPY_BYTES, "fetch_kib": cf,
                             "write_kib": cw, "fetch_factor": f_scale,
                             "write_factor": w_scale},
        "fetch_bytes_per_launch": fetch_b,
        "write_bytes_per_launch": write_b,
        "hbm_bytes_per_launch": fetch_b + write_b,
        "ray_steps_per_launch": ray_steps,
        "fetch_bytes_per_ray_step": fetch_b / ray_steps,
        "commands": commands or {},
    }


def load_traffic(path: str) -> dict | None:
    """The summary at `path` if it was measured on the current sources."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("source_hash") != source_hash(t.get("sources", TRACE_SOURCES)):
        return None
    return t


def main(argv=None):
    p = argparse.ArgumentParser(prog="python -m mantaray_amd.profiling")
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic", help="HBM bytes per launch from FETCH/WRITE_SIZE passes")
    for name in ("fetch", "write", "copy-fetch", "copy-write"):
        t.add_argument(f"--{name}", required=True)
    t.add_argument("--ray-steps", type=int, required=True)
    t.add_argument("--kernel", default="trace_kernel")
    t.add_argument("--command", action="append", default=[],
                   help="NAME=COMMAND line that made a pass (repeatable)")
    t.add_argument("--out", required=True)
    h = sub.add_parser("hash", help="print the trace kernel's source hash")
    del h
    a = p.parse_args(argv)
    if a.cmd == "hash":
        print(source_hash())
        return
    cmds = dict(c.split("=", 1) for c in a.command)
    s = traffic_summary(a.fetch, a.write, a.copy_fetch, a.copy_write, a.ray_steps,
                        kernel=a.kernel, commands=cmds)
    with open(a.out, "w") as f:
        json.dump(s, f, indent=1)
        f.write("\n")
    print(json.dumps({k: s[k] for k in ("source_hash", "hbm_bytes_per_launch",
                                        "fetch_bytes_per_ray_step")}))


if __name__ == "__main__":
    main()
