"""Build recipe for libmantaray_hip.so (gfx950), in-tree.

    python -m mantaray_amd.build

compiles the HIP kernels, the NetCDF-3 reader and the C ABI with hipcc into
mantaray_amd/libmantaray_hip.so.  Floating-point flags: -ffp-contract=off on
host and device (no FMA contraction, as rustc never contracts), no fast-math,
IEEE f32 division/sqrt and f32 denormals (the HIP defaults).  Packed FP32
instructions are disabled for the device code.  The machine scheduler uses the
max-ilp strategy: the integrators run one wave per SIMD, so nothing but the
wave's own instruction-level parallelism hides its latencies.  The post-RA
list scheduler is turned off (-disable-post-ra): it reorders the max-ilp
schedule after register allocation, and without it the seamount-jet forward
trace runs about 4 % faster (10.4 against 10.85 ms per launch), the adjoint
pair and the batch kernel 1-2 %.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmantaray_hip.so")
SOURCES = ["mr_kernels.hip", "mr_adjoint.hip", "mr_capi.cpp", "mr_netcdf3.cpp", "mr_results.cpp"]
ARCH = os.environ.get("MANTARAY_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-mcode-object-version=5", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-mllvm", "-amdgpu-sched-strategy=max-ilp", "-mllvm", "-disable-post-ra",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    ]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *flags(), "-shared", "-o", tmp,
           *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = [ln for ln in res.stderr.splitlines()
            if "unknown target feature" not in ln and "packed-fp32-ops" not in ln
            and ln.strip() and not ln.startswith("1 warning")]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    if verbose and errs:
        print("\n".join(errs), file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
