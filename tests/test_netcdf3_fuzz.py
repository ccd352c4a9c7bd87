"""The NetCDF-3 reader on randomly damaged files, under sanitizers.

The reader (mantaray_amd/csrc/mr_netcdf3.cpp, plain C++) is built here with
g++ and -fsanitize=address,undefined into a small driver
(tests/nc3/fuzz_driver.cpp) that opens each file and reads every variable a
field open reads.  Three well-formed files (a CDF-1 bathymetry, a CDF-2
current, and a file with record variables) are damaged 1,000 times each:
bytes replaced (mostly in the header, where the sizes and offsets are),
header words set to extreme sizes, files cut short or extended.  Every file
must end in a status (success or a format or I/O error) with no sanitizer
report, no exception and no abort.  Host code only; no GPU is involved.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from _ncfix import create_bathymetry, create_current, four_depth_fn, simple_current

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mantaray_amd", "csrc")


def _record_file(path):
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "w", version=1) as f:
        f.createDimension("y", None)
        f.createDimension("x", 5)
        f.createVariable("x", "f4", ("x",))[:] = np.arange(5, dtype=np.float32) * 10
        yv = f.createVariable("y", "f4", ("y",))
        dv = f.createVariable("depth", "f8", ("y", "x"))
        yv[:4] = np.arange(4, dtype=np.float32) * 10
        dv[:4, :] = np.arange(20, dtype=np.float64).reshape(4, 5) + 10


def _mutate(data: bytes, rng) -> bytes:
    b = bytearray(data)
    head = min(len(b), 200)
    kind = rng.integers(0, 5)
    if kind == 0:                                  # a few bytes replaced
        for _ in range(int(rng.integers(1, 5))):
            i = int(rng.integers(0, head if rng.random() < 0.8 else len(b)))
            b[i] = int(rng.integers(0, 256))
    elif kind == 1:                                # a header word set to an extreme size
        i = int(rng.integers(0, max(1, head // 4))) * 4
        word = int(rng.choice([0xFFFFFFFF, 0x7FFFFFFF, 0x80000000, 0x10000000, 1 << 20,
                               int(rng.integers(0, 1 << 32))]))
        b[i:i + 4] = word.to_bytes(4, "big")
    elif kind == 2:                                # cut short
        b = b[:int(rng.integers(0, len(b)))]
    elif kind == 3:                                # extended with garbage
        b += bytes(rng.integers(0, 256, int(rng.integers(1, 64)), dtype=np.uint8))
    else:                                          # one bit flipped in the header
        i = int(rng.integers(0, head))
        b[i] ^= 1 << int(rng.integers(0, 8))
    return bytes(b)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_damaged_files_under_sanitizers(tmp_path):
    exe = tmp_path / "fuzz_driver"
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I", CSRC,
         os.path.join(CSRC, "mr_netcdf3.cpp"), os.path.join(ROOT, "tests", "nc3", "fuzz_driver.cpp"),
         "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    create_bathymetry(tmp_path / "b1.nc", 6, 5, 10.0, 20.0, four_depth_fn, version=1)
    create_current(tmp_path / "c2.nc", 5, 4, 10.0, 10.0, simple_current, version=2)
    _record_file(tmp_path / "r1.nc")
    bases = [(tmp_path / n).read_bytes() for n in ("b1.nc", "c2.nc", "r1.nc")]
    rng = np.random.default_rng(20240601)
    files = [str(tmp_path / n) for n in ("b1.nc", "c2.nc", "r1.nc")]
    for k, base in enumerate(bases):
        for i in range(1000):
            p = tmp_path / f"m{k}_{i}.nc"
            p.write_bytes(_mutate(base, rng))
            files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(exe), *files], capture_output=True, text=True, timeout=600,
                         env=env)
    lines = res.stdout.strip().splitlines()
    assert res.returncode == 0, (f"stopped after {len(lines)} of {len(files)} files "
                                 f"(next: {files[len(lines)] if len(lines) < len(files) else '-'}): "
                                 f"{res.stderr[-2000:]}")
    assert "runtime error" not in res.stderr and "Sanitizer" not in res.stderr, res.stderr[-2000:]
    assert len(lines) == len(files)
    status = [int(ln.split()[1]) for ln in lines]
    assert status[:3] == [0, 0, 0]                 # the undamaged files read
    assert set(status) <= {0, -2, -3}
    reads = [int(ln.split()[2]) for ln in lines[:3]]
    assert reads == [6, 8, 6]                      # x, y and depth / x, y, u and v, twice each
    print(f"{len(files)} files: {status.count(0)} read, {status.count(-3)} format errors, "
          f"{status.count(-2)} I/O errors")
