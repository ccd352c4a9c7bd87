"""The build's packed-FP32 guard: the device pass honours the target feature
-packed-fp32-ops, and the post-compile disassembly check finds packed FP32
arithmetic when it is present.  A small kernel whose float2 arithmetic the
compiler packs is built with and without the feature."""
import os
import subprocess

import pytest

from mantaray_amd import build as B

KERNEL = r"""
#include <hip/hip_runtime.h>
__global__ void k(float2 *a, float s)
{
    float2 v = a[threadIdx.x];
    v.x = v.x * s + 1.0f;
    v.y = v.y * s + 2.0f;
    a[threadIdx.x] = v;
}
"""


def _object(tmp_path, with_feature: bool) -> str:
    src = tmp_path / "k.hip"
    src.write_text(KERNEL)
    obj = str(tmp_path / ("with.o" if with_feature else "without.o"))
    fl = [f"--offload-arch={B.ARCH}", "-O3", "-c", "-o", obj, str(src)]
    if with_feature:
        fl = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + fl
    subprocess.run([B.hipcc(), *fl], check=True, capture_output=True, text=True)
    return obj


def test_guard_detects_packed_fp32(tmp_path):
    obj = _object(tmp_path, with_feature=False)
    with pytest.raises(RuntimeError, match="packed FP32"):
        B.check_no_packed_fp32(obj)


def test_feature_removes_packed_fp32(tmp_path):
    obj = _object(tmp_path, with_feature=True)
    assert B.check_no_packed_fp32(obj) > 0


def test_product_flags_carry_the_feature():
    fl = B.flags()
    i = fl.index("-packed-fp32-ops")
    assert fl[i - 3:i] == ["-Xclang", "-target-feature", "-Xclang"]


@pytest.mark.skipif(not os.path.isdir(B.OBJ), reason="library not built in this tree")
def test_built_objects_have_no_packed_fp32():
    objs = [os.path.join(B.OBJ, f) for f in os.listdir(B.OBJ) if f.endswith(".o")]
    assert objs
    assert sum(B.check_no_packed_fp32(o) for o in objs) > 0
