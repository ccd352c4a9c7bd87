"""The torch-facing entry points check every tensor whose pointer they pass
to a kernel (trace_many_device's initial conditions and `out` buffers,
trace_fp64's initial conditions, adjoint's forward record and cotangents).
A wrong size, dtype, device or layout raises before anything is launched."""
import numpy as np
import pytest
import torch

from mantaray_amd import adjoint, core, workloads


def _fields():
    cfg = workloads.linear_beach(n_rays=8)
    return cfg, workloads.make_fields(cfg)


def test_host_tensors_are_refused():
    cfg, (b, c) = _fields()
    ics = [torch.tensor(cfg[k][:8]) for k in ("x0", "y0", "kx0", "ky0")]
    with pytest.raises(TypeError, match="GPU"):
        core.trace_many_device(b, c, *ics, 0.0, 10.0, 1.0)
    with pytest.raises(TypeError, match="GPU"):
        adjoint.trace_fp64(b, c, *ics, 0.0, 10.0, 1.0)


def test_non_tensor_initial_conditions_are_refused():
    cfg, (b, c) = _fields()
    with pytest.raises(TypeError):
        core.trace_many_device(b, c, *[cfg[k] for k in ("x0", "y0", "kx0", "ky0")], 0.0, 10.0, 1.0)


@pytest.mark.gpu
def test_wrong_out_buffers_are_refused():
    cfg, (b, c) = _fields()
    dev = torch.device("cuda:0")
    ics = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    nr, n = ics[0].numel(), 10
    good = {"t": torch.empty(n + 1, dtype=torch.float64, device=dev),
            "len": torch.empty(nr, dtype=torch.int32, device=dev),
            "end": torch.empty((nr, 4), dtype=torch.float64, device=dev),
            "x": torch.empty((n + 1, nr), dtype=torch.float64, device=dev)}
    cases = [
        ("x", torch.empty((n, nr), dtype=torch.float64, device=dev), ValueError),     # a row short
        ("x", torch.empty((n + 1, nr - 1), dtype=torch.float64, device=dev), ValueError),
        ("x", torch.empty((n + 1, nr), dtype=torch.float32, device=dev), TypeError),
        ("x", torch.empty((nr, n + 1), dtype=torch.float64, device=dev).t(), ValueError),
        ("len", torch.empty(nr, dtype=torch.int64, device=dev), TypeError),
        ("end", torch.empty((nr, 4), dtype=torch.float64), TypeError),                # host
        ("t", torch.empty(n, dtype=torch.float64, device=dev), ValueError),
    ]
    sentinel = torch.full((n + 1, nr), 7.0, dtype=torch.float64, device=dev)
    for k, bad, exc in cases:
        out = dict(good)
        out["y"] = sentinel
        out[k] = bad
        with pytest.raises(exc):
            core.trace_many_device(b, c, *ics, 0.0, float(n), 1.0, out=out)
    with pytest.raises(ValueError, match="unknown output"):
        core.trace_many_device(b, c, *ics, 0.0, float(n), 1.0, out={**good, "z": good["x"]})
    torch.cuda.synchronize()
    # nothing was launched: the sentinel buffer is untouched
    assert bool((sentinel == 7.0).all())
    # the good buffers work
    r = core.trace_many_device(b, c, *ics, 0.0, float(n), 1.0, out=dict(good))
    torch.cuda.synchronize()
    assert np.all(r["len"].cpu().numpy() >= 1)


@pytest.mark.gpu
def test_adjoint_checks_its_forward_record():
    cfg, (b, c) = _fields()
    dev = torch.device("cuda:0")
    ics = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    f = adjoint.trace_fp64(b, c, *ics, 0.0, 20.0, 1.0, stride=5)
    with pytest.raises(ValueError):
        adjoint.adjoint(b, c, f, 0.0, 20.0, 1.0, 4)          # another stride, other ckpt count
    bad = dict(f)
    bad["end"] = f["end"][:-1]
    with pytest.raises(ValueError):
        adjoint.adjoint(b, c, bad, 0.0, 20.0, 1.0, 5)
    with pytest.raises(ValueError):
        adjoint.adjoint(b, c, f, 0.0, 20.0, 1.0, 5,
                        g_end=torch.zeros((ics[0].numel() + 1, 4), dtype=torch.float64, device=dev))
    g = adjoint.adjoint(b, c, f, 0.0, 20.0, 1.0, 5, g_end=torch.ones_like(f["end"]))
    torch.cuda.synchronize()
    assert torch.isfinite(g["x0"]).all()
