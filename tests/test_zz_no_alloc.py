"""mr_adjoint_many_noalloc against mr_adjoint_many, through the C ABI.

The no-allocation variant runs the same reverse sweep with its work buffer
in the library's static device array, builds no field tables, and refuses
launches whose work buffer would not fit.  Two calls in a row, the second
with fewer rays and another stride (the array keeps the first call's
contents), must give the initial-state gradients bitwise and the grid
gradients (sums of f64 atomics, whose order varies) within 1e-12 relative of
mr_adjoint_many's on the same inputs, and device memory in use must not
change across the calls.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return None if t is None else t.data_ptr()


def _case(torch, n_rays, n_steps, seed):
    from mantaray_amd import workloads
    cfg = workloads.eddy_adjoint(n_rays=n_rays, n_steps=n_steps, seed=seed)
    b, c = workloads.make_fields(cfg)
    dev = torch.device("cuda:0")
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    u = torch.tensor(cfg["current"]["u"].ravel(), device=dev)
    v = torch.tensor(cfg["current"]["v"].ravel(), device=dev)
    return cfg, b, c, ic, u, v


def _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc, block_size=0, used=None):
    """One call of either entry point; returns the gradients.  With `used`,
    appends how far the device's free memory fell across the call itself
    (the outputs are allocated before it)."""
    from mantaray_amd.adjoint import _device_opts
    dev = fwd["end"].device
    nr = fwd["len"].numel()
    f64 = dict(dtype=torch.float64, device=dev)
    out = {k: torch.full((nr,), np.nan, **f64) for k in ("x0", "y0", "kx0", "ky0")}
    out["u"] = torch.full((u.numel(),), np.nan, **f64)
    out["v"] = torch.full((v.numel(),), np.nan, **f64)
    o = _device_opts(dev, None, "fast", block_size)
    common = [b.handle, c.handle, None, _ptr(u), _ptr(v), nr, 0.0, float(T), 1.0, int(stride),
              ctypes.byref(o), _ptr(fwd["ckpt"]), _ptr(fwd["len"]), _ptr(fwd["end"]),
              _ptr(g_end), None, None, None, None]
    grads = [_ptr(out[k]) for k in ("x0", "y0", "kx0", "ky0")] + [None, _ptr(out["u"]),
                                                                  _ptr(out["v"])]
    if noalloc:
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        C.check(C.lib().mr_adjoint_many_noalloc(*common, *grads))
        torch.cuda.synchronize()
        if used is not None:
            used.append(free0 - torch.cuda.mem_get_info()[0])
    else:
        nbytes = ctypes.c_int64()
        C.check(C.lib().mr_adjoint_workspace_size(nr, stride, ctypes.byref(nbytes)))
        work = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=dev)
        C.check(C.lib().mr_adjoint_many(*common, work.data_ptr(), nbytes.value, *grads))
        torch.cuda.synchronize()
        del work
    torch.cuda.synchronize()
    return out


def _agree(a, b):
    for k in ("x0", "y0", "kx0", "ky0"):
        assert torch_equal_bits(a[k], b[k]), k
    for k in ("u", "v"):
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.isfinite(x).all() and np.isfinite(y).all(), k
        d = np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)
        assert d <= 1e-12, (k, d)


def torch_equal_bits(a, b):
    x, y = a.cpu().numpy(), b.cpu().numpy()
    return np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_noalloc_matches_the_allocating_sweep():
    import torch

    from mantaray_amd import _capi as C
    from mantaray_amd import adjoint
    runs = [(2000, 300, 50, 3), (700, 200, 25, 5)]     # the second smaller, another stride
    used = []
    for n_rays, n_steps, stride, seed in runs:
        need = ctypes.c_int64()
        C.check(C.lib().mr_adjoint_workspace_size(n_rays, stride, ctypes.byref(need)))
        assert need.value <= C.MR_ADJOINT_SCRATCH_BYTES
        cfg, b, c, ic, u, v = _case(torch, n_rays, n_steps, seed)
        T = float(n_steps)
        fwd = adjoint.trace_fp64(b, c, *ic, 0.0, T, 1.0, stride=stride, u=u, v=v)
        g_end = torch.zeros_like(fwd["end"])
        g_end[:, 0] = 2.0 * (fwd["end"][:, 0] - torch.tensor(cfg["target_x"], device="cuda:0"))
        g_end[:, 1] = 2.0 * (fwd["end"][:, 1] - torch.tensor(cfg["target_y"], device="cuda:0"))
        torch.cuda.synchronize()
        ref = _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc=False)
        got = _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc=True, used=used)
        again = _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc=True, used=used)
        _agree(got, ref)
        _agree(again, ref)
        # every sweep kernel through the static buffer
        for block in (64, 128, 256):
            r1 = _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc=False,
                        block_size=block)
            r2 = _sweep(torch, C, b, c, fwd, T, stride, g_end, u, v, noalloc=True,
                        block_size=block, used=used)
            _agree(r2, r1)
    # nothing was allocated by the no-allocation calls: the device's free
    # memory did not shrink across any of them
    assert len(used) == 10 and all(u_ <= 0 for u_ in used), used


def test_noalloc_refuses_fields_without_tables():
    """Fresh fields hold no fp64 tables on the device; the variant builds
    none, so it refuses (every buffer below is sized for the call, so the
    arguments are valid apart from that)."""
    import torch

    from mantaray_amd import _capi as C
    from mantaray_amd import adjoint, workloads
    from mantaray_amd.adjoint import _device_opts
    nr, T, stride = 64, 20.0, 5
    cfg = workloads.eddy_adjoint(n_rays=nr, n_steps=int(T), seed=6)
    b, c = workloads.make_fields(cfg)
    dev = torch.device("cuda:0")
    o = _device_opts(dev, None, "fast", 0)
    n_ck = adjoint.checkpoint_count(0.0, T, 1.0, stride)
    ckpt = torch.zeros((n_ck, 4, nr), dtype=torch.float64, device=dev)
    ln = torch.ones(nr, dtype=torch.int32, device=dev)
    end = torch.zeros((nr, 4), dtype=torch.float64, device=dev)
    gx0 = torch.full((nr,), 7.0, dtype=torch.float64, device=dev)
    rc = C.lib().mr_adjoint_many_noalloc(b.handle, c.handle, None, None, None, nr, 0.0, T, 1.0,
                                         stride, ctypes.byref(o), ckpt.data_ptr(), ln.data_ptr(),
                                         end.data_ptr(), None, None, None, None, None,
                                         gx0.data_ptr(), None, None, None, None, None, None)
    assert rc == C.MR_ERR_INVALID_ARGUMENT
    assert "mr_trace_fp64" in C.lib().mr_last_error().decode()
    torch.cuda.synchronize()
    assert bool((gx0 == 7.0).all())                   # nothing ran
