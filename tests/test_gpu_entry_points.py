"""Entry points and kernel variants that the other GPU tests reach only
indirectly, each against the oracle:

  - mr_single_ray, the C-ABI form of _mantaray.single_ray (src/ffi.rs:25-49),
    called through ctypes on every single-ray behaviour case of src/ray.rs:
    rows (t, x, y, kx, ky) against oracle.trace_single;
  - ray_tracing end to end on rays of different lengths
    (python/tests/test_core.py:81-101): the NaN padding of core.py:114-125
    (time included) against the oracle's row counts, and the rows;
  - the launch bench.py times, on the seamount jet at full size: every
    trajectory component written (the OUT = 1 kernel), the tile shape the
    bench's autotuner picks (32 x 32 cells) and the default, with every
    256th ray's rows compared with the oracle's.
"""
import ctypes

import numpy as np
import pytest

import ray_cases as RC
from _helpers import oracle_fields
from test_gpu_ray_cases import hip_fields
from test_ray_cases import oracle_fields as case_oracle_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("math", ["fast", "literal"])
@pytest.mark.parametrize("case", RC.CASES, ids=[c[0] for c in RC.CASES])
def test_mr_single_ray_c_abi(oracle_mod, case, math):
    from mantaray_amd import _capi as C
    from mantaray_amd.core import _opts, num_steps
    name, bspec, cspec, s0, (t0, t1, dt), checks, _ = case
    b, c = hip_fields(bspec, cspec)
    n = num_steps(t0, t1, dt)
    rows = np.full((n + 1, 5), -1.0)
    n_rows = ctypes.c_int64(-7)
    o = _opts(math)
    C.check(C.lib().mr_single_ray(b.handle, c.handle, *[float(v) for v in s0], float(t0),
                                  float(t1), float(dt), ctypes.byref(o), rows.ctypes.data,
                                  ctypes.byref(n_rows)))
    got = rows[:n_rows.value]
    assert np.all(rows[n_rows.value:] == -1.0)             # nothing written past the rows
    ob, oc = case_oracle_fields(oracle_mod, bspec, cspec)
    ref = oracle_mod.trace_single(ob, oc, list(s0), t0, t1, dt)
    assert got.shape == ref.shape
    assert np.array_equal(got[:, 0], ref[:, 0])            # t += dt, bitwise
    RC.check_rows(got[:, 1:], checks)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    scale = np.maximum(np.nanmax(np.abs(ref[:, 1:]), axis=0), 1e-300)
    assert np.nanmax(np.abs(got[:, 1:] - ref[:, 1:]) / scale) <= 1e-9


def test_ray_tracing_ragged_rays(oracle_mod, tmp_path):
    """python/tests/test_core.py:81-101: two deep-water rays from x = -1 km,
    one heading -x and one +x on a 3 x 3 grid [-10, 0, 10] km, 1e6 s in
    steps of 20 s: each leaves the grid after a different number of rows."""
    from scipy.io import netcdf_file

    import mantaray_amd as mr
    with netcdf_file(tmp_path / "island.nc", "w") as f:
        f.createDimension("x", 3)
        f.createDimension("y", 3)
        f.createVariable("x", "f8", ("x",))[:] = [-1e4, 0, 1e4]
        f.createVariable("y", "f8", ("y",))[:] = [-1e4, 0, 1e4]
        f.createVariable("depth", "f8", ("x", "y"))[:] = 10_000 * np.ones((3, 3))
    with netcdf_file(tmp_path / "current.nc", "w") as f:
        f.createDimension("x", 3)
        f.createDimension("y", 3)
        f.createVariable("x", "f8", ("x",))[:] = [-1e8, 0, 1e8]
        f.createVariable("y", "f8", ("y",))[:] = [-1e8, 0, 1e8]
        f.createVariable("u", "f8", ("x", "y"))[:] = 0.01 * np.ones((3, 3))
        f.createVariable("v", "f8", ("x", "y"))[:] = 0.01 * np.ones((3, 3))
    x0, y0, kx0, ky0 = [-1e3, -1e3], [0.0, 0.0], [-0.01, 0.01], [0.0, 0.0]
    ds = mr.ray_tracing(x0, y0, kx0, ky0, 1e6, 20, str(tmp_path / "island.nc"),
                        str(tmp_path / "current.nc"))
    # the oracle on the same files' fields
    b = mr.Bathymetry.from_netcdf(tmp_path / "island.nc")
    c = mr.Current.from_netcdf(tmp_path / "current.nc")
    bx, by, bd = b.grid()
    cx, cy, cu, cv = c.grid()
    ob = oracle_mod.Bathymetry.grid(bx.astype(np.float32), by.astype(np.float32), bd)
    oc = oracle_mod.Current.grid(cx, cy, cu, cv)
    lens, refs = [], []
    for r in range(2):
        ref = oracle_mod.trace_single(ob, oc, [x0[r], y0[r], kx0[r], ky0[r]], 0.0, 1e6, 20.0)
        lens.append(ref.shape[0])
        refs.append(ref)
    assert lens[0] != lens[1] and max(lens) < 50001        # ragged, both leave early
    longest = max(lens)
    assert ds.sizes["time_step"] == longest and ds.sizes["ray"] == 2
    assert list(ds["time_step"].values) == list(range(longest))
    assert list(ds["ray"].values) == [0, 1]
    for r in range(2):
        n = lens[r]
        for q, name in enumerate(("time", "x", "y", "kx", "ky")):
            col = np.asarray(ds[name].values)[:, r]
            assert np.all(np.isnan(col[n:])), (name, r)        # the padding, time too
            ref = refs[r][:, q]
            assert np.array_equal(np.isnan(col[:n]), np.isnan(ref)), (name, r)
            fin = ~np.isnan(ref)
            assert np.allclose(col[:n][fin], ref[fin], rtol=1e-12, atol=1e-9), (name, r)
        # the terminal row: time valid, state all-NaN
        assert not np.isnan(np.asarray(ds["time"].values)[n - 1, r])
        assert all(np.isnan(np.asarray(ds[k].values)[n - 1, r]) for k in ("x", "y", "kx", "ky"))


@pytest.fixture(scope="module")
def seamount():
    from mantaray_amd import workloads
    return workloads.seamount_jet()


@pytest.mark.parametrize("record_tile", [5, -1])
def test_benched_launch_against_the_oracle(oracle_mod, seamount, record_tile):
    """bench.py's launch: trajectories of all 65,536 rays x 5001 rows written
    to device buffers by the OUT = 1 kernel; every 256th ray's full
    trajectory against the oracle's (row counts equal, positions within
    1e-6 m, wavenumbers within 1e-9 relative, NaN padding identical)."""
    import torch

    from mantaray_amd import core, workloads
    cfg = seamount
    b, c = workloads.make_fields(cfg)
    dev = torch.device("cuda:0")
    ic = [torch.tensor(cfg[k], dtype=torch.float64, device=dev) for k in ("x0", "y0", "kx0", "ky0")]
    out = core.trace_many_device(b, c, *ic, cfg["t0"], cfg["t_end"], cfg["dt"],
                                 record_tile=record_tile)
    torch.cuda.synchronize()
    sub = slice(0, None, 256)
    got = {k: out[k][:, sub].cpu().numpy() for k in ("x", "y", "kx", "ky")}
    ln = out["len"].cpu().numpy()[sub]
    t = out["t"].cpu().numpy()
    del out
    torch.cuda.empty_cache()
    ob, oc = oracle_fields(oracle_mod, cfg)
    ref = oracle_mod.trace_many(ob, oc, cfg["x0"][sub], cfg["y0"][sub], cfg["kx0"][sub],
                                cfg["ky0"][sub], cfg["t0"], cfg["t_end"], cfg["dt"])
    assert np.array_equal(t, ref["t"])
    assert np.array_equal(ln, ref["len"])
    assert int((ln < 5001).sum()) > 30                         # rays leave the grid
    for k in ("x", "y", "kx", "ky"):
        assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), k
    dxy = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("x", "y"))
    kscale = np.nanmax(np.abs(ref["kx"]))
    dk = max(np.nanmax(np.abs(got[k] - ref[k])) for k in ("kx", "ky")) / kscale
    print(f"record tile {record_tile}: max |d(x, y)| {dxy:.3e} m, relative |dk| {dk:.3e}")
    assert dxy <= 1e-6 and dk <= 1e-9


@pytest.mark.parametrize("precision", ["reference", "fp64"])
def test_too_many_rays_for_one_launch(precision):
    """A ray count whose grid of workgroups would not fit a launch is refused
    with MR_ERR_UNSUPPORTED before anything is launched (the pointers are
    never read: a few real device words stand in for them)."""
    import torch

    from mantaray_amd import _capi as C
    L = C.lib()
    b, c = ctypes.c_void_p(), ctypes.c_void_p()
    C.check(L.mr_bathy_constant(30.0, ctypes.byref(b)))
    C.check(L.mr_current_constant(0.0, 0.0, ctypes.byref(c)))
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = dummy.data_ptr()
    lenbuf = torch.zeros(8, dtype=torch.int32, device="cuda:0").data_ptr()
    prec = C.MR_PREC_REFERENCE if precision == "reference" else C.MR_PREC_FP64_INTERP
    o = C.options(memory=C.MR_MEM_DEVICE, device=0, precision=prec, block_size=64)
    rc = L.mr_trace_many(b, c, 1 << 40, p, p, p, p, 0.0, 2.0, 1.0, ctypes.byref(o),
                         None, None, None, None, None, lenbuf, None)
    assert rc == C.MR_ERR_UNSUPPORTED
    assert "too many rays" in L.mr_last_error().decode()
    torch.cuda.synchronize()
    L.mr_field_free(b)
    L.mr_field_free(c)
