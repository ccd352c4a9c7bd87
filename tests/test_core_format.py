"""The Python layer's output format (mantaray.core, python/mantaray/core.py)
from a synthetic trace result, on the CPU: variables, dimensions, coords,
the NaN padding of ragged rays (time included) and the no-ray error."""
import numpy as np
import pytest

from mantaray_amd import core


def fake_result(lens, n=6):
    nr = len(lens)
    rng = np.random.default_rng(0)
    r = {"t": np.arange(n + 1, dtype=np.float64) * 2.0, "len": np.array(lens, dtype=np.int32)}
    for k in ("x", "y", "kx", "ky"):
        a = rng.normal(size=(n + 1, nr))
        for j, ln in enumerate(lens):
            a[ln:, j] = np.nan
            a[ln - 1, j] = np.nan if ln < n + 1 else a[ln - 1, j]   # terminal NaN row
        r[k] = a
    return r


def test_rays_dataset_layout():
    """core.py:114-132: pad every ray to the longest with NaN rows, dims
    (time_step, ray), coords time, x, y; time_step = range, ray = arange."""
    r = fake_result([4, 7, 2])
    ds = core.rays_dataset(r)
    assert ds.sizes["time_step"] == 7 and ds.sizes["ray"] == 3
    for k in ("time", "x", "y", "kx", "ky"):
        assert ds[k].dims == ("time_step", "ray")
    assert {"time", "x", "y"} <= set(ds.coords)
    assert list(np.asarray(ds["time_step"])) == list(range(7))
    assert list(np.asarray(ds["ray"])) == [0, 1, 2]
    t = np.asarray(ds["time"])
    assert np.array_equal(t[:4, 0], r["t"][:4]) and np.isnan(t[4:, 0]).all()
    assert np.array_equal(t[:, 1], r["t"][:7])
    assert np.isnan(t[2:, 2]).all()
    x = np.asarray(ds["x"])
    assert np.array_equal(x[:3, 0], r["x"][:3, 0]) and np.isnan(x[3:, 0]).all()
    assert "date_created" in ds.attrs


def test_rays_dataset_no_rays():
    """core.py:111: max() of an empty sequence."""
    with pytest.raises(ValueError):
        core.rays_dataset({"t": np.zeros(3), "len": np.zeros(0, dtype=np.int32),
                           "x": np.zeros((3, 0)), "y": np.zeros((3, 0)),
                           "kx": np.zeros((3, 0)), "ky": np.zeros((3, 0))})


def test_single_dataset_layout():
    """core.py:56-63: one dimension time_step with every row of the ray."""
    r = fake_result([5])
    ds = core.single_dataset(r)
    assert ds.sizes["time_step"] == 5
    assert {"time", "x", "y"} <= set(ds.coords)
    assert np.array_equal(np.asarray(ds["time"]), r["t"][:5])
    assert np.array_equal(np.asarray(ds["kx"]), r["kx"][:5, 0], equal_nan=True)
