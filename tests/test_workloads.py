"""The synthetic workloads: SplitMix64 against its published test vectors,
the dispersion solve, and the shapes and ranges of the two configurations."""
import numpy as np

from mantaray_amd import workloads as W


def test_splitmix64_known_answers():
    # the reference sequence of SplitMix64 (Steele, Lea, Flood 2014; Vigna's
    # splitmix64.c): seed 0 and seed 1234567
    assert [int(v) for v in W.splitmix64(0, 0, 3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [int(v) for v in W.splitmix64(1234567, 0, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_stream_is_counter_based():
    a = W.Stream(7).uniform(10)
    s = W.Stream(7)
    b = np.concatenate([s.uniform(3), s.uniform(7)])
    assert np.array_equal(a, b) and (a >= 0).all() and (a < 1).all()


def test_wavenumber_solves_dispersion():
    sigma = 2 * np.pi / np.array([4.0, 8.0, 12.0, 16.0])
    for h in (1.0, 20.0, 200.0, 4000.0):
        k = W.wavenumber(sigma, np.full(4, h))
        res = W.G * k * np.tanh(k * h) - sigma ** 2
        assert np.all(np.abs(res) <= 1e-14 * sigma ** 2)


def test_linear_beach_shape():
    c = W.linear_beach()
    assert c["bathy"]["depth"].shape == (501, 501) and c["bathy"]["x"].dtype == np.float32
    assert c["x0"].shape == (128,) and np.all(c["x0"] == 1000.0)
    assert np.all((c["y0"] >= 5000) & (c["y0"] <= 45000))
    th = np.degrees(np.arctan2(c["ky0"], c["kx0"]))
    assert np.all(np.abs(th) <= 30.0 + 1e-9)


def test_seamount_jet_shape():
    c = W.seamount_jet(n_rays=4096, n_steps=10)
    assert c["bathy"]["depth"].shape == (1024, 1024)
    assert c["current"]["u"].shape == (1024, 1024)
    x = c["bathy"]["x"]
    assert np.all((c["x0"] > x[0]) & (c["x0"] < x[-1]))
    assert np.all((c["y0"] > x[0]) & (c["y0"] < x[-1]))
    assert c["t_end"] == 10.0 and c["dt"] == 1.0
    d = c["bathy"]["depth"]
    assert 190.0 <= d.min() < 300.0 and d.max() <= 4000.0
