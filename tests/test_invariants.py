"""Closed forms and invariants of the ray equations, checked on the CPU
oracle.  Each one fails for a plausible slip in the restatement (a dropped
current term, a swapped gradient, a wrong RK4 weight or sign)."""
import math

import numpy as np
import pytest

G = 9.8


def sigma(k, h):
    return math.sqrt(G * k * math.tanh(k * h))


def cg_exact(k, h):
    kh = k * h
    return 0.5 * sigma(k, h) / k * (1.0 + 2.0 * kh / math.sinh(2.0 * kh))


@pytest.mark.parametrize("h,k,theta", [(10.0, 0.05, 0.3), (1000.0, 0.01, 2.1),
                                       (50.0, 0.2, -1.2)])
def test_straight_ray_constant_depth(oracle_mod, h, k, theta):
    """Constant depth, no current: k is bitwise constant and the ray is a
    straight line at the closed-form group velocity."""
    O = oracle_mod
    kx, ky = k * math.cos(theta), k * math.sin(theta)
    rows = O.trace_single(O.Bathymetry.constant(h), O.Current.constant(), [0.0, 0.0, kx, ky],
                          0.0, 200.0, 1.0)
    assert np.all(rows[:, 3] == kx) and np.all(rows[:, 4] == ky)
    hf = float(np.float32(h))                       # the depth is an f32
    t = rows[:, 0]
    c = cg_exact(k, hf)
    assert np.allclose(rows[:, 1], c * math.cos(theta) * t, rtol=1e-12, atol=1e-9)
    assert np.allclose(rows[:, 2], c * math.sin(theta) * t, rtol=1e-12, atol=1e-9)


def test_constant_current_advects(oracle_mod):
    """A uniform current U adds U t to the position and leaves k unchanged."""
    O = oracle_mod
    s0 = [0.0, 0.0, 0.03, 0.04]
    a = O.trace_single(O.Bathymetry.constant(30.0), O.Current.constant(), s0, 0.0, 50.0, 1.0)
    b = O.trace_single(O.Bathymetry.constant(30.0), O.Current.constant(0.7, -0.4), s0,
                       0.0, 50.0, 1.0)
    t = a[:, 0]
    assert np.allclose(b[:, 1] - a[:, 1], 0.7 * t, atol=1e-9)
    assert np.allclose(b[:, 2] - a[:, 2], -0.4 * t, atol=1e-9)
    assert np.array_equal(a[:, 3:], b[:, 3:])


def _beach(O, n=200, dx=50.0):
    x = (np.arange(n) * dx).astype(np.float32)
    X, Y = np.meshgrid(x.astype(np.float64), x.astype(np.float64))
    return x, X, Y


def test_snell_ky_constant(oracle_mod):
    """h = h(x), no current: the cell-difference dh/dy is exactly 0, so ky is
    bitwise constant while kx changes (refraction)."""
    O = oracle_mod
    x, X, Y = _beach(O)
    b = O.Bathymetry.grid(x, x, 5.0 + 0.01 * X + 3.0 * np.sin(X / 700.0))
    rows = O.trace_single(b, O.Current.constant(), [9000.0, 2000.0, -0.02, 0.015],
                          0.0, 300.0, 1.0)
    fin = rows[~np.isnan(rows[:, 1])]
    assert len(fin) > 100
    assert np.all(fin[:, 4] == 0.015)
    assert np.ptp(fin[:, 3]) > 1e-4


def test_shear_current_kx_constant(oracle_mod):
    """Constant depth with U = U(y), V = 0: du/dx = dv/dx = 0 exactly, so kx
    is bitwise constant, and ky changes by -kx dU/dy."""
    O = oracle_mod
    x, X, Y = _beach(O)
    u = 0.8 * np.tanh((Y - 5000.0) / 1500.0)
    c = O.Current.grid(x.astype(np.float64), x.astype(np.float64), u, np.zeros_like(u))
    rows = O.trace_single(O.Bathymetry.constant(40.0), c, [1000.0, 4000.0, 0.05, 0.02],
                          0.0, 300.0, 1.0)
    fin = rows[~np.isnan(rows[:, 1])]
    assert len(fin) > 100
    assert np.all(fin[:, 3] == 0.05)
    assert np.ptp(fin[:, 4]) > 1e-4


def test_rk4_fourth_order(oracle_mod):
    """Halving dt shrinks the end-point error by about 2^4 on a smooth field
    (the analytic slope and a uniform current); a wrong RK4 weight or stage
    makes the method first or second order.  Below dt = 8 the f32 rounding
    of the position in the depth query sets the floor (about 1e-7 m)."""
    O = oracle_mod
    b = O.Bathymetry.slope(h0=60.0, x0=0.0, y0=0.0, dhdx=-0.02, dhdy=0.01)
    c = O.Current.constant(0.2, 0.1)
    s0 = [100.0, 50.0, 0.04, 0.03]

    def end(dt):
        return O.trace_single(b, c, s0, 0.0, 128.0, dt)[-1, 1:]
    ref = end(1.0 / 32.0)
    e1 = np.abs(end(16.0) - ref).max()
    e2 = np.abs(end(8.0) - ref).max()
    assert e1 > 0 and 10.0 < e1 / e2 < 24.0, (e1, e2)
