"""Deterministic synthetic workloads of the many-ray integrator.

The measured and parity-tested configurations, with every random draw taken
from SplitMix64 (a counter-based generator: draw i of seed s is
mix64(s + (i + 1) * 0x9E3779B97F4A7C15)), so the inputs are identical on any
machine with no dependence on numpy's RNG.

  linear_beach()   501 x 501 grid at 100 m, depth max(200 - 0.004 x, 0.1) m,
                   zero current on the same grid; 128 rays from x = 1 km,
                   y ~ U[5, 45] km, period 10 s, direction ~ U[-30, 30] deg;
                   dt = 1 s for 1000 s.  (The reference's CPU-runnable case.)
  seamount_jet()   1024 x 1024 grid at 200 m, depth 4000 m minus a Gaussian
                   seamount (3800 m, sigma 15 km) at the centre; zonal jet
                   U(y) = 1.5 sech^2((y - y_c)/20 km), V = 0; 65,536 rays at
                   uniform interior points, period ~ U[6, 16] s, direction
                   ~ U[0, 2 pi); dt = 1 s, 5000 steps; rays that leave the
                   grid end early.
  eddy_adjoint()   512 x 512 grid at 250 m, constant depth 50 m, current the
                   sum of 20 Gaussian eddies (peak speed ~ U[0, 2] m/s,
                   radius ~ U[10, 50] km, random sense, centres uniform over
                   the grid); 16,384 rays at uniform points at least 35 km
                   inside the edges, period ~ U[8, 14] s, direction
                   ~ U[0, 2 pi); dt = 1 s, 2000 steps; endpoint targets
                   launch point + U[-20, 20] km per axis.  The adjoint
                   workload: loss sum_r |(x, y)_r(T) - target_r|^2.
  batched_scenes() 64 scenes of the batched workload (see the function).
  eddy_inversion() 2048 x 2048 grid at 100 m, constant depth 30 m, true
                   current the sum of 20 Gaussian eddies (peak speed
                   ~ U[0.5, 1.5] m/s, radius ~ U[10, 50] km); 262,144 rays at
                   uniform points at least 50 km inside the edges, period
                   ~ U[8, 14] s, direction ~ U[0, 2 pi); dt = 1 s, 3000
                   steps.  The inversion workload: recover the current from
                   the rays' endpoints, starting from zero.
The wavenumber magnitude solves the linear dispersion relation
sigma^2 = g k tanh(k h) (g = 9.8, sigma = 2 pi / T) at the launch depth by
Newton's method.
"""
from __future__ import annotations

import numpy as np

G = 9.8
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Draws start .. start+count-1 of SplitMix64 seeded with `seed` (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Sequential uniform draws in [0, 1) (53-bit) from SplitMix64."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        z = splitmix64(self.seed, self.pos, n)
        self.pos += n
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


def wavenumber(sigma, h, iters: int = 60) -> np.ndarray:
    """|k| with sigma^2 = g k tanh(k h), Newton from the larger of the deep-
    and shallow-water limits (the dispersion function is concave in k)."""
    sigma = np.asarray(sigma, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    k = np.maximum(sigma * sigma / G, sigma / np.sqrt(G * h))
    for _ in range(iters):
        th = np.tanh(k * h)
        f = G * k * th - sigma * sigma
        df = G * th + G * k * h * (1.0 - th * th)
        k_new = k - f / df
        if np.all(np.abs(k_new - k) <= 1e-15 * np.abs(k)):
            k = k_new
            break
        k = k_new
    return k


def linear_beach(n_rays: int = 128, seed: int = 0) -> dict:
    n = 501
    x = (np.arange(n) * 100.0).astype(np.float32)
    y = x.copy()
    xd = x.astype(np.float64)
    depth_row = np.maximum(200.0 - 0.004 * xd, 0.1)
    depth = np.broadcast_to(depth_row[None, :], (n, n)).copy()      # depth[j, i]
    zeros = np.zeros((n, n))
    rs = Stream(seed)
    y0 = rs.uniform(n_rays, 5000.0, 45000.0)
    theta = np.deg2rad(rs.uniform(n_rays, -30.0, 30.0))
    x0 = np.full(n_rays, 1000.0)
    sigma = 2.0 * np.pi / 10.0
    k = wavenumber(np.full(n_rays, sigma), np.maximum(200.0 - 0.004 * x0, 0.1))
    return {
        "name": "linear_beach",
        "bathy": {"x": x, "y": y, "depth": depth},
        "current": {"x": xd, "y": xd.copy(), "u": zeros, "v": zeros.copy()},
        "x0": x0, "y0": y0, "kx0": k * np.cos(theta), "ky0": k * np.sin(theta),
        "t0": 0.0, "t_end": 1000.0, "dt": 1.0,
    }


def seamount_jet(n_rays: int = 65536, seed: int = 1, n_steps: int = 5000) -> dict:
    n = 1024
    dx = 200.0
    x = (np.arange(n) * dx).astype(np.float32)
    xd = x.astype(np.float64)
    xc = 0.5 * xd[-1]
    r2 = (xd[None, :] - xc) ** 2 + (xd[:, None] - xc) ** 2
    depth = 4000.0 - 3800.0 * np.exp(-r2 / (2.0 * 15000.0 ** 2))    # depth[j, i]
    u_row = 1.5 / np.cosh((xd - xc) / 20000.0) ** 2                 # U(y)
    u = np.broadcast_to(u_row[:, None], (n, n)).copy()              # u[j, i] = U(y_j)
    v = np.zeros((n, n))
    rs = Stream(seed)
    lo, hi = dx, xd[-1] - dx
    x0 = rs.uniform(n_rays, lo, hi)
    y0 = rs.uniform(n_rays, lo, hi)
    period = rs.uniform(n_rays, 6.0, 16.0)
    theta = rs.uniform(n_rays, 0.0, 2.0 * np.pi)
    h0 = 4000.0 - 3800.0 * np.exp(-((x0 - xc) ** 2 + (y0 - xc) ** 2) / (2.0 * 15000.0 ** 2))
    k = wavenumber(2.0 * np.pi / period, h0)
    return {
        "name": "seamount_jet",
        "bathy": {"x": x, "y": x.copy(), "depth": depth},
        "current": {"x": xd, "y": xd.copy(), "u": u, "v": v},
        "x0": x0, "y0": y0, "kx0": k * np.cos(theta), "ky0": k * np.sin(theta),
        "t0": 0.0, "t_end": float(n_steps), "dt": 1.0,
    }


def eddies(x: np.ndarray, y: np.ndarray, rs: Stream, n_eddies: int = 20,
           peak=(0.0, 2.0), radius=(10000.0, 50000.0)):
    """u, v [j, i] of a sum of Gaussian eddies: stream function
    psi = A exp(-r^2 / (2 R^2)), (u, v) = (-dpsi/dy, dpsi/dx), with
    A = +-peak R e^(1/2) so each eddy's speed peaks at r = R."""
    u = np.zeros((y.size, x.size))
    v = np.zeros((y.size, x.size))
    xc = rs.uniform(n_eddies, x[0], x[-1])
    yc = rs.uniform(n_eddies, y[0], y[-1])
    pk = rs.uniform(n_eddies, *peak)
    R = rs.uniform(n_eddies, *radius)
    sense = np.where(rs.uniform(n_eddies) < 0.5, -1.0, 1.0)
    for e in range(n_eddies):
        # the Gaussian factors into exp(-dx^2/2R^2) exp(-dy^2/2R^2): outer products
        A = sense[e] * pk[e] * R[e] * np.exp(0.5) / R[e] ** 2
        gx = np.exp(-(x - xc[e]) ** 2 / (2.0 * R[e] ** 2))
        gy = np.exp(-(y - yc[e]) ** 2 / (2.0 * R[e] ** 2))
        u += np.outer(A * (y - yc[e]) * gy, gx)
        v -= np.outer(A * gy, (x - xc[e]) * gx)
    return u, v


def eddy_adjoint(n_rays: int = 16384, seed: int = 3, n_steps: int = 2000) -> dict:
    n = 512
    dx = 250.0
    xd = np.arange(n) * dx
    rs = Stream(seed)
    u, v = eddies(xd, xd, rs)
    margin = 35000.0
    x0 = rs.uniform(n_rays, margin, xd[-1] - margin)
    y0 = rs.uniform(n_rays, margin, xd[-1] - margin)
    period = rs.uniform(n_rays, 8.0, 14.0)
    theta = rs.uniform(n_rays, 0.0, 2.0 * np.pi)
    h = 50.0
    k = wavenumber(2.0 * np.pi / period, np.full(n_rays, h))
    tx = x0 + rs.uniform(n_rays, -20000.0, 20000.0)
    ty = y0 + rs.uniform(n_rays, -20000.0, 20000.0)
    return {
        "name": "eddy_adjoint",
        "bathy": {"constant": h},
        "current": {"x": xd, "y": xd.copy(), "u": u, "v": v},
        "x0": x0, "y0": y0, "kx0": k * np.cos(theta), "ky0": k * np.sin(theta),
        "target_x": tx, "target_y": ty,
        "t0": 0.0, "t_end": float(n_steps), "dt": 1.0, "stride": 50,
    }


def eddy_inversion(n_rays: int = 262144, seed: int = 4, n_steps: int = 3000) -> dict:
    n = 2048
    dx = 100.0
    xd = np.arange(n) * dx
    rs = Stream(seed)
    u, v = eddies(xd, xd, rs, peak=(0.5, 1.5))
    margin = 50000.0
    x0 = rs.uniform(n_rays, margin, xd[-1] - margin)
    y0 = rs.uniform(n_rays, margin, xd[-1] - margin)
    period = rs.uniform(n_rays, 8.0, 14.0)
    theta = rs.uniform(n_rays, 0.0, 2.0 * np.pi)
    h = 30.0
    k = wavenumber(2.0 * np.pi / period, np.full(n_rays, h))
    return {
        "name": "eddy_inversion",
        "bathy": {"constant": h},
        "current": {"x": xd, "y": xd.copy(), "u": u, "v": v},
        "x0": x0, "y0": y0, "kx0": k * np.cos(theta), "ky0": k * np.sin(theta),
        "t0": 0.0, "t_end": float(n_steps), "dt": 1.0, "stride": 50,
    }


def normals(rs: Stream, n: int) -> np.ndarray:
    """n standard normal draws (Box-Muller on pairs of uniforms)."""
    m = (n + 1) // 2
    u1 = 1.0 - rs.uniform(m)                           # in (0, 1]
    u2 = rs.uniform(m)
    r = np.sqrt(-2.0 * np.log(u1))
    return np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]


def gaussian_random_field(n: int, dx: float, rs: Stream, mean: float, sd: float,
                          corr: float) -> np.ndarray:
    """White noise on an n x n grid filtered by a Gaussian of standard
    deviation corr (periodic, by FFT), scaled to the given mean and sd."""
    w = normals(rs, n * n).reshape(n, n)
    f = np.fft.fftfreq(n, d=dx) * 2.0 * np.pi
    tf = np.exp(-0.5 * (corr * f) ** 2)
    g = np.fft.ifft2(np.fft.fft2(w) * tf[:, None] * tf[None, :]).real
    g = (g - g.mean()) / g.std()
    return mean + sd * g


def _scene_draws(rs: Stream, n_scenes: int, total_rays: int, max_steps: int):
    """Per scene of the batched workload: grid side, ray count, step, steps."""
    sides = np.array([256, 512, 1024, 2048])[
        np.minimum((rs.uniform(n_scenes) * 4).astype(int), 3)]
    raw = np.exp(rs.uniform(n_scenes, np.log(1024.0), np.log(65536.0)))
    counts = np.maximum(1, np.floor(raw / raw.sum() * total_rays)).astype(np.int64)
    counts[np.argmax(counts)] += total_rays - counts.sum()
    dts = rs.uniform(n_scenes, 0.5, 2.0)
    steps = np.floor(rs.uniform(n_scenes, 1000.0, max_steps + 1.0)).astype(np.int64)
    return sides, counts, dts, steps


def batched_scene_plan(n_scenes: int = 64, total_rays: int = 1 << 20, seed: int = 2,
                       max_steps: int = 10000) -> dict:
    """The batched workload's per-scene grid sides, ray counts, steps and
    step counts, without building the fields (batched_scenes draws these
    first, from the same stream)."""
    sides, counts, dts, steps = _scene_draws(Stream(seed), n_scenes, total_rays, max_steps)
    return {"sides": sides, "rays": counts, "dt": dts, "steps": steps}


def batched_scenes(n_scenes: int = 64, total_rays: int = 1 << 20, seed: int = 2,
                   max_steps: int = 10000) -> list:
    """Scenes of the batched workload: grid side ~ {256, 512, 1024, 2048} at
    250 m; depth a Gaussian random field (mean 2000 m, sd 800 m, correlation
    length 30 km, clipped at 1 m); current the sum of 20 Gaussian eddies
    (peak 0-2 m/s, radius 10-50 km); rays per scene log-uniform in
    [1024, 65536], rescaled to total_rays; step ~ U[0.5, 2] s; steps
    ~ U[1000, max_steps]; rays at uniform interior points with period
    ~ U[6, 16] s and direction ~ U[0, 2 pi).  Rays leave small scenes early,
    so lifetimes are uneven."""
    rs = Stream(seed)
    sides, counts, dts, steps = _scene_draws(rs, n_scenes, total_rays, max_steps)
    dx = 250.0
    scenes = []
    for i in range(n_scenes):
        n = int(sides[i])
        xd = np.arange(n) * dx
        depth = np.maximum(gaussian_random_field(n, dx, rs, 2000.0, 800.0, 30000.0), 1.0)
        u, v = eddies(xd, xd, rs)
        nr = int(counts[i])
        x0 = rs.uniform(nr, dx, xd[-1] - dx)
        y0 = rs.uniform(nr, dx, xd[-1] - dx)
        period = rs.uniform(nr, 6.0, 16.0)
        theta = rs.uniform(nr, 0.0, 2.0 * np.pi)
        fx = np.clip(np.rint(x0 / dx).astype(int), 0, n - 1)
        fy = np.clip(np.rint(y0 / dx).astype(int), 0, n - 1)
        k = wavenumber(2.0 * np.pi / period, depth[fy, fx])
        dt = float(dts[i])
        scenes.append({
            "name": f"batched_{i}",
            "bathy": {"x": xd.astype(np.float32), "y": xd.astype(np.float32), "depth": depth},
            "current": {"x": xd, "y": xd.copy(), "u": u, "v": v},
            "x0": x0, "y0": y0, "kx0": k * np.cos(theta), "ky0": k * np.sin(theta),
            "t0": 0.0, "t_end": float(steps[i]) * dt, "dt": dt,
        })
    return scenes


def make_fields(cfg: dict):
    """Product field handles for a workload."""
    from .fields import Bathymetry, Current
    if "constant" in cfg["bathy"]:
        b = Bathymetry.constant(cfg["bathy"]["constant"])
    else:
        b = Bathymetry.from_arrays(cfg["bathy"]["x"], cfg["bathy"]["y"], cfg["bathy"]["depth"])
    c = Current.from_arrays(cfg["current"]["x"], cfg["current"]["y"], cfg["current"]["u"],
                            cfg["current"]["v"])
    return b, c
