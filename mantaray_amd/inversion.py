"""Current inversion by gradient descent over the device adjoint.

Given rays' launch states and observed endpoints, recover a gridded current
(u, v) that reproduces them: minimise
    L(u, v) = sum_r |(x, y)_r(T; u, v) - observed_r|^2
over the grid values, starting from a guess, with every forward and
gradient on the GPU (mantaray_amd.adjoint: the fp64-interpolation trace with
checkpoints, then the reverse sweep).  Rays that leave the grid in the
current iterate drop out of the loss.

The descent direction is the gradient smoothed by a separable Gaussian
filter (a Sobolev-type preconditioner: the endpoints see the current only
along the rays, and the filter spreads each update over the scale the
eddies vary on).  The first step is chosen so that the largest update is
`first_step` m/s; later steps keep it ("gd") or take the Barzilai-Borwein
length ("bb").
"""
from __future__ import annotations

import math
import time

from . import adjoint


class GaussSmoother:
    """Separable Gaussian filter (standard deviation sigma_cells) of an
    (ny, nx) grid with zeros outside, applied by FFT on a grid padded by
    3 sigma: two real FFTs of about (ny + 6 sigma) x (nx + 6 sigma) values."""

    def __init__(self, torch, ny: int, nx: int, sigma_cells: float, dev):
        self.torch = torch
        r = max(1, int(math.ceil(3.0 * sigma_cells)))
        self.r, self.ny, self.nx = r, ny, nx
        self.py, self.px = ny + 2 * r, nx + 2 * r
        k = torch.exp(-0.5 * (torch.arange(-r, r + 1, dtype=torch.float64, device=dev)
                              / sigma_cells) ** 2)
        k = k / k.sum()

        def transfer(n, rfft):
            z = torch.zeros(n, dtype=torch.float64, device=dev)
            z[:r + 1] = k[r:]                # centred kernel, wrapped
            z[n - r:] = k[:r]
            return torch.fft.rfft(z) if rfft else torch.fft.fft(z)

        self.ky = transfer(self.py, False)[:, None]
        self.kx = transfer(self.px, True)[None, :]

    def __call__(self, g2d):
        torch = self.torch
        a = torch.zeros((self.py, self.px), dtype=torch.float64, device=g2d.device)
        a[:self.ny, :self.nx] = g2d
        f = torch.fft.rfft2(a) * self.ky * self.kx
        return torch.fft.irfft2(f, s=(self.py, self.px))[:self.ny, :self.nx]


def invert(bathymetry, current, ic, observed, t_end: float, dt: float, *, u0, v0,
           iterations: int = 20, stride: int = 50, sigma_cells: float = 60.0,
           first_step: float = 0.5, method: str = "bb", u_true=None, v_true=None,
           math_mode: str = "fast", log=None) -> dict:
    """Run `iterations` descent steps.  ic: four float64 device tensors
    (n_rays,); observed: (n_rays, 2) device tensor; u0, v0: starting values
    (device tensors of the current's value-array shape, (ny, nx)).

    method "gd": the fixed step of the first iteration throughout; "bb": the
    Barzilai-Borwein step of the smoothed problem from the second iteration
    on, alpha_k = -alpha_{k-1} s.g_{k-1} / s.(g_k - g_{k-1}) with s the last
    update (kept within [alpha_{k-1}/10, 10 alpha_{k-1}], and the previous
    step where it is not positive).

    Returns the recovered u, v and per-iteration records (loss, rays used,
    seconds, RMS error against u_true, v_true when given)."""
    import torch
    dev = observed.device
    u = u0.clone()
    v = v0.clone()
    ny, nx = u.shape
    nr = observed.shape[0]
    S = GaussSmoother(torch, ny, nx, sigma_cells, dev)
    S(torch.zeros_like(u))             # builds the FFT plans before the first iteration
    work = torch.empty(adjoint.workspace_bytes(nr, stride), dtype=torch.uint8, device=dev)
    step = None
    prev = None                        # (g_u, g_v, s_u, s_v) of the last update
    hist = []
    for it in range(iterations + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd = adjoint.trace_fp64(bathymetry, current, *ic, 0.0, t_end, dt, u=u.view(-1),
                                 v=v.view(-1), stride=stride, math=math_mode)
        e = fwd["end"][:, :2]
        res = e - observed
        alive = ~torch.isnan(res).any(dim=1)
        res = torch.where(alive[:, None], res, torch.zeros_like(res))
        loss = float((res * res).sum())
        rec = {"iteration": it, "loss": loss, "rays_used": int(alive.sum())}
        if u_true is not None:
            rec["rms_error"] = float(torch.sqrt(((u - u_true) ** 2 + (v - v_true) ** 2).mean()))
        if it == iterations:
            torch.cuda.synchronize()
            rec["seconds"] = time.perf_counter() - t0
            hist.append(rec)
            break
        g_end = torch.zeros_like(fwd["end"])
        g_end[:, :2] = 2.0 * res
        g = adjoint.adjoint(bathymetry, current, fwd, 0.0, t_end, dt, stride, g_end=g_end,
                            u=u.view(-1), v=v.view(-1), want=("u", "v"), math=math_mode,
                            work=work)
        gu, gv = g["u"].view(ny, nx), g["v"].view(ny, nx)
        du, dv = S(gu), S(gv)
        if step is None:
            step = first_step / float(torch.maximum(du.abs().max(), dv.abs().max()))
        elif method == "bb":
            pgu, pgv, su, sv = prev
            num = -float((su * pgu).sum() + (sv * pgv).sum())
            den = float((su * (gu - pgu)).sum() + (sv * (gv - pgv)).sum())
            if num > 0.0 and den > 0.0:
                step = min(max(step * num / den, step / 10.0), step * 10.0)
        su, sv = -step * du, -step * dv
        u = u + su
        v = v + sv
        prev = (gu, gv, su, sv)
        torch.cuda.synchronize()
        rec["seconds"] = time.perf_counter() - t0
        rec["device_bytes_in_use"] = int(torch.cuda.mem_get_info(dev)[1]
                                         - torch.cuda.mem_get_info(dev)[0])
        hist.append(rec)
        if log:
            log(rec)
    return {"u": u, "v": v, "history": hist, "step": step}
