"""Checks on the torch tensors whose device pointers go to the kernels.

A kernel writes every row and component a launch covers, and reads every
initial condition, with no bounds of its own: a tensor of the wrong size,
dtype, device or layout would have it read or write past an allocation.  So
every tensor is checked here, on the host, before its pointer is taken.
"""
from __future__ import annotations


def check_device_tensor(t, name: str, shape: tuple, dtype, device) -> None:
    """Raise TypeError or ValueError unless t is a contiguous tensor of
    exactly `shape` and `dtype` on the GPU `device`."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor, not {type(t).__name__}")
    if not t.is_cuda:
        raise TypeError(f"{name} must be on a GPU, not on {t.device}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}; the trace runs on {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, not {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def check_initial_conditions(ics, names=("x0", "y0", "kx0", "ky0")):
    """The four initial-condition tensors: float64 on one GPU.  Returns the
    device."""
    import torch
    dev = None
    for a, name in zip(ics, names):
        if not isinstance(a, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if not a.is_cuda:
            raise TypeError(f"{name} must be on a GPU, not on {a.device}")
        if a.dtype != torch.float64:
            raise TypeError(f"{name} must be float64, not {a.dtype}")
        if dev is None:
            dev = a.device
        elif a.device != dev:
            raise ValueError("initial conditions must be on one device")
    return dev
