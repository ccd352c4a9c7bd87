"""mantaray_amd: the many-ray wave integrator of mantaray on MI355X (gfx950).

Drop-in for `mantaray.single_ray` / `mantaray.ray_tracing` (the reference's
python/mantaray/__init__.py:1-3) on hand-written HIP kernels; see core.py.
"""
from .core import num_steps, ray_tracing, rhs, single_ray, trace_many, trace_many_device
from .fields import Bathymetry, Current

__all__ = ["single_ray", "ray_tracing", "trace_many", "trace_many_device", "rhs",
           "num_steps", "Bathymetry", "Current"]
__version__ = "0.1.0"
