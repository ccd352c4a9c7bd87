"""`import mantaray` for callers of the reference's Python package
(python/mantaray/__init__.py:1-3): `single_ray` and `ray_tracing`, with the
reference's arguments and Dataset results, served by mantaray_amd's HIP
kernels (mantaray_amd/core.py)."""
from .core import ray_tracing, single_ray

__all__ = ["single_ray", "ray_tracing"]
