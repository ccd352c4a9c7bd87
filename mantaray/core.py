"""`mantaray.core` (python/mantaray/core.py): the same two functions, from
mantaray_amd.core."""
from mantaray_amd.core import ray_tracing, single_ray

__all__ = ["single_ray", "ray_tracing"]
