"""MI355X (gfx950) HIP implementation of the per-frame hair compute chain of
Scthe/frostbitten-hair-webgpu: PBD strand simulation, strand-space shading,
tile pass and fine OIT pass behind the reference's src/passes API.

The directory name is not a Python identifier; load it with
`__graft_entry__.load_package()`, which registers it as
`frostbitten_hair_webgpu_amd`.
"""
from . import formats, layouts, matrices  # noqa: F401

__all__ = ["formats", "layouts", "matrices"]
