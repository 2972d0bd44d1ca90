"""pyro_amd -- MI355X-native per-timestep fire solver (HIP on gfx950).

Drop-in for the wgpu dispatch of RuifengFu/pyro's ``Sim`` (src/sim.rs):
``Scene`` (TOML -> Params + SOL, src/scene.rs) and ``Sim`` (step /
read_stats / read_scalars) over the C ABI of ``include/pyro_hip.h``.
"""
from ._native import NativeLibraryMissing, PyroError
from .scene import Params, Scene, material_index, scene_hash
from .sim import FIELDS, STAGES, PinnedBuffer, Sim, Stats, level_field, shared_read

__all__ = [
    "NativeLibraryMissing", "PyroError", "Params", "Scene", "material_index", "scene_hash",
    "FIELDS", "STAGES", "PinnedBuffer", "Sim", "Stats", "level_field", "shared_read",
]
