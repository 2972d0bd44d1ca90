"""ctypes loader for the in-tree native libraries.

``libpyro_hip.so`` (the gfx950 solver) and ``libpyro_host.so`` (scene
plumbing) are built by ``pyro_amd/build.py`` into ``pyro_amd/lib``.  There is
no fallback: if a library is missing, loading raises ``NativeLibraryMissing``
so that nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"


class NativeLibraryMissing(RuntimeError):
    pass


class PyroError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_hip = None
_host = None

P = C.c_void_p
f32 = C.c_float
u32 = C.c_uint32
i32 = C.c_int
sz = C.c_size_t


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python pyro_amd/build.py`"
        )
    return C.CDLL(str(path))


def hip() -> C.CDLL:
    """The gfx950 solver library (include/pyro_hip.h)."""
    global _hip
    if _hip is None:
        lib = _load("libpyro_hip.so")
        lib.pyro_last_error.restype = C.c_char_p
        lib.pyro_sim_create.argtypes = [P, sz, P, P, u32, f32, u32, C.POINTER(P)]
        lib.pyro_sim_destroy.argtypes = [P]
        lib.pyro_sim_step.argtypes = [P, f32, i32]
        lib.pyro_sim_read_stats.argtypes = [P, P, sz]
        lib.pyro_sim_read_field.argtypes = [P, i32, P, sz]
        lib.pyro_sim_write_field.argtypes = [P, i32, P, sz]
        lib.pyro_sim_info.argtypes = [P, P, P, P, P, P]
        lib.pyro_sim_level_dims.argtypes = [P, i32, P]
        lib.pyro_sim_sync.argtypes = [P]
        lib.pyro_voxelize.argtypes = [P, sz, P, f32, f32, u32, P, i32]
        lib.pyro_sim_voxelize.argtypes = [P, P, sz, f32, u32]
        lib.pyro_voxelize_noalloc.argtypes = [P, sz, P, f32, f32, u32, P]
        lib.pyro_sim_render_pack.argtypes = [P, P, sz]
        lib.pyro_sim_render_wait.argtypes = [P]
        lib.pyro_sim_render_export.argtypes = [P, C.POINTER(i32), C.POINTER(C.c_uint64)]
        lib.pyro_shared_read.argtypes = [i32, C.c_uint64, P, sz]
        lib.pyro_host_alloc.argtypes = [sz, C.POINTER(P)]
        lib.pyro_host_free.argtypes = [P]
        lib.pyro_sim_device_ptr.argtypes = [P, i32, C.POINTER(P), C.POINTER(sz)]
        lib.pyro_sim_begin_step.argtypes = [P, f32, i32]
        lib.pyro_sim_run_stage.argtypes = [P, i32]
        lib.pyro_sim_end_step.argtypes = [P]
        lib.pyro_sim_abort_step.argtypes = [P]
        lib.pyro_sim_time_stage.argtypes = [P, i32, i32, C.POINTER(f32)]
        lib.pyro_sim_time_kernel.argtypes = [P, i32, i32, i32, C.POINTER(f32)]
        lib.pyro_sim_add_velocity_noise.argtypes = [P, f32, u32]
        lib.pyro_sim_check_finite.argtypes = [P]
        lib.pyro_sim_save.argtypes = [P, C.c_char_p]
        lib.pyro_sim_load.argtypes = [P, C.c_char_p]
        lib.pyro_debug_math.argtypes = [i32, P, P, P, sz]
        lib.pyro_debug_copy.argtypes = [sz, i32, C.POINTER(f32)]
        _hip = lib
    return _hip


def host() -> C.CDLL:
    """The scene plumbing library (include/pyro_scene.h)."""
    global _host
    if _host is None:
        lib = _load("libpyro_host.so")
        lib.pyro_scene_last_error.restype = C.c_char_p
        lib.pyro_scene_load.argtypes = [C.c_char_p, C.POINTER(P)]
        lib.pyro_scene_parse.argtypes = [C.c_char_p, C.POINTER(P)]
        lib.pyro_scene_free.argtypes = [P]
        lib.pyro_scene_get.argtypes = [P, C.c_char_p, C.POINTER(C.c_double)]
        lib.pyro_scene_set.argtypes = [P, C.c_char_p, C.c_double]
        lib.pyro_scene_has_ignition.argtypes = [P]
        lib.pyro_scene_grid.argtypes = [P, P, C.POINTER(f32)]
        lib.pyro_scene_params.argtypes = [P, P, sz]
        lib.pyro_scene_voxelize.argtypes = [P, P, sz]
        lib.pyro_material_index.argtypes = [C.c_char_p]
        lib.pyro_scene_hash.argtypes = [u32]
        lib.pyro_scene_hash.restype = f32
        lib.pyro_scene_shapes.argtypes = [P, P, sz, C.POINTER(sz)]
        lib.pyro_scene_voxel_spec.argtypes = [P, P, C.POINTER(f32), C.POINTER(f32), C.POINTER(u32)]
        _host = lib
    return _host


def check_hip(rc: int) -> None:
    if rc != 0:
        raise PyroError(rc, hip().pyro_last_error().decode(errors="replace"))


def check_host(rc: int) -> None:
    if rc < 0:
        raise PyroError(rc, host().pyro_scene_last_error().decode(errors="replace"))
