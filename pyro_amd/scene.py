"""Scene plumbing: TOML scene -> Params blob + voxelized SOL field.

Python face of ``libpyro_host.so`` (include/pyro_scene.h), which restates the
reference's ``src/scene.rs``: ``load`` (:211-215), ``Scene::grid``
(:485-493), ``Scene::params`` (:523-618), ``Scene::voxelize`` (:621-660),
``material`` (:408-421).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native

# Params field order (scene.rs:426-477); 48 scalars then mats[8][12].
PARAM_FIELDS = [
    "nx", "ny", "nz", "frame", "dx", "dt", "time", "t0", "rho0", "g", "cp", "yo2",
    "vort", "cs", "buoy", "smoke_w", "wind_x", "wind_z", "wind_relax", "sponge",
    "t_ign", "s_o2", "dhc", "tau_min", "soot_form", "soot_ox", "soot_ox_t", "smoke_conv",
    "kap_gas", "h_conv", "ign_on", "ign_temp", "ign_x", "ign_y", "ign_z", "ign_r",
    "ign_flux", "ign_fuel", "char_a", "char_e", "dh_char", "burnout", "layer", "rad_len",
    "h_mass", "vort_t", "rad_frac", "seed",
]
PARAMS_BYTES = 576


class Params:
    """The 576-byte Params blob with named access (u32 for nx..frame, f32 else)."""

    def __init__(self, blob: bytes):
        if len(blob) != PARAMS_BYTES:
            raise ValueError("Params blob must be 576 bytes")
        self.blob = bytearray(blob)

    def _idx(self, name: str) -> int:
        return PARAM_FIELDS.index(name)

    def __getitem__(self, name: str):
        i = self._idx(name)
        if i < 4:
            return int(np.frombuffer(self.blob, dtype=np.uint32, count=1, offset=4 * i)[0])
        return float(np.frombuffer(self.blob, dtype=np.float32, count=1, offset=4 * i)[0])

    def __setitem__(self, name: str, value) -> None:
        i = self._idx(name)
        dt = np.uint32 if i < 4 else np.float32
        self.blob[4 * i: 4 * i + 4] = np.array([value], dtype=dt).tobytes()

    @property
    def mats(self) -> np.ndarray:
        return np.frombuffer(bytes(self.blob), dtype=np.float32, offset=192).reshape(8, 12)

    def bytes(self) -> bytes:
        return bytes(self.blob)


@dataclass
class Grid:
    dims: tuple
    dx: float


# PyroShape (pyro_scene.h): kind, slot, inert, density, c[3], h[3], sin_yaw,
# cos_yaw, a[3], b[3], r, depth
SHAPE_DTYPE = np.dtype([("kind", np.uint32), ("slot", np.uint32), ("inert", np.uint32), ("density", np.float32),
                        ("c", np.float32, 3), ("h", np.float32, 3), ("sin_yaw", np.float32),
                        ("cos_yaw", np.float32), ("a", np.float32, 3), ("b", np.float32, 3), ("r", np.float32),
                        ("depth", np.float32)])


class Scene:
    """A parsed scene (scene.rs:8-25)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def load(cls, path: str) -> "Scene":
        h = C.c_void_p()
        _native.check_host(_native.host().pyro_scene_load(str(path).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def parse(cls, text: str) -> "Scene":
        h = C.c_void_p()
        _native.check_host(_native.host().pyro_scene_parse(text.encode(), C.byref(h)))
        return cls(h.value)

    def __del__(self):
        try:
            if self._h:
                _native.host().pyro_scene_free(self._h)
                self._h = None
        except Exception:
            pass

    def get(self, key: str) -> float:
        v = C.c_double()
        _native.check_host(_native.host().pyro_scene_get(self._h, key.encode(), C.byref(v)))
        return v.value

    def set(self, key: str, value: float) -> None:
        _native.check_host(_native.host().pyro_scene_set(self._h, key.encode(), float(value)))

    @property
    def has_ignition(self) -> bool:
        return bool(_native.host().pyro_scene_has_ignition(self._h))

    def grid(self) -> Grid:
        d = (C.c_uint32 * 3)()
        dx = C.c_float()
        _native.check_host(_native.host().pyro_scene_grid(self._h, d, C.byref(dx)))
        return Grid((d[0], d[1], d[2]), dx.value)

    def params(self) -> Params:
        buf = (C.c_uint8 * PARAMS_BYTES)()
        _native.check_host(_native.host().pyro_scene_params(self._h, buf, PARAMS_BYTES))
        return Params(bytes(buf))

    def voxelize(self) -> np.ndarray:
        nx, ny, nz = self.grid().dims
        out = np.zeros((nz, ny, nx, 4), dtype=np.float32)
        _native.check_host(
            _native.host().pyro_scene_voxelize(self._h, out.ctypes.data, out.size)
        )
        return out

    def shapes(self) -> np.ndarray:
        """The expanded voxelization primitives (PyroShape records, pyro_scene.h)."""
        cnt = C.c_size_t()
        _native.check_host(_native.host().pyro_scene_shapes(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=SHAPE_DTYPE)
        if cnt.value:
            _native.check_host(_native.host().pyro_scene_shapes(self._h, out.ctypes.data, cnt.value, C.byref(cnt)))
        return out

    def voxel_spec(self):
        """(dims, dx, t0, seed) of Scene::voxelize."""
        d = (C.c_uint32 * 3)()
        dx, t0, seed = C.c_float(), C.c_float(), C.c_uint32()
        _native.check_host(_native.host().pyro_scene_voxel_spec(self._h, d, C.byref(dx), C.byref(t0), C.byref(seed)))
        return (d[0], d[1], d[2]), dx.value, t0.value, seed.value

    def voxelize_device(self) -> np.ndarray:
        """Scene::voxelize computed on the GPU (pyro_voxelize), copied to the host."""
        sh = self.shapes()
        dims, dx, t0, seed = self.voxel_spec()
        nx, ny, nz = dims
        out = np.zeros((nz, ny, nx, 4), dtype=np.float32)
        d = (C.c_uint32 * 3)(*dims)
        _native.check_hip(_native.hip().pyro_voxelize(sh.ctypes.data if sh.size else None, sh.size, d, dx, t0, seed,
                                                     out.ctypes.data, 0))
        return out


def material_index(name: str) -> int:
    r = _native.host().pyro_material_index(name.encode())
    _native.check_host(r)
    return r


def scene_hash(v: int) -> float:
    return float(_native.host().pyro_scene_hash(v & 0xFFFFFFFF))
