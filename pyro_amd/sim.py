"""Sim: the per-timestep fire solver on the GPU, with the reference's API.

Python face of ``libpyro_hip.so`` (include/pyro_hip.h).  Method names and
meaning follow the reference's ``src/sim.rs``: ``Sim.new`` (:108),
``step`` (:246), ``read_stats`` (:510), ``read_scalars`` (:494), ``cells``
(:234), ``Stats.flame_height`` (:91-104).
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .scene import Params, Scene

STATS_N = 1040

# Field ids (include/pyro_hip.h PyroField)
FIELDS = {
    "s0": 0, "s1": 1, "s0b": 2, "s1b": 3, "h0": 4, "h1": 5, "sol": 6, "sol_out": 7,
    "pyr": 8, "flags": 9, "curl": 10, "u": 11, "v": 12, "w": 13, "ub": 14, "vb": 15,
    "wb": 16, "hu": 17, "hv": 18, "hw": 19, "render": 20,
}
LEVEL_BASE = 32
STAGES = [
    "solid", "curl", "adv_vel1", "adv_vel2", "adv_s1", "adv_s2", "react", "forces",
    "divergence", "vcycles", "resid", "project",
]
CREATE_NO_GRAPH = 1


def level_field(level: int, what: str) -> int:
    return LEVEL_BASE + 4 * level + {"phi": 0, "b": 1, "flags": 2}[what]


@dataclass
class Stats:
    """Per-step statistics (sim.rs:73-87), FP64."""

    hrr: float = 0.0
    rad: float = 0.0
    pyro: float = 0.0
    soot: float = 0.0
    max_t: float = 0.0
    max_u: float = 0.0
    fuel: float = 0.0
    solid: float = 0.0
    smoke: float = 0.0
    area: float = 0.0
    r2: float = 0.0
    b2: float = 0.0
    resid: float = 0.0
    layers: np.ndarray = field(default_factory=lambda: np.zeros(0))
    raw: np.ndarray = field(default_factory=lambda: np.zeros(STATS_N))

    @classmethod
    def from_raw(cls, v: np.ndarray, ny: int) -> "Stats":
        resid = float(np.sqrt(v[10] / v[11])) if v[11] > 0 else 0.0  # sim.rs:523-527
        return cls(hrr=v[0], rad=v[1], pyro=v[2], soot=v[3], max_t=v[4], max_u=v[5], fuel=v[6],
                   solid=v[7], smoke=v[8], area=v[9], r2=v[10], b2=v[11], resid=resid,
                   layers=v[16:16 + min(ny, 1024)].copy(), raw=v.copy())

    def flame_height(self, dx: float, frac: float = 0.95) -> float:
        """Height where cumulative per-layer HRR reaches frac of the total (sim.rs:91-104)."""
        tot = float(np.sum(self.layers))
        if tot <= 0.0:
            return 0.0
        acc = 0.0
        for j, v in enumerate(self.layers):
            acc += v
            if acc >= frac * tot:
                return (j + 1.0) * dx
        return len(self.layers) * dx


class Sim:
    """GPU simulation state (sim.rs:42-68) behind the C ABI."""

    def __init__(self, params: Params, sol_init: np.ndarray, s0_init=None, pressure_cycles: int = 4,
                 ign_until: float = 0.0, use_graph: bool = True):
        lib = _native.hip()
        self.p = params
        self.dims = (params["nx"], params["ny"], params["nz"])
        self.dx = params["dx"]
        n = self.dims[0] * self.dims[1] * self.dims[2]
        sol = None if sol_init is None else np.ascontiguousarray(sol_init, dtype=np.float32)
        if sol is not None and sol.size != 4 * n:
            raise ValueError("sol_init must hold N x 4 floats")
        s0p = None
        if s0_init is not None:
            s0 = np.ascontiguousarray(s0_init, dtype=np.float32)
            if s0.size != 4 * n:
                raise ValueError("s0_init must hold N x 4 floats")
            s0p = s0.ctypes.data
        h = C.c_void_p()
        blob = (C.c_uint8 * 576).from_buffer_copy(params.bytes())
        _native.check_hip(lib.pyro_sim_create(blob, 576, None if sol is None else sol.ctypes.data, s0p,
                                              int(pressure_cycles),
                                              float(ign_until), 0 if use_graph else CREATE_NO_GRAPH,
                                              C.byref(h)))
        self._h = h
        self.cycles = int(pressure_cycles)
        self.nv = (self.dims[0] + 1) * (self.dims[1] + 1) * (self.dims[2] + 1)

    @classmethod
    def new(cls, scene: Scene, use_graph: bool = True, device_voxelize: bool = False) -> "Sim":
        """Sim::new(gpu, scene) (sim.rs:108-232).  device_voxelize: SOL is
        voxelized on the GPU (pyro_sim_voxelize; the same bits) instead of on
        the host and uploaded."""
        p = scene.params()
        if not device_voxelize:
            return cls(p, scene.voxelize(), None, int(scene.get("sim.pressure_cycles")),
                       scene.get("ignition.until"), use_graph=use_graph)
        sim = cls(p, None, None, int(scene.get("sim.pressure_cycles")), scene.get("ignition.until"),
                  use_graph=use_graph)
        sim.voxelize(scene)
        return sim

    def voxelize(self, scene: Scene) -> None:
        """SOL of `scene` (same grid) voxelized on the device (pyro_sim_voxelize)."""
        sh = scene.shapes()
        dims, _, t0, seed = scene.voxel_spec()
        if tuple(dims) != tuple(self.dims):
            raise ValueError("scene grid differs from the simulation's")
        _native.check_hip(_native.hip().pyro_sim_voxelize(self._h, sh.ctypes.data if sh.size else None, sh.size,
                                                          t0, seed))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _native.hip().pyro_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- step API -------------------------------------------------------
    def step(self, dt: float, want_stats: bool = False) -> None:
        _native.check_hip(_native.hip().pyro_sim_step(self._h, float(dt), int(bool(want_stats))))

    def read_stats(self) -> Stats:
        out = np.zeros(STATS_N, dtype=np.float64)
        _native.check_hip(_native.hip().pyro_sim_read_stats(self._h, out.ctypes.data, STATS_N))
        return Stats.from_raw(out, self.dims[1])

    def read_scalars(self):
        """(s0, s1) as (nz, ny, nx, 4) float32 arrays (sim.rs:494-508)."""
        return self.read_field("s0"), self.read_field("s1")

    def cells(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    @property
    def time(self) -> float:
        t = C.c_double()
        _native.check_hip(_native.hip().pyro_sim_info(self._h, None, None, C.byref(t), None, None))
        return t.value

    @property
    def levels(self) -> int:
        n = C.c_uint32()
        _native.check_hip(_native.hip().pyro_sim_info(self._h, None, None, None, None, C.byref(n)))
        return n.value

    def level_dims(self, level: int):
        d = (C.c_uint32 * 3)()
        _native.check_hip(_native.hip().pyro_sim_level_dims(self._h, level, d))
        return (d[0], d[1], d[2])

    # -- fields ---------------------------------------------------------
    def _shape(self, fid: int):
        nx, ny, nz = self.dims
        if fid in (0, 1, 2, 3, 4, 5, 6, 7, 8, 10):
            return (nz, ny, nx, 4), np.float32
        if fid == 9:
            return (nz, ny, nx), np.uint32
        if 11 <= fid <= 19:
            return (nz + 1, ny + 1, nx + 1), np.float32
        if fid == 20:
            return (3, nz, ny, nx, 4), np.float16
        l, q = divmod(fid - LEVEL_BASE, 4)
        d = self.level_dims(l)
        return (d[2], d[1], d[0]), (np.uint32 if q == 2 else np.float32)

    def read_field(self, name) -> np.ndarray:
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        shape, dt = self._shape(fid)
        out = np.zeros(shape, dtype=dt)
        _native.check_hip(_native.hip().pyro_sim_read_field(self._h, fid, out.ctypes.data, out.nbytes))
        return out

    def save(self, path) -> None:
        """Checkpoint everything a step reads (pyro_sim_save)."""
        _native.check_hip(_native.hip().pyro_sim_save(self._h, str(path).encode()))

    def load(self, path) -> None:
        """Restore a checkpoint of a simulation on the same grid (pyro_sim_load)."""
        _native.check_hip(_native.hip().pyro_sim_load(self._h, str(path).encode()))

    def write_field(self, name, arr: np.ndarray) -> None:
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        shape, dt = self._shape(fid)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.size != int(np.prod(shape)):
            raise ValueError(f"field {name}: expected shape {shape}")
        _native.check_hip(_native.hip().pyro_sim_write_field(self._h, fid, a.ctypes.data, a.nbytes))

    # -- hooks ------------------------------------------------------------
    def begin_step(self, dt: float, want_stats: bool = False) -> None:
        _native.check_hip(_native.hip().pyro_sim_begin_step(self._h, float(dt), int(bool(want_stats))))

    def run_stage(self, stage) -> None:
        sid = STAGES.index(stage) if isinstance(stage, str) else int(stage)
        _native.check_hip(_native.hip().pyro_sim_run_stage(self._h, sid))

    def end_step(self) -> None:
        _native.check_hip(_native.hip().pyro_sim_end_step(self._h))

    def abort_step(self) -> None:
        """Close an open step without ending it (roles, time and frame unchanged)."""
        _native.check_hip(_native.hip().pyro_sim_abort_step(self._h))

    def time_stage(self, stage, reps: int) -> float:
        sid = STAGES.index(stage) if isinstance(stage, str) else int(stage)
        ms = C.c_float()
        _native.check_hip(_native.hip().pyro_sim_time_stage(self._h, sid, int(reps), C.byref(ms)))
        return ms.value

    def time_kernel(self, kernel: str, level: int, reps: int) -> float:
        kid = {"mg_smooth": 0, "mg_restrict": 1, "mg_prolong": 2, "mg_coarse": 3, "render_pack": 4, "adv_pass1": 5,
               "adv_pass2": 6, "solid_curl": 7}[kernel]
        ms = C.c_float()
        _native.check_hip(_native.hip().pyro_sim_time_kernel(self._h, kid, int(level), int(reps), C.byref(ms)))
        return ms.value

    def add_velocity_noise(self, amp: float, seed: int) -> None:
        _native.check_hip(_native.hip().pyro_sim_add_velocity_noise(self._h, float(amp), int(seed)))

    def check_finite(self) -> None:
        _native.check_hip(_native.hip().pyro_sim_check_finite(self._h))

    def sync(self) -> None:
        _native.check_hip(_native.hip().pyro_sim_sync(self._h))

    # -- renderer hand-off (render.rs:255-277, render.wgsl:16-37) ----------
    def render_pack(self, dst: "PinnedBuffer | None" = None):
        """Pack the current state into the renderer's volumes (vol, sol,
        vel: (3, nz, ny, nx, 4) float16).  Without `dst` the call waits and
        returns a new array; with a PinnedBuffer the copy is asynchronous and
        overlaps the steps enqueued next: call render_wait() before reading
        dst.array."""
        nx, ny, nz = self.dims
        nbytes = 24 * nx * ny * nz
        lib = _native.hip()
        if dst is None:
            out = np.zeros((3, nz, ny, nx, 4), np.float16)
            _native.check_hip(lib.pyro_sim_render_pack(self._h, out.ctypes.data, nbytes))
            _native.check_hip(lib.pyro_sim_render_wait(self._h))
            return out
        if dst.ptr is None:
            raise ValueError("render_pack: the buffer is closed")
        if dst.nbytes != nbytes:
            raise ValueError(f"render_pack: the buffer must hold {nbytes} bytes")
        _native.check_hip(lib.pyro_sim_render_pack(self._h, dst.ptr, nbytes))
        return dst

    def render_wait(self) -> None:
        _native.check_hip(_native.hip().pyro_sim_render_wait(self._h))

    def render_pack_device(self) -> None:
        """Pack into the device buffer only (the exported one after
        render_export), stream-ordered; render_wait() before it is read."""
        _native.check_hip(_native.hip().pyro_sim_render_pack(self._h, None, 0))

    def render_export(self) -> tuple[int, int]:
        """A new file descriptor of the exportable allocation that holds the
        packed volumes (pyro_sim_render_export) and the allocation's size in
        bytes; the caller closes the descriptor."""
        fd, nbytes = C.c_int(-1), C.c_uint64(0)
        _native.check_hip(_native.hip().pyro_sim_render_export(self._h, C.byref(fd), C.byref(nbytes)))
        return fd.value, nbytes.value


def shared_read(fd: int, alloc_bytes: int, dims) -> np.ndarray:
    """The packed volumes read back through an exported descriptor
    (pyro_shared_read: import, map, copy), as (3, nz, ny, nx, 4) float16."""
    nx, ny, nz = dims
    out = np.zeros((3, nz, ny, nx, 4), np.float16)
    _native.check_hip(_native.hip().pyro_shared_read(fd, alloc_bytes, out.ctypes.data, out.nbytes))
    return out


class PinnedBuffer:
    """Page-locked host memory (pyro_host_alloc) for the renderer hand-off,
    viewed as the (3, nz, ny, nx, 4) float16 volumes of one simulation.  The
    memory is released when the last array viewing it is gone, so a view
    kept past close() stays valid.  Call Sim.render_wait() before reading
    the array or dropping the buffer while a copy into it may be running."""

    def __init__(self, dims):
        nx, ny, nz = dims
        self.nbytes = 24 * nx * ny * nz
        p = C.c_void_p()
        _native.check_hip(_native.hip().pyro_host_alloc(self.nbytes, C.byref(p)))
        raw = (C.c_uint8 * self.nbytes).from_address(p.value)
        weakref.finalize(raw, _host_free, p.value)
        self.ptr = p
        self.array = np.frombuffer(raw, dtype=np.float16).reshape(3, nz, ny, nx, 4)

    def close(self) -> None:
        self.array = None
        self.ptr = None


def _host_free(addr: int) -> None:
    _native.hip().pyro_host_free(C.c_void_p(addr))


def debug_math(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    yp = None
    if y is not None:
        y = np.ascontiguousarray(y, dtype=np.float32)
        yp = y.ctypes.data
    _native.check_hip(_native.hip().pyro_debug_math(fn, x.ctypes.data, yp, out.ctypes.data, x.size))
    return out


def debug_copy(n_floats: int, reps: int) -> float:
    """Average ms of a plain device copy of n_floats floats (calibration)."""
    ms = C.c_float()
    _native.check_hip(_native.hip().pyro_debug_copy(int(n_floats), int(reps), C.byref(ms)))
    return ms.value
