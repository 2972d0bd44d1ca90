"""CPU oracle of pyro's per-timestep solver -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/liboracle.so`` (built from ``pyro_oracle.c`` by
``oracle/build.py``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker / CPU baseline; the product (``pyro_amd``) never does.

``OSim`` exposes the same methods as ``pyro_amd.Sim`` so parity tests drive
both with identical calls.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None

FIELDS = {
    "s0": 0, "s1": 1, "s0b": 2, "s1b": 3, "h0": 4, "h1": 5, "sol": 6, "sol_out": 7,
    "pyr": 8, "flags": 9, "curl": 10, "u": 11, "v": 12, "w": 13, "ub": 14, "vb": 15,
    "wb": 16, "hu": 17, "hv": 18, "hw": 19,
}
LEVEL_BASE = 32
STAGES = [
    "solid", "curl", "adv_vel1", "adv_vel2", "adv_s1", "adv_s2", "react", "forces",
    "divergence", "vcycles", "resid", "project",
]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `python oracle/build.py`")
        L = C.CDLL(str(path))
        V = C.c_void_p
        L.osim_create.restype = V
        L.osim_create.argtypes = [V, C.c_size_t, V, V, C.c_uint32, C.c_float]
        L.osim_destroy.argtypes = [V]
        L.osim_levels.argtypes = [V, V]
        L.osim_step.argtypes = [V, C.c_float, C.c_int]
        L.osim_begin_step.argtypes = [V, C.c_float, C.c_int]
        L.osim_run_stage.argtypes = [V, C.c_int]
        L.osim_end_step.argtypes = [V]
        L.osim_read_field.argtypes = [V, C.c_int, V, C.c_size_t]
        L.osim_write_field.argtypes = [V, C.c_int, V, C.c_size_t]
        L.osim_read_stats.argtypes = [V, V]
        L.osim_time.argtypes = [V]
        L.osim_time.restype = C.c_double
        L.osim_add_velocity_noise.argtypes = [V, C.c_float, C.c_uint32]
        L.osim_set_clock.argtypes = [V, C.c_double, C.c_uint32, C.c_float]
        L.osim_get_params.argtypes = [V, V]
        L.osim_set_threads.argtypes = [C.c_int]
        L.osim_render_pack.argtypes = [V, V]
        L.oracle_expf.argtypes = [C.c_float]
        L.oracle_expf.restype = C.c_float
        L.oracle_noise.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_noise.restype = C.c_float
        L.omg_create.restype = V
        L.omg_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, V]
        L.omg_destroy.argtypes = [V]
        L.omg_level_dims.argtypes = [V, C.c_int, V]
        L.omg_set.argtypes = [V, C.c_int, V, V]
        L.omg_get.argtypes = [V, C.c_int, V, V, V]
        L.omg_smooth.argtypes = [V, C.c_int, C.c_uint32, C.c_float]
        L.omg_restrict.argtypes = [V, C.c_int]
        L.omg_prolong.argtypes = [V, C.c_int]
        L.omg_vcycles.argtypes = [V, C.c_uint32, C.c_uint32]
        L.omg_resid_norms.argtypes = [V, V, V]
        L.oracle_poisson_cg64.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, V, V, V,
                                          C.c_double, C.c_int, V]
        L.oracle_residual64.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, V, V, V, V]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().osim_set_threads(int(n))


def expf(x: np.ndarray) -> np.ndarray:
    f = lib().oracle_expf
    x = np.asarray(x, dtype=np.float32).ravel()
    return np.array([f(float(v)) for v in x], dtype=np.float32)


def noise(comp: int, idx: int, seed: int) -> float:
    return lib().oracle_noise(comp, idx, seed)


# Checkpoint files of pyro_sim_save (pyro_amd/csrc/sim.hip): this fixed
# header, the 576-byte Params, then seven raw little-endian FP32 fields in the
# device layout -- u, v, w of the current velocity (padded (nz+1, ny+1, nx+1)
# each), s0, s1, the current SOL ((nz, ny, nx, 4) each) and level-0 phi
# ((nz, ny, nx)).  Read and written here independently of the product code.
CKPT_MAGIC = b"PYROCKPT"
CKPT_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("nx", "<u4"), ("ny", "<u4"), ("nz", "<u4"),
                        ("params_bytes", "<u4"), ("time", "<f4"), ("ign_until", "<f4"), ("frame", "<u4"),
                        ("cycles", "<u4"), ("nfields", "<u4"), ("pad", "<u4")])
CKPT_FIELDS = ("u", "v", "w", "s0", "s1", "sol", "phi")


def _ckpt_shapes(nx, ny, nz):
    fv, c4 = (nz + 1, ny + 1, nx + 1), (nz, ny, nx, 4)
    return (fv, fv, fv, c4, c4, c4, (nz, ny, nx))


def read_checkpoint(path) -> dict:
    """Header values, the Params blob and the seven fields of a checkpoint."""
    raw = Path(path).read_bytes()
    h = np.frombuffer(raw, CKPT_HEADER, count=1)[0]
    if h["magic"] != CKPT_MAGIC or int(h["version"]) != 1 or int(h["params_bytes"]) != 576:
        raise ValueError(f"{path}: not a version-1 checkpoint")
    nx, ny, nz = int(h["nx"]), int(h["ny"]), int(h["nz"])
    out = {k: h[k].item() for k in ("time", "ign_until", "frame", "cycles")}
    off = CKPT_HEADER.itemsize
    out["params"] = raw[off:off + 576]
    off += 576
    if int(h["nfields"]) != len(CKPT_FIELDS):
        raise ValueError(f"{path}: {int(h['nfields'])} fields")
    for name, shape in zip(CKPT_FIELDS, _ckpt_shapes(nx, ny, nz)):
        cnt = int(np.prod(shape))
        if off + 4 * cnt > len(raw):
            raise ValueError(f"{path}: truncated")
        out[name] = np.frombuffer(raw, "<f4", count=cnt, offset=off).reshape(shape).copy()
        off += 4 * cnt
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    out["dims"] = (nx, ny, nz)
    return out


class OSim:
    """CPU restatement of one Sim (same methods as pyro_amd.Sim)."""

    @classmethod
    def from_checkpoint(cls, path) -> "OSim":
        """Continue a saved run: Params, clock, V-cycle count, velocity, s0,
        s1, SOL and the level-0 phi warm start from the file."""
        ck = read_checkpoint(path)
        o = cls(ck["params"], ck["sol"], ck["s0"], ck["cycles"], ck["ign_until"])
        for name in ("u", "v", "w", "s1"):
            o.write_field(name, ck[name])
        o.write_field(LEVEL_BASE, ck["phi"])
        lib().osim_set_clock(o._h, float(ck["time"]), int(ck["frame"]), float(ck["ign_until"]))
        return o

    def params_blob(self) -> bytes:
        buf = (C.c_uint8 * 576)()
        lib().osim_get_params(self._h, buf)
        return bytes(buf)

    def save_checkpoint(self, path) -> None:
        """Write the current state in the checkpoint format (between steps)."""
        nx, ny, nz = self.dims
        blob = self.params_blob()
        frame = int(np.frombuffer(blob, np.uint32, count=1, offset=12)[0])
        h = np.zeros(1, CKPT_HEADER)
        h["magic"], h["version"], h["nx"], h["ny"], h["nz"] = CKPT_MAGIC, 1, nx, ny, nz
        h["params_bytes"], h["time"], h["ign_until"] = 576, self.time, self.ign_until
        h["frame"], h["cycles"], h["nfields"] = frame, self.cycles, len(CKPT_FIELDS)
        parts = [h.tobytes(), blob]
        for name in CKPT_FIELDS:
            a = self.read_field(LEVEL_BASE if name == "phi" else name)
            parts.append(np.ascontiguousarray(a, "<f4").tobytes())
        Path(path).write_bytes(b"".join(parts))

    def __init__(self, params, sol_init, s0_init=None, pressure_cycles=4, ign_until=0.0):
        blob = params.bytes() if hasattr(params, "bytes") else bytes(params)
        self.dims = tuple(int(v) for v in np.frombuffer(blob, dtype=np.uint32, count=3))
        self.dx = float(np.frombuffer(blob, dtype=np.float32, count=1, offset=16)[0])
        sol = np.ascontiguousarray(sol_init, dtype=np.float32)
        s0p = None
        if s0_init is not None:
            self._s0 = np.ascontiguousarray(s0_init, dtype=np.float32)
            s0p = self._s0.ctypes.data
        buf = (C.c_uint8 * 576).from_buffer_copy(blob)
        self._h = lib().osim_create(buf, 576, sol.ctypes.data, s0p, int(pressure_cycles), float(ign_until))
        if not self._h:
            raise RuntimeError("osim_create failed")
        self.cycles = pressure_cycles
        self.ign_until = float(np.float32(ign_until))

    def close(self):
        if getattr(self, "_h", None):
            lib().osim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt, want_stats=False):
        lib().osim_step(self._h, float(dt), int(bool(want_stats)))

    def begin_step(self, dt, want_stats=False):
        lib().osim_begin_step(self._h, float(dt), int(bool(want_stats)))

    def run_stage(self, stage):
        sid = STAGES.index(stage) if isinstance(stage, str) else int(stage)
        if lib().osim_run_stage(self._h, sid) != 0:
            raise ValueError(stage)

    def end_step(self):
        lib().osim_end_step(self._h)

    @property
    def time(self):
        return lib().osim_time(self._h)

    @property
    def levels(self):
        return lib().osim_levels(self._h, None)

    def level_dims(self, level):
        d = (C.c_uint32 * 24)()
        lib().osim_levels(self._h, d)
        return (d[3 * level], d[3 * level + 1], d[3 * level + 2])

    def _shape(self, fid):
        nx, ny, nz = self.dims
        if fid in (0, 1, 2, 3, 4, 5, 6, 7, 8, 10):
            return (nz, ny, nx, 4), np.float32
        if fid == 9:
            return (nz, ny, nx), np.uint32
        if 11 <= fid <= 19:
            return (nz + 1, ny + 1, nx + 1), np.float32
        l, q = divmod(fid - LEVEL_BASE, 4)
        d = self.level_dims(l)
        return (d[2], d[1], d[0]), (np.uint32 if q == 2 else np.float32)

    def read_field(self, name):
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        shape, dt = self._shape(fid)
        out = np.zeros(shape, dtype=dt)
        if lib().osim_read_field(self._h, fid, out.ctypes.data, out.nbytes) != 0:
            raise ValueError(f"bad field {name}")
        return out

    def write_field(self, name, arr):
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        shape, dt = self._shape(fid)
        a = np.ascontiguousarray(arr, dtype=dt)
        if lib().osim_write_field(self._h, fid, a.ctypes.data, a.nbytes) != 0:
            raise ValueError(f"bad field {name}")

    def read_stats_raw(self):
        out = np.zeros(1040, dtype=np.float64)
        lib().osim_read_stats(self._h, out.ctypes.data)
        return out

    def render_pack(self):
        """render.wgsl:16-37: (vol, sol, vel), each (nz, ny, nx, 4) f32, the
        values before the rgba16float store."""
        nx, ny, nz = self.dims
        out = np.zeros((3, nz, ny, nx, 4), np.float32)
        lib().osim_render_pack(self._h, out.ctypes.data)
        return out

    def add_velocity_noise(self, amp, seed):
        lib().osim_add_velocity_noise(self._h, float(amp), int(seed))


class OMg:
    """Stand-alone multigrid hierarchy of the oracle (mg.wgsl operator)."""

    def __init__(self, dims, h, levels, flags0=None):
        nx, ny, nz = dims
        self.dims = dims
        fp = None
        if flags0 is not None:
            self._f = np.ascontiguousarray(flags0, dtype=np.uint32)
            fp = self._f.ctypes.data
        self._h = lib().omg_create(nx, ny, nz, float(h), int(levels), fp)

    def __del__(self):
        try:
            lib().omg_destroy(self._h)
        except Exception:
            pass

    @property
    def levels(self):
        d = (C.c_uint32 * 3)()
        return lib().omg_level_dims(self._h, -1, d)

    def level_dims(self, l):
        d = (C.c_uint32 * 3)()
        lib().omg_level_dims(self._h, l, d)
        return (d[0], d[1], d[2])

    def set(self, l, phi=None, b=None):
        pp = None if phi is None else np.ascontiguousarray(phi, dtype=np.float32)
        bp = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
        lib().omg_set(self._h, l, None if pp is None else pp.ctypes.data, None if bp is None else bp.ctypes.data)

    def get(self, l):
        nx, ny, nz = self.level_dims(l)
        phi = np.zeros((nz, ny, nx), np.float32)
        b = np.zeros_like(phi)
        f = np.zeros((nz, ny, nx), np.uint32)
        lib().omg_get(self._h, l, phi.ctypes.data, b.ctypes.data, f.ctypes.data)
        return phi, b, f

    def smooth(self, l, iters, omega):
        lib().omg_smooth(self._h, l, iters, omega)

    def restrict(self, l):
        lib().omg_restrict(self._h, l)

    def prolong(self, l):
        lib().omg_prolong(self._h, l)

    def vcycles(self, cycles, coarse_iters=80):
        lib().omg_vcycles(self._h, cycles, coarse_iters)

    def resid_norms(self):
        r2 = C.c_double()
        b2 = C.c_double()
        lib().omg_resid_norms(self._h, C.byref(r2), C.byref(b2))
        return r2.value, b2.value


def poisson_cg64(flags, b, h, tol=1e-12, max_iter=100000, phi0=None):
    """FP64 Jacobi-PCG solve of the mg.wgsl operator; returns (phi, iters, rel_resid)."""
    nz, ny, nx = b.shape
    f = np.ascontiguousarray(flags, dtype=np.uint32)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    phi = np.zeros_like(bb) if phi0 is None else np.ascontiguousarray(phi0, dtype=np.float64).copy()
    rel = C.c_double()
    it = lib().oracle_poisson_cg64(nx, ny, nz, float(h), f.ctypes.data, bb.ctypes.data, phi.ctypes.data,
                                   float(tol), int(max_iter), C.byref(rel))
    return phi, it, rel.value


def residual64(flags, b, phi, h):
    nz, ny, nx = b.shape
    f = np.ascontiguousarray(flags, dtype=np.uint32)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    pp = np.ascontiguousarray(phi, dtype=np.float64)
    r = np.zeros_like(bb)
    lib().oracle_residual64(nx, ny, nz, float(h), f.ctypes.data, bb.ctypes.data, pp.ctypes.data, r.ctypes.data)
    return r
