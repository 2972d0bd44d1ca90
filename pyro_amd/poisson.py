"""Stand-alone pressure solves on the GPU (include/pyro_hip.h, pyro_poisson_*).

Methods: ``vcycle`` (the reference V-cycles, sim.rs:451-491), ``refine``
(FP64-residual iterative refinement around FP32 V-cycles) and ``mgpcg``
(flexible CG with FP64 vectors and dot products, preconditioned by an FP32
V-cycle with level-consistent Dirichlet weights).

``make_c3_problem`` builds the standalone pressure benchmark case: a cubic
grid with seeded random axis-aligned solid boxes (~5 % of the cells) and a
right-hand side of 32 Gaussian blobs (sigma 4-16 cells at 256^3, scaled with
the grid; amplitude U[0.1, 5] 1/s with a random sign), zero on solids and
made zero-mean on every face-connected fluid component (required where a
component has no open face -- a pure-Neumann pocket has no solution
otherwise -- and harmless elsewhere).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native

METHODS = {"vcycle": 0, "refine": 1, "mgpcg": 2, "learned": 3}


class _Info(C.Structure):
    _fields_ = [("iterations", C.c_int), ("vcycles", C.c_int), ("rel_resid", C.c_double), ("ms", C.c_float),
                ("fallbacks", C.c_int), ("diverged", C.c_int)]


@dataclass
class SolveInfo:
    iterations: int
    vcycles: int
    rel_resid: float
    ms: float
    fallbacks: int = 0
    diverged: bool = False

    @classmethod
    def _of(cls, info: "_Info") -> "SolveInfo":
        return cls(info.iterations, info.vcycles, info.rel_resid, info.ms, info.fallbacks, bool(info.diverged))


def _lib():
    lib = _native.hip()
    if not hasattr(lib, "_poisson_ready"):
        P = C.c_void_p
        lib.pyro_poisson_last_error.restype = C.c_char_p
        lib.pyro_poisson_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, P, C.POINTER(P)]
        lib.pyro_poisson_destroy.argtypes = [P]
        lib.pyro_poisson_levels.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.pyro_poisson_set.argtypes = [P, P, P]
        lib.pyro_poisson_get.argtypes = [P, P]
        lib.pyro_poisson_residual.argtypes = [P, C.POINTER(C.c_double)]
        lib.pyro_poisson_solve.argtypes = [P, C.c_int, C.c_double, C.c_int, C.POINTER(_Info)]
        lib.pyro_poisson_attach_net.argtypes = [P, P, C.c_int]
        lib.pyro_poisson_set_sealed_projection.argtypes = [P, C.c_int]
        lib.pyro_poisson_sealed_info.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.pyro_poisson_get_rhs.argtypes = [P, P]
        lib.pyro_poisson_set_flags_from_sim.argtypes = [P, P]
        lib.pyro_poisson_set_from_sim.argtypes = [P, P, C.c_int]
        lib.pyro_sim_step_solve.argtypes = [P, P, C.c_int, C.c_double, C.c_int, C.c_float, C.c_int,
                                            C.POINTER(_Info)]
        lib._poisson_ready = True
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise _native.PyroError(rc, _lib().pyro_poisson_last_error().decode(errors="replace"))


class Poisson:
    """A pressure problem on an (nz, ny, nx) grid with solid flags (1 = solid)."""

    def __init__(self, flags: np.ndarray, h: float):
        f = np.ascontiguousarray(flags, dtype=np.uint32)
        nz, ny, nx = f.shape
        self.shape = f.shape
        self.h = float(h)
        h_ = C.c_void_p()
        _check(_lib().pyro_poisson_create(nx, ny, nz, float(h), f.ctypes.data, C.byref(h_)))
        self._h = h_

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().pyro_poisson_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def levels(self):
        a, b = C.c_int(), C.c_int()
        _check(_lib().pyro_poisson_levels(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set(self, b: np.ndarray | None = None, phi: np.ndarray | None = None) -> None:
        bp = pp = None
        if b is not None:
            bb = np.ascontiguousarray(b, dtype=np.float64)
            assert bb.shape == self.shape
            bp = bb.ctypes.data
        if phi is not None:
            ph = np.ascontiguousarray(phi, dtype=np.float64)
            assert ph.shape == self.shape
            pp = ph.ctypes.data
        _check(_lib().pyro_poisson_set(self._h, bp, pp))

    def get(self) -> np.ndarray:
        out = np.zeros(self.shape, np.float64)
        _check(_lib().pyro_poisson_get(self._h, out.ctypes.data))
        return out

    def get_rhs(self) -> np.ndarray:
        """b as the solver holds it (after the sealed-component projection)."""
        out = np.zeros(self.shape, np.float64)
        _check(_lib().pyro_poisson_get_rhs(self._h, out.ctypes.data))
        return out

    def sealed_info(self) -> tuple[int, int]:
        """(components, cells) of the sealed fluid components (pure-Neumann pockets)."""
        k, c = C.c_uint32(), C.c_uint32()
        _check(_lib().pyro_poisson_sealed_info(self._h, C.byref(k), C.byref(c)))
        return k.value, c.value

    def set_sealed_projection(self, mode: str = "solvers") -> None:
        """Subtract b's mean over each sealed component before a solve:
        "never", "solvers" (every method but the reference V-cycles; default)
        or "all"."""
        _check(_lib().pyro_poisson_set_sealed_projection(self._h, {"never": 0, "solvers": 1, "all": 2}[mode]))

    def residual(self) -> float:
        r = C.c_double()
        _check(_lib().pyro_poisson_residual(self._h, C.byref(r)))
        return r.value

    def solve(self, method: str = "mgpcg", tol: float = 1e-6, max_iters: int = 200) -> SolveInfo:
        if method not in METHODS:
            raise ValueError(f"unknown method {method!r}; choices: {sorted(METHODS)}")
        info = _Info()
        _check(_lib().pyro_poisson_solve(self._h, METHODS[method], float(tol), int(max_iters), C.byref(info)))
        return SolveInfo._of(info)

    LEARNED_MODES = {"every": 0, "first": 1, "coarse-every": 2, "coarse-first": 3}

    def attach_net(self, net, mode: str = "every") -> None:
        """Use `net` (a pyro_amd.net.Net) in the "learned" method: in every
        preconditioner application ("every") or only the first ("first");
        "coarse-every" / "coarse-first" apply it to the residual restricted to
        level 1 of the preconditioner hierarchy and prolong its output."""
        self._net = net
        _check(_lib().pyro_poisson_attach_net(self._h, None if net is None else net._h,
                                              self.LEARNED_MODES[mode]))

    def set_from_sim(self, sim, warm: bool = True, flags: bool = True) -> None:
        """b = the simulation's current level-0 right-hand side; phi = its
        level-0 phi (warm) or 0; solid flags from the simulation (flags)."""
        if flags:
            _check(_lib().pyro_poisson_set_flags_from_sim(self._h, sim._h))
        _check(_lib().pyro_poisson_set_from_sim(self._h, sim._h, 1 if warm else 0))


def fluid_components(flags: np.ndarray):
    """Face-connected fluid components and whether each touches an open face.

    Open (Dirichlet, phi = 0) faces are the sides and the top of the box
    (mg.wgsl:29); the ground (y = 0 face) and solids are Neumann.
    Returns (labels, n, open_mask[n+1]).
    """
    from scipy import ndimage

    fluid = np.asarray(flags) == 0
    lab, n = ndimage.label(fluid)
    touch = np.zeros(n + 1, bool)
    for sl in (lab[0, :, :], lab[-1, :, :], lab[:, -1, :], lab[:, :, 0], lab[:, :, -1]):
        touch[np.unique(sl)] = True
    touch[0] = False
    return lab, n, touch


def make_c3_problem(n: int = 256, seed: int = 0, h: float = 1.0 / 256, solid_frac: float = 0.05,
                    blobs: int = 32):
    """(flags, b) of the standalone pressure case (see the module docstring)."""
    rng = np.random.default_rng(seed)
    nz = ny = nx = n
    flags = np.zeros((nz, ny, nx), np.uint32)
    target = solid_frac * flags.size
    while flags.sum() < target:
        sz = rng.integers(max(2, n // 64), max(3, n // 10), size=3)
        lo = [rng.integers(0, n - s) for s in sz]
        flags[lo[0]:lo[0] + sz[0], lo[1]:lo[1] + sz[1], lo[2]:lo[2] + sz[2]] = 1
    b = np.zeros((nz, ny, nx), np.float64)
    for _ in range(blobs):
        c = rng.uniform(0, n, 3)
        sig = rng.uniform(4.0, 16.0) * n / 256.0
        amp = rng.uniform(0.1, 5.0) * rng.choice([-1.0, 1.0])
        # each blob is summed over its 6-sigma window (the tail beyond is < 2e-8 of the peak)
        lo = np.maximum(np.floor(c - 6 * sig).astype(int), 0)
        hi = np.minimum(np.ceil(c + 6 * sig).astype(int) + 1, n)
        g = [np.exp(-((np.arange(lo[a], hi[a]) - c[a]) ** 2) / (2 * sig * sig)) for a in range(3)]
        b[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += amp * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    b[flags == 1] = 0.0
    lab, ncomp, open_ = fluid_components(flags)
    if ncomp:
        sums = np.bincount(lab.ravel(), weights=b.ravel(), minlength=ncomp + 1)
        cnts = np.bincount(lab.ravel(), minlength=ncomp + 1)
        mean = sums / np.maximum(cnts, 1)
        mean[0] = 0.0
        b -= mean[lab]
    return flags, b


def step_solve(sim, P: Poisson, dt: float, method: str = "mgpcg", tol: float = 1e-4, max_iters: int = 100,
               want_stats: bool = False) -> SolveInfo:
    """One simulation step whose pressure solve runs `method` on P to `tol`.

    Raises PyroError on a bad argument before the step begins; on a failure
    inside the step (HIP error, non-finite or diverged solve) the step is
    aborted (pyro_sim_abort_step) and PyroError is raised."""
    if method not in METHODS:
        raise ValueError(f"unknown method {method!r}; choices: {sorted(METHODS)}")
    info = _Info()
    _check(_lib().pyro_sim_step_solve(sim._h, P._h, METHODS[method], float(tol), int(max_iters), float(dt),
                                      int(bool(want_stats)), C.byref(info)))
    return SolveInfo._of(info)
