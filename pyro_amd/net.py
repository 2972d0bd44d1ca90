"""The learned pressure preconditioner (include/pyro_hip.h, pyro_net_*).

A 3-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions, BF16 on MFMA with
FP32 accumulation) trained on the GPU with Adam on pressure pairs that the
multigrid solver produces; see pyro_amd/csrc/net.hpp for the topology and
pyro_amd/csrc/net.hip for the parameter order.  ``oracle.unet_ref`` is its
CPU restatement.

Arrays are NumPy with shape (nz, ny, nx[, c]); the network input x0 has two
channels (scaled right-hand side, fluid mask).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .poisson import Poisson, SolveInfo, _Info

LAYERS = [  # (name, cout, cin) in parameter order; conv weights are [cout][cin][27]
    ("0a", 16, 2), ("0b", 16, 16), ("1a", 32, 16), ("1b", 32, 32), ("2a", 64, 32),
    ("2b", 64, 64), ("1c", 32, 96), ("0c", 16, 48),
]


def _lib():
    lib = _native.hip()
    if not hasattr(lib, "_net_ready"):
        P = C.c_void_p
        F = C.POINTER(C.c_float)
        U = C.c_uint32
        lib.pyro_net_last_error.restype = C.c_char_p
        lib.pyro_net_param_count.restype = C.c_size_t
        lib.pyro_net_create.argtypes = [C.c_uint64, C.POINTER(P)]
        lib.pyro_net_destroy.argtypes = [P]
        for f in ("pyro_net_get_params", "pyro_net_set_params", "pyro_net_get_grads"):
            getattr(lib, f).argtypes = [P, P, C.c_size_t]
        lib.pyro_net_get_output_scale.argtypes = [P, F]
        lib.pyro_net_set_output_scale.argtypes = [P, C.c_float]
        lib.pyro_net_forward.argtypes = [P, U, U, U, P, P]
        lib.pyro_net_backward.argtypes = [P, U, U, U, P, P]
        lib.pyro_net_dataset_clear.argtypes = [P]
        lib.pyro_net_dataset_size.argtypes = [P, C.POINTER(C.c_int)]
        lib.pyro_net_dataset_add.argtypes = [P, U, U, U, P, P]
        lib.pyro_net_dataset_get.argtypes = [P, C.c_int, P, P]
        lib.pyro_net_collect.argtypes = [P, P, P, C.c_double, C.POINTER(_Info)]
        lib.pyro_net_train.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, P]
        lib.pyro_net_eval.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_double)]
        lib.pyro_net_time_forward.argtypes = [P, U, U, U, C.c_int, F]
        lib.pyro_debug_mfma.argtypes = [P, P, P]
        lib.pyro_debug_conv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, U, U, U, P, P, P, P, P]
        lib.pyro_debug_conv_dw.argtypes = [C.c_int, C.c_int, C.c_int, U, U, U, P, P, P, P, P]
        lib._net_ready = True
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise _native.PyroError(rc, _lib().pyro_net_last_error().decode(errors="replace"))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def param_count() -> int:
    return int(_lib().pyro_net_param_count())


def split_params(flat: np.ndarray) -> dict:
    """Views of the flat parameter vector: {'W0a': [16,2,27], 'b0a': [16], ..., 'wh', 'bh'}."""
    out, o = {}, 0
    for name, co, ci in LAYERS:
        out["W" + name] = flat[o:o + co * ci * 27].reshape(co, ci, 27)
        o += co * ci * 27
        out["b" + name] = flat[o:o + co]
        o += co
    out["wh"] = flat[o:o + 16]
    out["bh"] = flat[o + 16:o + 17]
    assert o + 17 == flat.size
    return out


class Net:
    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        _check(_lib().pyro_net_create(int(seed), C.byref(h)))
        self._h = h
        self.n = param_count()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().pyro_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        _check(_lib().pyro_net_get_params(self._h, out.ctypes.data, self.n))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = _f32(p)
        assert p.size == self.n
        _check(_lib().pyro_net_set_params(self._h, p.ctypes.data, self.n))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        _check(_lib().pyro_net_get_grads(self._h, out.ctypes.data, self.n))
        return out

    @property
    def output_scale(self) -> float:
        s = C.c_float()
        _check(_lib().pyro_net_get_output_scale(self._h, C.byref(s)))
        return s.value

    @output_scale.setter
    def output_scale(self, v: float) -> None:
        _check(_lib().pyro_net_set_output_scale(self._h, float(v)))

    def forward(self, x0: np.ndarray) -> np.ndarray:
        x = _f32(x0)
        nz, ny, nx, c = x.shape
        assert c == 2
        y = np.zeros((nz, ny, nx), np.float32)
        _check(_lib().pyro_net_forward(self._h, nx, ny, nz, x.ctypes.data, y.ctypes.data))
        return y

    def backward(self, x0: np.ndarray, gy: np.ndarray) -> np.ndarray:
        """Gradient of sum(gy * y(x0)) for every parameter (flat vector)."""
        x = _f32(x0)
        g = _f32(gy)
        nz, ny, nx, _ = x.shape
        assert g.shape == (nz, ny, nx)
        _check(_lib().pyro_net_backward(self._h, nx, ny, nz, x.ctypes.data, g.ctypes.data))
        return self.grads()

    # training set
    def dataset_clear(self) -> None:
        _check(_lib().pyro_net_dataset_clear(self._h))

    def dataset_size(self) -> int:
        n = C.c_int()
        _check(_lib().pyro_net_dataset_size(self._h, C.byref(n)))
        return n.value

    def dataset_add(self, x0: np.ndarray, target: np.ndarray) -> None:
        x = _f32(x0)
        t = _f32(target)
        nz, ny, nx, _ = x.shape
        _check(_lib().pyro_net_dataset_add(self._h, nx, ny, nz, x.ctypes.data, t.ctypes.data))

    def dataset_get(self, k: int, shape) -> tuple[np.ndarray, np.ndarray]:
        nz, ny, nx = shape
        x = np.zeros((nz, ny, nx, 2), np.float32)
        t = np.zeros((nz, ny, nx), np.float32)
        _check(_lib().pyro_net_dataset_get(self._h, int(k), x.ctypes.data, t.ctypes.data))
        return x, t

    def collect(self, sim, poisson: Poisson, tol: float = 1e-6) -> SolveInfo:
        """Add the simulation's current pressure problem, solved to `tol`."""
        info = _Info()
        rc = _lib().pyro_net_collect(self._h, sim._h, poisson._h, float(tol), C.byref(info))
        if rc != 0:
            raise _native.PyroError(rc, _native.hip().pyro_poisson_last_error().decode(errors="replace"))
        return SolveInfo(info.iterations, info.vcycles, info.rel_resid, info.ms, info.fallbacks)

    def train(self, steps: int, lr: float = 1e-3, seed: int = 0, first: int = 0, count: int | None = None) -> np.ndarray:
        """Adam steps on random pairs of [first, first + count); per-step losses."""
        if count is None:
            count = self.dataset_size() - first
        losses = np.zeros(max(steps, 1), np.float32)
        _check(_lib().pyro_net_train(self._h, int(first), int(count), int(steps), float(lr), int(seed),
                                     losses.ctypes.data))
        return losses[:steps]

    def time_forward(self, shape, reps: int = 10) -> float:
        """average device ms of one forward pass on an (nz, ny, nx) grid"""
        nz, ny, nx = shape
        ms = C.c_float()
        _check(_lib().pyro_net_time_forward(self._h, nx, ny, nz, int(reps), C.byref(ms)))
        return ms.value

    def eval(self, first: int = 0, count: int | None = None) -> float:
        if count is None:
            count = self.dataset_size() - first
        v = C.c_double()
        _check(_lib().pyro_net_eval(self._h, int(first), int(count), C.byref(v)))
        return v.value


# test hooks of the kernels
def debug_mfma(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = _f32(a).reshape(64, 8)
    b = _f32(b).reshape(64, 8)
    d = np.zeros((64, 4), np.float32)
    _check(_lib().pyro_debug_mfma(a.ctypes.data, b.ctypes.data, d.ctypes.data))
    return d


def debug_conv(ca, cb, cout, a, b, W, bias=None, relu=False, transposed=False):
    """One conv on device: forward (b: (nz,ny,nx,cb), a: half grid x ca) or its
    adjoint applied to b: (nz,ny,nx,cout).  W: [cout][ca+cb][27]."""
    b = _f32(b)
    nz, ny, nx, _ = b.shape
    W = _f32(W)
    ap = None
    if ca and not transposed:
        a = _f32(a)
        ap = a.ctypes.data
    bp = None
    if bias is not None and not transposed:
        bias = _f32(bias)
        bp = bias.ctypes.data
    oc = (ca + cb) if transposed else cout
    out = np.zeros((nz, ny, nx, oc), np.float32)
    _check(_lib().pyro_debug_conv(ca, cb, cout, int(transposed), int(relu), nx, ny, nz, ap, b.ctypes.data,
                                  W.ctypes.data, bp, out.ctypes.data))
    return out


def debug_conv_dw(ca, cb, cout, a, b, g):
    b = _f32(b)
    g = _f32(g)
    nz, ny, nx, _ = b.shape
    ap = None
    if ca:
        a = _f32(a)
        ap = a.ctypes.data
    dW = np.zeros((cout, ca + cb, 27), np.float32)
    db = np.zeros(cout, np.float32)
    _check(_lib().pyro_debug_conv_dw(ca, cb, cout, nx, ny, nz, ap, b.ctypes.data, g.ctypes.data, dW.ctypes.data,
                                     db.ctypes.data))
    return dW, db
