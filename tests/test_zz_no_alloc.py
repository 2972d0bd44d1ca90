"""pyro_voxelize_noalloc (no allocation; shapes staged in the library's
fixed per-GPU array) against pyro_voxelize, through the C ABI: two calls in a
row, the second with fewer shapes and a smaller grid, each bit-identical to
the allocating entry point's output; oversized inputs and a host destination
are rejected before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

import pyro_amd as pa
from conftest import SCENES
from pyro_amd import _native

pytestmark = pytest.mark.gpu

MAX_SHAPES = 4096


def _device_sol(sim):
    ptr, nbytes = C.c_void_p(), C.c_size_t()
    _native.check_hip(_native.hip().pyro_sim_device_ptr(sim._h, pa.FIELDS["sol"], C.byref(ptr), C.byref(nbytes)))
    return ptr, nbytes.value


def _reference(sh, dims, dx, t0, seed):
    out = np.zeros((dims[2], dims[1], dims[0], 4), np.float32)
    d = (C.c_uint32 * 3)(*dims)
    _native.check_hip(_native.hip().pyro_voxelize(sh.ctypes.data if sh.size else None, sh.size, d, dx, t0, seed,
                                                  out.ctypes.data, 0))
    return out


def _noalloc_into_sim(scene):
    lib = _native.hip()
    sh = scene.shapes()
    dims, dx, t0, seed = scene.voxel_spec()
    sim = pa.Sim(scene.params(), None, None, 4, 0.0)  # SOL = 0
    ptr, nbytes = _device_sol(sim)
    assert nbytes == 16 * dims[0] * dims[1] * dims[2]
    d = (C.c_uint32 * 3)(*dims)
    _native.check_hip(lib.pyro_voxelize_noalloc(sh.ctypes.data if sh.size else None, sh.size, d, dx, t0, seed, ptr))
    got = sim.read_field("sol")
    ref = _reference(sh, dims, dx, t0, seed)
    return sh.size, got, ref, sim


def test_noalloc_voxelize_matches_pyro_voxelize_twice():
    big = pa.Scene.load(str(SCENES / "bonfire.toml"))
    big.set("domain.cell", 0.25)
    small = pa.Scene.load(str(SCENES / "pool.toml"))
    small.set("domain.cell", 0.1)
    n1, got1, ref1, s1 = _noalloc_into_sim(big)
    n2, got2, ref2, s2 = _noalloc_into_sim(small)  # the array still holds the bonfire's shapes past n2
    assert n1 > 20 and n2 < n1 and got2.size < got1.size
    assert np.array_equal(got1.view(np.uint32), ref1.view(np.uint32))
    assert np.array_equal(got2.view(np.uint32), ref2.view(np.uint32))
    assert (ref1[..., 1] > 0).sum() > 100 and (ref2[..., 1] > 0).sum() > 10
    # the first result is untouched by the second call
    assert np.array_equal(s1.read_field("sol").view(np.uint32), ref1.view(np.uint32))


def test_noalloc_voxelize_rejects_bad_inputs():
    lib = _native.hip()
    sc = pa.Scene.load(str(SCENES / "pool.toml"))
    sc.set("domain.cell", 0.1)
    dims, dx, t0, seed = sc.voxel_spec()
    d = (C.c_uint32 * 3)(*dims)
    sim = pa.Sim(sc.params(), None, None, 4, 0.0)
    ptr, _ = _device_sol(sim)
    many = np.zeros(MAX_SHAPES + 1, dtype=pa.scene.SHAPE_DTYPE)
    rc = lib.pyro_voxelize_noalloc(many.ctypes.data, many.size, d, dx, t0, seed, ptr)
    assert rc < 0
    assert b"at most 4096" in lib.pyro_last_error()
    host = np.zeros((dims[2], dims[1], dims[0], 4), np.float32)
    sh = sc.shapes()
    rc = lib.pyro_voxelize_noalloc(sh.ctypes.data, sh.size, d, dx, t0, seed, host.ctypes.data)
    assert rc < 0 and b"not device memory" in lib.pyro_last_error()
    assert not np.any(host)
    # the library is still usable afterwards
    _native.check_hip(lib.pyro_voxelize_noalloc(sh.ctypes.data, sh.size, d, dx, t0, seed, ptr))
    assert np.array_equal(sim.read_field("sol").view(np.uint32), _reference(sh, dims, dx, t0, seed).view(np.uint32))
