"""The C-ABI libraries load on the CPU and export every symbol their headers declare."""
import ctypes
import re
from pathlib import Path

import pytest

from conftest import ROOT

HEADERS = {
    "pyro_hip.h": "pyro_amd/lib/libpyro_hip.so",
    "pyro_scene.h": "pyro_amd/lib/libpyro_host.so",
}


def declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pyro_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header, lib", sorted(HEADERS.items()))
def test_library_exports_every_declared_symbol(header, lib):
    names = declared(ROOT / "include" / header)
    assert len(names) >= 8, names
    so = ctypes.CDLL(str(ROOT / lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_no_device_reported_without_gpu():
    """On a host without a GPU, creation fails with PYRO_E_NODEVICE instead of crashing."""
    import os

    import numpy as np

    import pyro_amd as pa

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU node is present")
    s = pa.Scene.parse('[domain]\nsize = [0.8, 0.8, 0.8]\ncell = 0.05\n[fuel]\nmaterial = "wood"\n')
    with pytest.raises(pa.PyroError) as e:
        pa.Sim(s.params(), s.voxelize())
    assert e.value.code == -6
    assert np.isfinite(1.0)


def test_bad_params_size_rejected():
    import ctypes as C

    from pyro_amd import _native

    lib = _native.hip()
    h = C.c_void_p()
    buf = (C.c_uint8 * 100)()
    sol = (C.c_float * 4)()
    rc = lib.pyro_sim_create(buf, 100, sol, None, 4, 0.0, 0, C.byref(h))
    assert rc == -1 and b"576" in lib.pyro_last_error()


def test_render_export_argument_checks_without_gpu():
    """The exported hand-off rejects null arguments before touching a device,
    and the in-process import reports that it is switched off."""
    import ctypes as C

    from pyro_amd import _native

    lib = _native.hip()
    fd, nbytes = C.c_int(5), C.c_uint64(7)
    assert lib.pyro_sim_render_export(None, C.byref(fd), C.byref(nbytes)) == -1
    assert b"null argument" in lib.pyro_last_error()
    buf = (C.c_uint8 * 16)()
    assert lib.pyro_shared_read(3, 16, buf, 16) == -5
    assert b"disabled" in lib.pyro_last_error()
