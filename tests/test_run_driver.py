"""The simulation-only frame loop (pyro_amd/run.py) against main.rs's
arithmetic: CFL substeps in FP32, frame counts, heat-release heights,
Heskestad, and the slice panels / PNG writer."""
import struct
import zlib

import numpy as np

from pyro_amd import run


def test_substeps_match_main_rs():
    # main.rs:140-142: dt_cfl = cfl*dx/max(u,1); n = ceil(frame_dt/dt_cfl).max(1); dt = frame_dt/n
    n, dt = run.substeps(np.float32(1 / 30), 0.5, 0.02, 2.0)
    assert n == 7 and dt == np.float32(np.float32(1 / 30) / np.float32(7))
    n, dt = run.substeps(np.float32(1 / 30), 0.5, 0.04, 0.1)  # u below 1 m/s is clamped to 1
    assert n == 2
    n, dt = run.substeps(np.float32(1 / 30), 10.0, 0.1, 1.0)
    assert n == 1 and dt == np.float32(1 / 30)


def test_frame_count_rounds_half_away():
    assert run.frame_count(np.float32(2.0), np.float32(30)) == 60
    assert run.frame_count(np.float32(0.05), np.float32(30)) == 2  # 1.5 -> 2
    assert run.frame_count(np.float32(0.0), np.float32(30)) == 0


def test_heat_heights_and_heskestad():
    layers = np.array([0.0, 1.0, 1.0, 1.0, 1.0, 0.0])
    assert run.heat_heights(layers, 0.5) == [1.5, 2.5, 2.5]
    L, f = run.heskestad(1000e3, 0.5)  # 1 MW over a 1 m pool
    assert abs(L - (0.235 * 1000 ** 0.4 - 1.02)) < 1e-12 and abs(f - 1.5) < 1e-12


def test_slice_png_roundtrip(tmp_path):
    a = np.zeros((4, 6, 5, 4), np.float32)
    b = np.zeros_like(a)
    a[..., 0] = 2000.0  # T at the top of its scale -> white
    img = run.slice_image(a, b)
    assert img.shape == (6, 5 * 5 + 8, 3)
    assert np.all(img[:, :5] == 255)
    run._png(tmp_path / "s.png", img)
    data = (tmp_path / "s.png").read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h = struct.unpack(">II", data[16:24])
    assert (w, h) == (33, 6)
    idat_len = struct.unpack(">I", data[33:37])[0]
    raw = zlib.decompress(data[41:41 + idat_len])
    assert len(raw) == h * (1 + 3 * w)
