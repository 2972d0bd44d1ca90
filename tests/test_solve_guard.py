"""Residual watch of the iterative pressure solves (pyro_amd/csrc/solve_guard.hpp,
used by every loop of pyro_poisson_solve in poisson.hip): NaN/Inf residuals
stop a solve with PYRO_E_NONFINITE instead of passing as converged, and a
residual above 1e3 x the best one so far stops it as diverged.  The host
program drives residual sequences through the loops' control flow; a planted
mistake (the finiteness test or the divergence test left out) must fail it."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "native" / "solve_guard_check.cpp"
HDR = ROOT / "pyro_amd" / "csrc" / "solve_guard.hpp"


def _run(tmp_path, header_text=None):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    inc = HDR.parent
    if header_text is not None:
        inc = tmp_path / "inc"
        inc.mkdir()
        (inc / "solve_guard.hpp").write_text(header_text)
    exe = tmp_path / "solve_guard_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(inc), str(SRC), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)


def test_residual_watch(tmp_path):
    r = _run(tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("0 wrong")


@pytest.mark.parametrize("old,new", [
    ("    if (!std::isfinite(rel)) return Watch::NonFinite;\n", ""),
    ("    if (rel > kDivergeFactor * best) return Watch::Diverged;\n", ""),
    ("    if (k == 0 || rel < best) best = rel;\n", "    if (k == 0) best = rel;\n"),
])
def test_residual_watch_catches_planted_mistakes(tmp_path, old, new):
    text = HDR.read_text()
    assert old in text
    r = _run(tmp_path, text.replace(old, new))
    assert r.returncode != 0, r.stdout
