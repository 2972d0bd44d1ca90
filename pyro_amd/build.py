"""Build the native libraries of pyro_amd in-tree.

- ``libpyro_hip.so``: the gfx950 solver (HIP, ``hipcc --offload-arch=gfx950``),
  C ABI in ``include/pyro_hip.h``.
- ``libpyro_host.so``: host-side scene plumbing (C++, ``g++``), C ABI in
  ``include/pyro_scene.h``.

Both are written to ``pyro_amd/lib/``.  Every HIP source is compiled with
``-ffp-contract=off`` (bit-exact operation order) and with packed-FP32
instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "pyro_amd"
LIB = PKG / "lib"
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["k_fluid.hip", "k_advect.hip", "k_combust.hip", "k_mg.hip", "k_solve.hip", "k_net.hip", "sim.hip", "poisson.hip",
               "net.hip", "voxel.hip", "k_render.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpyro_hip.so)")


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln or "warning" not in ln
    )
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(map(str, cmd))}")


def build_hip(jobs: int = 8) -> Path:
    """Compile the HIP sources (one object each, in parallel) and link the .so."""
    LIB.mkdir(parents=True, exist_ok=True)
    obj_dir = ROOT / "_build" / "hip"
    obj_dir.mkdir(parents=True, exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIP_FLAGS, "-I", str(INCLUDE), "-I", str(CSRC), "-c", str(CSRC / src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    out = LIB / "libpyro_hip.so"
    tmp = LIB / "libpyro_hip.so.tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def _drain(procs) -> None:
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            print("\n".join(lines), flush=True)
        if p.returncode != 0:
            errors.append(src)
    procs.clear()
    if errors:
        raise RuntimeError(f"hipcc failed for {errors}")


def build_host() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libpyro_host.so"
    tmp = LIB / "libpyro_host.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          "-I", str(INCLUDE), str(HOST / "scene.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host}[w]()
