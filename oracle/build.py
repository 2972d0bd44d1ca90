"""Build recipe of the CPU oracle (TEST INFRASTRUCTURE).

Compiles ``oracle/pyro_oracle.c`` into ``oracle/liboracle.so`` with gcc:
FP32 arithmetic in source order (``-ffp-contract=off``), OpenMP for speed.
``-march=x86-64-v3`` (AVX2/FMA) keeps the binary portable across the x86
hosts the tests run on.

The reference itself (Rust + wgpu + WGSL) cannot be built in this
environment (no Rust toolchain, its crates are not vendored, no Vulkan
driver), so there is no ``oracle/_ref`` build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "liboracle.so"
    tmp = HERE / "liboracle.so.tmp"
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fno-math-errno",
           "-fPIC", "-shared", "-Wall", "-Wextra", str(HERE / "pyro_oracle.c"), "-o", str(tmp), "-lm"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
