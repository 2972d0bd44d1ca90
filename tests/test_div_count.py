"""The integer-count division of the multigrid kernels (k_mg.hip div_count):
Markstein's correction with r = RN(1/c), c = 1..6, guarded by a subnormal
test, must equal the IEEE quotient x / c for every finite x != -0.  The C
program checks a sample here (every 3rd input below 2^-124 for c = 6, mantissas
of 12 exponents, random patterns); ``div_count_check --all`` covers all 2^32
bit patterns (run once: 0 wrong).  Planted mistakes (the guard left out, the
correction step left out) must be caught by the same sample."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "div_count_check.c"


def _build(tmp_path, extra=()):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = tmp_path / "div_count_check"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", *extra, str(SRC), "-o", str(exe), "-lm"], check=True)
    return exe


def test_div_count_equals_ieee_division(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 wrong, 0 unguarded" in r.stdout


@pytest.mark.parametrize("old,new", [
    ("  if (sub) q = x / c;\n", "\n"),                       # no subnormal guard
    ("  float q = fmaf(e, r, q0);", "  float q = q0 + 0.0f * e;"),  # no correction step
])
def test_div_count_check_catches_planted_mistakes(tmp_path, old, new):
    src = SRC.read_text()
    assert old in src
    planted = tmp_path / "planted.c"
    planted.write_text(src.replace(old, new))
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = tmp_path / "planted"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", str(planted), "-o", str(exe), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode != 0, r.stdout
