"""Golden checksums (tests/golden/oracle_checksums.json, made by make_golden.py).

CPU: the oracle still reproduces every checksum (regression pin of the
checker itself, and of the Params blob of scenes/pool.toml).
GPU: the HIP path reproduces the same checksums bit for bit, without running
the oracle on the device host.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
import make_golden as mg  # noqa: E402

GOLD = json.loads((HERE / "golden" / "oracle_checksums.json").read_text())


def test_params_blob_checksum():
    import pyro_amd as pa

    blob = pa.Scene.load(str(HERE.parent / "pyro_amd/scenes/pool.toml")).params().bytes()
    assert hashlib.sha256(blob).hexdigest() == GOLD["params_pool_4cm"]


@pytest.mark.parametrize("name, build, steps", mg.cases(), ids=[c[0] for c in mg.cases()])
def test_oracle_reproduces_golden(name, build, steps):
    got = mg.run_case(build, steps, mg.make_oracle, lambda s: s.read_stats_raw())
    for f in mg.FIELDS:
        assert got[f] == GOLD[name][f], f"{name}: {f}"
    np.testing.assert_allclose(got["_stats_head"], GOLD[name]["_stats_head"], rtol=1e-12)


def make_gpu(scene, seed, noise=None):
    import harness
    import pyro_amd as pa

    p = scene.params()
    sol = scene.voxelize()
    g = pa.Sim(p, sol, None, int(scene.get("sim.pressure_cycles")), scene.get("ignition.until"))
    if seed is not None:
        st = harness.perturbed_state(scene, seed)
        for n in ("s0", "s1", "u", "v", "w", "sol"):
            g.write_field(n, st[n])
        g.write_field(pa.level_field(0, "phi"), st["phi"])
    if noise:
        g.add_velocity_noise(*noise)
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("name, build, steps", mg.cases(), ids=[c[0] for c in mg.cases()])
def test_gpu_reproduces_golden(name, build, steps):
    got = mg.run_case(build, steps, make_gpu, lambda s: s.read_stats().raw)
    bad = [f for f in mg.FIELDS if got[f] != GOLD[name][f]]
    assert not bad, f"{name}: fields differ from the golden checksums: {bad}"
    np.testing.assert_allclose(got["_stats_head"], GOLD[name]["_stats_head"], rtol=1e-9, atol=1e-12)
