"""Write tests/golden/oracle_checksums.json: checksums of oracle outputs.

Each case is a seeded state (tests/harness.py) advanced by the oracle; for
every output field the file stores the SHA-256 of its bytes (signed zeros
folded to +0.0) and a few sample values.  The GPU path must reproduce the
same checksums bit for bit; the CPU suite checks that the oracle still does.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

FIELDS = ("s0", "s1", "sol", "pyr", "curl", "u", "v", "w", "phi0", "flags")


def digest(a: np.ndarray) -> str:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        a = a + np.float32(0.0)  # -0.0 -> +0.0
    return hashlib.sha256(a.tobytes()).hexdigest()


def read(sim, name):
    import pyro_amd as pa

    return sim.read_field(pa.level_field(0, "phi") if name == "phi0" else name)


def cases():
    """(name, builder, steps); builder(sim_cls) -> sim with the seeded state."""
    import harness

    def mini(seed):
        def build(make):
            return make(harness.mini_scene(), seed)
        return build

    def pool_noise(make):
        return make(harness.pool_scene(0.08), None, noise=(1e-3, 0))

    return [("mini_seed0_2steps", mini(0), 2), ("mini_seed4_5steps", mini(4), 5),
            ("pool8cm_noise_3steps", pool_noise, 3)]


def make_oracle(scene, seed, noise=None):
    import harness
    import oracle
    import pyro_amd as pa

    p = scene.params()
    sol = scene.voxelize()
    o = oracle.OSim(p, sol, None, int(scene.get("sim.pressure_cycles")), scene.get("ignition.until"))
    if seed is not None:
        st = harness.perturbed_state(scene, seed)
        for n in ("s0", "s1", "u", "v", "w", "sol"):
            o.write_field(n, st[n])
        o.write_field(pa.level_field(0, "phi"), st["phi"])
    if noise:
        o.add_velocity_noise(*noise)
    return o


def run_case(build, steps, make, stats_of):
    sim = build(make)
    for k in range(steps):
        sim.step(0.002, k == steps - 1)
    out = {n: digest(read(sim, n)) for n in FIELDS}
    s0 = read(sim, "s0")
    out["_sample_T"] = [float(v) for v in s0[..., 0].ravel()[:: max(1, s0[..., 0].size // 7)][:7]]
    out["_stats_head"] = [float(v) for v in stats_of(sim)[:12]]
    return out


def main():
    import pyro_amd as pa

    res = {"params_pool_4cm": hashlib.sha256(
        pa.Scene.load(str(HERE.parent.parent / "pyro_amd/scenes/pool.toml")).params().bytes()).hexdigest()}
    for name, build, steps in cases():
        res[name] = run_case(build, steps, make_oracle, lambda s: s.read_stats_raw())
        print(name, res[name]["s0"][:16], flush=True)
    (HERE / "oracle_checksums.json").write_text(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    main()
