"""The step's fused launches against the single stages, on the GPU.

The step runs solid with curl (k_solid_curl) and the four advection stages
as two launches (adv_vel1 with adv_s1, adv_vel2 with adv_s2); the
stage-by-stage API runs one kernel per reference pass.  A step by graph
replay and the same step stage by stage, from the same state, must agree
bit for bit, statistics included; the timing entry points for the fused
launches return finite times and leave a usable simulation.
"""
import numpy as np
import pytest

import pyro_amd as pa
from harness import VEC4, VEL, mini_scene, perturbed_state, pool_scene

pytestmark = pytest.mark.gpu


def _sim(scene, seed, use_graph=True):
    sim = pa.Sim(scene.params(), scene.voxelize(), None, 4, scene.get("ignition.until"), use_graph=use_graph)
    st = perturbed_state(scene, seed)
    for name in ("s0", "s1", "u", "v", "w", "sol"):
        sim.write_field(name, st[name])
    sim.write_field(pa.level_field(0, "phi"), st["phi"])
    return sim


@pytest.mark.parametrize("stats", [False, True])
def test_fused_step_equals_stage_by_stage(stats):
    scene = mini_scene()
    a, b = _sim(scene, 7), _sim(scene, 7)
    a.step(0.002, stats)
    b.begin_step(0.002, stats)
    for stg in pa.STAGES:
        if stg != "resid" or stats:
            b.run_stage(stg)
    b.end_step()
    for n in VEC4 + VEL + ("flags", "curl", pa.level_field(0, "phi")):
        assert np.array_equal(a.read_field(n), b.read_field(n)), n
    if stats:
        assert np.array_equal(a.read_stats().raw, b.read_stats().raw)


def test_fused_launch_timing_entry_points():
    scene = pool_scene(0.08)
    sim = _sim(scene, 3)
    sim.step(0.002, False)
    for kern in ("adv_pass1", "adv_pass2", "solid_curl"):
        ms = sim.time_kernel(kern, 0, 3)
        assert np.isfinite(ms) and ms > 0.0, kern
    sim.step(0.002, True)
    sim.check_finite()
    assert np.isfinite(sim.read_stats().raw).all()
