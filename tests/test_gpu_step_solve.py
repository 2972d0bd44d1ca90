"""A simulation step whose pressure solve runs to a tolerance
(pyro_sim_step_solve).

- with the reference V-cycles, tol = 0 and 4 iterations it is the plain
  step, bit for bit (every field, every statistic);
- with MGPCG every step reaches its tolerance and the step's FP64 residual
  statistic agrees.
"""
import numpy as np
import pytest

import pyro_amd as pa
from harness import VEC4, VEL, mini_scene, perturbed_state
from pyro_amd.poisson import Poisson, step_solve

pytestmark = pytest.mark.gpu


def _sim(scene, seed):
    p = scene.params()
    sim = pa.Sim(p, scene.voxelize(), None, 4, scene.get("ignition.until"))
    st = perturbed_state(scene, seed)
    for name in ("s0", "s1", "u", "v", "w", "sol"):
        sim.write_field(name, st[name])
    sim.write_field(pa.level_field(0, "phi"), st["phi"])
    return sim


def test_vcycle_step_solve_is_the_plain_step():
    scene = mini_scene()
    dt = 0.004
    a, b = _sim(scene, 3), _sim(scene, 3)
    P = Poisson(b.read_field(pa.level_field(0, "flags")), scene.grid().dx)
    for k in range(3):
        stats = k == 2
        a.step(dt, stats)
        info = step_solve(b, P, dt, "vcycle", tol=0.0, max_iters=4, want_stats=stats)
        assert info.vcycles == 4
    for n in VEC4 + VEL + (pa.level_field(0, "phi"), "flags"):
        assert np.array_equal(a.read_field(n), b.read_field(n)), n
    sa, sb = a.read_stats(), b.read_stats()
    assert sa.hrr == sb.hrr and sa.resid == sb.resid and sa.max_t == sb.max_t


def test_mgpcg_step_solve_reaches_tolerance():
    scene = mini_scene()
    dt = 0.004
    sim = _sim(scene, 5)
    P = Poisson(sim.read_field(pa.level_field(0, "flags")), scene.grid().dx)
    for k in range(4):
        info = step_solve(sim, P, dt, "mgpcg", tol=1e-6, max_iters=100, want_stats=True)
        assert info.rel_resid <= 1e-6, info
        st = sim.read_stats()
        # the step's own residual statistic (FP32 phi, FP64 sums) at the same level
        assert st.resid < 1e-5, st.resid
    sim.check_finite()
