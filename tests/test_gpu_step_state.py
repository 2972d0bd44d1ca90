"""Step state after failures and partial steps (pyro_sim_step_solve,
pyro_sim_abort_step, pyro_sim_load, stage-by-stage stats):

- step_solve checks method, tol, max_iters, the attached network and the
  grid before the step begins: a rejected call leaves the simulation as it
  was, and the next plain step equals that of a twin that never made the
  call, bit for bit;
- a checkpoint loaded into a simulation that has already stepped by graph
  replay, with a different V-cycle count, continues as the saved run does
  (the captured graphs are dropped);
- a checkpoint whose cell size differs from the simulation's is rejected;
- a stats step run stage by stage without the residual stage reports 0 in the
  residual slots, as the oracle does, and every other slot as the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import pyro_amd as pa
from harness import VEC4, VEL, mini_scene, perturbed_state, twin
from pyro_amd.poisson import Poisson, _Info, step_solve, _lib as plib

pytestmark = pytest.mark.gpu


def _sim(scene, seed=None, cycles=4):
    sim = pa.Sim(scene.params(), scene.voxelize(), None, cycles, scene.get("ignition.until"))
    if seed is not None:
        st = perturbed_state(scene, seed)
        for name in ("s0", "s1", "u", "v", "w", "sol"):
            sim.write_field(name, st[name])
        sim.write_field(pa.level_field(0, "phi"), st["phi"])
    return sim


# entries of the padded (nz+1, ny+1, nx+1) velocity arrays that are a
# component of the state (compValid, common.wgsl:66-72); a step neither reads
# nor writes the others, so a simulation keeps whatever its buffers held there
_VALID = {"u": (0, 0, 1), "v": (0, 1, 0), "w": (1, 0, 0)}


def _state(sim, n):
    a = sim.read_field(n)
    if n in _VALID:
        ez, ey, ex = _VALID[n]
        nz, ny, nx = (d - 1 for d in a.shape)
        a = a[:nz + ez, :ny + ey, :nx + ex]
    return a


def _same(a, b, names):
    for n in names:
        assert np.array_equal(_state(a, n), _state(b, n)), n


def test_rejected_step_solve_leaves_the_sim_usable():
    scene = mini_scene()
    dt = 0.004
    a, b = _sim(scene, 11), _sim(scene, 11)
    P = Poisson(a.read_field(pa.level_field(0, "flags")), scene.grid().dx)
    t0 = a.time
    with pytest.raises(pa.PyroError, match="no network attached"):
        step_solve(a, P, dt, "learned", tol=1e-4)
    info = _Info()
    L = plib()
    for method, tol, iters in ((7, 1e-4, 10), (2, -1.0, 10), (2, float("nan"), 10), (2, 1e-4, 0)):
        rc = L.pyro_sim_step_solve(a._h, P._h, method, tol, iters, dt, 0, C.byref(info))
        assert rc == -1, (method, tol, iters, rc)
    other = Poisson(np.zeros((8, 8, 8), np.uint32), 0.05)
    with pytest.raises(pa.PyroError, match="grid mismatch"):
        step_solve(a, other, dt, "mgpcg", tol=1e-4)
    assert a.time == t0
    # the simulation is not left inside a step: a plain step, a save and a
    # render pack all work, and the step equals the untouched twin's
    a.step(dt, True)
    b.step(dt, True)
    _same(a, b, VEC4 + VEL + ("flags", pa.level_field(0, "phi")))
    assert a.read_stats().raw.tolist() == b.read_stats().raw.tolist()
    a.render_pack()


def test_abort_step_keeps_roles_and_time():
    scene = mini_scene()
    a, b = _sim(scene, 12), _sim(scene, 12)
    u0 = a.read_field("u")
    a.begin_step(0.004, True)
    for st in ("solid", "curl", "adv_vel1", "adv_vel2"):
        a.run_stage(st)
    a.abort_step()
    assert a.time == b.time
    assert np.array_equal(a.read_field("u"), u0)  # vel[cur] untouched
    a.abort_step()  # no open step: a no-op
    # velocities and SOL are those before the step: the next step matches a
    # twin on them (s0/s1 were not yet written by react)
    a.step(0.004, False)
    b.step(0.004, False)
    _same(a, b, ("u", "v", "w", "s0", "s1", "sol", "flags"))


def test_checkpoint_into_a_stepped_sim_with_other_cycles(tmp_path):
    scene = mini_scene()
    dt = 0.004
    a = _sim(scene, 21, cycles=4)
    for k in range(2):
        a.step(dt, k == 1)
    ck = tmp_path / "a.ckpt"
    a.save(ck)
    b = _sim(scene, 22, cycles=2)
    for k in range(3):  # graphs captured with 2 V-cycles, both roles, stats and not
        b.step(dt, k == 2)
    b.load(ck)
    for k in range(3):
        a.step(dt, k == 2)
        b.step(dt, k == 2)
    _same(a, b, ("s0", "s1", "sol", "flags", "u", "v", "w", pa.level_field(0, "phi")))
    assert a.read_stats().raw.tolist() == b.read_stats().raw.tolist()


def test_checkpoint_with_other_cell_size_is_rejected(tmp_path):
    scene = mini_scene()
    a = _sim(scene, 3)
    ck = tmp_path / "a.ckpt"
    a.save(ck)
    # the same 32 x 48 x 32 cells at 10 cm instead of 5 cm
    big = pa.Scene.parse("[domain]\nsize = [3.2, 4.8, 3.2]\ncell = 0.1\n[fuel]\nmaterial = \"heptane\"\n")
    assert big.grid().dims == scene.grid().dims
    b = pa.Sim(big.params(), big.voxelize(), None, 4, 0.0)
    s0 = b.read_field("s0")
    with pytest.raises(pa.PyroError, match="cell size"):
        b.load(ck)
    assert np.array_equal(b.read_field("s0"), s0)
    b.step(0.004, False)


def test_stats_step_without_resid_stage_matches_oracle():
    o, g = twin(mini_scene(), seed=4)
    for sim in (o, g):
        sim.begin_step(0.002, True)
        for st in ("solid", "curl", "adv_vel1", "adv_vel2", "adv_s1", "adv_s2", "react", "forces", "divergence",
                   "vcycles", "project"):
            sim.run_stage(st)
        sim.end_step()
    so = o.read_stats_raw()
    sg = g.read_stats().raw
    assert so[10] == 0.0 and so[11] == 0.0
    assert sg[10] == 0.0 and sg[11] == 0.0
    assert np.allclose(sg, so, rtol=1e-9, atol=1e-12)
    # a stats step with only the react stage: the solid, curl and residual
    # slots are 0 as well
    for sim in (o, g):
        sim.begin_step(0.002, True)
        sim.run_stage("react")
        sim.end_step()
    so = o.read_stats_raw()
    sg = g.read_stats().raw
    for slot in (2, 5, 7, 9, 10, 11):
        assert so[slot] == 0.0 and sg[slot] == 0.0, slot
    assert np.allclose(sg, so, rtol=1e-9, atol=1e-12)
