"""Checkpoint save/load (pyro_sim_save / pyro_sim_load): a run restored from
a checkpoint continues bit-identically; bad files are rejected before the
simulation is touched."""
import numpy as np
import pytest

import pyro_amd as pa
from harness import VEC4, VEL, mini_scene, perturbed_state

pytestmark = pytest.mark.gpu


def _sim(scene, seed=None):
    sim = pa.Sim(scene.params(), scene.voxelize(), None, 4, scene.get("ignition.until"))
    if seed is not None:
        st = perturbed_state(scene, seed)
        for name in ("s0", "s1", "u", "v", "w", "sol"):
            sim.write_field(name, st[name])
        sim.write_field(pa.level_field(0, "phi"), st["phi"])
    return sim


def test_restored_run_continues_bitwise(tmp_path):
    scene = mini_scene()
    dt = 0.004
    a = _sim(scene, 7)
    for k in range(3):
        a.step(dt, False)
    ck = tmp_path / "mini.ckpt"
    a.save(ck)
    b = _sim(scene)  # fresh state, then the checkpoint
    b.load(ck)
    assert b.time == a.time
    for k in range(4):
        a.step(dt, k == 3)
        b.step(dt, k == 3)
    for n in ("s0", "s1", "sol", "flags", "u", "v", "w", pa.level_field(0, "phi")):
        assert np.array_equal(a.read_field(n), b.read_field(n)), n
    sa, sb = a.read_stats(), b.read_stats()
    assert sa.hrr == sb.hrr and sa.resid == sb.resid and sa.max_t == sb.max_t


def test_bad_checkpoints_are_rejected(tmp_path):
    scene = mini_scene()
    a = _sim(scene, 1)
    before = a.read_field("s0")
    junk = tmp_path / "junk.ckpt"
    junk.write_bytes(b"not a checkpoint at all" * 10)
    with pytest.raises(pa.PyroError, match="not a checkpoint"):
        a.load(junk)
    ck = tmp_path / "ok.ckpt"
    a.save(ck)
    trunc = tmp_path / "trunc.ckpt"
    trunc.write_bytes(ck.read_bytes()[:-100])
    with pytest.raises(pa.PyroError, match="truncated"):
        a.load(trunc)
    other = pa.Scene.parse(scene_text_other_grid())
    b = pa.Sim(other.params(), other.voxelize(), None, 4, 0.0)
    with pytest.raises(pa.PyroError, match="does not match"):
        b.load(ck)
    with pytest.raises(pa.PyroError, match="cannot open"):
        a.load(tmp_path / "missing.ckpt")
    assert np.array_equal(a.read_field("s0"), before)


def scene_text_other_grid():
    return """
[domain]
size = [0.8, 0.8, 0.8]
cell = 0.05
[[object]]
kind = "pool"
material = "heptane"
center = [0.4, 0.0, 0.4]
radius = 0.2
depth = 0.1
"""


def test_checkpoint_is_read_by_the_oracle_and_written_for_it(tmp_path):
    """The file format against the oracle's independent reader and writer:
    a GPU checkpoint holds exactly the state of the oracle twin that ran the
    same steps, the oracle continues from it bit-identically to the GPU, and
    the GPU continues from a checkpoint the oracle wrote."""
    import oracle
    from harness import compare, twin

    dt = 0.004
    o, g = twin(mini_scene(), seed=13)
    for k in range(2):
        o.step(dt, False)
        g.step(dt, False)
    ck = tmp_path / "gpu.ckpt"
    g.save(ck)
    ro = oracle.OSim.from_checkpoint(ck)
    names = ("s0", "s1", "sol", "u", "v", "w", pa.level_field(0, "phi"))
    for n in names:
        assert np.array_equal(ro.read_field(n), o.read_field(n)), n
    assert ro.time == o.time and ro.params_blob()[:16] == o.params_blob()[:16]
    for k in range(3):
        ro.step(dt, k == 2)
        g.step(dt, k == 2)
    for n, (mx, rl2, nmis) in compare(ro, g, names + ("pyr", "flags", "curl")).items():
        assert nmis == 0, f"{n}: {nmis} values differ"
    np.testing.assert_allclose(g.read_stats().raw, ro.read_stats_raw(), rtol=1e-9, atol=1e-12)
    # the other direction: the oracle's file into a fresh GPU simulation
    ck2 = tmp_path / "oracle.ckpt"
    ro.save_checkpoint(ck2)
    g2 = _sim(mini_scene())
    g2.load(ck2)
    for k in range(2):
        ro.step(dt, k == 1)
        g2.step(dt, k == 1)
    for n, (mx, rl2, nmis) in compare(ro, g2, names + ("pyr", "flags")).items():
        assert nmis == 0, f"{n}: {nmis} values differ"
    np.testing.assert_allclose(g2.read_stats().raw, ro.read_stats_raw(), rtol=1e-9, atol=1e-12)
