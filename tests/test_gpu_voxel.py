"""Scene voxelization on the GPU (pyro_voxelize / pyro_sim_voxelize) against
the host loop of Scene::voxelize (scene.rs:621-660), bit for bit: every
primitive kind (yawed boxes, cylinders, spheres, pools), every object kind
that expands into them (crib, pallets, bonfire), the last-shape-wins overlap
rule, inert materials and the per-cell density jitter from the scene seed.
A simulation whose SOL was voxelized on the device steps identically to one
built from the host array."""
import numpy as np
import pytest

import pyro_amd as pa
from harness import SCENES, VEC4, VEL, compare, mini_scene

pytestmark = pytest.mark.gpu

MIXED = """
[domain]
size = [3.2, 2.4, 2.4]
cell = 0.04
[sim]
seed = 7
ambient_temp = 301.5
[[object]]
kind = "box"
material = "wood"
center = [0.7, 0.4, 0.8]
size = [0.5, 0.3, 0.7]
yaw = 33
[[object]]
kind = "cylinder"
material = "pmma"
a = [1.2, 0.1, 0.4]
b = [2.1, 1.3, 1.9]
radius = 0.13
[[object]]
kind = "sphere"
material = "rubber"
center = [2.5, 0.9, 1.6]
radius = 0.35
[[object]]
kind = "pool"
material = "diesel"
center = [2.4, 0.0, 0.6]
radius = 0.4
depth = 0.08
[[object]]
kind = "crib"
material = "wood"
center = [1.6, 0.0, 1.5]
size = [0.6, 0.0, 0.0]
count = 4
sticks = 5
stick = 0.05
yaw = -20
[[object]]
kind = "pallets"
material = "wood"
center = [0.9, 1.2, 1.6]
count = 2
yaw = 75
[[object]]
kind = "box"
material = "steel"
center = [1.6, 0.2, 1.5]
size = [0.4, 0.4, 0.4]
[[object]]
kind = "box"
material = "concrete"
center = [1.6, 2.2, 1.2]
size = [3.2, 0.1, 0.3]
"""


def _scenes():
    bon = pa.Scene.load(str(SCENES / "bonfire.toml"))
    bon.set("domain.cell", 0.1)
    pool = pa.Scene.load(str(SCENES / "pool.toml"))
    pool.set("domain.cell", 0.04)
    return {"pool_4cm": pool, "bonfire_10cm": bon, "mini": mini_scene(), "mixed": pa.Scene.parse(MIXED)}


@pytest.mark.parametrize("name", ["pool_4cm", "bonfire_10cm", "mini", "mixed"])
def test_device_voxelize_bitwise(name):
    sc = _scenes()[name]
    print(f"[voxel] {name} dims={sc.grid().dims} shapes={len(sc.shapes())}", flush=True)
    host = sc.voxelize()
    dev = sc.voxelize_device()
    assert host.shape == dev.shape
    solid = host[..., 1] > 0
    assert solid.sum() > 20
    bad = host.view(np.uint32) != dev.view(np.uint32)
    assert not bad.any(), f"{int(bad.any(-1).sum())} cells differ, first at {np.argwhere(bad.any(-1))[0]}"


def test_mixed_scene_covers_every_material_slot():
    sc = pa.Scene.parse(MIXED)
    sh = sc.shapes()
    assert set(sh["kind"].tolist()) == {0, 1, 2, 3}  # box, cylinder, sphere, pool
    sol = sc.voxelize_device()
    slots = set(np.unique(sol[..., 0]).tolist())
    assert len(slots) >= 6
    inert = sh["inert"] == 1
    assert inert.any() and (~inert).any()


def test_sim_device_voxelize_steps_identically():
    sc = mini_scene()
    a = pa.Sim.new(sc, device_voxelize=False)
    b = pa.Sim.new(sc, device_voxelize=True)
    assert np.array_equal(a.read_field("sol").view(np.uint32), b.read_field("sol").view(np.uint32))
    for k in range(3):
        a.step(0.004, k == 2)
        b.step(0.004, k == 2)
    res = compare(a, b, VEC4 + VEL + ("flags",))
    for f, (mx, _, nmis) in res.items():
        assert nmis == 0, f"{f}: {nmis} values differ (max {mx:.3e})"
    np.testing.assert_array_equal(a.read_stats().raw, b.read_stats().raw)


def test_scene_without_objects():
    """No shapes: SOL = (0, 0, 0, T0) everywhere on both paths, and a
    simulation created without SOL starts the same way."""
    sc = pa.Scene.parse("[domain]\nsize = [0.8, 0.8, 0.8]\ncell = 0.05\n[sim]\nambient_temp = 288.0\n"
                        "[fuel]\nmaterial = \"heptane\"\n")
    assert sc.shapes().size == 0
    host = sc.voxelize()
    dev = sc.voxelize_device()
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))
    assert (host[..., 3] == 288.0).all() and not host[..., :3].any()
    sim = pa.Sim.new(sc, device_voxelize=True)
    assert np.array_equal(sim.read_field("sol").view(np.uint32), host.view(np.uint32))
    assert not sim.read_field("flags").any()


def test_sim_voxelize_rejects_other_grid():
    sc = mini_scene()
    sim = pa.Sim.new(sc)
    other = mini_scene()
    other.set("domain.cell", 0.1)
    with pytest.raises(ValueError):
        sim.voxelize(other)
