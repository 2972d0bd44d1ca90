"""Scene plumbing (host C++ restatement of src/scene.rs).

The first four tests are the reference's own unit tests (scene.rs:881-932),
restated against the C ABI; the rest pin Scene::params (scene.rs:523-618)
by its closed forms and the grid/voxelization of the benchmark scenes.
"""
import math

import numpy as np
import pytest

import pyro_amd as pa
from conftest import SCENES


def scene(objects: str = "") -> pa.Scene:
    # scene.rs:885-890
    return pa.Scene.parse(f"[domain]\nsize = [1.6, 1.6, 1.6]\ncell = 0.05\n{objects}")


def test_grid_dims_are_multiples_of_16():
    # scene.rs:892-897
    assert scene().grid().dims == (32, 32, 32)


def test_box_voxel_volume():
    # scene.rs:899-908
    s = scene('[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.8, 0.4, 0.8]\nsize = [0.4, 0.4, 0.4]\n')
    sol = s.voxelize()
    n = int((sol[..., 1] > 0.0).sum())
    assert 512 <= n <= 729, n


def test_thin_parts_and_shallow_pools_keep_one_cell():
    # scene.rs:910-925
    s = scene(
        '[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.8, 0.8, 0.8]\nsize = [0.5, 0.01, 0.5]\n'
        '[[object]]\nkind = "pool"\nmaterial = "heptane"\ncenter = [0.8, 0.0, 0.8]\nradius = 0.3\ndepth = 0.001\n'
    )
    sol = s.voxelize()  # (z, y, x, 4)

    def at(i, j, k):
        return sol[k, j, i]

    assert at(16, 16, 16)[1] > 0.0 or at(16, 15, 16)[1] > 0.0, "thin plate dropped"
    assert at(16, 0, 16)[1] > 0.0, "shallow pool dropped"
    assert at(16, 1, 16)[1] == 0.0


def test_unknown_material_lists_choices():
    # scene.rs:927-931
    with pytest.raises(pa.PyroError) as e:
        pa.material_index("lava")
    assert "heptane" in str(e.value)


def test_benchmark_grids():
    pool = pa.Scene.load(str(SCENES / "pool.toml"))
    assert pool.grid().dims == (80, 160, 80)  # C1: 4 cm
    pool.set("domain.cell", 0.02)
    assert pool.grid().dims == (160, 320, 160)  # C2: 2 cm, 8.19 M cells
    bon = pa.Scene.load(str(SCENES / "bonfire.toml"))
    assert bon.grid().dims == (160, 240, 160)  # C5: 10 cm


def test_pool_params_closed_forms():
    p = pa.Scene.load(str(SCENES / "pool.toml")).params()
    f32 = np.float32
    assert p["nx"] == 80 and p["ny"] == 160 and p["nz"] == 80 and p["frame"] == 0
    assert p["dx"] == f32(0.04)
    assert p["t0"] == 293.0 and p["g"] == f32(9.81) and p["yo2"] == f32(0.233)
    assert p["rho0"] == f32(f32(1.204) * f32(293.0)) / f32(293.0)
    assert p["vort"] == 6.0 and p["cs"] == f32(0.17)
    assert p["s_o2"] == f32(f32(44.6e6) / f32(13.1e6))  # heptane dhc / 13.1 MJ/kg O2
    assert p["dhc"] == f32(44.6e6)
    assert p["soot_form"] == f32(f32(0.037) * f32(2.5))
    assert p["soot_ox"] == 4.0 and p["wind_x"] == 0.0 and p["wind_relax"] == 0.0
    assert p["ign_on"] == 1.0 and p["ign_temp"] == 1300.0 and p["ign_flux"] == 30000.0
    assert (p["ign_x"], p["ign_y"], p["ign_z"], p["ign_r"]) == (f32(1.6), f32(0.12), f32(1.6), f32(0.3))
    assert p["rad_len"] == 0.5  # 0.5 * max(0.5, 2 r) with r = 0.5
    assert p["rad_frac"] == f32(0.33)
    assert p["seed"] == 1.0
    m = p.mats
    assert list(m[0][:4]) == [0.0, 1.0, 1e4, 2.0]  # air, inert
    assert m[1][0] == 684.0 and m[1][2] == 371.0 and m[1][3] == 1.0  # heptane, liquid pool
    assert m[1][9] == 1.0 and np.all(m[2:] == 0.0)


def test_bonfire_params_wind_and_fuel_overrides():
    p = pa.Scene.load(str(SCENES / "bonfire.toml")).params()
    f32 = np.float32
    wr = f32(f32(20.0) * f32(f32(math.pi) / f32(180.0)))
    assert abs(p["wind_x"] - 0.8 * math.cos(wr)) < 1e-6
    assert abs(p["wind_z"] - 0.8 * math.sin(wr)) < 1e-6
    assert p["wind_relax"] == f32(0.3)
    assert p["soot_form"] == f32(0.015)  # wood soot yield x ratio 1.0
    assert p["soot_ox"] == 10.0
    assert p["rad_frac"] == f32(0.30)  # wood: default 0.30
    assert p.mats[1][3] == 0.0  # wood: charring solid


def test_pool_voxel_count_matches_disc_area():
    s = pa.Scene.load(str(SCENES / "pool.toml"))
    sol = s.voxelize()
    solid = sol[..., 1] > 0
    # depth 0.08 m at 4 cm: the layers with centres 0.02 and 0.06 m
    assert solid[:, :2, :].any() and not solid[:, 2:, :].any()
    per_layer = solid[:, 0, :].sum()
    area_cells = math.pi * 0.5 ** 2 / 0.04 ** 2
    assert abs(per_layer - area_cells) < 0.03 * area_cells
    # heptane mass jitter in density * [0.85, 1.15]
    mass = sol[..., 1][solid]
    assert mass.min() >= 684.0 * 0.85 - 1e-3 and mass.max() <= 684.0 * 1.15 + 1e-3
    assert np.all(sol[..., 0][solid] == 1.0) and np.all(sol[..., 3] == 293.0)


def test_bonfire_voxelizes_logs():
    s = pa.Scene.parse(open(SCENES / "bonfire.toml").read().replace("cell = 0.1", "cell = 0.2"))
    sol = s.voxelize()
    wood = sol[..., 1] > 0
    n = int(wood.sum())
    assert 150 < n < 5000, n
    ys = np.nonzero(wood)[1]
    assert ys.max() * 0.2 > 3.0  # logs reach ~4 m


def test_hash_is_uniform_on_unit_interval():
    v = np.array([pa.scene_hash(i) for i in range(20000)])
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert abs(v.mean() - 0.5) < 0.01
    assert abs(np.var(v) - 1.0 / 12.0) < 0.005


@pytest.mark.parametrize(
    "text, frag",
    [
        ("[domain]\nsize = [1, 1]\ncell = 0.1\n", "array of 3"),
        ("[sim]\nfps = 24\n", "domain"),
        ("[domain]\nsize = [1.6, 1.6, 1.6]\ncell = 0.05\n[[object]]\nkind = \"torus\"\nmaterial = \"wood\"\n",
         "unknown object kind"),
        ("[domain]\nsize = [1.6, 1.6, 1.6]\ncell = 0.05\n[[object]]\nkind = \"box\"\nmaterial = \"steel\"\n",
         "no combustible"),
        ("[domain]\nsize = [1.6, 1.6, 1.6]\ncell = \"x\"\n", "invalid type"),
        ("[domain\n", "format error"),
    ],
)
def test_scene_errors(text, frag):
    with pytest.raises(pa.PyroError) as e:
        s = pa.Scene.parse(text)
        s.params()
        s.voxelize()
    assert frag in str(e.value)


def test_too_many_materials():
    mats = ["wood", "cardboard", "pu_foam", "pmma", "rubber", "heptane", "gasoline", "diesel"]
    objs = "".join(f'[[object]]\nkind = "box"\nmaterial = "{m}"\ncenter = [0.8, 0.8, 0.8]\nsize = [0.1, 0.1, 0.1]\n'
                   for m in mats)
    with pytest.raises(pa.PyroError) as e:
        scene(objs).params()
    assert "at most 7" in str(e.value)
