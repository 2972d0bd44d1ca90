"""Parity harness: build identical seeded states in the oracle and the GPU sim.

A case is a scene (Params + SOL) plus a seeded perturbation of the state
(temperature/species fields, velocity, warm-start pressure).  ``twin()``
returns an (oracle, gpu) pair holding bit-identical state; parity tests then
drive both with the same calls and compare every field.
"""
from __future__ import annotations

import numpy as np

import oracle
import pyro_amd as pa
from conftest import SCENES

# 32 x 48 x 32 cells at 5 cm: heptane pool, charring wood block, inert steel
# block, ignition and wind -- every branch of solid/react/forces is live.
MINI_SCENE = """
[domain]
size = [1.6, 2.4, 1.6]
cell = 0.05
[sim]
vorticity = 6
[[object]]
kind = "pool"
material = "heptane"
center = [0.55, 0.0, 0.8]
radius = 0.3
depth = 0.1
[[object]]
kind = "box"
material = "wood"
center = [1.15, 0.35, 0.8]
size = [0.3, 0.3, 0.4]
[[object]]
kind = "box"
material = "steel"
center = [0.8, 1.5, 0.3]
size = [0.2, 0.1, 0.2]
[ignition]
pos = [0.55, 0.15, 0.8]
radius = 0.3
flux = 30000
temp = 1300
until = 4.0
fuel = 0.5
[wind]
speed = 1.5
dir = 30
"""


def mini_scene() -> pa.Scene:
    return pa.Scene.parse(MINI_SCENE)


def pool_scene(cell: float = 0.04) -> pa.Scene:
    s = pa.Scene.load(str(SCENES / "pool.toml"))
    s.set("domain.cell", cell)
    return s


def perturbed_state(scene: pa.Scene, seed: int, vel_amp: float = 8.0, hot: bool = True):
    """Seeded, physically plausible state arrays (s0, s1, u, v, w, sol, phi)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = scene.grid().dims
    p = scene.params()
    sol = scene.voxelize()
    solid = sol[..., 1] > 0.0
    s0 = np.zeros((nz, ny, nx, 4), np.float32)
    s1 = np.zeros_like(s0)
    if hot:
        # smooth hot plume plus noise
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        r2 = ((x - nx / 3) ** 2 + (z - nz / 2) ** 2) / (0.2 * nx) ** 2 + ((y - ny / 4) / (0.5 * ny)) ** 2
        plume = np.exp(-r2)
        s0[..., 0] = 293.0 + 1500.0 * plume + rng.uniform(0, 60, plume.shape)
        s0[..., 1] = 0.08 * plume * rng.uniform(0.5, 1.0, plume.shape)
        s0[..., 2] = 0.233 * (1.0 - 0.7 * plume) * rng.uniform(0.9, 1.0, plume.shape)
        s0[..., 3] = 2e-3 * plume * rng.uniform(0, 1, plume.shape)
        s1[..., 0] = 1e-3 * rng.uniform(0, 1, plume.shape)
        s1[..., 2] = rng.uniform(-0.5, 2.0, plume.shape)
    else:
        s0[..., 0] = p["t0"]
        s0[..., 2] = p["yo2"]
    # hot solid surfaces so that pyrolysis is active
    sol[..., 3] = np.where(solid, rng.uniform(300.0, 700.0, solid.shape), sol[..., 3]).astype(np.float32)
    vel = []
    for k in range(3):
        a = rng.normal(0.0, 1.0, (nz + 1, ny + 1, nx + 1))
        # cheap smoothing for a correlated field
        for ax in range(3):
            a = 0.25 * np.roll(a, 1, ax) + 0.5 * a + 0.25 * np.roll(a, -1, ax)
        vel.append((vel_amp * a / max(a.std(), 1e-6)).astype(np.float32))
    phi = (rng.normal(0.0, 1e-3, (nz, ny, nx))).astype(np.float32)
    return dict(s0=s0, s1=s1, u=vel[0], v=vel[1], w=vel[2], sol=sol, phi=phi)


def twin(scene: pa.Scene, seed: int | None = 0, use_graph: bool = True, **kw):
    """(oracle, gpu) sims with identical state."""
    p = scene.params()
    sol = scene.voxelize()
    cyc = int(scene.get("sim.pressure_cycles"))
    until = scene.get("ignition.until")
    o = oracle.OSim(p, sol, None, cyc, until)
    g = pa.Sim(p, sol, None, cyc, until, use_graph=use_graph)
    if seed is not None:
        st = perturbed_state(scene, seed, **kw)
        for sim in (o, g):
            for name in ("s0", "s1", "u", "v", "w", "sol"):
                sim.write_field(name, st[name])
            sim.write_field(pa.level_field(0, "phi"), st["phi"])
    return o, g


VEC4 = ("s0", "s1", "s0b", "s1b", "h0", "h1", "sol", "sol_out", "pyr", "curl")
VEL = ("u", "v", "w", "ub", "vb", "wb", "hu", "hv", "hw")


def compare(o, g, names, rtol=0.0, report=None):
    """Return {name: (max_abs, rel_l2, n_mismatch)} for the listed fields."""
    out = {}
    for n in names:
        a = g.read_field(n)
        b = o.read_field(n)
        a64 = a.astype(np.float64)
        b64 = b.astype(np.float64)
        d = np.abs(a64 - b64)
        den = np.linalg.norm(b64.ravel())
        rl2 = float(np.linalg.norm(d.ravel()) / den) if den > 0 else float(np.linalg.norm(d.ravel()))
        out[n] = (float(d.max()) if d.size else 0.0, rl2, int(np.count_nonzero(a != b)))
    if report is not None:
        for n, v in out.items():
            report(f"{str(n):>10s}: max|d|={v[0]:.3e} relL2={v[1]:.3e} mismatches={v[2]}")
    return out
