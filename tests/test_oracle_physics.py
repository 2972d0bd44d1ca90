"""Closed-form pins of the oracle's physics stages: forces, react, solid and
the statistics slots (oracle/pyro_oracle.c stage_forces, stage_react,
stage_solid).

Each case builds a state in which one term of the model acts alone, or in
which the terms that act have an answer in closed form, and checks the
oracle's FP32 output against that answer evaluated in float64:

- forces (fluid.wgsl:31-109): Boussinesq-free ideal-gas buoyancy of a
  stratified column with particle weight; vorticity confinement of a linear
  |omega| ramp (direction N x omega, magnitude vort*dx*gate*|omega|, 60 m/s^2
  cap, zero for a uniform field); exponential relaxation to the wind with the
  side-sponge and hot-gas terms; top-sponge damping of downflow only; the
  +-40 m/s clamp;
- react (combust.wgsl:139-273): Smagorinsky diffusion of spikes (interior,
  ground, open side, next to a solid; clamped coefficient); mass-weighted
  mixing of pyrolysis vapour, wall convection and char-oxidation heat at a
  solid face; the ignition source; soot growth and oxidation; condensation;
  optically-thin and escape-factor radiation with its cap; the sponge;
- solid (combust.wgsl:3-137): irradiance of an isothermal absorbing medium
  (transmittance exp(-kappa L) over the ray's marched length), evaporation
  energy balance of a liquid at and below boiling, the charring split of
  pyrolysed wood, Arrhenius and diffusion-limited char oxidation, and
  conduction into an enclosed solid cell;
- statistics (stats.wgsl:2-3; combust.wgsl:115-116,127,262-272): every
  summed slot recomputed from the stage's output fields in float64.

A wrong sign, a dropped or swapped factor, a wrong neighbour or a wrong
boundary rule in the oracle changes one of these numbers by far more than
the tolerances, which cover FP32 rounding only.
"""
import math

import numpy as np

import oracle
import pyro_amd as pa

F32 = np.float32
SIGMA = 5.670374e-8
RGAS = 8.314
STRIDES = (1, 1, 1, 2, 2, 3, 4, 5, 7, 9)
DIRS = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))


# ------------------------------------------------------------------ helpers
def sstep(e0, e1, x):
    t = np.clip((np.asarray(x, np.float64) - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def rho_of(p, T):
    return p["rho0"] * p["t0"] / np.maximum(T, 200.0)


def cp_of(T):
    return np.clip(1000.0 + 0.24 * (np.asarray(T, np.float64) - 300.0), 1000.0, 1450.0)


def side_sponge(p, dims):
    """1 - smoothstep(0, sponge, cells to the nearest side wall), (nz, nx)."""
    nx, ny, nz = dims
    x, z = np.arange(nx), np.arange(nz)
    ex = np.minimum(x, nx - 1 - x)[None, :]
    ez = np.minimum(z, nz - 1 - z)[:, None]
    return 1.0 - sstep(0.0, p["sponge"], np.minimum(ex, ez))


def top_sponge(p, dims):
    nx, ny, nz = dims
    return sstep(ny - 1.5 * p["sponge"], ny - 1.0, np.arange(ny))


def blank(size=(1.6, 2.4, 1.6), cell=0.05, extra=""):
    return pa.Scene.parse(
        f"[domain]\nsize = [{size[0]}, {size[1]}, {size[2]}]\ncell = {cell}\n"
        f'[fuel]\nmaterial = "heptane"\n{extra}'
    )


def gas(dims, p, T=None, yf=0.0, yo=None, ys=0.0, ym=0.0):
    nx, ny, nz = dims
    s0 = np.zeros((nz, ny, nx, 4), np.float64)
    s1 = np.zeros((nz, ny, nx, 4), np.float64)
    s0[..., 0] = p["t0"] if T is None else T
    s0[..., 1] = yf
    s0[..., 2] = p["yo2"] if yo is None else yo
    s0[..., 3] = ys
    s1[..., 0] = ym
    return s0.astype(F32), s1.astype(F32)


# ------------------------------------------------------------------ forces
def run_forces(sc, p, s0, s1, curl=None, vel=None, dt=0.005):
    """fluid.wgsl:63-109 on the given inputs: the post-advection velocity
    (ub, vb, wb) is updated in place; returns it in float64."""
    nx, ny, nz = sc.grid().dims
    o = oracle.OSim(p, sc.voxelize(), None, 4, 0.0)
    o.begin_step(dt, False)
    o.write_field("s0", s0)
    o.write_field("s1", s1)
    o.write_field("curl", np.zeros((nz, ny, nx, 4), F32) if curl is None else curl)
    for k, name in enumerate(("ub", "vb", "wb")):
        o.write_field(name, np.zeros((nz + 1, ny + 1, nx + 1), F32) if vel is None else vel[k])
    o.run_stage("forces")
    return [o.read_field(n).astype(np.float64) for n in ("ub", "vb", "wb")]


def test_buoyancy_of_a_stratified_column_with_particle_weight():
    """v += dt (g buoy (Tf - T0)/Tf - g smoke_w m), Tf and m the means of the
    two cells sharing the face; ground faces fixed; the top face takes the
    top cell's values; horizontal faces untouched without wind."""
    sc = blank()
    p = sc.params()
    p["buoy"], p["smoke_w"] = 0.8, 2.0
    nx, ny, nz = dims = sc.grid().dims
    y = np.arange(ny, dtype=np.float64)
    T = 300.0 + 15.0 * y
    ys, ym = 1e-3 + 1e-5 * y, 2e-3 * np.ones(ny)
    s0, s1 = gas(dims, p, T=T[None, :, None], ys=ys[None, :, None], ym=ym[None, :, None])
    dt = 0.005
    u, v, w = run_forces(sc, p, s0, s1, dt=dt)
    T, m = s0[0, :, 0, 0].astype(np.float64), (s0[0, :, 0, 3] + s1[0, :, 0, 0]).astype(np.float64)
    Tf = np.concatenate([[np.nan], 0.5 * (T[:-1] + T[1:]), [T[-1]]])
    mf = np.concatenate([[np.nan], 0.5 * (m[:-1] + m[1:]), [m[-1]]])
    g, b, sw, t0 = p["g"], p["buoy"], p["smoke_w"], p["t0"]
    want = dt * (g * b * (Tf - t0) / Tf - g * sw * mf)
    assert np.all(want[1:] > 0)  # upflow everywhere: the top sponge leaves it alone
    got = v[:nz, :, :nx]
    np.testing.assert_allclose(got[:, 1:, :], np.broadcast_to(want[None, 1:, None], got[:, 1:, :].shape), rtol=2e-6)
    assert not np.any(got[:, 0, :])
    assert not np.any(u) and not np.any(w)


def test_top_sponge_damps_downflow_only():
    """Cold heavy gas sinks; in the top sponge the downflow decays by
    exp(-8 st dt) with st = smoothstep(ny - 1.5 sponge, ny - 1, y)."""
    sc = blank()
    p = sc.params()
    nx, ny, nz = dims = sc.grid().dims
    s0, s1 = gas(dims, p, T=250.0)
    dt = 0.05
    _, v, _ = run_forces(sc, p, s0, s1, dt=dt)
    t0, g = p["t0"], p["g"]
    f = dt * g * (250.0 - t0) / 250.0
    st = top_sponge(p, dims)
    yf = np.arange(1, ny + 1)
    want = f * np.exp(-8.0 * st[np.minimum(yf, ny - 1)] * dt)
    assert st[ny - 1] == 1.0 and st[ny - 10] == 0.0
    np.testing.assert_allclose(v[:nz, 1:, :nx], np.broadcast_to(want[None, :, None], (nz, ny, nx)), rtol=3e-6)
    # the same fields with warm gas rising through the layer: no damping
    s0, s1 = gas(dims, p, T=400.0)
    _, v, _ = run_forces(sc, p, s0, s1, dt=dt)
    np.testing.assert_allclose(v[:nz, 1:, :nx], dt * g * (400.0 - t0) / 400.0, rtol=2e-6)


def test_wind_relaxation_with_side_sponge_and_hot_gas():
    """u -> u + (wind - u)(1 - exp(-kr dt)), kr = relax (1 - 0.85 hot) +
    8 side(c): hot = smoothstep(350, 900, T of the cell on the face's + side,
    ambient beyond the domain); v is not relaxed."""
    sc = blank()
    p = sc.params()
    p["wind_x"], p["wind_z"], p["wind_relax"] = 3.0, -2.0, 0.5
    nx, ny, nz = dims = sc.grid().dims
    T = np.full((nz, ny, nx), p["t0"])
    T[:, :, nx // 2:] = 1000.0
    T[:, 20:30, :] = 600.0
    s0, s1 = gas(dims, p, T=T)
    rng = np.random.default_rng(3)
    vel0 = [rng.uniform(-2, 2, (nz + 1, ny + 1, nx + 1)).astype(F32) for _ in range(3)]
    dt = 0.02
    u, v, w = run_forces(sc, p, s0, s1, vel=vel0, dt=dt)
    Tc = s0[..., 0].astype(np.float64)
    ss = side_sponge(p, dims)
    for k, got, wind in ((0, u, p["wind_x"]), (2, w, p["wind_z"])):
        sh = (nz, ny, nx + 1) if k == 0 else (nz + 1, ny, nx)
        Tl = np.full(sh, p["t0"])
        if k == 0:
            Tl[:, :, :nx] = Tc
            side = ss[:, np.minimum(np.arange(nx + 1), nx - 1)][:, None, :]
        else:
            Tl[:nz] = Tc
            side = ss[np.minimum(np.arange(nz + 1), nz - 1), :][:, None, :]
        hot = sstep(350.0, 900.0, Tl)
        kr = p["wind_relax"] * (1.0 - 0.85 * hot) + 8.0 * side
        a = 1.0 - np.exp(-kr * dt)
        u0 = vel0[k][:sh[0], :sh[1], :sh[2]].astype(np.float64)
        want = u0 + (wind - u0) * a
        np.testing.assert_allclose(got[:sh[0], :sh[1], :sh[2]], want, rtol=1e-5, atol=2e-6)
    # v: ambient cells have no buoyancy and the warm ones only rise
    t0, g = p["t0"], p["g"]
    Tf = 0.5 * (Tc[:, :-1, :] + Tc[:, 1:, :])
    want_v = vel0[1][:nz, 1:ny, :nx].astype(np.float64) + dt * g * (Tf - t0) / Tf
    st = top_sponge(p, dims)[1:ny][None, :, None]
    want_v = np.where(want_v < 0, want_v * np.exp(-8.0 * st * dt), want_v)
    np.testing.assert_allclose(v[:nz, 1:ny, :nx], want_v, rtol=1e-5, atol=2e-6)


def test_velocity_clamp():
    sc = blank()
    p = sc.params()
    nx, ny, nz = dims = sc.grid().dims
    s0, s1 = gas(dims, p)
    vel0 = [np.zeros((nz + 1, ny + 1, nx + 1), F32) for _ in range(3)]
    c = (nz // 2, ny // 2, nx // 2)
    for k in range(3):
        vel0[k][c] = 55.0
        vel0[k][c[0] + 1, c[1], c[2]] = -70.0
        vel0[k][c[0] + 2, c[1], c[2]] = 39.5
    u, v, w = run_forces(sc, p, s0, s1, vel=vel0)
    for got in (u, v, w):
        assert got[c] == 40.0 and got[c[0] + 1, c[1], c[2]] == -40.0 and got[c[0] + 2, c[1], c[2]] == F32(39.5)


def confinement_field(dims, a=12.0, b=7.0, base=40.0):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    W = base + a * x + b * y
    curl = np.zeros((nz, ny, nx, 4))
    curl[..., 2] = W
    curl[..., 3] = W
    return curl.astype(F32), W


def test_vorticity_confinement_of_a_linear_vorticity_ramp():
    """omega = (0, 0, W) with W linear in x and y: the unit gradient N of |omega|
    is (a, b, 0)/sqrt(a^2 + b^2) at interior cells, the cell force is
    vort dx gate (N x omega) capped at 60 m/s^2, and a face takes the mean of
    its two cells.  gate = max(smoothstep(vort_t, vort_t + 500, T),
    0.35 smoothstep(0, 2e-4, Y_s + Y_m))."""
    sc = blank()
    p = sc.params()
    nx, ny, nz = dims = sc.grid().dims
    a, b = 12.0, 7.0
    curl, W = confinement_field(dims, a, b)
    T = np.full((nz, ny, nx), p["t0"])
    T[:, :, :10] = 1200.0
    T[:, :, 10:14] = 700.0
    sm = np.zeros((nz, ny, nx))
    sm[:, :, 14:20] = 1e-4
    s0, s1 = gas(dims, p, T=T, ys=sm)
    dt = 0.002
    u, v, w = run_forces(sc, p, s0, s1, curl=curl, dt=dt)
    Tc = s0[..., 0].astype(np.float64)
    gate = np.maximum(sstep(p["vort_t"], p["vort_t"] + 500.0, Tc), 0.35 * sstep(0.0, 2e-4, s0[..., 3]))
    n = np.array([a, b]) / math.hypot(a, b)
    mag = p["vort"] * p["dx"] * gate * W
    scale = np.where(mag > 60.0, 60.0 / np.maximum(mag, 1e-30), 1.0)
    fx = p["vort"] * p["dx"] * gate * n[1] * W * scale  # (N x omega)_x = N_y W
    fy = -p["vort"] * p["dx"] * gate * n[0] * W * scale  # (N x omega)_y = -N_x W
    assert (mag > 60.0).any() and ((mag > 0) & (mag < 60.0)).any() and (gate == 0).any()
    ss = side_sponge(p, dims)
    st = top_sponge(p, dims)
    # interior u faces: both cells have central gradients
    fu = 0.5 * (fx[1:-1, 1:-1, :-1] + fx[1:-1, 1:-1, 1:])[:, :, 1:]  # faces x = 2..nx-2
    fu = fu[:, :, :-1]
    x = np.arange(2, nx - 1)
    side = ss[1:-1][:, None, x]
    want_u = dt * fu * np.exp(-8.0 * side * dt)
    np.testing.assert_allclose(u[1:nz - 1, 1:ny - 1, 2:nx - 1], want_u, rtol=2e-5, atol=1e-7)
    # interior v faces: buoyancy of the hot cells plus the y component,
    # downflow damped by the top sponge
    t0, g = p["t0"], p["g"]
    Tf = 0.5 * (Tc[:, :-1, :] + Tc[:, 1:, :])
    fv = g * (Tf - t0) / Tf + 0.5 * (fy[:, :-1, :] + fy[:, 1:, :])
    vv = dt * fv[1:-1, 1:-1, 1:-1]  # faces y = 2..ny-2, cells x, z interior
    yidx = np.arange(2, ny - 1)
    damp = np.where(vv < 0, np.exp(-8.0 * st[yidx][None, :, None] * dt), 1.0)
    np.testing.assert_allclose(v[1:nz - 1, 2:ny - 1, 1:nx - 1], vv * damp, rtol=2e-5, atol=1e-7)
    # w: omega_z gives no z force
    assert np.abs(w).max() == 0.0
    # a uniform |omega| has no gradient and gives no force at all
    curl[..., 3] = 300.0
    u2, _, w2 = run_forces(sc, p, s0, s1, curl=curl, dt=dt)
    assert not np.any(u2) and not np.any(w2)


# ------------------------------------------------------------------ react
def run_react(sc, p, s0b, s1b, curl=None, dt=0.005, flags=None, pyr=None, sol=None, ign_until=0.0, stats=True):
    nx, ny, nz = sc.grid().dims
    o = oracle.OSim(p, sc.voxelize(), None, 4, ign_until)
    o.begin_step(dt, stats)
    o.write_field("s0b", s0b)
    o.write_field("s1b", s1b)
    o.write_field("curl", np.zeros((nz, ny, nx, 4), F32) if curl is None else curl)
    if flags is not None:
        o.write_field("flags", flags)
    if pyr is not None:
        o.write_field("pyr", pyr)
    if sol is not None:
        o.write_field("sol_out", sol)
    o.run_stage("react")
    return o, o.read_field("s0").astype(np.float64), o.read_field("s1").astype(np.float64)


def test_subgrid_diffusion_of_spikes():
    """a += kd sum_n (a_n - a), kd = min(2 nu_t dt/dx^2, 0.15), nu_t =
    (cs dx)^2 |omega|: six neighbours inside, five at the ground (no flux),
    an open side counts ambient and the side sponge then decays the cell,
    solid neighbours are skipped."""
    sc = blank()
    p = sc.params()
    nx, ny, nz = dims = sc.grid().dims
    dt = 0.005
    for Wm in (200.0, 2000.0):
        nut = (p["cs"] * p["dx"]) ** 2 * Wm
        kd = min(nut / 0.5 * dt / p["dx"] ** 2, 0.15)
        s0b, s1b = gas(dims, p)
        s0b[16, 24, 16, 3] = 0.01     # soot spike inside
        s1b[16, 24, 10, 0] = 0.02     # smoke spike inside
        s0b[10, 24, 16, 0] = 303.0    # temperature spike inside (no radiation: kappa = 0)
        s0b[16, 0, 16, 3] = 0.01      # soot spike on the ground
        s0b[16, 30, 0, 3] = 0.01      # soot spike on the open x = 0 side
        s1b[10, 30, 0, 0] = 0.02      # smoke spike on the open x = 0 side
        flags = np.zeros((nz, ny, nx), np.uint32)
        flags[22, 10, 17] = 1          # a solid beside a spike
        s0b[22, 10, 16, 3] = 0.01
        curl = np.zeros((nz, ny, nx, 4), F32)
        curl[..., 3] = Wm
        sol = np.zeros((nz, ny, nx, 4), F32)
        sol[22, 10, 17, 3] = 293.0
        o, s0, s1 = run_react(sc, p, s0b, s1b, curl=curl, dt=dt, flags=flags, sol=sol)
        tol = dict(rtol=3e-6, atol=1e-12)
        np.testing.assert_allclose(s0[16, 24, 16, 3], 0.01 * (1 - 6 * kd), **tol)
        for dz, dy, dx_ in DIRS:
            np.testing.assert_allclose(s0[16 + dz, 24 + dy, 16 + dx_, 3], 0.01 * kd, **tol)
        np.testing.assert_allclose(s1[16, 24, 10, 0], 0.02 * (1 - 6 * kd), **tol)
        np.testing.assert_allclose(s1[16, 24, 11, 0], 0.02 * kd, **tol)
        np.testing.assert_allclose(s0[10, 24, 16, 0], 303.0 - 60.0 * kd, rtol=1e-7)
        np.testing.assert_allclose(s0[10, 25, 16, 0], 293.0 + 10.0 * kd, rtol=1e-7)
        np.testing.assert_allclose(s0[16, 0, 16, 3], 0.01 * (1 - 5 * kd), **tol)
        dec = 1.0 - math.exp(-dt * 4.0 * 1.0)  # side sponge weight 1 on the outermost cells
        np.testing.assert_allclose(s0[16, 30, 0, 3], 0.01 * (1 - 6 * kd) * (1 - dec), **tol)
        np.testing.assert_allclose(s1[10, 30, 0, 0], 0.02 * (1 - 6 * kd) * (1 - dec), **tol)
        np.testing.assert_allclose(s0[22, 10, 16, 3], 0.01 * (1 - 5 * kd), **tol)
        # the solid cell carries its surface temperature and no species
        assert list(s0[22, 10, 17]) == [293.0, 0.0, float(F32(p["yo2"])), 0.0] and not np.any(s1[22, 10, 17])
        assert kd == 0.15 or kd < 0.1


def test_solid_face_injection_convection_and_char_heat():
    """Vapour of flux m'' at Ts mixes by mass, dm = m'' dt/(rho dx):
    Y -> (Y + dm [fuel])/(1 + dm), T -> (T + dm Ts)/(1 + dm); then wall
    convection h (Ts - T) dt/(rho cp dx); char oxidation takes 1.33 m_c of O2
    and leaves half of dh_char m_c in the gas; S = m''/(rho dx)."""
    sc = blank()
    p = sc.params()
    p["kap_gas"], p["smoke_conv"] = 0.0, 0.0
    nx, ny, nz = dims = sc.grid().dims
    T, Ts, mv, mc = 400.0, 500.0, 0.02, 0.004
    s0b, s1b = gas(dims, p, T=T)
    c = (16, 10, 16)
    flags = np.zeros((nz, ny, nx), np.uint32)
    flags[c] = 1
    pyr = np.zeros((nz, ny, nx, 4), F32)
    pyr[c] = (mv, mc, 6.0, 0.0)
    sol = np.zeros((nz, ny, nx, 4), F32)
    sol[c] = (1.0, 100.0, 10.0, Ts)
    dt = 0.005
    o, s0, s1 = run_react(sc, p, s0b, s1b, dt=dt, flags=flags, pyr=pyr, sol=sol)
    dx, yo2 = p["dx"], p["yo2"]
    rho = rho_of(p, T)
    dm = mv * dt / (rho * dx)
    T1 = (T + dm * Ts) / (1 + dm)
    T2 = T1 + p["h_conv"] * (Ts - T1) * dt / (rho * cp_of(T1) * dx)
    T3 = T2 + 0.5 * p["dh_char"] * mc * dt / (rho * cp_of(T2) * dx)
    yo = yo2 / (1 + dm) - 1.33 * mc * dt / (rho * dx)
    for dz, dy, dx_ in DIRS:
        j = (c[0] + dz, c[1] + dy, c[2] + dx_)
        np.testing.assert_allclose(s0[j], [T3, dm / (1 + dm), yo, 0.0], rtol=2e-6, atol=1e-9)
        np.testing.assert_allclose(s1[j], [0.0, 0.0, mv / (rho * dx), 0.0], rtol=2e-6, atol=1e-9)
    assert list(s0[c]) == [Ts, 0.0, float(F32(yo2)), 0.0]
    # a cell two away sees nothing
    assert s0[c[0], c[1] + 2, c[2], 0] == F32(T)


def test_ignition_source():
    """T -> max(T, T0 + (T_ign - T0) smoothstep(0, 0.5, w)) and Y_F += fuel w dt,
    w = 1 - |x - x_ign|/r_ign, while the ignition is on."""
    sc = blank()
    p = sc.params()
    p["smoke_conv"] = 0.0
    nx, ny, nz = dims = sc.grid().dims
    p["ign_x"], p["ign_y"], p["ign_z"], p["ign_r"] = 0.8, 0.3, 0.8, 0.4
    p["ign_temp"], p["ign_fuel"] = 600.0, 0.0
    s0b, s1b = gas(dims, p)
    s0b[:, :, :8, 0] = 500.0  # a warm band that the source must not cool
    o, s0, _ = run_react(sc, p, s0b, s1b, ign_until=1.0)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dx = p["dx"]
    d = np.sqrt(((x + 0.5) * dx - 0.8) ** 2 + ((y + 0.5) * dx - 0.3) ** 2 + ((z + 0.5) * dx - 0.8) ** 2)
    w = np.clip(1.0 - d / 0.4, 0.0, 1.0)
    t0 = p["t0"]
    want = np.maximum(s0b[..., 0], t0 + (600.0 - t0) * sstep(0.0, 0.5, w))
    inner = side_sponge(p, dims)[:, None, :] == 0.0
    inner = inner & (top_sponge(p, dims)[None, :, None] == 0.0)
    np.testing.assert_allclose(s0[..., 0][inner], want[inner], rtol=2e-6)
    assert (s0[..., 0][inner] > 400.0).sum() > 50
    # the same with fuel injection and a source at ambient temperature
    p["ign_temp"], p["ign_fuel"] = t0, 0.5
    s0b, s1b = gas(dims, p)
    o, s0, _ = run_react(sc, p, s0b, s1b, ign_until=1.0)
    # w = 1 - d/r loses relative precision where d is close to r: absolute tolerance
    np.testing.assert_allclose(s0[..., 1][inner], (0.5 * w * 0.005)[inner], rtol=2e-6, atol=5e-9)
    # off after ign_until
    o, s0, _ = run_react(sc, p, s0b, s1b, ign_until=0.0)
    assert not np.any(s0[..., 1])


def burn_case(T, yf, yo, ys=0.0, ym=0.0, dt=0.005, params=None):
    sc = blank()
    p = sc.params()
    for k, v in (params or {}).items():
        p[k] = v
    dims = sc.grid().dims
    s0b, s1b = gas(dims, p, T=T, yf=yf, yo=yo, ys=ys, ym=ym)
    o, s0, s1 = run_react(sc, p, s0b, s1b, dt=dt)
    c = (dims[2] // 2, dims[1] // 2, dims[0] // 2)
    return p, o, s0[c], s1[c]


def test_soot_growth_in_rich_and_lean_burning():
    """Y_s += soot_form dY (0.25 + 0.75 smoothstep(0, 0.03, Y_F after
    burning)), with no oxidation below soot_ox_t; dY = Y_F,in - Y_F,out."""
    for yf in (0.2, 0.0165):
        p, o, a, b = burn_case(T=900.0, yf=yf, yo=0.03, ys=1e-3)
        dY = yf - a[1]
        assert dY > 1e-4 and a[0] < p["soot_ox_t"]
        tau = max(math.sqrt(2 * p["dx"] / p["g"]), p["tau_min"])
        # dY read back as a difference of FP32 mass fractions: absolute tolerance
        np.testing.assert_allclose(dY, min(yf, 0.03 / p["s_o2"]) * (1 - math.exp(-0.005 / tau)), rtol=2e-5, atol=5e-8)
        rich = sstep(0.0, 0.03, a[1])
        np.testing.assert_allclose(a[3], 1e-3 + p["soot_form"] * dY * (0.25 + 0.75 * rich), rtol=2e-5)
        assert (rich == 1.0) == (yf == 0.2)


def test_soot_oxidation():
    """Without fuel: Y_s -> Y_s exp(-soot_ox oxg dt), oxg = smoothstep(soot_ox_t,
    soot_ox_t + 300, T) Y_O2 / yo2."""
    dt = 0.005
    for T, frac in ((1300.0, 1.0), (1300.0, 0.5), (1450.0, 1.0), (1100.0, 1.0)):
        p, o, a, b = burn_case(T=T, yf=0.0, yo=frac * 0.233, ys=2e-3, dt=dt)
        yo2 = p["yo2"]
        oxg = sstep(p["soot_ox_t"], p["soot_ox_t"] + 300.0, T) * (F32(frac * 0.233) / yo2)
        np.testing.assert_allclose(a[3], 2e-3 * math.exp(-p["soot_ox"] * oxg * dt), rtol=2e-6)


def test_condensation_of_unburnt_vapour():
    """Cold vapour condenses: cv = Y_F min(smoke_conv cold dt, 1), cold =
    1 - smoothstep(450, 650, T); Y_F -= cv, Y_m += cv."""
    # 650 K: no condensation and no burning yet (the ignition gate starts there)
    for T, conv, dt in ((400.0, 0.5, 0.005), (550.0, 0.5, 0.005), (400.0, 300.0, 0.005), (650.0, 0.5, 0.005)):
        p, o, a, b = burn_case(T=T, yf=0.02, yo=0.2, ym=1e-3, dt=dt, params={"smoke_conv": conv})
        cold = 1.0 - sstep(450.0, 650.0, T)
        cv = 0.02 * min(conv * cold * dt, 1.0)
        np.testing.assert_allclose(a[1], 0.02 - cv, rtol=2e-6, atol=1e-12)
        np.testing.assert_allclose(b[0], 1e-3 + cv, rtol=2e-6)


def radiation_expect(p, T, ys, dt):
    """Emission 4 kappa sigma (T^4 - T0^4) times the escape probability
    (1 - exp(-tau))/tau over the radiation length (1 for tau <= 1e-3),
    capped at half the excess temperature."""
    rho = rho_of(p, T)
    fv = rho * ys / 1800.0
    kap = 1862.0 * fv * T + p["kap_gas"]  # Y_O2 = 0: the gas term is fully on
    tl = kap * p["rad_len"]
    esc = (1 - math.exp(-tl)) / tl if tl > 1e-3 else 1.0
    prad = 4 * kap * esc * SIGMA * (T ** 4 - p["t0"] ** 4)
    dTr = min(prad * dt / (rho * cp_of(T)), 0.5 * max(T - p["t0"], 0.0))
    return dTr, rho, tl


def test_radiation_loss_thin_thick_and_capped():
    sc = blank()
    dims = sc.grid().dims
    n = dims[0] * dims[1] * dims[2]
    cases = (
        (1200.0, 0.0, 0.005, {"kap_gas": 1e-4}),    # optically thin, escape factor 1
        (1200.0, 5e-3, 0.005, {}),                   # tau about 0.5
        (1200.0, 0.05, 0.005, {}),                   # thick: tau > 1
        (1800.0, 2e-3, 0.005, {}),
        (600.0, 0.05, 1.0, {}),                      # capped at 0.5 (T - T0)
    )
    for T, ys, dt, extra in cases:
        p, o, a, b = burn_case(T=T, yf=0.0, yo=0.0, ys=ys, dt=dt, params=extra)
        dTr, rho, tl = radiation_expect(p, T, ys, dt)
        np.testing.assert_allclose(a[0], T - dTr, rtol=2e-6)
        Tn = T - dTr
        prr = dTr * rho * cp_of(Tn) / dt
        np.testing.assert_allclose(b[2], -prr / (rho * cp_of(Tn) * Tn), rtol=2e-5)
        assert (tl <= 1e-3) == (extra.get("kap_gas") == 1e-4)
        assert (tl > 1.0) == (ys >= 0.05)
        if dt == 1.0:
            assert dTr == 0.5 * (T - p["t0"])
        # every cell loses the same before the sponge acts: slot 1 is n times
        # the one-cell loss
        np.testing.assert_allclose(o.read_stats_raw()[1], prr * F32(p["dx"]) ** 3 * n, rtol=1e-5)


def test_sponge_relaxes_to_ambient():
    """T -> T0 + (T - T0) e, Y_O2 -> yo2 + (Y_O2 - yo2) e, Y_F, Y_s, Y_m -> x e,
    e = exp(-4 sp dt), sp = max(side sponge, top sponge)."""
    sc = blank()
    p = sc.params()
    dims = nx, ny, nz = sc.grid().dims
    dt = 0.01
    s0b, s1b = gas(dims, p, T=340.0, yo=p["yo2"], ym=0.01)  # kappa = 0: no radiation
    o, s0, s1 = run_react(sc, p, s0b, s1b, dt=dt)
    sp = np.maximum(side_sponge(p, dims)[:, None, :], top_sponge(p, dims)[None, :, None])
    e = np.exp(-4.0 * sp * dt)
    t0 = p["t0"]
    np.testing.assert_allclose(s0[..., 0], t0 + (340.0 - t0) * e, rtol=2e-6)
    np.testing.assert_allclose(s1[..., 0], 0.01 * e, rtol=2e-6)
    assert (sp == 0).sum() > 1000 and (sp == 1).sum() > 1000


def test_react_statistics_from_output_fields():
    """Slots 0, 1, 3, 4, 6, 8 and 16 + y recomputed from the outputs of a hot
    random state: heat release q dV, soot rho Y_s dV, fuel rho Y_F dV, smoke
    rho Y_m dV (rho of the final T), maximum T, and per layer the heat
    release of cells with q dV > 1 W."""
    sc = blank()
    p = sc.params()
    dims = nx, ny, nz = sc.grid().dims
    rng = np.random.default_rng(7)
    sh = (nz, ny, nx)
    s0b = np.stack([rng.uniform(300, 1800, sh), rng.uniform(0, 0.1, sh), rng.uniform(0, 0.233, sh),
                    rng.uniform(0, 0.01, sh)], -1).astype(F32)
    s1b = np.stack([rng.uniform(0, 0.01, sh), np.zeros(sh), np.zeros(sh), np.zeros(sh)], -1).astype(F32)
    curl = np.zeros(sh + (4,), F32)
    curl[..., 3] = rng.uniform(0, 400, sh)
    s0b[:, :, :4, 1] = rng.uniform(0, 1e-7, (nz, ny, 4))  # burning below 1 W per cell
    dt = 0.005
    o, s0, s1 = run_react(sc, p, s0b, s1b, curl=curl, dt=dt)
    st = o.read_stats_raw()
    dV = F32(p["dx"]) ** 3
    rho = rho_of(p, s0[..., 0])
    q = s1[..., 1]
    assert ((q * dV > 0) & (q * dV <= 1.0)).sum() > 100
    np.testing.assert_allclose(st[0], (q * dV).sum(), rtol=1e-6)
    np.testing.assert_allclose(st[3], (rho * s0[..., 3] * dV).sum(), rtol=1e-6)
    np.testing.assert_allclose(st[6], (rho * s0[..., 1] * dV).sum(), rtol=1e-6)
    np.testing.assert_allclose(st[8], (rho * s1[..., 0] * dV).sum(), rtol=1e-6)
    assert st[4] == s0[..., 0].max()
    layer = np.where(q * dV > 1.0, q * dV, 0.0).sum(axis=(0, 2))
    np.testing.assert_allclose(st[16:16 + ny], layer, rtol=1e-6)
    assert not np.any(st[16 + ny:])
    # radiation: the energy the gas lost, dTr rho(T before loss) cp(T after)
    # per cell, is bounded by the emitted power 4 kappa sigma (T^4 - T0^4)
    assert 0.0 < st[1] < 4.0 * 1e4 * SIGMA * 1800.0 ** 4 * dV * nx * ny * nz


def test_radiation_statistic_is_the_sum_of_cell_losses():
    """Slot 1 = sum of dTr rho cp(T - dTr)/dt dV over the cells; on a uniform
    sooty gas without sponge contributions it is the single-cell value times
    the number of cells with sponge weight 0 plus the sponge cells' share,
    recomputed from each cell's own closed form."""
    sc = blank()
    p = sc.params()
    dims = nx, ny, nz = sc.grid().dims
    T, ys, dt = 1200.0, 3e-3, 0.005
    s0b, s1b = gas(dims, p, T=T, yo=0.0, ys=ys)
    o, s0, s1 = run_react(sc, p, s0b, s1b, dt=dt)
    dTr, rho, _ = radiation_expect(p, T, ys, dt)
    prr = dTr * rho * cp_of(T - dTr) / dt
    dV = F32(p["dx"]) ** 3
    np.testing.assert_allclose(o.read_stats_raw()[1], prr * dV * nx * ny * nz, rtol=1e-5)


# ------------------------------------------------------------------ solid
def solid_scene(obj, size=0.8, cell=0.05, fuel="wood"):
    return pa.Scene.parse(f"[domain]\nsize = [{size}, {size}, {size}]\ncell = {cell}\n"
                          f'[fuel]\nmaterial = "{fuel}"\n' + obj)


ONE_CELL = '[[object]]\nkind = "box"\nmaterial = "{m}"\ncenter = [0.425, 0.425, 0.425]\nsize = [0.01, 0.01, 0.01]\n'


def run_solid(sc, p, sol, s0, dt=0.005, ign_until=0.0):
    o = oracle.OSim(p, sol, None, 4, ign_until)
    o.begin_step(dt, True)
    o.write_field("s0", s0)
    o.run_stage("solid")
    return o, o.read_field("sol_out").astype(np.float64), o.read_field("pyr").astype(np.float64)


def ray_length(dims, start, d):
    """Marched length (cells along d) of a ray from `start`: each in-domain
    sample adds its whole stride (1,1,1,2,2,3,4,5,7,9)."""
    nx, ny, nz = dims
    x, y, z = start
    L = 0
    for s in STRIDES:
        if not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz):
            break
        L += s
        x, y, z = x + d[0] * s, y + d[1] * s, z + d[2] * s
    return L


def isothermal_irradiance(p, dims, cell, kap, Tg):
    """Mean over the six faces and five rays per face of
    sigma [(1 - t) Tg^4 + t T0^4], t = exp(-kap |d| dx L)."""
    t0 = p["t0"]
    tot = 0.0
    for d in DIRS:
        t1 = (1, 0, 0) if d[1] != 0 else (0, 1, 0)
        t2 = (1, 0, 0) if d[2] != 0 else (0, 0, 1)
        start = (cell[0] + d[0], cell[1] + d[1], cell[2] + d[2])
        face = 0.0
        for r in (d, tuple(a + b for a, b in zip(d, t1)), tuple(a - b for a, b in zip(d, t1)),
                  tuple(a + b for a, b in zip(d, t2)), tuple(a - b for a, b in zip(d, t2))):
            L = ray_length(dims, start, r)
            tr = math.exp(-kap * math.sqrt(sum(v * v for v in r)) * p["dx"] * L)
            face += SIGMA * ((1 - tr) * Tg ** 4 + tr * t0 ** 4)
        tot += 0.2 * face
    return tot / 6.0


def test_irradiance_of_an_isothermal_absorbing_gas():
    """A lone wood cell below its pyrolysis window in hot sooty gas: the
    surface temperature rises by dt qnet/cap with qnet = eps (q_in - sigma
    Ts^4) + h (Tg - Ts), q_in the isothermal-medium irradiance."""
    sc = solid_scene(ONE_CELL.format(m="wood"), size=0.8)
    p = sc.params()
    dims = sc.grid().dims
    sol = sc.voxelize()
    cell = tuple(int(v) for v in np.argwhere(sol[..., 1] > 0)[0][::-1])  # (x, y, z)
    Ts, Tg, dt = 400.0, 1200.0, 0.005
    sol[cell[2], cell[1], cell[0], 3] = Ts
    m = p.mats[1]
    cap = m[0] * m[6] * min(p["layer"], p["dx"])
    for ys in (1e-5, 2e-4, 0.5):
        s0, _ = gas(dims, p, T=Tg, ys=ys)
        o, so, pyr = run_solid(sc, p, sol, s0, dt)
        kap = 1862.0 * (p["rho0"] * p["t0"] * ys / 1800.0)  # f_v T: T cancels
        qin = isothermal_irradiance(p, dims, cell, kap, Tg)
        qnet = m[5] * (qin - SIGMA * Ts ** 4) + p["h_conv"] * (Tg - Ts)
        got = so[cell[2], cell[1], cell[0]]
        np.testing.assert_allclose(got[3], Ts + dt * qnet / cap, rtol=2e-6)
        assert list(pyr[cell[2], cell[1], cell[0]]) == [0.0, 0.0, 6.0, 0.0]
        if ys == 0.5:  # optically thick: the black-body flux of the gas
            np.testing.assert_allclose(qin, SIGMA * Tg ** 4, rtol=1e-6)


def test_liquid_evaporation_energy_balance():
    """A liquid cell at its boiling point puts the whole net flux into
    evaporation (m'' L = qnet, Ts stays at Tb); 30 K below, smoothstep gives
    half of it and the rest heats the 1 mm layer; the liquid loses
    m'' nF dt/dx per volume; statistics slots 2, 7 and 9."""
    sc = solid_scene(ONE_CELL.format(m="heptane"), size=0.8, fuel="heptane")
    p = sc.params()
    dims = sc.grid().dims
    sol0 = sc.voxelize()
    z, y, x = np.argwhere(sol0[..., 1] > 0)[0]
    m = p.mats[int(sol0[z, y, x, 0])]
    assert m[3] == 1.0
    Tb, L, mult, dt, dx = m[2], m[1], m[9], 0.005, p["dx"]
    cap = m[0] * m[6] * 0.001
    Tg = 1100.0
    s0, _ = gas(dims, p, T=Tg, ys=0.5)  # optically thick: q_in = sigma Tg^4
    for Ts in (Tb, Tb - 30.0, Tb - 60.0):
        sol = sol0.copy()
        sol[z, y, x, 3] = Ts
        o, so, pyr = run_solid(sc, p, sol, s0, dt)
        qin = isothermal_irradiance(p, dims, (x, y, z), 1862.0 * p["rho0"] * p["t0"] * 0.5 / 1800.0, Tg)
        qnet = m[5] * (qin - SIGMA * Ts ** 4) + p["h_conv"] * (Tg - Ts)
        frac = sstep(Tb - 60.0, Tb, Ts)
        mv = frac * qnet / L * mult
        np.testing.assert_allclose(pyr[z, y, x], [mv, 0.0, 6.0, 0.0], rtol=3e-6, atol=1e-12)
        np.testing.assert_allclose(so[z, y, x, 3], min(Ts + dt * (qnet - mv * L) / cap, Tb), rtol=1e-6)
        np.testing.assert_allclose(so[z, y, x, 1], sol0[z, y, x, 1] - mv * 6 / dx * dt, rtol=1e-6)
        if Ts == Tb:
            np.testing.assert_allclose(mv * L, qnet * mult, rtol=1e-6)
            assert so[z, y, x, 3] == F32(Tb)
        st = o.read_stats_raw()
        np.testing.assert_allclose(st[2], mv * 6 * dx * dx, rtol=3e-6, atol=1e-15)
        np.testing.assert_allclose(st[7], (so[z, y, x, 1] + so[z, y, x, 2]) * F32(dx) ** 3, rtol=1e-6)
        np.testing.assert_allclose(st[9], 6 * dx * dx if mv > 1e-4 else 0.0, rtol=1e-6)


def test_charring_pyrolysis_conserves_mass():
    """Wood inside its pyrolysis window: m'' = smoothstep(Tp - 20, Tp + 30, Ts)
    qnet/L; the virgin mass consumed splits into volatiles (1 - chi), which
    leave as m'' nF dt/dx, and char (chi)."""
    sc = solid_scene(ONE_CELL.format(m="wood"), size=0.8)
    p = sc.params()
    dims = sc.grid().dims
    sol = sc.voxelize()
    z, y, x = np.argwhere(sol[..., 1] > 0)[0]
    m = p.mats[int(sol[z, y, x, 0])]
    assert m[3] == 0.0
    Ts, Tg, dt, dx = 610.0, 1200.0, 0.005, p["dx"]
    sol[z, y, x, 2:4] = (0.0, Ts)
    s0, _ = gas(dims, p, T=Tg, ys=0.5)
    o, so, pyr = run_solid(sc, p, sol, s0, dt)
    qin = isothermal_irradiance(p, dims, (x, y, z), 1862.0 * p["rho0"] * p["t0"] * 0.5 / 1800.0, Tg)
    qnet = m[5] * (qin - SIGMA * Ts ** 4) + p["h_conv"] * (Tg - Ts)
    mv = sstep(m[2] - 20.0, m[2] + 30.0, Ts) * qnet / m[1] * m[9]
    np.testing.assert_allclose(pyr[z, y, x, 0], mv, rtol=3e-6)
    chi = m[4]
    released = mv * 6 * dt / dx  # volatiles per volume
    np.testing.assert_allclose(so[z, y, x, 2], chi / (1 - chi) * released, rtol=1e-5)
    # virgin mass lost = volatiles + char (the difference of two FP32 masses
    # near 500 kg/m^3 carries about 3e-5 of rounding)
    lost = float(sol[z, y, x, 1]) - so[z, y, x, 1]
    np.testing.assert_allclose(lost, released + so[z, y, x, 2], atol=1e-4)
    assert lost > 100 * 1e-4
    cap = m[0] * m[6] * min(p["layer"], dx)
    np.testing.assert_allclose(so[z, y, x, 3], Ts + dt * (qnet - mv * m[1]) / cap, rtol=1e-6)


def test_char_oxidation_kinetic_and_diffusion_limits():
    """m_c = min(A exp(-E/(R Ts)) char layer, h_m rho(Tg) Y_O2/1.33): kinetic
    at 700 K, oxygen-diffusion limited at 900 K; half of dh_char m_c heats the
    surface; char falls by m_c nF dt/dx."""
    sc = solid_scene(ONE_CELL.format(m="wood"), size=0.8)
    p = sc.params()
    dims = sc.grid().dims
    sol0 = sc.voxelize()
    z, y, x = np.argwhere(sol0[..., 1] > 0)[0]
    m = p.mats[int(sol0[z, y, x, 0])]
    dt, dx, chr_ = 0.005, p["dx"], 50.0
    layer = min(p["layer"], dx)
    cap = m[0] * m[6] * layer
    s0, _ = gas(dims, p)  # ambient, transparent
    limits = []
    for Ts in (700.0, 900.0):
        sol = sol0.copy()
        sol[z, y, x, 2:4] = (chr_, Ts)
        o, so, pyr = run_solid(sc, p, sol, s0, dt)
        kin = p["char_a"] * math.exp(-p["char_e"] / (RGAS * Ts)) * chr_ * layer
        dif = p["h_mass"] * rho_of(p, p["t0"]) * p["yo2"] / 1.33
        mc = min(kin, dif)
        limits.append(kin < dif)
        qnet = m[5] * (SIGMA * p["t0"] ** 4 - SIGMA * Ts ** 4) + p["h_conv"] * (p["t0"] - Ts)
        assert qnet < 0
        np.testing.assert_allclose(pyr[z, y, x], [0.0, mc, 6.0, 0.0], rtol=1e-5)
        np.testing.assert_allclose(so[z, y, x, 2], chr_ - mc * 6 * dt / dx, rtol=1e-6)
        np.testing.assert_allclose(so[z, y, x, 3], Ts + dt * (qnet + 0.5 * mc * p["dh_char"]) / cap, rtol=1e-6)
    assert limits == [True, False]


def test_conduction_into_an_enclosed_solid_cell():
    """The centre of a 3x3x3 wood block has no gas face: Ts += dt sum_n
    k/dx (T_n - Ts) / (rho c dx), the explicit heat equation."""
    obj = '[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.425, 0.425, 0.425]\nsize = [0.15, 0.15, 0.15]\n'
    sc = solid_scene(obj, size=0.8)
    p = sc.params()
    dims = sc.grid().dims
    sol = sc.voxelize()
    assert (sol[..., 1] > 0).sum() == 27
    c = (8, 8, 8)
    m = p.mats[int(sol[c][0])]
    # wood conducts slowly (rho c dx^2 / k = 3.3 h at 5 cm): large contrasts
    # and a long step make the change resolvable in FP32
    Tn = (1350.0, 1420.0, 390.0, 300.0, 1960.0, 1330.0)
    sol[c + (3,)] = 400.0
    for (dz, dy, dx_), t in zip(DIRS, Tn):
        sol[c[0] + dz, c[1] + dy, c[2] + dx_, 3] = t
    s0, _ = gas(dims, p)
    dt, dx = 5.0, p["dx"]
    o, so, pyr = run_solid(sc, p, sol, s0, dt)
    q = sum(m[7] / dx * (t - 400.0) for t in Tn)
    rise = dt * q / (m[0] * m[6] * dx)
    assert rise > 1.0
    np.testing.assert_allclose(so[c + (3,)] - 400.0, rise, rtol=1e-4)
    assert not np.any(pyr[c])


def test_solid_statistics_on_a_burning_scene():
    """Slots 2 (pyrolysis rate), 7 (combustible solid mass) and 9 (pyrolysing
    area) recomputed from pyr and the updated solid of a scene with a pool,
    wood and an inert block under an ignition flux."""
    obj = ('[[object]]\nkind = "pool"\nmaterial = "heptane"\ncenter = [0.3, 0.0, 0.3]\nradius = 0.2\ndepth = 0.1\n'
           '[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.55, 0.3, 0.55]\nsize = [0.2, 0.2, 0.2]\n'
           '[[object]]\nkind = "box"\nmaterial = "steel"\ncenter = [0.2, 0.6, 0.6]\nsize = [0.1, 0.1, 0.1]\n')
    sc = solid_scene(obj, size=0.8)
    p = sc.params()
    dims = sc.grid().dims
    p["ign_x"], p["ign_y"], p["ign_z"], p["ign_r"], p["ign_flux"] = 0.4, 0.2, 0.4, 0.4, 60000.0
    sol = sc.voxelize()
    rng = np.random.default_rng(5)
    solid = sol[..., 1] > 0
    sol[..., 3] = np.where(solid, rng.uniform(330.0, 650.0, solid.shape), sol[..., 3]).astype(F32)
    s0, _ = gas(dims, p, T=rng.uniform(293, 1400, dims[::-1]), ys=rng.uniform(0, 1e-3, dims[::-1]))
    o, so, pyr = run_solid(sc, p, sol, s0, 0.005, ign_until=1.0)
    st = o.read_stats_raw()
    dx = F32(p["dx"])
    kinds = p.mats[:, 3][sol[..., 0].astype(int)]
    burnable = solid & (kinds != 2.0)
    np.testing.assert_allclose(st[2], (pyr[..., 0] * pyr[..., 2]).sum() * dx * dx, rtol=1e-6)
    np.testing.assert_allclose(st[7], ((so[..., 1] + so[..., 2]) * burnable).sum() * dx ** 3, rtol=1e-6)
    np.testing.assert_allclose(st[9], (pyr[..., 2] * (pyr[..., 0] > 1e-4)).sum() * dx * dx, rtol=1e-6)
    assert (pyr[..., 0] > 1e-4).sum() > 10 and st[7] > 0
