"""Pinning the CPU oracle (oracle/pyro_oracle.c) independently of the GPU.

The reference (Rust/wgpu/WGSL) cannot run here and its own tests pin no
kernel output, so each oracle function is checked against what the
mathematics fixes: closed forms, exact invariants, and an independent FP64
assembly of the pressure operator with scipy.  Each check is chosen so that
a plausible restatement mistake (wrong sign, index, neighbour rule, missing
term) fails it.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
import pyro_amd as pa

F32 = np.float32


def blank(size=(1.6, 2.4, 1.6), cell=0.05, extra=""):
    return pa.Scene.parse(
        f"[domain]\nsize = [{size[0]}, {size[1]}, {size[2]}]\ncell = {cell}\n"
        f'[fuel]\nmaterial = "heptane"\n{extra}'
    )


def make(scene, sol=None, params=None):
    p = params or scene.params()
    if sol is None:
        sol = scene.voxelize()
    return oracle.OSim(p, sol, None, 4, 0.0), p


def faces(shape_cells):
    nx, ny, nz = shape_cells
    return np.zeros((nz + 1, ny + 1, nx + 1), F32)


# ------------------------------------------------------------------ exp
def test_expf_max_two_ulp():
    x = np.concatenate([np.linspace(-87.0, 88.0, 200001), np.linspace(-1.0, 1.0, 20001)]).astype(F32)
    got = oracle.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(F32)).astype(np.float64)
    err = np.abs(got - ref) / ulp
    assert err.max() <= 2.0, err.max()
    assert oracle.expf(np.array([0.0]))[0] == 1.0
    assert oracle.expf(np.array([-200.0]))[0] == 0.0
    assert np.isinf(oracle.expf(np.array([100.0]))[0])
    sub = oracle.expf(np.array([-95.0], F32))[0]  # subnormal range
    assert abs(sub - math.exp(-95.0)) <= 2 * 1.4e-45 + 1e-6 * math.exp(-95.0)


# ------------------------------------------------------------------ advection
def adv_case(u=None, v=None, w=None, s0=None, s1=None, scene=None, dt=0.002):
    sc = scene or blank()
    o, p = make(sc)
    nx, ny, nz = sc.grid().dims
    for name, arr in (("u", u), ("v", v), ("w", w), ("s0", s0), ("s1", s1)):
        if arr is not None:
            o.write_field(name, arr)
    o.begin_step(dt, False)
    for st in ("solid", "adv_vel1", "adv_vel2", "adv_s1", "adv_s2"):
        o.run_stage(st)
    return o, (nx, ny, nz), p


def rand_state(dims, seed, amp=5.0):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    vel = [rng.uniform(-amp, amp, (nz + 1, ny + 1, nx + 1)).astype(F32) for _ in range(3)]
    s0 = np.stack([rng.uniform(300, 1800, (nz, ny, nx)), rng.uniform(0, 0.1, (nz, ny, nx)),
                   rng.uniform(0, 0.233, (nz, ny, nx)), rng.uniform(0, 0.01, (nz, ny, nx))], -1).astype(F32)
    s1 = np.stack([rng.uniform(0, 0.01, (nz, ny, nx)), rng.uniform(0, 1e6, (nz, ny, nx)),
                   rng.uniform(-1, 2, (nz, ny, nx)), np.zeros((nz, ny, nx))], -1).astype(F32)
    return vel, s0, s1


def test_zero_velocity_advection_is_identity():
    sc = blank()
    dims = sc.grid().dims
    vel, s0, s1 = rand_state(dims, 0)
    zero = faces(dims)
    o, _, _ = adv_case(zero, zero, zero, s0, s1, sc)
    # exact except on the upper faces, where the sampler's base index is
    # clamped to n-2 and mix(a, b, 1) = a + (b - a) rounds at the scale of
    # the larger of the two neighbours
    inner = (slice(0, -1),) * 3
    for name, ref in (("h0", s0), ("s0b", s0), ("h1", s1), ("s1b", s1)):
        got = o.read_field(name)
        assert np.array_equal(got[inner], ref[inner]), name
        scale = np.spacing(np.abs(ref).reshape(-1, 4).max(0).astype(F32))
        assert np.all(np.abs(got - ref).reshape(-1, 4) <= 2 * scale), name
    for k in ("hu", "hv", "hw", "ub", "wb"):
        assert not np.any(o.read_field(k)), k
    vb = o.read_field("vb")
    assert not np.any(vb)


def test_constant_field_is_preserved():
    sc = blank()
    p = sc.params()
    dims = sc.grid().dims
    vel, _, _ = rand_state(dims, 1, amp=8.0)
    nx, ny, nz = dims
    amb = np.zeros((nz, ny, nx, 4), F32)
    amb[..., 0] = p["t0"]
    amb[..., 2] = p["yo2"]
    o, _, _ = adv_case(*vel, amb, np.zeros_like(amb), sc)
    assert np.array_equal(o.read_field("s0b"), amb)
    assert not np.any(o.read_field("s1b"))


def test_uniform_velocity_one_cell_shift():
    sc = blank()
    dims = nx, ny, nz = sc.grid().dims
    dt = 0.002
    U = 0.05 / dt  # one cell per step along +x
    u = faces(dims) + F32(U)
    zero = faces(dims)
    rng = np.random.default_rng(2)
    s0 = rng.uniform(0.0, 1.0, (nz, ny, nx, 4)).astype(F32)
    o, _, _ = adv_case(u, zero, zero, s0, None, sc, dt)
    got = o.read_field("s0b")
    # interior: away from the sides (outside-blend) and the x-clamp
    np.testing.assert_allclose(got[3:-3, 3:-3, 3:-3], s0[3:-3, 3:-3, 2:-4], rtol=0, atol=2e-6)
    # the u field itself is transported unchanged
    ub = o.read_field("ub")
    np.testing.assert_allclose(ub[2:-2, 2:-3, 2:-2], U, rtol=1e-6)


def test_linear_field_exact_under_fractional_shift():
    sc = blank()
    dims = nx, ny, nz = sc.grid().dims
    dt = 0.002
    shift = (0.3, -0.45, 0.2)  # cells per step
    vel = [faces(dims) + F32(s * 0.05 / dt) for s in shift]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lin = (1.0 + 0.01 * x - 0.02 * y + 0.015 * z).astype(F32)
    s0 = np.repeat(lin[..., None], 4, -1)
    o, _, _ = adv_case(*vel, s0, None, sc, dt)
    got = o.read_field("s0b")[..., 1]
    exp = 1.0 + 0.01 * (x - shift[0]) - 0.02 * (y - shift[1]) + 0.015 * (z - shift[2])
    sl = (slice(3, -3),) * 3
    np.testing.assert_allclose(got[sl], exp[sl], rtol=0, atol=3e-6)


def test_maccormack_limiter_max_principle():
    sc = blank()
    p = sc.params()
    dims = sc.grid().dims
    vel, s0, s1 = rand_state(dims, 3, amp=15.0)
    o, _, _ = adv_case(*vel, s0, s1, sc)
    got = o.read_field("s0b")
    amb = np.array([p["t0"], 0.0, p["yo2"], 0.0], F32)
    lo = np.minimum(s0.reshape(-1, 4).min(0), amb)
    hi = np.maximum(s0.reshape(-1, 4).max(0), amb)
    assert np.all(got.reshape(-1, 4) >= lo - 1e-6) and np.all(got.reshape(-1, 4) <= hi + 1e-6)
    # without the limiter MacCormack overshoots on this rough data
    h0 = o.read_field("h0")
    assert not np.array_equal(got, h0)


# ------------------------------------------------------------------ curl, divergence
def test_curl_of_solid_body_rotation():
    sc = blank()
    dims = nx, ny, nz = sc.grid().dims
    dx = 0.05
    om = 2.5
    xc, zc = nx * dx / 2, nz * dx / 2
    u = faces(dims)
    w = faces(dims)
    Z, Y, X = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    u[:] = om * ((Z + 0.5) * dx - zc)   # u-faces at (i, j+1/2, k+1/2)
    w[:] = -om * ((X + 0.5) * dx - xc)  # w-faces at (i+1/2, j+1/2, k)
    o, _ = make(sc)
    o.write_field("u", u)
    o.write_field("w", w)
    o.begin_step(0.002, True)
    o.run_stage("curl")
    c = o.read_field("curl")[2:-2, 2:-2, 2:-2]
    np.testing.assert_allclose(c[..., 1], 2 * om, rtol=1e-5)
    assert np.abs(c[..., 0]).max() < 1e-4 and np.abs(c[..., 2]).max() < 1e-4
    np.testing.assert_allclose(c[..., 3], 2 * om, rtol=1e-5)
    o.end_step()
    # stats slot 5: largest |cell-centred velocity| = om * r_max at a corner
    umax = o.read_stats_raw()[5]
    rmax = math.hypot(xc - 0.5 * dx, zc - 0.5 * dx)
    assert abs(umax - om * rmax) < 1e-3 * om * rmax


def test_divergence_of_affine_velocity():
    sc = blank(extra='[[object]]\nkind = "box"\nmaterial = "steel"\ncenter = [0.8, 1.2, 0.8]\nsize = [0.2, 0.2, 0.2]\n')
    dims = nx, ny, nz = sc.grid().dims
    dx = 0.05
    a, b, c, S = 1.5, -0.7, 0.4, 0.25
    Z, Y, X = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    o, _ = make(sc)
    o.write_field("ub", (a * X * dx).astype(F32))
    o.write_field("vb", (b * Y * dx).astype(F32))
    o.write_field("wb", (c * Z * dx).astype(F32))
    s1 = np.zeros((nz, ny, nx, 4), F32)
    s1[..., 2] = S
    o.write_field("s1", s1)
    o.begin_step(0.002, False)
    o.run_stage("solid")  # flags from the steel block
    o.run_stage("divergence")
    div = o.read_field(pa.level_field(0, "b"))
    fl = o.read_field("flags")
    assert fl.sum() > 0
    np.testing.assert_allclose(div[fl == 0], a + b + c - S, rtol=0, atol=2e-5)
    assert np.all(div[fl == 1] == 0.0)


# ------------------------------------------------------------------ pressure operator
def assemble(flags, h):
    """Independent FP64 assembly of the mg.wgsl operator: A phi = (sum_fluid phi_n - cnt phi)/h^2."""
    nz, ny, nx = flags.shape
    idx = lambda x, y, z: x + nx * (y + ny * z)  # noqa: E731
    rows, cols, vals = [], [], []
    fluid = flags == 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = idx(x, y, z)
                if not fluid[z, y, x]:
                    rows.append(i); cols.append(i); vals.append(1.0)  # identity on solids
                    continue
                cnt = 0
                for dx_, dy_, dz_ in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    a, b, c = x + dx_, y + dy_, z + dz_
                    if b < 0:
                        continue  # ground: Neumann
                    if a < 0 or c < 0 or a >= nx or b >= ny or c >= nz:
                        cnt += 1  # sides/top: phi = 0 outside
                        continue
                    if not fluid[c, b, a]:
                        continue  # solid: Neumann
                    rows.append(i); cols.append(idx(a, b, c)); vals.append(1.0 / h ** 2)
                    cnt += 1
                rows.append(i); cols.append(i); vals.append(-cnt / h ** 2)
    n = nx * ny * nz
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def box_flags(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((nz, ny, nx), np.uint32)
    for _ in range(4):
        x0, y0, z0 = rng.integers(0, nx - 3), rng.integers(0, ny - 3), rng.integers(0, nz - 3)
        f[z0:z0 + rng.integers(1, 4), y0:y0 + rng.integers(1, 4), x0:x0 + rng.integers(1, 4)] = 1
    return f


def test_operator_symmetric_and_residual64():
    nx, ny, nz, h = 10, 12, 9, 0.1
    f = box_flags(nx, ny, nz, 4)
    A = assemble(f, h)
    fluid = (f.ravel() == 0)
    Af = A[fluid][:, fluid]
    assert abs(Af - Af.T).max() < 1e-9
    # negative definite (every fluid component touches an open side/top here)
    ev = spla.eigsh(-Af, k=1, sigma=0, which="LM", return_eigenvectors=False)
    assert ev.min() > 0
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(nz, ny, nx))
    b = rng.normal(size=(nz, ny, nx))
    r = oracle.residual64(f, b, phi, h)
    ref = (b.ravel() - A @ phi.ravel())
    ref[~fluid] = 0.0
    np.testing.assert_allclose(r.ravel(), ref, rtol=1e-12, atol=1e-9)


def test_one_red_black_sweep_matches_matrix_gauss_seidel():
    nx, ny, nz, h = 12, 16, 10, 0.08
    f = box_flags(nx, ny, nz, 6)
    A = assemble(f, h).tolil()
    rng = np.random.default_rng(7)
    phi = rng.normal(size=(nz, ny, nx)).astype(F32)
    b = rng.normal(size=(nz, ny, nx)).astype(F32)
    m = oracle.OMg((nx, ny, nz), h, 1, f)
    m.set(0, phi, b)
    m.smooth(0, 1, 1.0)
    got, _, _ = m.get(0)
    # matrix form: red then black, phi_i <- (b_i - sum_{j != i} A_ij phi_j) / A_ii
    ref = phi.astype(np.float64).ravel().copy()
    bb = b.astype(np.float64).ravel()
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    colour = ((X + Y + Z) & 1).ravel()
    Acsr = A.tocsr()
    for col in (0, 1):
        new = ref.copy()
        for i in np.nonzero((colour == col) & (f.ravel() == 0))[0]:
            row = Acsr.getrow(i)
            aii = Acsr[i, i]
            if aii == 0.0:
                continue
            off = row.dot(ref)[0] - aii * ref[i]
            new[i] = (bb[i] - off) / aii
        ref = new
    np.testing.assert_allclose(got.ravel(), ref, rtol=1e-5, atol=1e-5)
    # solids untouched
    assert np.array_equal(got[f == 1], phi[f == 1])


def test_vcycles_converge_to_direct_solution():
    nx, ny, nz, h = 32, 32, 32, 0.05
    f = box_flags(nx, ny, nz, 8)
    A = assemble(f, h)
    fluid = f.ravel() == 0
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = np.exp(-((x - 12) ** 2 + (y - 20) ** 2 + (z - 15) ** 2) / 30.0) - 0.5 * np.exp(
        -((x - 22) ** 2 + (y - 8) ** 2 + (z - 20) ** 2) / 20.0)
    b = b * (f == 0)
    exact = np.zeros(nx * ny * nz)
    exact[fluid] = spla.spsolve(A[fluid][:, fluid].tocsc(), b.ravel()[fluid])
    m = oracle.OMg((nx, ny, nz), h, 4, f)
    assert m.levels == 3  # 32 -> 16 -> 8 (stops below 16)
    m.set(0, np.zeros_like(b, dtype=F32), b.astype(F32))
    rel = []
    for k in range(40):
        m.vcycles(1)
        r2, b2 = m.resid_norms()
        rel.append(math.sqrt(r2 / b2))
    # the reference V(2,2) cycle contracts by ~0.55 per cycle from a cold start here
    assert all(rel[i + 1] < 0.6 * rel[i] for i in range(8)), rel[:9]
    phi, _, _ = m.get(0)
    err = np.linalg.norm(phi.ravel()[fluid] - exact[fluid]) / np.linalg.norm(exact[fluid])
    assert err < 1e-4, err


def test_vcycle_matches_independent_numpy_restatement():
    from mg_numpy import vcycle, resid

    nx, ny, nz, h = 24, 32, 20, 0.07
    f = box_flags(nx, ny, nz, 13)
    rng = np.random.default_rng(14)
    b = (rng.normal(size=(nz, ny, nx)) * (f == 0)).astype(F32)
    p0 = (rng.normal(size=(nz, ny, nx)) * 0.01).astype(F32)
    m = oracle.OMg((nx, ny, nz), h, 4, f)
    m.set(0, p0, b)
    m.vcycles(2)
    got, _, _ = m.get(0)
    ref = p0.astype(np.float64)
    for _ in range(2):
        ref = vcycle(ref, b.astype(np.float64), f, h, m.levels)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_restrict_and_prolong_of_constants():
    nx, ny, nz, h = 16, 20, 18, 0.1
    f = box_flags(nx, ny, nz, 10)
    f[0:2, 0:2, 0:2] = 1  # one coarse cell entirely solid
    m = oracle.OMg((nx, ny, nz), h, 2, f)
    # phi = 0, b = c  ->  residual = c on fluid cells -> coarse b = c where any child is fluid
    m.set(0, np.zeros((nz, ny, nx), F32), np.full((nz, ny, nx), 3.0, F32))
    m.restrict(0)
    pc, bc, fc = m.get(1)
    child_fluid = (f == 0).reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).any(axis=(1, 3, 5))
    assert np.array_equal(fc == 0, child_fluid)
    assert np.all(bc[child_fluid] == 3.0) and np.all(bc[~child_fluid] == 0.0) and not np.any(pc)
    # constant coarse correction -> every fluid fine cell gains exactly that constant
    m.set(1, np.full(pc.shape, 0.5, F32), None)
    m.set(0, np.zeros((nz, ny, nx), F32), None)
    m.prolong(0)
    pf, _, _ = m.get(0)
    assert np.all(pf[f == 0] == 0.5) and np.all(pf[f == 1] == 0.0)


def test_projection_removes_divergence():
    sc = blank(size=(0.8, 1.6, 0.8), cell=0.05,
               extra='[[object]]\nkind = "box"\nmaterial = "steel"\ncenter = [0.4, 0.6, 0.4]\nsize = [0.15, 0.2, 0.15]\n')
    dims = nx, ny, nz = sc.grid().dims
    h = 0.05
    o, _ = make(sc)
    rng = np.random.default_rng(12)
    for k in ("ub", "vb", "wb"):
        o.write_field(k, rng.normal(0, 1, (nz + 1, ny + 1, nx + 1)).astype(F32))
    o.begin_step(0.002, False)
    o.run_stage("solid")
    o.run_stage("project")  # phi = 0: zeroes the blocked faces (as adv_vel2 does in a step)
    o.run_stage("divergence")
    b = o.read_field(pa.level_field(0, "b")).astype(np.float64)
    f = o.read_field("flags")
    A = assemble(f, h)
    fluid = f.ravel() == 0
    phi = np.zeros(nx * ny * nz)
    phi[fluid] = spla.spsolve(A[fluid][:, fluid].tocsc(), b.ravel()[fluid])
    o.write_field(pa.level_field(0, "phi"), phi.reshape(nz, ny, nx).astype(F32))
    o.run_stage("project")
    o.run_stage("divergence")  # divergence of the projected field (S = 0)
    b2 = o.read_field(pa.level_field(0, "b")).astype(np.float64)
    assert np.linalg.norm(b2[f == 0]) < 2e-5 * np.linalg.norm(b[f == 0])


# ------------------------------------------------------------------ react / solid / whole step
def test_ambient_state_is_a_fixed_point_of_the_step():
    sc = blank()
    o, p = make(sc)
    s0 = o.read_field("s0")
    for k in range(3):
        o.step(0.002, True)
    assert np.array_equal(o.read_field("s0"), s0)
    assert not np.any(o.read_field("s1"))
    for k in ("u", "v", "w"):
        assert not np.any(o.read_field(k)), k
    st = o.read_stats_raw()
    assert st[0] == 0.0 and st[1] == 0.0 and st[4] == p["t0"] and st[5] == 0.0


def test_eddy_dissipation_closed_form():
    sc = blank()
    o, p = make(sc)
    nx, ny, nz = sc.grid().dims
    T, yf, yo = 1500.0, 0.05, 0.2
    s0 = np.zeros((nz, ny, nx, 4), F32)
    s0[..., 0], s0[..., 1], s0[..., 2] = T, yf, yo
    o.write_field("s0b", s0)
    o.write_field("s1b", np.zeros_like(s0))
    o.begin_step(0.002, True)
    o.run_stage("solid")   # no solids: flags 0
    o.run_stage("react")   # curl = 0 -> nu_t = 0, no diffusion at interior cells
    o.end_step()
    out0, out1 = o.read_field("s0"), o.read_field("s1")
    dt, dx, g = 0.002, 0.05, 9.81
    s_o2 = p["s_o2"]
    tau = max(math.sqrt(2 * dx / g), 2e-3)
    dY = min(yf, yo / s_o2) * (1.0 - math.exp(-dt / tau))  # gate = 1 at 1500 K
    rho = p["rho0"] * p["t0"] / T
    q = rho * dY * p["dhc"] / dt
    c = (nz // 2, ny // 2, nx // 2)
    assert abs(out1[c][1] - q) < 2e-5 * q
    assert abs(out0[c][1] - (yf - dY)) < 1e-6
    assert abs(out0[c][2] - (yo - s_o2 * dY)) < 1e-6
    # heat release stat = sum of q dV over cells (FP64)
    hrr = (out1[..., 1].astype(np.float64) * dx ** 3).sum()
    assert abs(o.read_stats_raw()[0] - hrr) < 1e-5 * hrr


def test_no_fuel_no_heat_release():
    sc = blank()
    o, p = make(sc)
    nx, ny, nz = sc.grid().dims
    s0 = np.zeros((nz, ny, nx, 4), F32)
    s0[..., 0], s0[..., 2] = 1600.0, 0.233
    o.write_field("s0b", s0)
    o.write_field("s1b", np.zeros_like(s0))
    o.begin_step(0.002, False)
    o.run_stage("solid")
    o.run_stage("react")
    assert not np.any(o.read_field("s1")[..., 1])


def solid_scene(kind_obj):
    return pa.Scene.parse(
        "[domain]\nsize = [0.8, 0.8, 0.8]\ncell = 0.05\n" + kind_obj)


def test_liquid_pool_below_boiling_window_does_not_evaporate():
    sc = solid_scene('[[object]]\nkind = "pool"\nmaterial = "heptane"\ncenter = [0.4, 0.0, 0.4]\nradius = 0.2\ndepth = 0.05\n')
    o, _ = make(sc)
    sol0 = o.read_field("sol")
    o.begin_step(0.002, True)
    o.run_stage("solid")
    pyr = o.read_field("pyr")
    assert not np.any(pyr[..., 0])  # Ts = 293 K < Tb - 60 = 311 K
    out = o.read_field("sol_out")
    solid = sol0[..., 1] > 0
    assert np.array_equal(out[solid][:, 1], sol0[solid][:, 1])
    assert np.all(o.read_field("flags")[solid] == 1)


def test_isolated_hot_wood_cell_cools_by_closed_form():
    sc = solid_scene('[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.425, 0.425, 0.425]\nsize = [0.01, 0.01, 0.01]\n')
    p = sc.params()
    sol = sc.voxelize()
    cell = np.argwhere(sol[..., 1] > 0)
    assert len(cell) == 1
    z, y, x = cell[0]
    Ts = 400.0
    sol[z, y, x, 3] = Ts
    o = oracle.OSim(p, sol, None, 4, 0.0)
    o.begin_step(0.002, False)
    o.run_stage("solid")
    out = o.read_field("sol_out")[z, y, x]
    sig, eps, h, t0 = 5.670374e-8, 0.9, 15.0, 293.0
    qnet = eps * (sig * t0 ** 4 - sig * Ts ** 4) + h * (t0 - Ts)  # ambient gas, transparent
    cap = 480.0 * 1500.0 * min(0.005, 0.05)
    assert abs(out[3] - (Ts + 0.002 * qnet / cap)) < 1e-4
    assert out[1] == sol[z, y, x, 1] and out[2] == 0.0  # below pyrolysis window, no char
    assert o.read_field("pyr")[z, y, x][2] == 6.0  # six exposed faces


def test_burnout_clears_the_cell():
    sc = solid_scene('[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.425, 0.425, 0.425]\nsize = [0.01, 0.01, 0.01]\n')
    p = sc.params()
    p["ign_on"] = 0.0
    sol = sc.voxelize()
    z, y, x = np.argwhere(sol[..., 1] > 0)[0]
    sol[z, y, x, 1:4] = (2.0001, 0.0, 650.0)  # just above burnout, inside the pyrolysis window
    # a strong ignition flux centred on the cell drives the pyrolysis
    c = ((x + 0.5) * 0.05, (y + 0.5) * 0.05, (z + 0.5) * 0.05)
    p["ign_on"], p["ign_flux"], p["ign_r"] = 1.0, 1e7, 0.1
    p["ign_x"], p["ign_y"], p["ign_z"] = c
    o = oracle.OSim(p, sol, None, 4, 1.0)
    o.begin_step(0.002, False)
    o.run_stage("solid")
    assert list(o.read_field("sol_out")[z, y, x]) == [0.0, 0.0, 0.0, 293.0]
    assert o.read_field("flags")[z, y, x] == 0
    assert not np.any(o.read_field("pyr")[z, y, x])


def test_inert_block_is_flagged_and_unchanged():
    sc = solid_scene('[[object]]\nkind = "box"\nmaterial = "steel"\ncenter = [0.4, 0.4, 0.4]\nsize = [0.2, 0.2, 0.2]\n'
                     '[fuel]\nmaterial = "wood"\n')
    o, _ = make(sc)
    sol0 = o.read_field("sol")
    o.begin_step(0.002, True)
    o.run_stage("solid")
    o.end_step()
    solid = sol0[..., 1] > 0.5
    assert np.array_equal(o.read_field("flags") == 1, solid)
    assert np.array_equal(o.read_field("sol"), sol0)  # after end_step the output is current
    assert o.read_stats_raw()[7] == 0.0  # inert mass is not counted


def test_noise_is_deterministic_and_bounded():
    v = np.array([oracle.noise(k, i, 0) for k in range(3) for i in range(3000)])
    assert v.min() >= -1.0 and v.max() < 1.0 and abs(v.mean()) < 0.05
    assert oracle.noise(1, 17, 0) != oracle.noise(1, 17, 1)


# ------------------------------------------------------------------ renderer pack
def test_render_pack_closed_forms():
    """render.wgsl:16-37 against its physical meaning, in float64: soot volume
    fraction f_v = rho Y_s / rho_soot (ppm), smoke rho Y_m (g/m3), heat
    release q (kW/m3) with rho = rho0 T0 / max(T, 200); solids carry only T;
    the solid volume is occupancy, surface temperature, char / (fuel + char)
    and the material slot; linear face velocities average to their exact
    value at the cell centre (which also pins the x-fastest layout)."""
    sc = pa.Scene.parse(
        "[domain]\nsize = [0.8, 0.8, 0.8]\ncell = 0.05\n"
        '[[object]]\nkind = "box"\nmaterial = "wood"\ncenter = [0.4, 0.2, 0.4]\nsize = [0.3, 0.2, 0.3]\n'
        '[[object]]\nkind = "box"\nmaterial = "steel"\ncenter = [0.2, 0.6, 0.2]\nsize = [0.15, 0.1, 0.15]\n')
    o, p = make(sc)
    nx, ny, nz = sc.grid().dims
    rng = np.random.default_rng(5)
    s0 = np.stack([rng.uniform(150, 2200, (nz, ny, nx)), rng.uniform(0, 0.1, (nz, ny, nx)),
                   rng.uniform(0, 0.23, (nz, ny, nx)), rng.uniform(0, 0.02, (nz, ny, nx))], -1).astype(F32)
    s1 = np.stack([rng.uniform(0, 0.05, (nz, ny, nx)), rng.uniform(0, 5e6, (nz, ny, nx)),
                   rng.uniform(-1, 1, (nz, ny, nx)), np.zeros((nz, ny, nx))], -1).astype(F32)
    sol = o.read_field("sol")
    wood = sol[..., 0] == 1.0
    sol[..., 2] = np.where(wood, rng.uniform(0, 0.5, wood.shape), 0.0)
    sol[..., 3] = np.where(sol[..., 1] > 0, rng.uniform(300, 900, wood.shape), sol[..., 3])
    o.write_field("s0", s0)
    o.write_field("s1", s1)
    o.write_field("sol", sol)
    # flags come from the first solid pass (Sim::new marks solids with it)
    o.begin_step(0.002, False)
    o.run_stage("solid")
    o.end_step()
    z, y, x = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    o.write_field("u", (3 * x - 2 * y + z).astype(F32))
    o.write_field("v", (x + 5 * y - 4 * z).astype(F32))
    o.write_field("w", (-x + y + 7 * z).astype(F32))
    s0, s1, sol = o.read_field("s0"), o.read_field("s1"), o.read_field("sol")
    solid = o.read_field("flags") == 1
    assert solid.sum() > 50 and (~solid).sum() > 1000
    vol, sl, vel = o.render_pack().astype(np.float64)
    T = s0[..., 0].astype(np.float64)
    rho = p["rho0"] * p["t0"] / np.maximum(T, 200.0)
    fl = ~solid
    np.testing.assert_allclose(vol[fl, 0], rho[fl] * s0[fl, 3] / 1800.0 * 1e6, rtol=3e-6, atol=1e-30)
    np.testing.assert_allclose(vol[fl, 2], rho[fl] * s1[fl, 0] * 1e3, rtol=3e-6, atol=1e-30)
    np.testing.assert_allclose(vol[fl, 3], s1[fl, 1] / 1e3, rtol=2e-7)
    assert np.array_equal(vol[..., 1], T)
    assert not vol[solid][:, [0, 2, 3]].any()
    np.testing.assert_array_equal(sl[..., 0], solid.astype(np.float64))
    assert np.array_equal(sl[solid, 1], sol[solid, 3]) and np.array_equal(sl[solid, 3], sol[solid, 0])
    tot = sol[solid, 1].astype(np.float64) + sol[solid, 2]
    np.testing.assert_allclose(sl[solid, 2], sol[solid, 2] / np.maximum(tot, 1e-6), rtol=3e-7, atol=1e-30)
    assert not sl[fl].any()
    zc, yc, xc = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    np.testing.assert_array_equal(vel[..., 0], 3 * xc - 2 * (yc - 0.5) + (zc - 0.5))
    np.testing.assert_array_equal(vel[..., 1], (xc - 0.5) + 5 * yc - 4 * (zc - 0.5))
    np.testing.assert_array_equal(vel[..., 2], -(xc - 0.5) + (yc - 0.5) + 7 * zc)
    assert not vel[..., 3].any()


# ------------------------------------------------------------------ checkpoints
def test_checkpoint_round_trip_in_the_oracle(tmp_path):
    """The oracle writes and reads the checkpoint format of pyro_sim_save:
    a run continued from its own file is bit-identical to the uninterrupted
    run; damaged files are refused."""
    from harness import mini_scene, perturbed_state

    sc = mini_scene()
    a = oracle.OSim(sc.params(), sc.voxelize(), None, 3, sc.get("ignition.until"))
    st = perturbed_state(sc, 9)
    for name in ("s0", "s1", "u", "v", "w", "sol"):
        a.write_field(name, st[name])
    a.write_field(pa.level_field(0, "phi"), st["phi"])
    for k in range(2):
        a.step(0.004, False)
    ck = tmp_path / "o.ckpt"
    a.save_checkpoint(ck)
    head = oracle.read_checkpoint(ck)
    assert head["dims"] == sc.grid().dims and head["cycles"] == 3 and head["frame"] == 2
    assert head["time"] == a.time and head["ign_until"] == F32(sc.get("ignition.until"))
    b = oracle.OSim.from_checkpoint(ck)
    assert b.time == a.time and b.params_blob() == a.params_blob()
    for k in range(2):
        a.step(0.004, k == 1)
        b.step(0.004, k == 1)
    for n in ("s0", "s1", "sol", "pyr", "flags", "u", "v", "w", pa.level_field(0, "phi")):
        assert np.array_equal(a.read_field(n), b.read_field(n)), n
    assert np.array_equal(a.read_stats_raw(), b.read_stats_raw())
    raw = ck.read_bytes()
    for bad in (raw[:-4], b"PYROCKPX" + raw[8:], raw + b"\0"):
        p = tmp_path / "bad.ckpt"
        p.write_bytes(bad)
        with pytest.raises(ValueError):
            oracle.read_checkpoint(p)
