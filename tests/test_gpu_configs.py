"""GPU vs oracle at the benchmark configurations' full sizes.

- pool at 2 cm (160 x 320 x 160 = 8.19 M cells): two full steps from a hot,
  moving state (the second with statistics), and a ten-step window from
  another, every field bit-identical to the oracle and the FP64 statistics
  to 1e-9;
- the standalone pressure case at 256^3 (5 % random solid boxes, 32
  Gaussian sources): FP64-residual refinement and MGPCG to 1e-6 within 1e-6
  (relative L2 over fluid cells) of the oracle's FP64 solution; the
  reference V-cycle, which grows without bound on this hierarchy, is stopped
  as diverged long before its iteration cap;
- the bonfire at 10 cm (160 x 240 x 160 = 6.14 M cells): a ten-step window
  from a hot state bit-identical to the oracle; three steps from another,
  bit-identical, then one step whose pressure solve
  is MGPCG to 1e-4, whose reported residual an FP64 CPU residual of the same
  phi and b confirms.
"""
import numpy as np
import pytest

import oracle
import pyro_amd as pa
from conftest import SCENES
from harness import VEC4, VEL, compare, pool_scene, twin
from pyro_amd.poisson import Poisson, make_c3_problem, step_solve

pytestmark = pytest.mark.gpu


def _bitwise(o, g, names):
    res = compare(o, g, names, report=print)
    for name, (mx, rl2, nmis) in res.items():
        assert nmis == 0, f"{name}: {nmis} values differ (max {mx:.3e}, rel L2 {rl2:.3e})"


def test_c2_pool_2cm_two_steps_bitwise():
    sc = pool_scene(0.02)
    assert sc.grid().dims == (160, 320, 160)
    o, g = twin(sc, seed=0, vel_amp=4.0)
    for k in range(2):
        o.step(0.002, k == 1)
        g.step(0.002, k == 1)
    _bitwise(o, g, VEC4 + VEL + ("flags", pa.level_field(0, "phi"), pa.level_field(0, "b")))
    np.testing.assert_allclose(g.read_stats().raw, o.read_stats_raw(), rtol=1e-9, atol=1e-12)


def test_c2_pool_2cm_ten_step_window():
    """Ten full steps at the benchmark size from another moving state, the
    last with statistics: every field bit-identical to the oracle after the
    window (the reference's 10-step tolerance is 1e-3 relative L2)."""
    sc = pool_scene(0.02)
    o, g = twin(sc, seed=1, vel_amp=2.0)
    for k in range(10):
        o.step(0.002, k == 9)
        g.step(0.002, k == 9)
    _bitwise(o, g, VEC4 + VEL + ("flags", pa.level_field(0, "phi"), pa.level_field(0, "b")))
    np.testing.assert_allclose(g.read_stats().raw, o.read_stats_raw(), rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def c3():
    h = 1.0 / 256
    flags, b = make_c3_problem(256, seed=0, h=h)
    P = Poisson(flags, h)
    P.set(b=b)
    bh = P.get_rhs()
    ref, its, rel = oracle.poisson_cg64(flags, bh, h, tol=1e-10, max_iter=20000)
    assert rel <= 1e-10, rel
    return flags, b, h, P, ref


@pytest.mark.parametrize("method", ["refine", "mgpcg"])
def test_c3_256_solvers_match_fp64_oracle(c3, method):
    flags, b, h, P, ref = c3
    P.set(b=b, phi=np.zeros_like(b))
    info = P.solve(method, tol=1e-6, max_iters=200)
    print(method, info)
    assert info.rel_resid <= 1e-6 and not info.diverged
    fluid = flags == 0
    x = P.get()
    err = float(np.linalg.norm((x - ref)[fluid]) / np.linalg.norm(ref[fluid]))
    print(f"{method}: |phi - phi64| / |phi64| = {err:.3e}")
    assert err <= 1e-6


def test_c3_256_reference_vcycle_stops_as_diverged(c3):
    flags, b, h, P, ref = c3
    P.set(b=b, phi=np.zeros_like(b))
    info = P.solve("vcycle", tol=1e-6, max_iters=400)
    print(info)
    assert info.diverged and info.iterations < 400
    assert np.isfinite(info.rel_resid)


def test_c5_bonfire_10cm_ten_step_window():
    """Ten full steps of the bonfire at 10 cm (dt = 5 ms) from a hot, moving
    state, the last with statistics: every field bit-identical to the
    oracle, the statistics to 1e-9."""
    sc = pa.Scene.load(str(SCENES / "bonfire.toml"))
    o, g = twin(sc, seed=2, vel_amp=3.0)
    for k in range(10):
        o.step(0.005, k == 9)
        g.step(0.005, k == 9)
    _bitwise(o, g, VEC4 + VEL + ("flags", pa.level_field(0, "phi"), pa.level_field(0, "b")))
    np.testing.assert_allclose(g.read_stats().raw, o.read_stats_raw(), rtol=1e-9, atol=1e-12)


def test_c5_bonfire_10cm_steps_bitwise_then_mgpcg_step():
    sc = pa.Scene.load(str(SCENES / "bonfire.toml"))
    assert sc.grid().dims == (160, 240, 160)
    o, g = twin(sc, seed=1, vel_amp=3.0)
    for k in range(3):
        o.step(0.005, k == 2)
        g.step(0.005, k == 2)
    _bitwise(o, g, ("s0", "s1", "sol", "pyr", "flags", "curl", "u", "v", "w", pa.level_field(0, "phi")))
    np.testing.assert_allclose(g.read_stats().raw, o.read_stats_raw(), rtol=1e-9, atol=1e-12)
    del o
    h = sc.grid().dx
    P = Poisson(g.read_field(pa.level_field(0, "flags")), h)
    info = step_solve(g, P, 0.005, "mgpcg", tol=1e-4, max_iters=100, want_stats=True)
    print(info, "sealed components/cells", P.sealed_info())
    assert info.rel_resid <= 1e-4 and not info.diverged
    # the solver's FP64 residual, recomputed on the CPU from its phi and b
    # under the solid flags of the step it solved
    flags = g.read_field(pa.level_field(0, "flags"))
    x, bh = P.get(), P.get_rhs()
    r = oracle.residual64(flags, bh, x, h)
    fluid = flags == 0
    rel = float(np.sqrt((r[fluid] ** 2).sum() / (bh[fluid] ** 2).sum()))
    print(f"CPU FP64 residual {rel:.4e} vs reported {info.rel_resid:.4e}")
    assert abs(rel - info.rel_resid) <= 1e-3 * info.rel_resid
    g.check_finite()
