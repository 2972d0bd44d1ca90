"""Pressure solvers on the GPU: the reference V-cycle (bit-equal to the
oracle), FP64-residual refinement and flexible MGPCG (checked against the
FP64 CPU solution of the oracle)."""
import numpy as np
import pytest

import oracle
from pyro_amd.poisson import Poisson, make_c3_problem

pytestmark = pytest.mark.gpu


def rel_err(a, b, mask):
    return float(np.linalg.norm((a - b)[mask]) / np.linalg.norm(b[mask]))


def test_reference_vcycles_bitwise_equal_oracle():
    flags, b = make_c3_problem(48, seed=1, h=0.05)
    P = Poisson(flags, 0.05)
    assert P.levels[0] == 3
    rng = np.random.default_rng(2)
    phi0 = (rng.normal(size=flags.shape) * 0.01).astype(np.float32)
    P.set(b=b.astype(np.float32).astype(np.float64), phi=phi0.astype(np.float64))
    info = P.solve("vcycle", tol=0.0, max_iters=3)
    assert info.vcycles == 3
    m = oracle.OMg(flags.shape[::-1], 0.05, 4, flags)
    assert m.levels == P.levels[0]
    m.set(0, phi0, b.astype(np.float32))
    m.vcycles(3)
    ref, _, _ = m.get(0)
    got = P.get()
    assert np.array_equal(got.astype(np.float32), ref)


@pytest.mark.parametrize("method", ["mgpcg", "refine"])
def test_solvers_reach_1e_minus_8_and_match_fp64_oracle(method):
    n, h = 64, 1.0 / 64
    flags, b = make_c3_problem(n, seed=3, h=h)
    P = Poisson(flags, h)
    P.set(b=b)
    info = P.solve(method, tol=1e-8, max_iters=200)
    assert info.rel_resid <= 1e-8, info
    ref, its, rel = oracle.poisson_cg64(flags, b, h, tol=1e-12, max_iter=20000)
    assert rel <= 1e-11
    fluid = flags == 0
    assert rel_err(P.get(), ref, fluid) < 1e-6
    print(method, info)


def test_mgpcg_beats_reference_vcycles_to_1e_minus_4():
    n, h = 64, 1.0 / 64
    flags, b = make_c3_problem(n, seed=4, h=h)
    P = Poisson(flags, h)
    P.set(b=b)
    plain = P.solve("vcycle", tol=1e-4, max_iters=400)
    P.set(phi=np.zeros_like(b))
    pcg = P.solve("mgpcg", tol=1e-4, max_iters=100)
    print("plain", plain, "mgpcg", pcg)
    assert pcg.rel_resid <= 1e-4
    assert pcg.vcycles * 3 < max(plain.vcycles, 1) or plain.rel_resid > 1e-4


def test_mgpcg_on_enclosed_pockets():
    """Fluid pockets sealed by solids (pure Neumann) with a zero-mean RHS."""
    n, h = 40, 0.05
    flags = np.zeros((n, n, n), np.uint32)
    flags[10:30, 0:20, 10:30] = 1
    flags[12:28, 0:18, 12:28] = 0  # hollow box on the ground: a sealed pocket
    flags[18:22, 6:10, 18:22] = 1
    rng = np.random.default_rng(6)
    b = rng.normal(size=flags.shape)
    b[flags == 1] = 0
    from pyro_amd.poisson import fluid_components

    lab, nc, open_ = fluid_components(flags)
    assert (~open_[1:]).sum() >= 1
    for k in range(1, nc + 1):
        if not open_[k]:
            b[lab == k] -= b[lab == k].mean()
    P = Poisson(flags, h)
    P.set(b=b)
    info = P.solve("mgpcg", tol=1e-7, max_iters=300)
    assert info.rel_resid <= 1e-7, info
