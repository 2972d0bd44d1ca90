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


def _vcycles_match_oracle(dims, levels, blocks=12):
    nx, ny, nz = dims
    rng = np.random.default_rng(4)
    flags = np.zeros((nz, ny, nx), np.uint32)
    for _ in range(blocks):
        x0, y0, z0 = rng.integers(0, nx - 8), rng.integers(0, ny - 8), rng.integers(0, nz - 4)
        flags[z0:z0 + 3, y0:y0 + 6, x0:x0 + 7] = 1
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = (np.sin(0.11 * x) * np.cos(0.07 * y) + 0.3 * np.sin(0.5 * z)).astype(np.float64)
    b[flags == 1] = 0.0
    h = 0.05
    P = Poisson(flags, h)
    assert P.levels[0] == levels
    phi0 = (rng.normal(size=flags.shape) * 0.01).astype(np.float32)
    P.set(b=b.astype(np.float32).astype(np.float64), phi=phi0.astype(np.float64))
    P.set_sealed_projection("never")
    info = P.solve("vcycle", tol=0.0, max_iters=2)
    assert info.vcycles == 2
    m = oracle.OMg((nx, ny, nz), h, 4, flags)
    assert m.levels == levels
    m.set(0, phi0, b.astype(np.float32))
    m.vcycles(2)
    ref, _, _ = m.get(0)
    got = P.get().astype(np.float32)
    bad = got.view(np.uint32) != ref.view(np.uint32)
    assert not bad.any(), f"{int(bad.sum())} values differ"


@pytest.mark.parametrize("dims,levels", [((160, 160, 14), 1), ((320, 320, 32), 3)])
def test_coarsest_level_beyond_lds_bitwise(dims, levels):
    """A coarsest level too large for LDS (here a single 160x160x14 level, and
    80x80x8 under two finer ones) is solved from global memory by one
    workgroup; its 80 SOR iterations and the V-cycles around them stay
    bit-identical to the oracle."""
    _vcycles_match_oracle(dims, levels)


@pytest.mark.parametrize("dims,levels", [((15, 52, 20), 1), ((30, 104, 40), 2), ((160, 320, 160), 4),
                                         ((14, 43, 20), 1), ((15, 52, 21), 1), ((12, 60, 21), 1)])
def test_coarse_column_layouts_bitwise(dims, levels):
    """The coarsest-level column solve in both of its layouts, with 21 and 22
    register slots per column (levels 20 and 21 cells deep).  780 columns
    (15x52) and 800 (20x40, the pool at 2 cm) run as 12 waves of whole
    columns with the leftover columns' cells taken one per thread and colour;
    602 columns (14x43) and 720 (12x60) run as waves of columns, which issue
    fewer slots on the busiest SIMD.  Random initial phi and solid blocks;
    bit-identical to the oracle's V-cycles."""
    _vcycles_match_oracle(dims, levels, blocks=6)


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


def _projected(flags, b):
    """b minus its mean over every sealed fluid component (scipy labels)."""
    from pyro_amd.poisson import fluid_components

    lab, n, open_ = fluid_components(flags)
    out = b.copy()
    sealed = 0
    cells = 0
    for k in range(1, n + 1):
        if not open_[k]:
            m = lab == k
            out[m] -= b[m].mean()
            sealed += 1
            cells += int(m.sum())
    return out, sealed, cells


def _pocket_geometry():
    nz, ny, nx = 36, 32, 40
    f = np.zeros((nz, ny, nx), np.uint32)
    f[4:16, 0:12, 4:18] = 1
    f[5:15, 0:11, 5:17] = 0          # hollow box on the ground: sealed pocket (ground is Neumann)
    f[20:30, 10:20, 20:32] = 1
    f[21:29, 11:19, 21:31] = 0       # floating hollow box: sealed
    f[24:26, 13:15, 24:26] = 1       # a solid inside it (pocket stays one component)
    f[2:5, 20:23, 30:33] = 1
    f[3, 21, 31] = 0                 # one trapped fluid cell
    f[10:14, 20:24, 0:6] = 1
    f[11:13, 21:23, 0:5] = 0         # box open to the x = 0 side: open
    return f


def test_sealed_components_match_scipy_and_make_b_consistent():
    flags = _pocket_geometry()
    rng = np.random.default_rng(11)
    b = rng.normal(size=flags.shape)
    b[flags == 1] = 0.0
    ref, n_sealed, n_cells = _projected(flags, b)
    assert n_sealed == 3
    P = Poisson(flags, 0.05)
    assert P.sealed_info() == (n_sealed, n_cells)
    P.set(b=b)
    info = P.solve("mgpcg", tol=1e-9, max_iters=300)
    assert info.rel_resid <= 1e-9, info
    got = P.get_rhs()
    fl = flags == 0
    assert np.abs(got - ref)[fl].max() <= 1e-12 * np.abs(ref).max()
    # a second solve of the same right-hand side does not project again
    P.solve("mgpcg", tol=1e-9, max_iters=300)
    assert np.array_equal(P.get_rhs(), got)


def test_reference_vcycles_keep_b_and_projection_modes():
    flags = _pocket_geometry()
    rng = np.random.default_rng(12)
    b = rng.normal(size=flags.shape)
    b[flags == 1] = 0.0
    P = Poisson(flags, 0.05)
    P.set(b=b)
    P.solve("vcycle", tol=0.0, max_iters=2)
    assert np.array_equal(P.get_rhs(), b)  # the reference method solves b as given
    P.set_sealed_projection("never")
    P.set(b=b)
    info = P.solve("mgpcg", tol=1e-9, max_iters=30)
    assert np.array_equal(P.get_rhs(), b)
    assert info.rel_resid > 1e-3  # inconsistent: the sealed part of b cannot be removed
    P.set_sealed_projection("all")
    P.set(b=b)
    P.solve("vcycle", tol=0.0, max_iters=1)
    assert not np.array_equal(P.get_rhs(), b)


@pytest.mark.parametrize("frac", [0.3, 0.55, 0.62])
def test_sealed_labels_on_random_solids(frac):
    # near the fluid percolation threshold: hundreds of pockets of up to ~25 cells
    rng = np.random.default_rng(13)
    flags = (rng.random((30, 34, 38)) < frac).astype(np.uint32)
    b = rng.normal(size=flags.shape)
    b[flags == 1] = 0.0
    ref, n_sealed, n_cells = _projected(flags, b)
    assert n_sealed >= 20
    P = Poisson(flags, 0.05)
    assert P.sealed_info() == (n_sealed, n_cells)
    P.set(b=b)
    P.solve("mgpcg", tol=0.0, max_iters=1)  # the projection happens at the start of a solve
    fl = flags == 0
    assert np.abs(P.get_rhs() - ref)[fl].max() <= 1e-12 * np.abs(ref).max()
