"""Pins of the U-Net oracle (oracle/unet_ref.py), CPU only.

- the hand-written backward equals torch autograd of the float64 network;
- tap orientation from first principles: a delta input reproduces the
  mirrored kernel at the documented offsets (plain NumPy, no torch conv);
- adjoint identities of pool/unpool, up/unup, conv/conv_adj and the weight
  gradient's linearity;
- with rounding, every stored tensor is BF16-representable.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
from oracle import unet_ref as U  # noqa: E402


def rand_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for _, co, ci in U.LAYERS:
        parts += [rng.normal(0, scale * np.sqrt(2.0 / (27 * ci)), co * ci * 27), rng.normal(0, 0.05, co)]
    parts += [rng.normal(0, 0.25, 16), rng.normal(0, 0.05, 1)]
    p = np.concatenate(parts)
    assert p.size == U.N_PARAMS
    return p


def rand_x0(shape, seed=1):
    rng = np.random.default_rng(seed)
    x = np.zeros(shape + (2,))
    m = (rng.random(shape) > 0.1).astype(float)
    x[..., 0] = rng.normal(size=shape) * m
    x[..., 1] = m
    return x


@pytest.mark.parametrize("shape", [(6, 10, 8), (5, 7, 9)])
def test_backward_equals_autograd(shape):
    p = rand_params(0)
    x0 = rand_x0(shape)
    gy = np.random.default_rng(2).normal(size=shape)
    g1 = U.backward(p, x0, gy, rounding=False)
    g2 = U.autograd_backward(p, x0, gy)
    rel = np.linalg.norm(g1 - g2) / np.linalg.norm(g2)
    assert rel < 1e-12, rel
    # every layer's block is nonzero (a dropped term would show as zeros)
    d = U.split(g1)
    for k, v in d.items():
        assert np.abs(v).max() > 0, k


def test_delta_input_gives_mirrored_kernel():
    """out[v] = sum_t W[t] in[v + off(t)], off(t) = (kx-1, ky-1, kz-1), t = kz*9+ky*3+kx"""
    rng = np.random.default_rng(3)
    W = rng.integers(-3, 4, size=(16, 16, 27)).astype(float)
    b = np.zeros((5, 6, 7, 16))
    z0, y0, x0, ci = 2, 3, 3, 5
    b[z0, y0, x0, ci] = 1.0
    out = U.conv_layer(None, b, W, rounding=False)
    expect = np.zeros_like(out)
    for t in range(27):
        kx, ky, kz = t % 3, (t // 3) % 3, t // 9
        expect[z0 - (kz - 1), y0 - (ky - 1), x0 - (kx - 1), :] = W[:, ci, t]
    assert np.array_equal(out, expect)


def test_pool_up_adjoints_and_constants():
    rng = np.random.default_rng(4)
    for shape in [(4, 6, 8), (5, 7, 3)]:
        x = torch.as_tensor(rng.normal(size=(1, 3) + shape))
        cs = tuple((s + 1) // 2 for s in shape)
        g = torch.as_tensor(rng.normal(size=(1, 3) + cs))
        assert np.isclose(float((U.pool(x) * g).sum()), float((x * U.unpool(g, shape)).sum()), rtol=1e-12)
        assert np.isclose(float((U.up(g, shape) * x).sum()), float((g * U.unup(x)).sum()), rtol=1e-12)
        c = torch.full((1, 2) + shape, 3.25, dtype=torch.float64)
        assert torch.all(U.pool(c) == 3.25)


def test_conv_adjoint_and_dw_linearity():
    rng = np.random.default_rng(5)
    shape = (5, 6, 7)
    for ca, cb, co in [(0, 16, 32), (32, 16, 16), (64, 32, 32)]:
        cs = tuple((s + 1) // 2 for s in shape)
        a = rng.normal(size=cs + (ca,)) if ca else None
        b = rng.normal(size=shape + (cb,))
        W = rng.normal(size=(co, ca + cb, 27))
        g = rng.normal(size=shape + (co,))
        y = U.conv_layer(a, b, W, rounding=False)
        # <conv(x), g> = <x, conv_adj(g)> with x = [up(a), b]
        xfull = b if a is None else np.concatenate([U._n(U.up(U._t(a), shape)), b], -1)
        lhs = float((y * g).sum())
        rhs = float((xfull * U.conv_layer_adj(g, W, rounding=False)).sum())
        assert np.isclose(lhs, rhs, rtol=1e-11)
        # <dW, V> = <conv(x; V), g>
        dW, db = U.conv_layer_dw(a, b, g, co, rounding=False)
        V = rng.normal(size=W.shape)
        assert np.isclose(float((dW * V).sum()), float((U.conv_layer(a, b, V, rounding=False) * g).sum()), rtol=1e-11)
        assert np.allclose(db, g.reshape(-1, co).sum(0))


def test_rounded_activations_are_bf16():
    p = rand_params(7)
    x0 = rand_x0((6, 8, 8), 8)
    _, A = U.forward(p, x0, rounding=True, keep=True)
    for k in ("x", "a0", "s0", "p1", "a1", "s1", "p2", "a2", "c2", "d1", "d0"):
        v = A[k].to(torch.float32)
        assert torch.equal(v.to(torch.bfloat16).to(torch.float32), v), k


def test_param_count_matches_library():
    import pyro_amd.net as N

    try:
        n = N.param_count()
    except OSError:
        pytest.skip("native library not built")
    assert n == U.N_PARAMS
