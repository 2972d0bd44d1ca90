"""CPU restatement of the learned pressure preconditioner (the U-Net of
pyro_amd/csrc/net.hpp) -- test infrastructure only, never the product path.

float64 torch on the CPU, with BF16 rounding (round to nearest even) at
exactly the points where the GPU stores BF16: the input, every conv weight,
every activation, every pooled tensor and every activation gradient that
feeds a convolution.  With ``rounding=False`` it is the exact float64
network, whose hand-written backward the tests check against torch
autograd.  No reference counterpart exists (the reference has no learned
part; it runs fixed V-cycles, sim.rs:405).

Arrays at the interface are NumPy, channels-last (nz, ny, nx, c); weights
W[cout][cin][27] with tap = kz*9 + ky*3 + kx at offset (kx-1, ky-1, kz-1).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [("0a", 16, 2), ("0b", 16, 16), ("1a", 32, 16), ("1b", 32, 32), ("2a", 64, 32),
          ("2b", 64, 64), ("1c", 32, 96), ("0c", 16, 48)]
N_PARAMS = sum(co * ci * 27 + co for _, co, ci in LAYERS) + 17


def split(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, co, ci in LAYERS:
        out["W" + name] = np.asarray(flat[o:o + co * ci * 27], np.float64).reshape(co, ci, 27)
        o += co * ci * 27
        out["b" + name] = np.asarray(flat[o:o + co], np.float64)
        o += co
    out["wh"] = np.asarray(flat[o:o + 16], np.float64)
    out["bh"] = np.asarray(flat[o + 16:o + 17], np.float64)
    return out


def join(p: dict) -> np.ndarray:
    parts = []
    for name, _, _ in LAYERS:
        parts += [p["W" + name].ravel(), p["b" + name].ravel()]
    parts += [p["wh"].ravel(), p["bh"].ravel()]
    return np.concatenate(parts)


class _R:
    def __init__(self, rounding: bool):
        self.on = rounding

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        if not self.on:
            return x
        return x.to(torch.float32).to(torch.bfloat16).to(torch.float64)


def _t(a: np.ndarray) -> torch.Tensor:
    """(nz, ny, nx, c) numpy -> (1, c, nz, ny, nx) float64 torch"""
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(np.asarray(a, np.float64), -1, 0)))[None]


def _n(t: torch.Tensor) -> np.ndarray:
    return np.moveaxis(t[0].detach().numpy(), 0, -1).copy()


def _w(W) -> torch.Tensor:
    W = torch.as_tensor(np.asarray(W, np.float64))
    return W.reshape(W.shape[0], W.shape[1], 3, 3, 3)


def up(a: torch.Tensor, shape) -> torch.Tensor:
    """nearest upsampling of (1, c, cz, cy, cx) to the fine grid `shape`"""
    u = a.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    return u[:, :, :shape[0], :shape[1], :shape[2]]


def _pad_even(x: torch.Tensor) -> torch.Tensor:
    nz, ny, nx = x.shape[2:]
    return F.pad(x, (0, nx % 2, 0, ny % 2, 0, nz % 2))


def _block_sum(x: torch.Tensor) -> torch.Tensor:
    x = _pad_even(x)
    n, c, z, y, w = x.shape
    return x.reshape(n, c, z // 2, 2, y // 2, 2, w // 2, 2).sum(dim=(3, 5, 7))


def child_count(shape) -> torch.Tensor:
    return _block_sum(torch.ones((1, 1) + tuple(shape), dtype=torch.float64))


def pool(x: torch.Tensor) -> torch.Tensor:
    """average over the (up to 8) children inside the grid"""
    return _block_sum(x) / child_count(x.shape[2:])


def unpool(g: torch.Tensor, shape) -> torch.Tensor:
    """adjoint of pool: each child receives g(parent) / count(parent)"""
    return up(g / child_count(shape), shape)


def unup(g: torch.Tensor) -> torch.Tensor:
    """adjoint of up: sum over the children"""
    return _block_sum(g)


def conv(x: torch.Tensor, W: torch.Tensor, b=None) -> torch.Tensor:
    return F.conv3d(x, W, None if b is None else torch.as_tensor(np.asarray(b, np.float64)), padding=1)


def conv_adj(g: torch.Tensor, W: torch.Tensor, cin: int) -> torch.Tensor:
    return torch.nn.grad.conv3d_input((1, cin) + tuple(g.shape[2:]), W, g, padding=1)


def conv_dw(x: torch.Tensor, g: torch.Tensor, wshape) -> torch.Tensor:
    return torch.nn.grad.conv3d_weight(x, tuple(wshape), g, padding=1)


def relu(x):
    return torch.clamp(x, min=0.0)


def forward(params: np.ndarray, x0: np.ndarray, rounding: bool = True, keep: bool = False):
    """y (nz, ny, nx) for input x0 (nz, ny, nx, 2); with keep, also the activations."""
    R = _R(rounding)
    p = split(params)
    W = {k: R(_w(p[k])) for k in p if k.startswith("W")}
    b = {k: torch.as_tensor(p[k]) for k in p if k.startswith("b")}
    x = R(_t(x0))
    g0 = x.shape[2:]
    a0 = R(relu(conv(x, W["W0a"], b["b0a"])))
    s0 = R(relu(conv(a0, W["W0b"], b["b0b"])))
    p1 = R(pool(s0))
    g1 = p1.shape[2:]
    a1 = R(relu(conv(p1, W["W1a"], b["b1a"])))
    s1 = R(relu(conv(a1, W["W1b"], b["b1b"])))
    p2 = R(pool(s1))
    a2 = R(relu(conv(p2, W["W2a"], b["b2a"])))
    c2 = R(relu(conv(a2, W["W2b"], b["b2b"])))
    d1 = R(relu(conv(torch.cat([up(c2, g1), s1], 1), W["W1c"], b["b1c"])))
    d0 = R(relu(conv(torch.cat([up(d1, g0), s0], 1), W["W0c"], b["b0c"])))
    wh = torch.as_tensor(p["wh"]).reshape(1, 16, 1, 1, 1)
    y = (d0 * wh).sum(1, keepdim=True) + float(p["bh"][0])
    yn = y[0, 0].numpy().copy()
    if not keep:
        return yn
    acts = dict(x=x, a0=a0, s0=s0, p1=p1, a1=a1, s1=s1, p2=p2, a2=a2, c2=c2, d1=d1, d0=d0, W=W, g0=g0, g1=g1)
    return yn, acts


def backward(params: np.ndarray, x0: np.ndarray, gy: np.ndarray, rounding: bool = True) -> np.ndarray:
    """Flat gradient of sum(gy * y(x0)) in the parameter order (same rounding
    points as the GPU backward)."""
    R = _R(rounding)
    p = split(params)
    _, A = forward(params, x0, rounding, keep=True)
    W, g0, g1 = A["W"], A["g0"], A["g1"]
    g = torch.as_tensor(np.asarray(gy, np.float64))[None, None]
    grads = {}
    wh = torch.as_tensor(p["wh"]).reshape(1, 16, 1, 1, 1)
    grads["wh"] = (g * A["d0"]).sum(dim=(0, 2, 3, 4)).numpy()
    grads["bh"] = np.array([g.sum().item()])
    pos = lambda t: (t > 0).to(torch.float64)  # noqa: E731

    def layer(name, x_in, gpre):
        grads["W" + name] = conv_dw(x_in, gpre, W["W" + name].shape).reshape(W["W" + name].shape[0], -1, 27).numpy()
        grads["b" + name] = gpre.sum(dim=(0, 2, 3, 4)).numpy()

    gd0 = R(g * wh) * pos(A["d0"])
    layer("0c", torch.cat([up(A["d1"], g0), A["s0"]], 1), gd0)
    gcat0 = R(conv_adj(gd0, W["W0c"], 48))
    gd1 = R(unup(gcat0[:, :32])) * pos(A["d1"])
    layer("1c", torch.cat([up(A["c2"], g1), A["s1"]], 1), gd1)
    gcat1 = R(conv_adj(gd1, W["W1c"], 96))
    gc2 = R(unup(gcat1[:, :64])) * pos(A["c2"])
    layer("2b", A["a2"], gc2)
    ga2 = R(conv_adj(gc2, W["W2b"], 64)) * pos(A["a2"])
    layer("2a", A["p2"], ga2)
    gp2 = R(conv_adj(ga2, W["W2a"], 32))
    gs1 = R(gcat1[:, 64:] + unpool(gp2, g1)) * pos(A["s1"])
    layer("1b", A["a1"], gs1)
    ga1 = R(conv_adj(gs1, W["W1b"], 32)) * pos(A["a1"])
    layer("1a", A["p1"], ga1)
    gp1 = R(conv_adj(ga1, W["W1a"], 16))
    gs0 = R(gcat0[:, 32:] + unpool(gp1, g0)) * pos(A["s0"])
    layer("0b", A["a0"], gs0)
    ga0 = R(conv_adj(gs0, W["W0b"], 16)) * pos(A["a0"])
    layer("0a", A["x"], ga0)
    return join(grads)


def autograd_backward(params: np.ndarray, x0: np.ndarray, gy: np.ndarray) -> np.ndarray:
    """torch autograd of the unrounded float64 network (the pin of `backward`)."""
    p = split(params)
    leaves = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    x = _t(x0)
    Wt = {k: leaves[k].reshape(leaves[k].shape[0], leaves[k].shape[1], 3, 3, 3) for k in leaves if k.startswith("W")}
    g0 = x.shape[2:]
    a0 = relu(F.conv3d(x, Wt["W0a"], leaves["b0a"], padding=1))
    s0 = relu(F.conv3d(a0, Wt["W0b"], leaves["b0b"], padding=1))
    p1 = pool(s0)
    g1 = p1.shape[2:]
    a1 = relu(F.conv3d(p1, Wt["W1a"], leaves["b1a"], padding=1))
    s1 = relu(F.conv3d(a1, Wt["W1b"], leaves["b1b"], padding=1))
    p2 = pool(s1)
    a2 = relu(F.conv3d(p2, Wt["W2a"], leaves["b2a"], padding=1))
    c2 = relu(F.conv3d(a2, Wt["W2b"], leaves["b2b"], padding=1))
    d1 = relu(F.conv3d(torch.cat([up(c2, g1), s1], 1), Wt["W1c"], leaves["b1c"], padding=1))
    d0 = relu(F.conv3d(torch.cat([up(d1, g0), s0], 1), Wt["W0c"], leaves["b0c"], padding=1))
    y = (d0 * leaves["wh"].reshape(1, 16, 1, 1, 1)).sum(1) + leaves["bh"]
    (y[0] * torch.as_tensor(np.asarray(gy, np.float64))).sum().backward()
    return join({k: v.grad.numpy() for k, v in leaves.items()})


# single layers (for the per-kernel tests)
def conv_layer(a, b, W, bias=None, relu_on=False, rounding=True):
    """forward of one conv: input [up(a), b] (a may be None), NumPy channels-last"""
    R = _R(rounding)
    xb = R(_t(b))
    x = xb if a is None else torch.cat([up(R(_t(a)), xb.shape[2:]), xb], 1)
    o = conv(x, R(_w(W)), bias)
    if relu_on:
        o = relu(o)
    return _n(R(o))


def conv_layer_adj(g, W, rounding=True):
    R = _R(rounding)
    Wt = R(_w(W))
    return _n(R(conv_adj(R(_t(g)), Wt, Wt.shape[1])))


def conv_layer_dw(a, b, g, cout, rounding=True):
    R = _R(rounding)
    xb = R(_t(b))
    x = xb if a is None else torch.cat([up(R(_t(a)), xb.shape[2:]), xb], 1)
    gg = R(_t(g))
    dW = conv_dw(x, gg, (cout, x.shape[1], 3, 3, 3)).reshape(cout, x.shape[1], 27).numpy()
    return dW, gg.sum(dim=(0, 2, 3, 4)).numpy()


def loss(y: np.ndarray, t: np.ndarray, mask: np.ndarray) -> float:
    """relative squared error over fluid cells"""
    m = mask > 0.5
    return float(((y - t)[m] ** 2).sum() / (t[m] ** 2).sum())
