"""Training pipeline of the learned pressure preconditioner.

``collect_pool_pairs``: run the pool scene (default 4 cm cells,
80 x 160 x 80, dt = 4 ms) and after every ``every``-th step store that step's
pressure problem, solved from zero by MGPCG to 1e-6, as a training pair on
the device (right-hand side as the divergence stage left it).
``train_pool_net``: collect, then train the U-Net with Adam on the GPU,
holding out the last 5 % of the pairs.  ``save``/``load`` keep the FP32
parameters and the output scale in an .npz of plain arrays (loaded with
allow_pickle=False).
"""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np

from .net import Net
from .poisson import Poisson

WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_pool_4cm.npz"


def pool_sim(cell: float, seed: int = 0):
    from .scene import Scene
    from .sim import Sim

    s = Scene.load(str(Path(__file__).resolve().parent / "scenes" / "pool.toml"))
    s.set("domain.cell", cell)
    sim = Sim(s.params(), s.voxelize(), None, 4, s.get("ignition.until"))
    sim.add_velocity_noise(1e-3, seed)
    return s, sim


def collect_pool_pairs(net: Net, cell: float = 0.04, frames: int = 2000, every: int = 1, dt: float = 0.004,
                       tol: float = 1e-6, seed: int = 0, log=None) -> dict:
    from .sim import level_field

    scene, sim = pool_sim(cell, seed)
    P = Poisson(sim.read_field(level_field(0, "flags")), scene.grid().dx)
    t0 = time.perf_counter()
    its = []
    for f in range(frames):
        for _ in range(every):
            sim.step(dt, False)
        info = net.collect(sim, P, tol)
        its.append(info.iterations)
        if log and (f % 250 == 0 or f == frames - 1):
            log(f"[learned] pair {f + 1}/{frames}: MGPCG {info.iterations} it to {info.rel_resid:.2e}, "
                f"{time.perf_counter() - t0:.1f} s")
    P.close()
    sim.close()
    return {"grid": list(scene.grid().dims), "cell": cell, "frames": frames, "every": every, "dt": dt,
            "tol": tol, "mean_mgpcg_iterations": float(np.mean(its)), "collect_seconds": time.perf_counter() - t0}


def train(net: Net, steps: int, lr: float, first: int, count: int, seed: int = 0, log=None) -> dict:
    """Adam over pairs [first, first + count): lr for 3/4 of the steps, lr/4 after."""
    t0 = time.perf_counter()
    losses = []
    done = 0
    chunk = 500
    while done < steps:
        k = min(chunk, steps - done)
        rate = lr if done < steps * 3 // 4 else lr / 4
        l = net.train(k, lr=rate, seed=seed * 7919 + done, first=first, count=count)
        losses.extend(l.tolist())
        done += k
        if log:
            log(f"[learned] step {done}/{steps}: loss {np.mean(l[-50:]):.4f}, {time.perf_counter() - t0:.1f} s")
    secs = time.perf_counter() - t0
    return {"steps": steps, "lr": lr, "train_seconds": secs, "ms_per_step": 1e3 * secs / max(steps, 1),
            "loss_first_50": float(np.mean(losses[:50])), "loss_last_50": float(np.mean(losses[-50:]))}


def train_pool_net(cell: float = 0.04, frames: int = 2000, steps: int = 4000, lr: float = 1e-3, every: int = 1,
                   dt: float = 0.004, seed: int = 0, log=None) -> tuple[Net, dict]:
    net = Net(seed=seed)
    rep = collect_pool_pairs(net, cell, frames, every, dt, seed=seed, log=log)
    n = net.dataset_size()
    hold = max(1, n // 20)
    rep.update(train(net, steps, lr, 0, n - hold, seed, log))
    rep["train_loss"] = net.eval(0, min(n - hold, 100))
    rep["heldout_loss"] = net.eval(n - hold, hold)
    rep["heldout_pairs"] = hold
    rep["output_scale"] = net.output_scale
    return net, rep


def save(net: Net, path: Path = WEIGHTS) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=net.params(), output_scale=np.float32(net.output_scale))


def load(path: Path = WEIGHTS) -> Net:
    z = np.load(Path(path), allow_pickle=False)
    net = Net(seed=0)
    net.set_params(z["params"])
    net.output_scale = float(z["output_scale"])
    return net


def trained(path: Path = WEIGHTS, frames: int = 2000, steps: int = 3000, log=None) -> Net:
    """The network saved at `path`; when the file is absent it is trained
    first (pool at 4 cm, `frames` pairs, `steps` Adam steps: about 20 s on the
    GPU with the defaults) and saved there."""
    path = Path(path)
    if path.exists():
        return load(path)
    net, rep = train_pool_net(frames=frames, steps=steps, lr=1e-3, log=log)
    save(net, path)
    if log:
        log(f"U-Net trained: held-out loss {rep['heldout_loss']:.4f}, saved to {path}")
    return net


def main(argv=None) -> None:
    """``python -m pyro_amd.learned [--out PATH] [--frames N] [--steps N]``:
    train the preconditioner on the pool at 4 cm and save its weights."""
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=3000)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    net, rep = train_pool_net(frames=args.frames, steps=args.steps, lr=1e-3, log=print)
    save(net, Path(args.out))
    print(f"held-out loss {rep['heldout_loss']:.4f} (train {rep['train_loss']:.4f}), "
          f"{time.perf_counter() - t0:.1f} s; saved to {args.out}", flush=True)


if __name__ == "__main__":
    main()
