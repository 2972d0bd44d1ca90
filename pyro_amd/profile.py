"""Short driver for rocprofv3 runs: ``python -m pyro_amd.profile [--steps N] [--sweeps R]``.

Builds the pool scene at 2 cm (160 x 320 x 160; ``--scene bonfire --cell
0.10 --dt 0.005`` for the bonfire at 10 cm), runs N full steps (every
kernel of the step appears in the trace), then R launches of the level-0
red-black sweep and R launches of the plain copy kernel over 64 Mi floats
(the calibration for FETCH_SIZE / WRITE_SIZE counters).  Optional phases:
``--solves K`` K MGPCG solves to 1e-4 of the next step's pressure problem
(warm start); ``--unet F`` F U-Net forward passes on the same grid;
``--train T`` T training steps of the U-Net on 8 pairs of the pool at 4 cm;
``--noise A`` the velocity noise amplitude of the start state (m/s).
Prints one line per phase so a profiled run is never silent.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def main() -> None:
    import pyro_amd as pa

    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", default="pool", help="scene file in pyro_amd/scenes (without .toml)")
    ap.add_argument("--cell", type=float, default=0.02)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--copies", type=int, default=10)
    ap.add_argument("--solves", type=int, default=0)
    ap.add_argument("--unet", type=int, default=0)
    ap.add_argument("--train", type=int, default=0)
    ap.add_argument("--noise", type=float, default=1e-3,
                    help="amplitude (m/s) of the uniform velocity noise the run starts from")
    args = ap.parse_args()
    s = pa.Scene.load(str(ROOT / "pyro_amd" / "scenes" / f"{args.scene}.toml"))
    s.set("domain.cell", args.cell)
    sim = pa.Sim(s.params(), s.voxelize(), None, 4, s.get("ignition.until"))
    sim.add_velocity_noise(args.noise, 0)
    for k in range(args.steps):
        sim.step(args.dt, k == args.steps - 1)
    sim.sync()
    print(f"[profile] {args.steps} steps done", flush=True)
    ms = sim.time_kernel("mg_smooth", 0, args.sweeps)
    print(f"[profile] {args.sweeps} level-0 sweeps: {ms * 1e3:.1f} us each", flush=True)
    ms = pa.sim.debug_copy(64 << 20, args.copies)
    print(f"[profile] {args.copies} copies of 256 MiB: {ms * 1e3:.1f} us each", flush=True)
    nx, ny, nz = s.grid().dims
    if args.solves:
        from pyro_amd.poisson import Poisson

        sim.begin_step(args.dt, False)
        for stg in pa.STAGES[: pa.STAGES.index("divergence") + 1]:
            sim.run_stage(stg)
        P = Poisson(sim.read_field(pa.level_field(0, "flags")), s.grid().dx)
        for k in range(args.solves):
            P.set_from_sim(sim, warm=True, flags=k == 0)
            info = P.solve("mgpcg", 1e-4, 100)
        print(f"[profile] {args.solves} MGPCG solves to 1e-4: {info}", flush=True)
    if args.unet or args.train:
        from pyro_amd.net import Net

        net = Net(seed=0)
        if args.unet:
            ms = net.time_forward((nz, ny, nx), args.unet)
            print(f"[profile] {args.unet} U-Net forwards on {nx}x{ny}x{nz}: {ms:.3f} ms each", flush=True)
        if args.train:
            from pyro_amd.learned import collect_pool_pairs

            collect_pool_pairs(net, frames=8)
            losses = net.train(args.train, lr=1e-3)
            print(f"[profile] {args.train} training steps at 4 cm: last loss {losses[-1]:.4f}", flush=True)


if __name__ == "__main__":
    main()
