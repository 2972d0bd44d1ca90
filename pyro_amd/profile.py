"""Short driver for rocprofv3 runs: ``python -m pyro_amd.profile [--steps N] [--sweeps R]``.

Builds the pool scene at 2 cm (160 x 320 x 160), runs N full steps (every
kernel of the step appears in the trace), then R launches of the level-0
red-black sweep and R launches of the plain copy kernel over 64 Mi floats
(the calibration for FETCH_SIZE / WRITE_SIZE counters).  Prints one line per
phase so a profiled run is never silent.
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
    ap.add_argument("--cell", type=float, default=0.02)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--copies", type=int, default=10)
    args = ap.parse_args()
    s = pa.Scene.load(str(ROOT / "pyro_amd" / "scenes" / "pool.toml"))
    s.set("domain.cell", args.cell)
    sim = pa.Sim(s.params(), s.voxelize(), None, 4, s.get("ignition.until"))
    sim.add_velocity_noise(1e-3, 0)
    for k in range(args.steps):
        sim.step(0.002, k == args.steps - 1)
    sim.sync()
    print(f"[profile] {args.steps} steps done", flush=True)
    ms = sim.time_kernel("mg_smooth", 0, args.sweeps)
    print(f"[profile] {args.sweeps} level-0 sweeps: {ms * 1e3:.1f} us each", flush=True)
    ms = pa.sim.debug_copy(64 << 20, args.copies)
    print(f"[profile] {args.copies} copies of 256 MiB: {ms * 1e3:.1f} us each", flush=True)


if __name__ == "__main__":
    main()
