"""The reference's simulation-only frame loop (main.rs:55-283 with --sim-only),
driving the MI355X solver: ``python -m pyro_amd.run SCENE [options]``.

Per frame (main.rs:139-147): the CFL substep count from the previous frame's
maximum speed, ``n = max(ceil(frame_dt / (cfl * dx / max(u_max, 1))), 1)``
equal substeps, statistics on the last one; a report line per frame; from
``max(preroll, 0.4 * total)`` on, running averages of HRR, radiation,
pyrolysis, burning area and the per-layer HRR profile; at the end the
averages, the heights below which 50/90/99 % of the heat is released and, for
pool objects, Heskestad's mean flame height ``0.235 Q^0.4 - 1.02 D`` (Q in kW)
with the puffing frequency ``1.5 / sqrt(D)`` (main.rs:233-280).  The
dt/substep arithmetic is FP32 as in the reference.

Additions: ``--solver`` runs each step's pressure solve to ``--tol`` with
MGPCG / refinement / learned MGPCG instead of the fixed V-cycle count;
``--load-state`` / ``--save-state`` restore and write a checkpoint (the
``--preroll`` reuse the reference lacks); ``--slice`` writes the centre
z-slice panels of main.rs:201-230 as a PNG.
"""
from __future__ import annotations

import argparse
import math
import struct
import sys
import time
import zlib
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

f32 = np.float32


def substeps(frame_dt, cfl, dx, max_u):
    """(n, dt) of one frame, in FP32 as main.rs:140-142."""
    dt_cfl = f32(cfl) * f32(dx) / max(f32(max_u), f32(1.0))
    n = max(math.ceil(float(f32(frame_dt) / dt_cfl)), 1)
    return n, f32(f32(frame_dt) / f32(n))


def frame_count(total, fps) -> int:
    """(total * fps).round() as u32, FP32 with half away from zero"""
    v = float(f32(total) * f32(fps))
    return int(math.floor(v + 0.5)) if v >= 0 else 0


def heat_heights(layer_avg: np.ndarray, dx: float, fracs=(0.5, 0.9, 0.99)):
    """(j + 1) dx at the first layer where the cumulative HRR reaches frac (main.rs:238-247)"""
    tot = float(np.sum(layer_avg))
    out = []
    for fr in fracs:
        acc, h = 0.0, 0.0
        for j, v in enumerate(layer_avg):
            acc += float(v)
            if acc >= fr * tot:
                h = (j + 1.0) * dx
                break
        out.append(h)
    return out


def heskestad(q_w: float, radius: float):
    d = 2.0 * radius
    return 0.235 * (q_w / 1e3) ** 0.4 - 1.02 * d, 1.5 / math.sqrt(d)


def _png(path: Path, rgb: np.ndarray) -> None:
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    path.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
                     + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def slice_image(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """five panels of the centre z-slice: T, Y_F, Y_O2, soot volume fraction, q (main.rs:205-230)"""
    nz, ny, nx, _ = a.shape
    k = nz // 2
    c, d = a[k], b[k]
    rho = 1.204 * 293.0 / np.maximum(c[..., 0], 200.0)
    vals = [(c[..., 0] - 300.0) / 1700.0, c[..., 1] / 0.3, c[..., 2] / 0.233,
            rho * c[..., 3] / 1800.0 * 1e6 / 3.0, d[..., 1] / 5e6]
    pw = nx * 5 + 8
    img = np.zeros((ny, pw, 3), np.uint8)
    for pi, v in enumerate(vals):
        v = np.clip(v, 0.0, 1.0)
        col = np.stack([np.minimum(v * 3.0, 1.0), np.clip(v * 3.0 - 1.0, 0.0, 1.0),
                        np.clip(v * 3.0 - 2.0, 0.0, 1.0)], -1)
        img[:, pi * (nx + 2):pi * (nx + 2) + nx] = (col * 255.0).astype(np.uint8)[::-1]
    return img


def main(argv=None) -> int:
    run_scene(argv)
    return 0


def run_scene(argv=None) -> dict:
    """Run the frame loop for the command line `argv`; returns the summary
    printed at the end (frames, seconds, and when averaged: hrr_kw, h50, h90,
    h99, heskestad_m, puffing_hz, burning_rate, pyrolysis_kg_s, radiative)."""
    import pyro_amd as pa

    ap = argparse.ArgumentParser(prog="pyro_amd.run", description="simulation-only frame loop of main.rs")
    ap.add_argument("scene")
    ap.add_argument("--duration", type=float)
    ap.add_argument("--cell", type=float)
    ap.add_argument("--preroll", type=float)
    ap.add_argument("--still", type=float)
    ap.add_argument("--out", default=None)
    ap.add_argument("--slice", action="store_true")
    ap.add_argument("--solver", choices=["vcycles", "mgpcg", "refine", "learned"], default="vcycles")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--weights", default=None, help="U-Net .npz for --solver learned")
    ap.add_argument("--load-state")
    ap.add_argument("--save-state")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args(argv)

    scene = pa.Scene.load(args.scene)
    if args.duration is not None:
        scene.set("sim.duration", args.duration)
    if args.cell is not None:
        scene.set("domain.cell", args.cell)
    if args.preroll is not None:
        scene.set("output.preroll", args.preroll)
    grid = scene.grid()
    nx, ny, nz = grid.dims
    print(f"grid {nx}x{ny}x{nz} ({nx * ny * nz / 1e6:.1f} M cells, {grid.dx * 100:.1f} cm)", flush=True)
    # voxelized on the device: the same bits as Scene::voxelize, without the
    # O(cells x shapes) host loop
    sim = pa.Sim.new(scene, device_voxelize=True)
    if args.load_state:
        sim.load(args.load_state)
    P = None
    if args.solver != "vcycles":
        from pyro_amd.poisson import Poisson, step_solve

        P = Poisson(sim.read_field(pa.level_field(0, "flags")), grid.dx)
        if args.solver == "learned":
            from pyro_amd import learned

            # the given weights, or the default file (trained and saved on first use)
            net = learned.load(args.weights) if args.weights else learned.trained(log=print)
            P.attach_net(net, "first")

    fps = f32(scene.get("sim.fps"))
    frame_dt = f32(f32(1.0) / fps)
    if args.still is not None:
        preroll, total = f32(args.still), f32(f32(args.still) + frame_dt * f32(0.5))
    else:
        preroll = f32(scene.get("output.preroll"))
        total = f32(preroll + f32(scene.get("sim.duration")))
    frames = frame_count(total, fps)
    cfl = scene.get("sim.cfl")
    max_u = f32(2.0)
    avg_from = max(float(preroll), 0.4 * float(total))
    n_avg, hrr_avg, rad_avg, pyro_avg, area_avg = 0, 0.0, 0.0, 0.0, 0.0
    layer_avg = np.zeros(ny)
    t_start = time.perf_counter()
    iters = []
    start_time = sim.time
    if args.load_state:  # continue from the checkpoint's time to the same end
        frames = frame_count(max(f32(total) - f32(start_time), f32(0.0)), fps)
    for f in range(frames):
        n, dt = substeps(frame_dt, cfl, grid.dx, max_u)
        for s in range(n):
            if P is None:
                sim.step(float(dt), s == n - 1)
            else:
                info = step_solve(sim, P, float(dt), args.solver, args.tol, 100, s == n - 1)
                iters.append(info.iterations)
        st = sim.read_stats()
        max_u = max(f32(st.max_u), f32(1.0))
        if sim.time >= avg_from:
            n_avg += 1
            hrr_avg += st.hrr
            rad_avg += st.rad
            pyro_avg += st.pyro
            area_avg += st.area
            layer_avg += st.layers[:ny]
        if not args.quiet:
            rad = 100.0 * st.rad / st.hrr if st.hrr > 0.0 else 0.0
            print(f"t={sim.time:5.2f}s substeps {n:2d} HRR {st.hrr / 1e3:7.1f} kW radiative {rad:4.0f}% "
                  f"pyrolysis {st.pyro:6.3f} kg/s Tmax {st.max_t:4.0f} K umax {st.max_u:4.1f} "
                  f"flame {st.flame_height(grid.dx):4.2f} m soot {st.soot * 1e3:5.1f} g fuel {st.solid:.1f} kg "
                  f"resid {st.resid:.0e}  [{time.perf_counter() - t_start:.0f}s]", flush=True)
    if args.slice:
        out = Path(args.out or ".")
        out.mkdir(parents=True, exist_ok=True)
        a, b = sim.read_scalars()
        _png(out / "slice.png", slice_image(a, b))
        print(f"slice: {out / 'slice.png'}")
    if args.save_state:
        sim.save(args.save_state)
        print(f"state saved: {args.save_state} (t = {sim.time:.3f} s)")
    summary = {"frames": frames, "seconds": time.perf_counter() - t_start}
    if iters:
        summary["mean_solver_iterations"] = float(np.mean(iters))
    if n_avg > 0:
        q = hrr_avg / n_avg
        h50, h90, h99 = heat_heights(layer_avg, grid.dx)
        print(f"--- averages (t >= {avg_from:.1f} s, {n_avg} frames) ---")
        print(f"HRR {q / 1e3:.0f} kW  radiative fraction {rad_avg / max(hrr_avg, 1.0):.2f}  "
              f"pyrolysis {pyro_avg / n_avg:.4f} kg/s  burning area {area_avg / n_avg:.3f} m2  -> "
              f"{pyro_avg / max(area_avg, 1e-9):.4f} kg/m2s")
        print(f"cumulative heat release 50% height {h50:.2f} m  90% {h90:.2f} m  99% {h99:.2f} m")
        summary.update(hrr_kw=q / 1e3, h50=h50, h90=h90, h99=h99, pyrolysis_kg_s=pyro_avg / n_avg,
                       burning_rate=pyro_avg / max(area_avg, 1e-9), radiative=rad_avg / max(hrr_avg, 1.0))
        for i in range(int(scene.get("objects"))):
            if scene.get(f"object.{i}.pool") > 0.5:
                L, fq = heskestad(q, scene.get(f"object.{i}.radius"))
                print(f"Heskestad mean flame height (at this HRR) {L:.2f} m, puffing frequency {fq:.2f} Hz")
                summary.update(heskestad_m=L, puffing_hz=fq)
    print(summary, flush=True)
    return summary


if __name__ == "__main__":
    sys.exit(main())
