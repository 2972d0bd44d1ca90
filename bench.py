"""Benchmark: cell-updates/s of the full fire-solver step on one MI355X.

Workload (``config.workload``): scenes/pool.toml (1 m n-heptane pool) at 2 cm
cells = 160 x 320 x 160 = 8.19 M cells, FP32, 4-level multigrid with 4
V-cycles per step, dt = 2 ms, every step one hipGraph replay; stats
(FP64 reductions) on every 10th step as the reference's frame loop requests
them once per frame (main.rs:143-145).  Synthetic data: the voxelized scene
plus U[-1e-3, 1e-3] m/s velocity noise (seed 0).  ``value`` = cells x steps /
wall seconds, inputs resident in HBM, timed between two device syncs.

Extra fields:
  roofline      -- the kernel with the largest share of the step time (average
                   launch time x launches per step), in the contract's form:
                   algorithmic bytes per launch (the reference layout's
                   compulsory bytes per cell, below, x the cells one launch
                   processes) / average launch time (HIP events on the sim
                   stream, repeated launches), against the 8 TB/s HBM peak;
                   ``traffic`` is the HBM bytes per launch measured by
                   rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/
                   pool_2cm_kernel_traffic.json, scaled by the copy-kernel
                   calibration), when present.
  roofline_kernels -- the same for every kernel of the step, sorted by share:
                   react 84, forces 76, divergence 36, project 32 B per
                   cell; the fused launches solid + curl 52 + 28 (solid's
                   bytes are the reference's dense pass; this one is
                   sparse), adv_vel1 + adv_s1 24 + 76 and adv_vel2 + adv_s2
                   40 + 108 B per cell; the level-0 red-black sweep 24 B per cell
                   (12 B per colour pass, two passes per launch, 16 launches
                   per step); restriction 13.5 and prolongation 13 B per
                   fine cell; the coarsest-level solve 80 x 24 B per coarse
                   cell (80 SOR iterations in one launch, 4 per step).  Timed
                   on a scratch simulation restored from a checkpoint of the
                   burning pool, so the timed runs are not disturbed.
  cpu_baseline  -- the CPU restatement in oracle/ (kind "port") timed on this
                   host on a bounded sample of the same workload, on the
                   host's CPU share (OMP_NUM_THREADS, 16 on the MI355X host,
                   ``host_nproc`` beside it) and on one core.
  pool_4cm      -- the small pool (80 x 160 x 80) on the GPU, 20 timed steps.
  stage_ms      -- per-stage time of one step (HIP events, direct launches;
                   the stages singly, and the fused launches the step
                   uses: solid_curl, adv_pass1 and adv_pass2).
  pressure      -- time to a relative residual of 1e-4 (sqrt of sum r^2 over
                   sum b^2 on fluid cells, FP64) on the next step's pressure
                   problem of this run, warm-started as the step would be,
                   by the reference V-cycles, FP64-residual refinement,
                   MGPCG and MGPCG with the learned U-Net (trained first, on
                   the GPU, from 4 cm pool frames); and the standalone
                   256^3 case (random solid boxes, 32 Gaussian sources):
                   V-cycles to 1e-4, FP64-residual refinement and MGPCG to
                   1e-6, with the error against an FP64 solution; the
                   bonfire at 10 cm stepped from one checkpointed mid-fire
                   state with each pressure method to 1e-4, and the FP64
                   mass balance of its fuel over those steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# one red-black colour pass of mg.wgsl moves 12 B per cell in the reference
# layout (phi 4, flags 4, half of b 2, half of phi written 2); one launch of
# the fused sweep runs both passes.  The HBM bytes it really moves (both
# passes fused, u8 codes) are the rocprof `traffic`, reported beside it.
SWEEP_BYTES_PER_CELL = 24.0
# renderer pack: read s0, s1, sol 48 + flags 4 + three velocity faces 12, write three half4 24
RENDER_PACK_BYTES_PER_CELL = 88.0
TRAFFIC_FILE = ROOT / "profiles" / "pool_2cm_kernel_traffic.json"
# compulsory bytes per cell of one launch in the reference layout, and
# launches per step, of the single-kernel stages (kernel names as rocprof
# shows them for a step without statistics)
STAGE_KERNELS = {
    "solid": ("k_solid<false>", 52.0), "curl": ("k_curl<false>", 28.0), "adv_vel1": ("k_adv_vel1", 24.0),
    "adv_vel2": ("k_adv_vel2", 40.0), "adv_s1": ("k_adv_s1", 76.0), "adv_s2": ("k_adv_s2", 108.0),
    "react": ("k_react<false>", 84.0), "forces": ("k_forces", 76.0), "divergence": ("k_divergence", 36.0),
    "project": ("k_project", 32.0),
}
RESTRICT_BYTES_PER_FINE_CELL = 13.5
PROLONG_BYTES_PER_FINE_CELL = 13.0
COARSE_ITERS = 80
COARSE_BYTES_PER_CELL_ITER = 24.0
# The coarse solve's own bound is the VALU issue rate of its one CU (one
# wave64 VALU instruction per cycle at 2.4 GHz): the counter run of
# profiles/pool_2cm_kernel_counters.md reads 25,550 VALU instructions per wave
# over 12 waves for the 20x40x20 level of the pool at 2 cm.
COARSE_VALU_PER_WAVE, COARSE_WAVES, CU_CLOCK_HZ = 25550, 12, 2.4e9
# U-Net FLOPs per level-0 voxel: 2*27*cin*cout per conv at its level's
# resolution (1/8 per halving) plus the 1x1x1 head: level 0 2*27*(2*16 + 16*16
# + 48*16) + 2*16, level 1 2*27*(16*32 + 32*32 + 96*32)/8, level 2
# 2*27*(32*64 + 64*64)/64
UNET_FLOP_PER_VOXEL = 2 * 27 * (2 * 16 + 16 * 16 + 48 * 16) + 2 * 16 + 2 * 27 * (16 * 32 + 32 * 32 + 96 * 32) / 8 \
    + 2 * 27 * (32 * 64 + 64 * 64) / 64


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_pool(cell: float):
    import pyro_amd as pa

    s = pa.Scene.load(str(ROOT / "pyro_amd" / "scenes" / "pool.toml"))
    s.set("domain.cell", cell)
    return s


def gpu_run(args):
    import pyro_amd as pa

    scene = make_pool(args.cell)
    p = scene.params()
    sol = scene.voxelize()
    nx, ny, nz = scene.grid().dims
    cells = nx * ny * nz
    log(f"[bench] pool {nx}x{ny}x{nz} = {cells / 1e6:.2f} M cells, dt={args.dt}, cycles=4")
    sim = pa.Sim(p, sol, None, 4, scene.get("ignition.until"))
    sim.add_velocity_noise(1e-3, 0)
    for k in range(args.warmup):
        sim.step(args.dt, k % 10 == 9)
    sim.sync()
    log(f"[bench] warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step(args.dt, k % 10 == 9)
    sim.sync()
    wall = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps: {wall:.3f} s")
    sim.check_finite()
    st = sim.read_stats()
    # achievable streaming bandwidth on this device (plain copy of 2 x 256 MiB)
    copy_ms = pa.sim.debug_copy(64 << 20, 20)
    copy_gbs = 2 * (64 << 20) * 4 / (copy_ms * 1e-3) / 1e9
    log(f"[bench] plain copy: {copy_gbs:.0f} GB/s")
    burning = None
    if args.burn_time > 0:
        # the same run continued to t = burn_time (the pool burns from about
        # 1.2 s), then timed again: state-dependent work (pyrolysis, the
        # solid pass, combustion) at full fire
        t0 = time.perf_counter()
        n_pre = 0
        while sim.time < args.burn_time:
            sim.step(args.dt, False)
            n_pre += 1
            if n_pre % 500 == 0:
                sim.sync()
                log(f"[bench] pool at t = {sim.time:.2f} s ({time.perf_counter() - t0:.1f} s)")
        sim.step(args.dt, True)
        sim.sync()
        t1 = time.perf_counter()
        for k in range(args.steps):
            sim.step(args.dt, k % 10 == 9)
        sim.sync()
        wall_b = time.perf_counter() - t1
        sb = sim.read_stats()
        burning = {"t_start_s": round(float(sim.time) - args.steps * args.dt, 3), "steps": args.steps,
                   "ms_per_step": 1e3 * wall_b / args.steps, "cell_updates_per_s": cells * args.steps / wall_b,
                   "hrr_W": sb.hrr, "pyrolysis_kg_s": sb.pyro, "max_T": sb.max_t, "max_u": sb.max_u,
                   "step_resid": sb.resid}
        log(f"[bench] pool from t = {burning['t_start_s']} s: {burning['ms_per_step']:.3f} ms/step, "
            f"HRR {sb.hrr / 1e3:.0f} kW")
    # per-kernel times on a scratch copy of this state (the runs above and
    # below are not disturbed)
    import tempfile

    ck = Path(tempfile.mkdtemp()) / "pool.ckpt"
    sim.save(ck)
    kern, stage_ms = kernel_rooflines(scene, ck, args)
    ck.unlink()
    handoff = render_handoff(sim, args) if args.burn_time > 0 else None
    net, learn = (None, None) if args.no_pressure else learned_net(args)
    frame = None if args.no_pressure else pressure_frame(sim, args, p["dx"], net)
    if learn is not None:
        learn["forward_ms_pool_2cm"] = net.time_forward((nz, ny, nx), 10)
        learn["forward_tflops_pool_2cm"] = UNET_FLOP_PER_VOXEL * cells / (learn["forward_ms_pool_2cm"] * 1e-3) / 1e12
        log(f"[bench] U-Net forward on {nx}x{ny}x{nz}: {learn['forward_ms_pool_2cm']:.3f} ms, "
            f"{learn['forward_tflops_pool_2cm']:.0f} TFLOP/s")
    return dict(cells=cells, wall=wall, stats=st, stage_ms=stage_ms, kernels=kern, dims=(nx, ny, nz),
                copy_gbs=copy_gbs, frame=frame, learned=learn, net=net, burning=burning, handoff=handoff)


def load_traffic():
    """{rocprof kernel name: HBM bytes per launch} from the committed
    FETCH_SIZE / WRITE_SIZE summary, when present."""
    if not TRAFFIC_FILE.exists():
        return {}
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
        return {k: v["hbm_bytes_per_launch"] for k, v in d.get("kernels", {}).items()}
    except (ValueError, OSError, KeyError, TypeError):
        return {}


def kernel_rooflines(scene, ck, args):
    """Average launch time of every kernel of the step, on a scratch
    simulation restored from checkpoint `ck`, against its compulsory bytes.
    Each stage is launched once to warm up, then `reps` times between two
    HIP events on the simulation's stream (the average includes the short
    gaps between back-to-back launches)."""
    import pyro_amd as pa

    tsim = pa.Sim(scene.params(), scene.voxelize(), None, 4, scene.get("ignition.until"))
    tsim.load(ck)
    nx, ny, nz = tsim.dims
    cells = nx * ny * nz
    reps = 5
    stage_ms = {}
    tsim.begin_step(args.dt, False)
    for stg in pa.STAGES:
        if stg == "resid":
            continue
        tsim.time_stage(stg, 1)
        stage_ms[stg] = tsim.time_stage(stg, 1 if stg == "vcycles" else reps)
    tsim.end_step()
    tsim.sync()
    # the step runs the four advection stages as two fused launches
    # and solid with curl as one
    pass_ms = {}
    for kn in ("adv_pass1", "adv_pass2", "solid_curl"):
        tsim.time_kernel(kn, 0, 1)  # warm-up
        pass_ms[kn] = tsim.time_kernel(kn, 0, reps)
    passes = {"k_adv_pass1": ("adv_vel1 24 + adv_s1 76 B per cell, one launch", 100.0, pass_ms["adv_pass1"]),
              "k_adv_pass2": ("adv_vel2 40 + adv_s2 108 B per cell, one launch", 148.0, pass_ms["adv_pass2"]),
              "k_solid_curl<false>": ("solid 52 (the reference's dense pass; sparse here: only solid cells move "
                                      "data, so frac can exceed 1) + curl 28 B per cell, one launch", 80.0,
                                      pass_ms["solid_curl"])}
    stage_ms.update(pass_ms)
    sweep = tsim.time_kernel("mg_smooth", 0, 200)
    restrict0 = tsim.time_kernel("mg_restrict", 0, 50)
    prolong0 = tsim.time_kernel("mg_prolong", 0, 50)
    coarse = tsim.time_kernel("mg_coarse", 0, 20)
    cd = tsim.level_dims(tsim.levels - 1)
    ncoarse = cd[0] * cd[1] * cd[2]
    tsim.close()
    traffic = load_traffic()
    ents = []

    def add(kernel, note, algo_bytes, avg_ms, per_step):
        ach = algo_bytes / (avg_ms * 1e-3) / 1e9
        tr = traffic.get(kernel)
        ents.append({"kernel": kernel, "note": note, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": tr,
                     "traffic_GBs": tr / (avg_ms * 1e-3) / 1e9 if tr else None,
                     "algo_bytes_per_launch": algo_bytes, "avg_launch_us": avg_ms * 1e3,
                     "launches_per_step": per_step, "ms_per_step": avg_ms * per_step})

    for kname, (note, bpc, ms) in passes.items():
        add(kname, note, bpc * cells, ms, 1)
    for stg, (kname, bpc) in STAGE_KERNELS.items():
        if stg.startswith("adv_") or stg in ("solid", "curl"):
            continue  # in the fused launches above
        add(kname, f"stage {stg}, {bpc:g} B per cell", bpc * cells, stage_ms[stg], 1)
    add("k_mg_sweep<true, true, true>", "level-0 red-black sweep, both colours, 24 B per cell", 24.0 * cells,
        sweep, 16)
    add("k_mg_restrict<true>" if nx % 4 == 0 else "k_mg_restrict<false>", "level 0 -> 1, 13.5 B per fine cell",
        RESTRICT_BYTES_PER_FINE_CELL * cells, restrict0, 4)
    add("k_mg_prolong", "level 1 -> 0, 13 B per fine cell", PROLONG_BYTES_PER_FINE_CELL * cells, prolong0, 4)
    coarse_name = next((k for k in traffic if k.startswith("k_mg_coarse_z")), "k_mg_coarse_z")
    add(coarse_name, f"coarsest level {cd[0]}x{cd[1]}x{cd[2]}: {COARSE_ITERS} SOR iterations in one "
        "workgroup (a dependency chain on one CU), 24 B per cell and iteration",
        COARSE_ITERS * COARSE_BYTES_PER_CELL_ITER * ncoarse, coarse, 4)
    if (cd[0], cd[1], cd[2]) == (20, 40, 20):
        # share of the launch its one CU spends issuing VALU instructions
        ents[-1]["one_cu_valu_issue_frac"] = COARSE_VALU_PER_WAVE * COARSE_WAVES / CU_CLOCK_HZ / (coarse * 1e-3)
    ents.sort(key=lambda e: -e["ms_per_step"])
    for e in ents:
        log(f"[bench] {e['kernel']}: {e['avg_launch_us']:.1f} us x {e['launches_per_step']} = "
            f"{e['ms_per_step']:.3f} ms/step, {e['achieved']:.0f} GB/s (frac {e['frac']:.3f})")
    return ents, {k: round(v, 4) for k, v in stage_ms.items()}


def pool_4cm(args):
    """The small pool (80 x 160 x 80 at 4 cm): 20 hipGraph steps after 10."""
    import pyro_amd as pa

    scene = make_pool(0.04)
    nx, ny, nz = scene.grid().dims
    sim = pa.Sim(scene.params(), scene.voxelize(), None, 4, scene.get("ignition.until"))
    sim.add_velocity_noise(1e-3, 0)
    for k in range(10):
        sim.step(args.dt, k % 10 == 9)
    sim.sync()
    t0 = time.perf_counter()
    for k in range(20):
        sim.step(args.dt, k % 10 == 9)
    sim.sync()
    wall = time.perf_counter() - t0
    sim.close()
    out = {"grid": [nx, ny, nz], "steps": 20, "ms_per_step": 1e3 * wall / 20,
           "cell_updates_per_s": nx * ny * nz * 20 / wall}
    log(f"[bench] pool 4 cm: {out}")
    return out


def render_handoff(sim, args):
    """The renderer hand-off on the burning pool (render.rs:255-277,
    render.wgsl:16-37): the pack kernel against its compulsory bytes (s0, s1,
    sol 48 B + flag 4 B + 3 velocity faces 12 B read, three half4 volumes
    24 B written per cell, reference layout); one blocking pack + copy into
    pinned memory; the
    raw host staging of the seven fields the reference binds
    (pyro_sim_read_field, pageable); and steps with a hand-off every 21 steps
    (one 24 fps frame at dt = 2 ms) against the same steps without, the copy
    of frame k running beside the steps of frame k+1; both from the same
    checkpointed state, alternated twice."""
    import tempfile

    import pyro_amd as pa

    nx, ny, nz = sim.dims
    cells = nx * ny * nz
    bufs = [pa.PinnedBuffer(sim.dims), pa.PinnedBuffer(sim.dims)]
    sim.render_pack()
    sim.time_kernel("render_pack", 0, 5)
    pack_ms = sim.time_kernel("render_pack", 0, 50)
    pack_gbs = RENDER_PACK_BYTES_PER_CELL * cells / (pack_ms * 1e-3) / 1e9
    ts = []
    for k in range(5):
        sim.sync()
        t = time.perf_counter()
        sim.render_pack(bufs[0])
        sim.render_wait()
        ts.append(time.perf_counter() - t)
    handoff_ms = 1e3 * min(ts)
    raw = ["s0", "s1", "sol", "flags", "u", "v", "w"]
    t = time.perf_counter()
    raw_bytes = sum(sim.read_field(f).nbytes for f in raw)
    raw_ms = 1e3 * (time.perf_counter() - t)
    frames, per = 6, 21

    def run(with_handoff):
        sim.sync()
        t0 = time.perf_counter()
        for f in range(frames):
            for k in range(per):
                sim.step(args.dt, False)
            if with_handoff:
                sim.render_wait()  # frame f-1 has landed: the renderer may take it
                sim.render_pack(bufs[f % 2])
        sim.sync()
        sim.render_wait()
        return 1e3 * (time.perf_counter() - t0) / (frames * per)

    ck = Path(tempfile.mkdtemp()) / "handoff.ckpt"
    sim.save(ck)
    times = {False: [], True: []}
    for rep in range(2):
        for mode in (False, True):
            sim.load(ck)
            times[mode].append(run(mode))
    sim.load(ck)
    ck.unlink()
    plain, overlapped = float(np.mean(times[False])), float(np.mean(times[True]))
    for bf in bufs:
        bf.close()
    out = {"volumes_MB": 24 * cells / 1e6, "pack_kernel_us": pack_ms * 1e3, "pack_kernel_GBs": pack_gbs,
           "pack_kernel_hbm_frac": pack_gbs / HBM_PEAK_GBS, "pack_and_pinned_copy_ms": handoff_ms,
           "pcie_GBs": 24 * cells / (handoff_ms * 1e-3) / 1e9, "raw_staging_MB": raw_bytes / 1e6,
           "raw_staging_ms": raw_ms, "steps_per_frame": per, "ms_per_step_plain": plain,
           "ms_per_step_with_handoff": overlapped,
           "handoff_cost_per_frame_ms": (overlapped - plain) * per}
    log(f"[bench] render hand-off: {out}")
    return out


def _info(i):
    return {"ms": round(i.ms, 3), "vcycles": i.vcycles, "iterations": i.iterations, "rel_resid": i.rel_resid}


def learned_net(args):
    """The U-Net trained on the GPU by this library's kernels: pairs from the
    pool at 4 cm (80 x 160 x 80) solved by MGPCG to 1e-6, then Adam."""
    from pyro_amd import learned

    t0 = time.perf_counter()
    net, rep = learned.train_pool_net(frames=args.train_pairs, steps=args.train_steps, lr=1e-3, log=log)
    rep["total_seconds"] = time.perf_counter() - t0
    rep["flop_per_voxel"] = UNET_FLOP_PER_VOXEL
    log(f"[bench] U-Net trained: held-out loss {rep['heldout_loss']:.4f} in {rep['total_seconds']:.1f} s")
    return net, rep


def pressure_frame(sim, args, dx, net=None):
    """Time to a 1e-4 relative residual on the pressure problem of the next
    step of the benchmark run (the burning pool at t = --burn-time unless that
    is 0): its right-hand side, solid flags and the warm start the step
    itself would use (phi of the previous step).  Methods:
    the reference V-cycles (plain MG as the reference runs it), FP64-residual
    refinement (plain MG iteration on the consistent hierarchy), MGPCG, and
    MGPCG with the learned network (every preconditioner application, or
    the first only; on the fine grid, or on level 1 of the preconditioner
    hierarchy: "coarse")."""
    import pyro_amd as pa
    from pyro_amd.poisson import Poisson

    sim.begin_step(args.dt, False)
    for stg in pa.STAGES[: pa.STAGES.index("divergence") + 1]:
        sim.run_stage(stg)
    P = Poisson(sim.read_field(pa.sim.level_field(0, "flags")), dx)
    P.set_from_sim(sim, warm=True)
    out = {"tol": 1e-4, "start": "warm (phi of the previous step)", "t_s": round(float(sim.time), 3),
           "initial_rel_resid": P.residual()}
    runs = [("vcycle", None, 400), ("refine", None, 100), ("mgpcg", None, 100)]
    if net is not None:
        runs += [("learned", "every", 100), ("learned", "first", 100), ("learned", "coarse-every", 100),
                 ("learned", "coarse-first", 100)]
    for m, mode, cap in runs:
        if mode:
            P.attach_net(net, mode)
        key = m if mode is None else f"{m}_{mode.replace('-', '_')}"
        P.set_from_sim(sim, warm=True, flags=False)
        P.solve(m, 1e-4, min(cap, 20))  # first call: allocation / code-object warm-up
        P.set_from_sim(sim, warm=True, flags=False)
        out[key] = _info(P.solve(m, 1e-4, cap))
        out[key]["reached"] = out[key]["rel_resid"] <= 1e-4
        log(f"[bench] frame pressure {key}: {out[key]}")
    for stg in pa.STAGES[pa.STAGES.index("divergence") + 1:]:
        sim.run_stage(stg)
    sim.end_step()
    P.close()
    return out


def pressure_c3(args):
    """The standalone 256^3 pressure case: plain V-cycles to 1e-4, FP32
    V-cycles in FP64-residual refinement and MGPCG to 1e-6; error of each
    against an FP64 solution (MGPCG to 1e-12 on the GPU, and the oracle's
    FP64 Jacobi-PCG on the host with --c3-cpu)."""
    from pyro_amd.poisson import Poisson, make_c3_problem

    n = args.c3_n
    t0 = time.perf_counter()
    flags, b = make_c3_problem(n, seed=0, h=1.0 / n)
    log(f"[bench] c3 problem {n}^3 built in {time.perf_counter() - t0:.1f} s")
    fluid = flags == 0
    P = Poisson(flags, 1.0 / n)
    zero = np.zeros_like(b)
    P.set(b=b, phi=zero)
    P.solve("mgpcg", 1e-12, 400)
    phi64 = P.get()
    out = {"grid": [n, n, n], "solid_frac": float(flags.mean()), "levels_ref_pc": list(P.levels)}
    for m, tol, cap in (("vcycle", 1e-4, 400), ("refine", 1e-4, 200), ("mgpcg", 1e-4, 200),
                         ("refine", 1e-6, 200), ("mgpcg", 1e-6, 200)):
        P.set(b=b, phi=zero)
        info = P.solve(m, tol, cap)
        d = _info(info)
        d["err_vs_fp64"] = float(np.linalg.norm((P.get() - phi64)[fluid]) / np.linalg.norm(phi64[fluid]))
        out[f"{m}_{tol:g}"] = d
        log(f"[bench] c3 {m} to {tol:g}: {d}")
    if args.c3_cpu:
        import oracle

        nproc = os.cpu_count() or 1
        oracle.set_threads(max(1, min(int(os.environ.get("OMP_NUM_THREADS", nproc)), nproc)))
        t0 = time.perf_counter()
        ref, its, rel = oracle.poisson_cg64(flags, b, 1.0 / n, tol=1e-10, max_iter=20000)
        P.set(b=b, phi=zero)
        P.solve("refine", 1e-6, 200)
        err = float(np.linalg.norm((P.get() - ref)[fluid]) / np.linalg.norm(ref[fluid]))
        out["cpu_fp64_oracle"] = {"iterations": its, "rel_resid": rel, "seconds": time.perf_counter() - t0,
                                  "refine_1e-6_err": err}
        log(f"[bench] c3 vs CPU FP64 oracle: {out['cpu_fp64_oracle']}")
    P.close()
    return out


def cpu_baseline(args, dims):
    """The oracle on this host: full steps of the same workload (bounded sample)."""
    import oracle

    nproc = os.cpu_count() or 1
    # this host's CPU share: OMP_NUM_THREADS where the host sets it (16 per
    # GPU on the MI355X hosts, whatever nproc shows), else every core
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", nproc)), nproc))
    oracle.set_threads(threads)
    scene = make_pool(args.cell)
    o = oracle.OSim(scene.params(), scene.voxelize(), None, 4, scene.get("ignition.until"))
    o.add_velocity_noise(1e-3, 0)
    o.step(args.dt, False)  # first-touch / warm
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(args.dt, (n % 10) == 9)
        n += 1
        el = time.perf_counter() - t0
        log(f"[bench] cpu baseline step {n}: {el:.1f} s")
        if el > args.cpu_seconds or n >= 50:
            break
    cells = dims[0] * dims[1] * dims[2]
    out = dict(value=cells * n / el, unit="cell-updates/s", cores=threads, kind="port", host_nproc=nproc,
               sample=f"{n} full steps (after 1 warm step) of the same {dims[0]}x{dims[1]}x{dims[2]} pool "
                      f"workload on the oracle (OpenMP, {threads} threads: the host's CPU share, "
                      f"nproc {nproc})")
    # the small pool (4 cm) on all threads and on one
    del o
    c1 = make_pool(0.04)
    d1 = c1.grid().dims
    n1cells = d1[0] * d1[1] * d1[2]
    extra = {}
    for th, budget in ((threads, 0.3 * args.cpu_seconds), (1, 0.5 * args.cpu_seconds)):
        oracle.set_threads(th)
        o = oracle.OSim(c1.params(), c1.voxelize(), None, 4, c1.get("ignition.until"))
        o.add_velocity_noise(1e-3, 0)
        o.step(args.dt, False)
        k = 0
        t0 = time.perf_counter()
        while True:
            o.step(args.dt, False)
            k += 1
            e = time.perf_counter() - t0
            if e > budget or k >= 50:
                break
        log(f"[bench] cpu baseline {d1[0]}x{d1[1]}x{d1[2]}, {th} thread(s): {k} steps in {e:.1f} s")
        extra[f"pool_4cm_{th}_threads"] = {"value": n1cells * k / e, "steps": k, "threads": th}
        del o
    oracle.set_threads(threads)
    out["pool_4cm"] = extra
    return out


def bonfire_c5(args, net):
    """The large-fire comparison: scenes/bonfire.toml at 10 cm (160 x 240 x
    160), pre-rolled with the reference step, checkpointed, then `c5_steps`
    steps per pressure method from that same state: the reference step
    (fixed 4 V-cycles), and pressure solves to 1e-4 by the reference V-cycles
    (at most 50) on b as the step forms it and on the consistent b (mean
    removed on sealed pockets inside the log pile, where heat release makes
    the system inconsistent), refinement, MGPCG and learned MGPCG (network in the first application
    on the fine grid, or in every application on level 1; consistent b); wall time per step, solver iterations and the FP64 statistics (HRR,
    pyrolysis, solid mass)."""
    import tempfile

    import pyro_amd as pa
    from pyro_amd.poisson import Poisson, step_solve

    s = pa.Scene.load(str(ROOT / "pyro_amd" / "scenes" / "bonfire.toml"))
    s.set("domain.cell", 0.1)
    nx, ny, nz = s.grid().dims
    dt = args.c5_dt
    sim = pa.Sim.new(s, device_voxelize=True)
    sim.add_velocity_noise(1e-3, 0)
    t0 = time.perf_counter()
    for k in range(args.c5_preroll):
        sim.step(dt, False)
    sim.sync()
    log(f"[bench] bonfire {nx}x{ny}x{nz}: {args.c5_preroll} pre-roll steps in {time.perf_counter() - t0:.1f} s")
    ck = Path(tempfile.mkdtemp()) / "bonfire.ckpt"
    sim.save(ck)
    out = {"grid": [nx, ny, nz], "cells": nx * ny * nz, "dt": dt, "preroll_steps": args.c5_preroll,
           "steps": args.c5_steps, "tol": 1e-4}
    P = Poisson(sim.read_field(pa.level_field(0, "flags")), s.grid().dx)
    runs = [("reference_step", None, None), ("vcycle", "vcycle", 50), ("vcycle_consistent", "vcycle", 50),
            ("refine", "refine", 100), ("mgpcg", "mgpcg", 100)]
    if net is not None:
        runs += [("learned_first", "learned", 100), ("learned_coarse_every", "learned", 100)]
    for key, method, cap in runs:
        if key.startswith("learned"):
            P.attach_net(net, key[len("learned_"):].replace("_", "-"))
        # every method but the reference V-cycles solves the consistent
        # problem (b minus its mean on sealed pockets); vcycle_consistent is
        # the reference V-cycle on that same problem
        P.set_sealed_projection("all" if key == "vcycle_consistent" else "solvers")
        sim.load(ck)
        its, worst, reached = [], 0.0, 0
        sim.sync()
        t0 = time.perf_counter()
        for k in range(args.c5_steps):
            stats = k % 10 == 9
            if method is None:
                sim.step(dt, stats)
            else:
                info = step_solve(sim, P, dt, method, 1e-4, cap, stats)
                its.append(info.iterations)
                worst = max(worst, info.rel_resid)
                reached += info.rel_resid <= 1e-4
        sim.sync()
        wall = time.perf_counter() - t0
        st = sim.read_stats()
        d = {"ms_per_step": 1e3 * wall / args.c5_steps, "hrr_W": st.hrr, "pyrolysis_kg_s": st.pyro,
             "solid_kg": st.solid, "max_T": st.max_t, "step_resid": st.resid}
        if its:
            d.update(mean_iterations=float(np.mean(its)), worst_rel_resid=worst,
                     steps_reaching_tol=int(reached), sealed_components=P.sealed_info()[0])
        out[key] = d
        log(f"[bench] bonfire {key}: {d}")
    P.close()
    out["fuel_balance"] = fuel_balance(sim, ck, s.params(), dt, args.c5_steps)
    ck.unlink()
    return out


def fuel_balance(sim, ck, p, dt, steps):
    """FP64 mass balance of the fuel over `steps` reference steps from the
    checkpoint, statistics on every step: the volatiles released by the
    solid (slot 2, kg/s, times dt) against the change of gas-phase fuel
    (slot 6), the fuel burnt (heat release / heat of combustion), and the
    change of condensed smoke (slot 8, formed from unburnt vapour).  The
    rest is fuel carried out of the domain or absorbed by the sponge layers
    (and the advection's non-conservation).  On the solid side: the change
    of solid mass (slot 7) against the volatiles; the rest is char burnt and
    the mass of cells that burnt out."""
    sim.load(ck)
    sim.step(dt, True)
    s0 = sim.read_stats()
    released = burnt = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        sim.step(dt, True)
        st = sim.read_stats()
        released += st.pyro * dt
        burnt += st.hrr / p["dhc"] * dt
    wall = time.perf_counter() - t0
    d_fuel, d_smoke, d_solid = st.fuel - s0.fuel, st.smoke - s0.smoke, st.solid - s0.solid
    rest = released - burnt - d_fuel - d_smoke
    out = {"steps": steps, "seconds_simulated": steps * dt, "volatiles_released_kg": released,
           "fuel_burnt_kg": burnt, "gas_fuel_change_kg": d_fuel, "smoke_change_kg": d_smoke,
           "outflow_and_sponge_kg": rest, "outflow_frac_of_released": rest / released if released else None,
           "solid_change_kg": d_solid, "char_and_burnout_kg": -d_solid - released,
           "wall_s_with_stats_every_step": wall}
    log(f"[bench] bonfire fuel balance: {out}")
    return out


def voxelize_leg():
    """Scene::voxelize (scene.rs:621-660) of the bonfire (57 shapes) on the
    device at 10 cm and 5 cm: pyro_voxelize into a simulation's SOL buffer
    (shape upload, kernel and synchronisation), and pyro_sim_voxelize (the
    same plus the second SOL buffer and the solid flags); the host loop at
    10 cm."""
    import ctypes as C

    import pyro_amd as pa
    from pyro_amd import _native

    out = {}
    lib = _native.hip()
    for cell in (0.1, 0.05):
        s = pa.Scene.load(str(ROOT / "pyro_amd" / "scenes" / "bonfire.toml"))
        s.set("domain.cell", cell)
        sh = s.shapes()
        dims, dx, t0, seed = s.voxel_spec()
        n = dims[0] * dims[1] * dims[2]
        sim = pa.Sim(s.params(), None, None, 4, s.get("ignition.until"))
        ptr, nbytes = C.c_void_p(), C.c_size_t()
        _native.check_hip(lib.pyro_sim_device_ptr(sim._h, pa.FIELDS["sol"], C.byref(ptr), C.byref(nbytes)))
        d = (C.c_uint32 * 3)(*dims)
        ts, tsim = [], []
        for _ in range(4):
            sim.sync()
            t = time.perf_counter()
            _native.check_hip(lib.pyro_voxelize(sh.ctypes.data, sh.size, d, dx, t0, seed, ptr, 1))
            ts.append(time.perf_counter() - t)
            t = time.perf_counter()
            sim.voxelize(s)
            tsim.append(time.perf_counter() - t)
        key = f"bonfire_{cell * 100:g}cm"
        out[key] = {"grid": list(dims), "cells": n, "shapes": int(sh.size), "device_ms": 1e3 * min(ts[1:]),
                    "sim_voxelize_ms": 1e3 * min(tsim[1:])}
        if cell == 0.1:
            t = time.perf_counter()
            s.voxelize()
            out[key]["host_ms"] = 1e3 * (time.perf_counter() - t)
        sim.close()
        log(f"[bench] voxelize {key}: {out[key]}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cell", type=float, default=0.02)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pressure", action="store_true", help="skip the pressure-solver legs")
    ap.add_argument("--c3-n", type=int, default=256)
    ap.add_argument("--c3-cpu", action="store_true", help="also solve C3 with the FP64 CPU oracle (~1 min)")
    ap.add_argument("--train-pairs", type=int, default=2000, help="U-Net training pairs (pool, 4 cm)")
    ap.add_argument("--train-steps", type=int, default=3000, help="U-Net Adam steps")
    ap.add_argument("--c5-steps", type=int, default=200, help="bonfire steps per pressure method (0: skip)")
    ap.add_argument("--c5-preroll", type=int, default=3600,
                    help="bonfire steps before the comparison (3600 x 5 ms: the logs burn from about 12 s)")
    ap.add_argument("--c5-dt", type=float, default=0.005)
    ap.add_argument("--burn-time", type=float, default=4.0,
                    help="continue the C2 run to this time and time it again (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU z-slab decomposition is a design, not built.", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        sys.exit(0)
    r = gpu_run(args)
    c1 = pool_4cm(args)
    c3 = None if args.no_pressure else pressure_c3(args)
    c5 = None if (args.no_pressure or args.c5_steps <= 0) else bonfire_c5(args, r.get("net"))
    vox = voxelize_leg()
    cpu = None if args.no_cpu else cpu_baseline(args, r["dims"])
    nx, ny, nz = r["dims"]
    value = r["cells"] * args.steps / r["wall"]
    top = dict(r["kernels"][0])
    out = {
        "metric": "cell-updates/s at 8.2M-cell step; time to 1e-4 pressure residual, learned vs plain MG",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: voxelized pool scene + U[-1e-3,1e-3] m/s velocity noise (seed 0)",
        "config": {
            "workload": f"pool 1 m heptane, {args.cell * 100:g} cm cells, {nx}x{ny}x{nz} "
                        f"({r['cells'] / 1e6:.2f} M cells), 4 V-cycles/step, dt={args.dt}s, hipGraph steps, "
                        "stats every 10th step",
            "grid": [nx, ny, nz],
            "pressure_cycles": 4,
            "dt": args.dt,
        },
        "roofline": top,
        "roofline_kernels": r["kernels"],
        "cpu_baseline": cpu,
        "stage_ms": r["stage_ms"],
        "pool_4cm": c1,
        "pressure": {"pool_frame": r["frame"], "standalone": c3, "learned_training": r["learned"],
                     "bonfire_10cm": c5},
        "pool_burning": r["burning"],
        "render_handoff": r["handoff"],
        "scene_voxelize": vox,
        "copy_kernel_GBs": r["copy_gbs"],
        "last_stats": {"hrr_W": r["stats"].hrr, "max_T": r["stats"].max_t, "resid": r["stats"].resid},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
