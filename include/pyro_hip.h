/*
 * pyro_hip.h -- C ABI of the MI355X (gfx950) per-timestep fire solver.
 *
 * Drop-in boundary for the wgpu dispatch of the reference's Sim
 * (src/sim.rs).  A Rust FFI crate wraps these calls as Sim::new / step /
 * read_stats / read_scalars with the reference's signatures; C++ and Python
 * hosts call them directly.
 *
 * Conventions
 *  - Return 0 on success, a negative PYRO_E_* code on failure; the message is
 *    available from pyro_last_error() (thread-local).
 *  - The library owns all device memory; host pointers are borrowed for the
 *    duration of a call only.
 *  - A PyroSim handle is single-threaded.  Work is ordered on the handle's
 *    own HIP stream; pyro_sim_step() returns once the step is enqueued
 *    (like queue.submit in sim.rs:418) and errors of an enqueued step are
 *    reported by the next synchronising call.
 *  - Field layouts are the reference's: vec4 fields N x 16 B, velocity
 *    components (nx+1)(ny+1)(nz+1) x f32 with x fastest, flags N x u32.
 */
#ifndef PYRO_HIP_H
#define PYRO_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PYRO_PARAMS_BYTES 576 /* scene.rs:426-477 */
#define PYRO_STATS_N 1040     /* sim.rs:12 STATS_N */

#define PYRO_OK 0
#define PYRO_E_INVALID (-1)   /* bad argument / size mismatch */
#define PYRO_E_OOM (-2)       /* device allocation failed */
#define PYRO_E_HIP (-3)       /* HIP runtime error (incl. a failed earlier launch) */
#define PYRO_E_NONFINITE (-4) /* a finiteness check found NaN/Inf */
#define PYRO_E_STATE (-5)     /* call not valid in the current state */
#define PYRO_E_NODEVICE (-6)  /* no usable GPU */
#define PYRO_E_DIVERGED (-7)  /* an iterative solve's residual grew past 1e3 x its best */

/* Field ids for read/write_field (superset of sim.rs:494 read_scalars). */
enum PyroField {
  PYRO_FIELD_S0 = 0,   /* (T, Y_F, Y_O2, Y_s)           N x vec4 */
  PYRO_FIELD_S1 = 1,   /* (Y_m, q, S, 0)                N x vec4 */
  PYRO_FIELD_S0B = 2,  /* advected s0 (before react)    N x vec4 */
  PYRO_FIELD_S1B = 3,
  PYRO_FIELD_H0 = 4,   /* first-order predictor of s0   N x vec4 */
  PYRO_FIELD_H1 = 5,
  PYRO_FIELD_SOL = 6,      /* sol[sol_cur]   (slot, fuel, char, Ts) */
  PYRO_FIELD_SOL_OUT = 7,  /* sol[1-sol_cur] */
  PYRO_FIELD_PYR = 8,      /* (m_v, m_c, exposed faces, 0) */
  PYRO_FIELD_FLAGS = 9,    /* N x u32, 1 = solid */
  PYRO_FIELD_CURL = 10,    /* (w, |w|) */
  PYRO_FIELD_U = 11, PYRO_FIELD_V = 12, PYRO_FIELD_W = 13,    /* vel[cur] */
  PYRO_FIELD_UB = 14, PYRO_FIELD_VB = 15, PYRO_FIELD_WB = 16, /* vel[1-cur] */
  PYRO_FIELD_HU = 17, PYRO_FIELD_HV = 18, PYRO_FIELD_HW = 19, /* predictor */
  PYRO_FIELD_RENDER = 20,  /* last pyro_sim_render_pack: vol | sol | vel, 3 x N x half4 */
  PYRO_FIELD_LEVEL_BASE = 32 /* + 4*level + {0 phi, 1 b, 2 flags(u32)} */
};

/* Stages of one step, in the dispatch order of sim.rs:396-413. */
enum PyroStage {
  PYRO_STAGE_SOLID = 0, PYRO_STAGE_CURL, PYRO_STAGE_ADV_VEL1, PYRO_STAGE_ADV_VEL2,
  PYRO_STAGE_ADV_S1, PYRO_STAGE_ADV_S2, PYRO_STAGE_REACT, PYRO_STAGE_FORCES,
  PYRO_STAGE_DIVERGENCE, PYRO_STAGE_VCYCLES, PYRO_STAGE_RESID, PYRO_STAGE_PROJECT,
  PYRO_STAGE_COUNT
};

/* pyro_sim_create flags */
#define PYRO_CREATE_NO_GRAPH 1u /* launch kernels directly instead of replaying hipGraphs */

typedef struct PyroSim PyroSim;

const char *pyro_last_error(void);

/* Sim::new (sim.rs:108-232).  params: the 576-byte Params blob of
 * Scene::params (scene.rs:523-618); sol_init: N x vec4 from Scene::voxelize
 * (scene.rs:621-660); s0_init: N x vec4 or NULL for ambient (sim.rs:134);
 * pressure_cycles: SimCfg.pressure_cycles (sim.rs:229); ign_until:
 * Ignition.until or 0 (sim.rs:204). */
int pyro_sim_create(const void *params, size_t params_bytes, const float *sol_init,
                    const float *s0_init, uint32_t pressure_cycles, float ign_until,
                    uint32_t flags, PyroSim **out);
/* Scene::voxelize (scene.rs:621-660) on the device: SOL (N x vec4) from the
 * expanded primitives of pyro_scene_shapes (pyro_scene.h), bit-identical to
 * pyro_scene_voxelize; written to dst in device memory when dst_on_device,
 * else copied to host memory.  No counterpart in the reference (its
 * voxelization is a host loop); for grids too large for a host SOL array. */
struct PyroShape;
int pyro_voxelize(const struct PyroShape *shapes, size_t count, const uint32_t dims[3], float dx, float t0,
                  uint32_t seed, float *dst, int dst_on_device);
/* pyro_voxelize into device memory without allocating: no counterpart in the
 * reference.  For callers that cannot call hipMalloc (code near graph
 * capture, serving loops, harnesses that forbid allocation).  The shape list
 * is staged in a fixed array of PYRO_VOXELIZE_NOALLOC_MAX_SHAPES entries that
 * each GPU holds until the process ends (320 KiB); more shapes are rejected
 * (PYRO_E_INVALID), as is a dst_device that is not device memory.  Same
 * kernel and bits as pyro_voxelize; runs on the null stream and returns
 * after the kernel.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other host threads) race. */
#define PYRO_VOXELIZE_NOALLOC_MAX_SHAPES 4096
int pyro_voxelize_noalloc(const struct PyroShape *shapes, size_t count, const uint32_t dims[3], float dx,
                          float t0, uint32_t seed, float *dst_device);
/* A simulation's SOL from primitives, on the device: both SOL buffers and the
 * solid flags (as sol_init does in pyro_sim_create, which also accepts
 * sol_init = NULL: no solids). */
int pyro_sim_voxelize(PyroSim *sim, const struct PyroShape *shapes, size_t count, float t0, uint32_t seed);
int pyro_sim_destroy(PyroSim *sim);

/* Sim::step (sim.rs:246-423): one timestep; stats are produced when
 * want_stats != 0 and read back by pyro_sim_read_stats. */
int pyro_sim_step(PyroSim *sim, float dt, int want_stats);
/* Sim::read_stats (sim.rs:510-530): the 1040 slots of the last stats step
 * in FP64 (slot map of stats.wgsl:2-3; 10 = sum r^2, 11 = sum b^2). */
int pyro_sim_read_stats(PyroSim *sim, double *out, size_t n);
/* Sim::read_scalars (sim.rs:494-508) and every other field, blocking.
 * Writing SOL (or SOL_OUT) sets both SOL buffers, as Sim::new does from one
 * voxelized array (sim.rs:136-140), and recomputes the solid flags.  Flags
 * follow SOL from creation on; the reference leaves them zero until the
 * first step's solid pass rewrites every cell (combust.wgsl:131), so they
 * differ only when read before the first step. */
int pyro_sim_read_field(PyroSim *sim, int field, void *dst, size_t bytes);
int pyro_sim_write_field(PyroSim *sim, int field, const void *src, size_t bytes);
/* Sim::cells (sim.rs:234), Sim::time / dims / cycles / levels. */
int pyro_sim_info(const PyroSim *sim, uint32_t dims[3], uint64_t *cells, double *time,
                  uint32_t *frame, uint32_t *levels);
int pyro_sim_level_dims(const PyroSim *sim, int level, uint32_t dims[3]);
int pyro_sim_sync(PyroSim *sim);
/* Device pointer of a field for interop (renderer hand-off); *bytes = size. */
int pyro_sim_device_ptr(PyroSim *sim, int field, void **dptr, size_t *bytes);

/* Renderer hand-off (replaces the pack pass that starts Renderer::render,
 * render.rs:255-277 binding s0, s1, sol[sol_cur], flags, vel[cur];
 * render.wgsl:16-37).  Packs the current state into the renderer's three
 * rgba16float volumes -- vol = (soot ppm, T, smoke g/m3, HRR kW/m3),
 * sol = (occupancy, Ts, char fraction, slot) premultiplied by occupancy,
 * vel = (cell-centre u, v, w, 0) -- each N x half4, x fastest, vol then sol
 * then vel: 24 bytes per cell, the rows queue.write_texture takes.
 * Stream-ordered after the steps already enqueued.  With host_dst (bytes =
 * 24 N) the volumes are then copied to host_dst on a second stream, which
 * overlaps the steps enqueued next; pyro_sim_render_wait blocks until that
 * copy has landed.  host_dst should be pinned (pyro_host_alloc) for the copy
 * to run asynchronously; pageable memory works but blocks the caller.  With
 * host_dst NULL only the device buffer (PYRO_FIELD_RENDER) is written.
 * host_dst must stay valid and untouched until pyro_sim_render_wait. */
int pyro_sim_render_pack(PyroSim *sim, void *host_dst, size_t bytes);
/* Blocks until the last hand-off has landed: its copy to host_dst, or with
 * host_dst NULL the pack itself (for a consumer of the device buffer). */
int pyro_sim_render_wait(PyroSim *sim);
/* The renderer's volumes as memory another API imports (the zero-copy form
 * of the same hand-off, render.rs:255-277: the renderer binds the packed
 * volumes instead of uploading them).  From the first call on, the packed
 * volumes (PYRO_FIELD_RENDER, 24 N bytes at offset 0) live in one
 * exportable device allocation of *bytes bytes (N x 24 rounded up to the
 * allocation granularity); each call returns a new POSIX file descriptor
 * for it (a dma-buf on this platform), which the caller owns and closes.
 * Vulkan imports it with VkImportMemoryFdInfoKHR.  Each
 * pyro_sim_render_pack(sim, NULL, 0) then rewrites it in stream order;
 * pyro_sim_render_wait before the importer reads it.  The allocation lives
 * until pyro_sim_destroy (importers keep their own reference). */
int pyro_sim_render_export(PyroSim *sim, int *fd, uint64_t *bytes);
/* Meant to import an allocation exported as above (fd, alloc_bytes as
 * returned), map it and copy its first `bytes` bytes to host_dst.  Disabled:
 * it returns PYRO_E_STATE (an in-process import ended in a host fault inside
 * the runtime; no cause has been shown). */
int pyro_shared_read(int fd, uint64_t alloc_bytes, void *host_dst, size_t bytes);
/* Pinned host memory for the hand-off (hipHostMalloc / hipHostFree). */
int pyro_host_alloc(size_t bytes, void **out);
int pyro_host_free(void *p);

/* Checkpoint (no reference counterpart): everything a step reads -- vel,
 * s0, s1, sol, the level-0 pressure warm start, Params, time, frame -- in one
 * binary file.  load requires a simulation of the same grid and restores the
 * saved Params; the continuation is bit-identical to the uninterrupted run. */
int pyro_sim_save(PyroSim *sim, const char *path);
int pyro_sim_load(PyroSim *sim, const char *path);

/* Fine-grained hooks (per-kernel parity tests, profiling). begin_step sets
 * dt/time/ign_on exactly as step does; run_stage enqueues one stage with the
 * current ping-pong roles; end_step finalises stats and swaps roles. */
int pyro_sim_begin_step(PyroSim *sim, float dt, int want_stats);
int pyro_sim_run_stage(PyroSim *sim, int stage);
int pyro_sim_end_step(PyroSim *sim);
/* Close an open step without ending it: the ping-pong roles, time and frame
 * stay those before begin_step (vel[cur] and sol[sol_cur] are unchanged), but
 * the stages already run have written their outputs -- after react, s0/s1
 * hold the step's scalars -- so a retry of the step needs the state restored
 * (pyro_sim_load).  Stats of the aborted step are not finalised.  Returns 0
 * when no step is open. */
int pyro_sim_abort_step(PyroSim *sim);
/* Time `reps` launches of one stage with HIP events on the sim stream
 * (average ms per launch; state is advanced reps times). */
int pyro_sim_time_stage(PyroSim *sim, int stage, int reps, float *avg_ms);
/* Kernel ids for pyro_sim_time_kernel. */
#define PYRO_KERNEL_MG_SMOOTH 0   /* one red-black sweep (both colours) on `level` */
#define PYRO_KERNEL_MG_RESTRICT 1 /* residual + restriction level -> level+1 */
#define PYRO_KERNEL_MG_PROLONG 2  /* prolongation level+1 -> level */
#define PYRO_KERNEL_MG_COARSE 3   /* all 80 SOR iterations of the coarsest level */
#define PYRO_KERNEL_RENDER_PACK 4 /* the renderer pack (after one pyro_sim_render_pack) */
#define PYRO_KERNEL_ADV_PASS1 5   /* the step's fused adv_vel1 + adv_s1 launch (level ignored) */
#define PYRO_KERNEL_ADV_PASS2 6   /* the step's fused adv_vel2 + adv_s2 launch (level ignored) */
#define PYRO_KERNEL_SOLID_CURL 7  /* the step's fused solid + curl launch (level ignored) */
/* Time `reps` launches of one multigrid kernel (HIP events on the sim
 * stream); the pressure state is modified. */
int pyro_sim_time_kernel(PyroSim *sim, int kernel, int level, int reps, float *avg_ms);
/* C1 harness: vel[cur] += amp * U[-1,1] (PCG hash of component, face, seed). */
int pyro_sim_add_velocity_noise(PyroSim *sim, float amp, uint32_t seed);
/* Scan the state for NaN/Inf; returns PYRO_E_NONFINITE if any. */
int pyro_sim_check_finite(PyroSim *sim);

/* Calibration: `reps` launches of a plain copy kernel over n_floats floats
 * (reads and writes 4 * n_floats bytes each); average ms per launch. */
int pyro_debug_copy(size_t n_floats, int reps, float *avg_ms);
/* ------------------------------------------------------------------------
 * Stand-alone pressure solve on one grid (no reference counterpart: the
 * reference only runs a fixed number of V-cycles per step, sim.rs:405).
 * Operator: level 0 of mg.wgsl (ground Neumann, phi = 0 outside the sides and
 * top, Neumann at solids); FP64 right-hand side, solution and residuals.
 * ------------------------------------------------------------------------ */
typedef struct PyroPoisson PyroPoisson;
enum PyroSolver {
  PYRO_SOLVER_VCYCLE = 0, /* the reference V-cycles (sim.rs:451-491) on FP32 */
  PYRO_SOLVER_REFINE = 1, /* FP64-residual iterative refinement, FP32 V-cycle corrections */
  PYRO_SOLVER_MGPCG = 2,  /* flexible CG, FP64 vectors/dots, FP32 V-cycle preconditioner */
  PYRO_SOLVER_LEARNED = 3 /* MGPCG whose preconditioner first applies the attached network
                             (pyro_poisson_attach_net), V-cycle fallback */
};
typedef struct {
  int iterations; /* outer iterations */
  int vcycles;    /* V-cycles applied */
  double rel_resid; /* sqrt(sum r^2 / sum b^2) over fluid cells, FP64 */
  float ms;       /* device time of the solve (HIP events) */
  int fallbacks;  /* LEARNED: 1 if the network's direction was rejected and plain V-cycles took over */
  int diverged;   /* 1 if the iteration stopped because the residual grew past 1e3 x its best */
} PyroSolveInfo;
const char *pyro_poisson_last_error(void);
/* flags: N x u32, 1 = solid (Neumann). */
int pyro_poisson_create(uint32_t nx, uint32_t ny, uint32_t nz, float h, const uint32_t *flags,
                        PyroPoisson **out);
int pyro_poisson_destroy(PyroPoisson *p);
int pyro_poisson_levels(const PyroPoisson *p, int *ref_levels, int *pc_levels);
/* Set the right-hand side b and/or the initial guess phi (either may be NULL). */
int pyro_poisson_set(PyroPoisson *p, const double *b, const double *phi);
int pyro_poisson_get(PyroPoisson *p, double *phi);
int pyro_poisson_residual(PyroPoisson *p, double *rel);
/* Iterate until the relative residual is <= tol (tol >= 0) or max_iters outer
 * iterations, or until the residual grows past 1e3 times the best one reached
 * (returns 0 with info->diverged = 1; the iterate is the last one).  A NaN or
 * Inf residual returns PYRO_E_NONFINITE.
 * Sealed fluid components (pure-Neumann pockets: face-connected fluid cells
 * none of which lies on a side or the top) make the system singular and, with
 * a right-hand side that does not sum to zero over them, inconsistent: no
 * method can then reduce the residual below that part of b.  Unless disabled
 * (pyro_poisson_set_sealed_projection), a solve first subtracts the mean of b
 * over each sealed component, once per right-hand side; the residual is
 * relative to that b.  The components are found on the device when the solid
 * flags change. */
int pyro_poisson_solve(PyroPoisson *p, int method, double tol, int max_iters, PyroSolveInfo *info);
/* mode 0: never project; 1 (default): for every method but PYRO_SOLVER_VCYCLE
 * (which stays the reference's); 2: for every method. */
int pyro_poisson_set_sealed_projection(PyroPoisson *p, int mode);
/* Number of sealed fluid components and of their cells under the current flags. */
int pyro_poisson_sealed_info(const PyroPoisson *p, uint32_t *components, uint32_t *cells);
/* The right-hand side as the solver holds it (after any projection). */
int pyro_poisson_get_rhs(PyroPoisson *p, double *b);

/* Pressure problem of a simulation (same grid): solid flags from its level-0
 * flags; b = its current level-0 right-hand side (set by the divergence
 * stage), phi = its level-0 phi when warm != 0, else 0. */
int pyro_poisson_set_flags_from_sim(PyroPoisson *p, PyroSim *sim);
int pyro_poisson_set_from_sim(PyroPoisson *p, PyroSim *sim, int warm);
/* One step (as pyro_sim_step) whose pressure solve runs `method` on p to a
 * relative residual of tol (at most max_iters outer iterations), warm-started
 * from the previous phi, instead of the fixed V-cycle count.  The arguments
 * (method, tol, max_iters, an attached network for LEARNED, the grid) are
 * checked before the step begins.  A failure after that -- a HIP error, a
 * non-finite residual (PYRO_E_NONFINITE), a diverged solve (PYRO_E_DIVERGED,
 * info filled) -- aborts the step as pyro_sim_abort_step does. */
int pyro_sim_step_solve(PyroSim *sim, PyroPoisson *p, int method, double tol, int max_iters, float dt,
                        int want_stats, PyroSolveInfo *info);

/* ------------------------------------------------------------------------
 * Learned preconditioner (no reference counterpart): a 3-level 3D U-Net,
 * 16/32/64 channels, 3x3x3 convolutions, BF16 on MFMA with FP32
 * accumulation, FP32 master weights, trained on the device with Adam on
 * (b, phi) pairs solved by the multigrid solver.  Parameter order: see
 * pyro_amd/csrc/net.hip.  Network input per voxel: (b / rms(b), fluid mask),
 * output: phi / (h^2 rms(b) output_scale).
 * ------------------------------------------------------------------------ */
typedef struct PyroNet PyroNet;
const char *pyro_net_last_error(void);
size_t pyro_net_param_count(void);
int pyro_net_create(uint64_t seed, PyroNet **out);
int pyro_net_destroy(PyroNet *net);
int pyro_net_get_params(PyroNet *net, float *dst, size_t n);
int pyro_net_set_params(PyroNet *net, const float *src, size_t n);
int pyro_net_get_grads(PyroNet *net, float *dst, size_t n);
int pyro_net_get_output_scale(PyroNet *net, float *scale);
int pyro_net_set_output_scale(PyroNet *net, float scale);
/* x0: N x 2 floats (input, mask), rounded to BF16; y: N floats. */
int pyro_net_forward(PyroNet *net, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, float *y);
/* Forward, then gradients of sum(gy * y) for all parameters (read with get_grads). */
int pyro_net_backward(PyroNet *net, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, const float *gy);
int pyro_net_dataset_clear(PyroNet *net);
int pyro_net_dataset_size(PyroNet *net, int *count);
/* Normalised pair on the host: x0 N x 2 (input, mask), target N. */
int pyro_net_dataset_add(PyroNet *net, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, const float *target);
int pyro_net_dataset_get(PyroNet *net, int index, float *x0, float *target);
/* Pair from the simulation's current pressure problem, solved from zero to
 * tol by MGPCG on p (which must have the simulation's grid). */
int pyro_net_collect(PyroNet *net, PyroSim *sim, PyroPoisson *p, double tol, PyroSolveInfo *info);
/* `steps` Adam steps, each on one random pair of [first, first + count);
 * losses[k] = relative squared error of step k (may be NULL). */
int pyro_net_train(PyroNet *net, int first, int count, int steps, float lr, uint64_t seed, float *losses);
int pyro_net_eval(PyroNet *net, int first, int count, double *mean_loss);
/* Average device time (HIP events) of `reps` forward passes on a grid. */
int pyro_net_time_forward(PyroNet *net, uint32_t nx, uint32_t ny, uint32_t nz, int reps, float *avg_ms);
/* mode 0: the network in every preconditioner application; 1: in the first
 * only (a learned warm start); 2 and 3: the same, with the network applied to
 * the residual restricted to level 1 of the preconditioner hierarchy and its
 * output prolonged back (1/8 of the voxels).  net = NULL detaches. */
int pyro_poisson_attach_net(PyroPoisson *p, PyroNet *net, int mode);

/* Test hooks of the network kernels.  mfma: one v_mfma_f32_16x16x32_bf16 of
 * lane fragments a, b (64 x 8 floats, rounded to BF16) -> d (64 x 4).
 * conv: one 3x3x3 conv of the network's kinds (forward, or its adjoint with
 * transposed = 1 applied to b = an N x cout gradient).  conv_dw: its weight
 * and bias gradients. */
int pyro_debug_mfma(const float *a, const float *b, float *d);
int pyro_debug_conv(int ca, int cb, int cout, int transposed, int relu, uint32_t nx, uint32_t ny, uint32_t nz,
                    const float *a, const float *b, const float *w, const float *bias, float *out);
int pyro_debug_conv_dw(int ca, int cb, int cout, uint32_t nx, uint32_t ny, uint32_t nz, const float *a,
                       const float *b, const float *g, float *dw, float *db);

/* Test hook: apply the device's exp (fn 0), sqrt (1) or division in/in2 (2). */
int pyro_debug_math(int fn, const float *in, const float *in2, float *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
