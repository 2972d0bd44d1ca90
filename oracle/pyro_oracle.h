/*
 * pyro_oracle.h -- CPU restatement of pyro's per-timestep solver (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle for the HIP implementation in pyro_amd/.
 * It restates, in plain C (FP32 arithmetic, operation order of the WGSL text,
 * no FMA contraction), the reference kernels:
 *   src/shaders/common.wgsl, tmpl_comp.wgsl, tmpl_scalar.wgsl, vel.wgsl,
 *   advect.wgsl, fluid.wgsl, combust.wgsl, mg.wgsl, stats.wgsl
 * and the dispatch order of src/sim.rs:246-491.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker / CPU baseline.  The product never links it.
 *
 * Stats are accumulated in FP64 in a fixed order (the reference uses f32 CAS
 * atomics in an undefined order, stats.wgsl:6-13).
 */
#ifndef PYRO_ORACLE_H
#define PYRO_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct OSim OSim;

/* Stage ids of one step, in the order of sim.rs:396-413. */
enum {
  O_STAGE_SOLID = 0, O_STAGE_CURL, O_STAGE_ADV_VEL1, O_STAGE_ADV_VEL2,
  O_STAGE_ADV_S1, O_STAGE_ADV_S2, O_STAGE_REACT, O_STAGE_FORCES,
  O_STAGE_DIVERGENCE, O_STAGE_VCYCLES, O_STAGE_RESID, O_STAGE_PROJECT,
  O_STAGE_COUNT
};

/* Field ids (same numbering as include/pyro_hip.h PYRO_FIELD_*). */
enum {
  O_F_S0 = 0, O_F_S1, O_F_S0B, O_F_S1B, O_F_H0, O_F_H1, O_F_SOL, O_F_SOL_OUT,
  O_F_PYR, O_F_FLAGS, O_F_CURL, O_F_U, O_F_V, O_F_W, O_F_UB, O_F_VB, O_F_WB,
  O_F_HU, O_F_HV, O_F_HW, O_F_LEVEL_BASE = 32 /* + 4*l + {0 phi, 1 b, 2 flags} */
};

OSim *osim_create(const void *params, size_t params_bytes, const float *sol_init,
                  const float *s0_init, uint32_t cycles, float ign_until);
void osim_destroy(OSim *s);
int osim_levels(const OSim *s, uint32_t *dims_out /* 3*levels */);
/* One full step: begin + all stages + end (sim.rs:246-423). */
void osim_step(OSim *s, float dt, int want_stats);
/* Fine-grained hooks for per-kernel parity tests. */
void osim_begin_step(OSim *s, float dt, int want_stats);
int osim_run_stage(OSim *s, int stage);
void osim_end_step(OSim *s);
int osim_read_field(const OSim *s, int field, void *dst, size_t bytes);
int osim_write_field(OSim *s, int field, const void *src, size_t bytes);
void osim_read_stats(const OSim *s, double *out /* 1040 */);
double osim_time(const OSim *s);
void osim_get_params(const OSim *s, void *out);
void osim_set_clock(OSim *s, double time, uint32_t frame, float ign_until);
/* C1 harness noise: u,v,w += amp*U[-1,1] on every valid face of vel[cur]. */
void osim_add_velocity_noise(OSim *s, float amp, uint32_t seed);
/* render.wgsl:16-37: vol | sol | vel (3 x N x vec4 f32, before the f16 store) */
void osim_render_pack(const OSim *s, float *out);
void osim_set_threads(int n);

/* Building blocks exposed for closed-form tests. */
float oracle_expf(float x);
float oracle_noise(uint32_t comp, uint32_t idx, uint32_t seed);

/* Stand-alone multigrid pieces on one level set (tests, C3 oracle). */
typedef struct OMg OMg;
OMg *omg_create(uint32_t nx, uint32_t ny, uint32_t nz, float h, uint32_t levels,
                const uint32_t *flags0);
void omg_destroy(OMg *m);
int omg_level_dims(const OMg *m, int l, uint32_t *d3);
void omg_set(OMg *m, int l, const float *phi, const float *b);
void omg_get(const OMg *m, int l, float *phi, float *b, uint32_t *flags);
void omg_smooth(OMg *m, int l, uint32_t iters, float omega);
void omg_restrict(OMg *m, int l);
void omg_prolong(OMg *m, int l);
void omg_vcycles(OMg *m, uint32_t cycles, uint32_t coarse_iters);
void omg_resid_norms(const OMg *m, double *r2, double *b2);

/* FP64 reference Poisson solve with the mg.wgsl operator: Jacobi-PCG on
 * -A phi = -b, fluid cells only; returns iterations, writes final rel. resid. */
int oracle_poisson_cg64(uint32_t nx, uint32_t ny, uint32_t nz, double h,
                        const uint32_t *flags, const double *b, double *phi,
                        double tol, int max_iter, double *rel_resid);
/* r = b - A phi in FP64 over fluid cells (solids give 0). */
void oracle_residual64(uint32_t nx, uint32_t ny, uint32_t nz, double h,
                       const uint32_t *flags, const double *b, const double *phi,
                       double *r);

#ifdef __cplusplus
}
#endif
#endif
