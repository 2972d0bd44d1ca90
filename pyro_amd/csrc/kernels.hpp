// kernels.hpp -- host-side launchers of the gfx950 kernels (one per reference
// compute pass).  All launches are stream-ordered and capturable in a
// hipGraph; none allocates.
#pragma once
#include "common.hpp"

struct PyroShape;  // pyro_scene.h

namespace pyro {

struct Fields {
  const Params *P;          // device copy of the Params blob
  Geo g;
  float *va[3], *vb[3];     // velocity ping-pong: va = vel[cur], vb = vel[1-cur]
  float *hat[3];
  float2 *mm[3];            // adv_vel1 -> adv_vel2: (min, max) of each face's back-trace corners
  float4 *s0, *s1, *s0b, *s1b, *h0, *h1, *sol_in, *sol_out, *pyr, *curl;
  uint8_t *flags;           // 1 = solid (u8 copy of the reference u32 flags)
  float *phi0, *b0;         // multigrid level 0
};

// Per-block FP64 partials of the step statistics (stats.wgsl slot map).
struct StatParts {
  double *solid;   // [blocks][3]: pyro, solid mass, area        (slots 2,7,9)
  float *curl;     // [blocks]: max speed                         (slot 5)
  double *react;   // [ny][gx][6]: H,R,S,F,M,layerH; Tmax below   (0,1,3,6,8,16+y)
  float *react_t;  // [ny][gx]: Tmax                              (slot 4)
  double *resid;   // [blocks][2]: sum r^2, sum b^2               (10,11)
  double *out;     // [1040]
  int n_solid, n_curl, react_gx, n_resid;
};

constexpr int NT = 256;  // threads per block of the element-wise kernels

void launch_solid(const Fields &f, StatParts *st, hipStream_t s);
void launch_curl(const Fields &f, StatParts *st, hipStream_t s);
// solid and curl in one launch (the step's order; same statistics partials)
void launch_solid_curl(const Fields &f, StatParts *st, hipStream_t s);
// flags[i] = 1 where sol[i] is solid (at creation and after SOL is written).
void launch_init_flags(const Params *P, const Geo &g, const float4 *sol, uint8_t *flags, hipStream_t s);
// Scene::voxelize on the device (voxel.hip): SOL of dims from the expanded
// primitives (device copy), bit-identical to the host voxelizer.
// render.wgsl:16-37: vol | sol | vel as three N x half4 volumes into `out`
void launch_render_pack(const Fields &f, void *out, hipStream_t s);
int launch_voxelize(const ::PyroShape *d_shapes, int count, const uint32_t dims[3], float dx, float t0,
                    uint32_t seed, float4 *out, hipStream_t s);
void launch_adv_vel1(const Fields &f, hipStream_t s);
void launch_adv_vel2(const Fields &f, hipStream_t s);
void launch_adv_s1(const Fields &f, hipStream_t s);
void launch_adv_s2(const Fields &f, hipStream_t s);
// adv_vel1 with adv_s1, and adv_vel2 with adv_s2, one launch each (the step's order)
void launch_adv_pass1(const Fields &f, hipStream_t s);
void launch_adv_pass2(const Fields &f, hipStream_t s);
void launch_react(const Fields &f, StatParts *st, hipStream_t s);
void launch_forces(const Fields &f, hipStream_t s);
void launch_divergence(const Fields &f, hipStream_t s);
void launch_project(const Fields &f, hipStream_t s);
// bits of the stats-producing stages that ran in a step (k_stats_finalize)
constexpr uint32_t STATS_RAN_SOLID = 1u, STATS_RAN_CURL = 2u, STATS_RAN_REACT = 4u, STATS_RAN_RESID = 8u;
constexpr uint32_t STATS_RAN_ALL = 15u;
void launch_stats_finalize(const Geo &g, const StatParts &st, uint32_t ran, hipStream_t s);
void launch_velocity_noise(const Fields &f, float amp, uint32_t seed, hipStream_t s);
void launch_set_step(Params *P, float dt, float time, float ign_on, uint32_t frame, hipStream_t s);

// Multigrid on one level hierarchy (mg.wgsl).
struct MgLevel {
  int d[3];
  float h, hh;   // spacing and h*h (f32)
  float wdir = 1.0f;  // weight of a neighbour outside the box (1 = mg.wgsl)
  float *phi, *b;
  float *tmp;    // second phi buffer of the ping-pong sweeps
  uint8_t *f;    // 1 = solid
  uint8_t *nc;   // neighbour codes (k_mg.hip setup pass)
  // fluid neighbours as a 6-bit mask (+x, -x, +y, -y, +z, -z; bit set when
  // the neighbour is inside the level and not solid), for the FP64 operator
  // of the pressure solvers; set up with the codes where allocated (level 0
  // of a Poisson handle), null elsewhere
  uint8_t *nm = nullptr;
  uint32_t n;
};
// Coarse flags and neighbour codes of every level (once per step).
void launch_mg_setup(const MgLevel *L, int nl, hipStream_t s);
// One red-black sweep (colour 0 then 1) reading phi from pin, writing pout.
void launch_mg_sweep(const MgLevel &L, const float *pin, float *pout, float omega, hipStream_t s, bool fine);
void launch_mg_restrict(const MgLevel &F, const MgLevel &C, hipStream_t s);
void launch_mg_prolong(const MgLevel &C, const MgLevel &F, hipStream_t s);
// All `iters` red-black sweeps of the coarsest level in one launch.
void launch_mg_coarse(const MgLevel &L, int iters, float omega, hipStream_t s);
bool mg_coarse_fits_lds(const MgLevel &L);
void launch_mg_resid(const MgLevel &L, StatParts *st, hipStream_t s);
int mg_resid_blocks(const MgLevel &L);

void launch_copy(const float4 *src, float4 *dst, uint32_t n4, hipStream_t s);
// V-cycle of sim.rs:451-491 over an arbitrary hierarchy (2 pre- and 2
// post-sweeps per level, `coarse_iters` SOR iterations at omega = 1.6).
void enqueue_vcycle(const MgLevel *lv, int nl, int coarse_iters, hipStream_t s);

// FP64 solver kernels (k_solve.hip).
void launch_resid64(const MgLevel &L, const double *x, const double *b, double *r, hipStream_t s);
void launch_apply64(const MgLevel &L, const double *x, double *y, hipStream_t s);
void launch_dots(const double *const *a, const double *const *b, int k, uint32_t n, double *part, double *out,
                 hipStream_t s);
constexpr int DOT_PART = 4 * 1024;  // doubles of scratch for launch_dots
void launch_axpy64(uint32_t n, double alpha, const double *x, double *y, hipStream_t s);
void launch_xpby64(uint32_t n, const double *z, double beta, double *p, hipStream_t s);
void launch_to_f32(uint32_t n, const double *a, float *b, hipStream_t s);
void launch_to_f64(uint32_t n, const float *a, double *b, hipStream_t s);
void launch_add_f32_to_f64(uint32_t n, const float *a, double *b, hipStream_t s);
void launch_zero32(uint32_t n, float *a, hipStream_t s);
void launch_zero64(uint32_t n, double *a, hipStream_t s);
// fused flexible-CG steps with the iteration's scalars on the device:
// dev[0] = p.q, dev[1] = p.r, dev[2] = r.r, dev[3] = z.q; part holds
// cg_part_blocks(n) partials per sum (2 sums at most)
uint32_t cg_part_blocks(uint32_t n);
void launch_apply_dots(const MgLevel &L, const double *p, const double *r, double *q, double *part, double *dev,
                       hipStream_t s);
void launch_update_rr(uint32_t n, const double *dev, const double *p, const double *q, double *x, double *r,
                      double *part, double *dev_out, hipStream_t s);
void launch_to_f32_zero(uint32_t n, const double *r, float *b, float *phi, hipStream_t s);
void launch_to_f64_dot(uint32_t n, const float *phi, const double *q, double *z, double *part, double *dev,
                       hipStream_t s);
void launch_xpby_dev(uint32_t n, const double *z, const double *dev, double *p, hipStream_t s);
// r = b - A x, dev[2] = r.r
// Sealed fluid components (k_solve.hip): union-find labels with every
// fluid cell's parent at its component's smallest index, open[root] = 1 for
// components touching a side or the top; (root, cell) pairs of the sealed
// ones; b -= mean over each sealed component; a flag-change detector.
void launch_cc_label(const MgLevel &L, uint32_t *parent, uint8_t *open, hipStream_t s);
void launch_cc_sealed(const MgLevel &L, const uint32_t *parent, const uint8_t *open, uint2 *out,
                      unsigned int *count, hipStream_t s);
void launch_seal_project(uint32_t ncomp, const uint32_t *off, const uint32_t *idx, double *b, hipStream_t s);
void launch_flags_changed(uint32_t n, const uint8_t *f, uint8_t *prev, unsigned int *changed, hipStream_t s);
void launch_resid_rr(const MgLevel &L, const double *x, const double *b, double *r, double *part, double *dev,
                     hipStream_t s);

// Debug / test: apply device math to an array (0 exp, 1 sqrt, 2 div by in2).
void launch_debug_math(int fn, const float *in, const float *in2, float *out, uint32_t n, hipStream_t s);

inline int nblocks(uint32_t n) { return (int)((n + NT - 1) / NT); }

}  // namespace pyro
