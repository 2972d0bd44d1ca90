// common.hpp -- device-side definitions shared by the gfx950 kernels.
//
// Params mirrors the 576-byte uniform of the reference (scene.rs:426-477,
// common.wgsl:14-28).  The builtin helpers fix the bits that WGSL leaves to
// the backend, in the forms of the reference's validated Metal backend
// (README.md:25): mix(a,b,t) = a + (b-a)*t, smoothstep = t*t*(3-2t).
// min/max are written as compare-select so that signed zeros resolve the
// same way on every device.  exp is computed by pexp() below, a fixed
// algorithm (not the device library), so results do not depend on libm bits.
// Every kernel is compiled with -ffp-contract=off: no a*b+c is fused unless
// written as fmaf().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pyro {

// Buffer resource over a read-only array below 4 GiB (pyro_sim_create limits
// the grid): loads take a 32-bit byte offset, and the range check of raw
// buffer loads returns 0 for an offset at or past `bytes` (RB_OOB), which
// replaces bounds branches where 0 is the value wanted outside.
struct RB {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ RB rb_of(const void *p, uint32_t bytes) {
  RB b;
  b.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
  return b;
}
constexpr uint32_t RB_OOB = 0x80000000u;  // reads 0 from every array below 2 GiB, which are the
                                           // only ones addressed with it (phi, b, codes, velocity)
__device__ __forceinline__ float4 rb_ld4(const RB &b, uint32_t byte_off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(b.r, byte_off, 0, 0));
}
__device__ __forceinline__ uint32_t rb_ld32(const RB &b, uint32_t byte_off) {
  return __builtin_amdgcn_raw_buffer_load_b32(b.r, byte_off, 0, 0);
}

struct Mat { float a[4], b[4], c[4]; };
struct Params {
  uint32_t nx, ny, nz, frame;
  float dx, dt, time, t0;
  float rho0, g, cp, yo2;
  float vort, cs, buoy, smoke_w;
  float wind_x, wind_z, wind_relax, sponge;
  float t_ign, s_o2, dhc, tau_min;
  float soot_form, soot_ox, soot_ox_t, smoke_conv;
  float kap_gas, h_conv, ign_on, ign_temp;
  float ign_x, ign_y, ign_z, ign_r;
  float ign_flux, ign_fuel, char_a, char_e;
  float dh_char, burnout, layer, rad_len;
  float h_mass, vort_t, rad_frac, seed;
  Mat mats[8];
};
static_assert(sizeof(Params) == 576, "Params layout must match scene.rs:426-477");

constexpr float SIGMA = 5.670374e-8f;
constexpr float RGAS = 8.314f;
constexpr float RHO_SOOT = 1800.0f;
constexpr int STATS_N = 1040;

// Exact unsigned division by a run-time constant (Granlund-Montgomery, valid
// for every 32-bit numerator): q = (t + ((n - t) >> 1)) >> (l - 1),
// t = umulhi(m, n).
struct FastDiv {
  uint32_t d, m, l;
};
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  f.l = l;
  f.m = l == 0 ? 0u : (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
  if (f.l == 0) return n;
  uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> 1)) >> (f.l - 1);
}

__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clampf_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
__device__ __forceinline__ float mixf_(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float smoothstep_(float e0, float e1, float x) {
  float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
__device__ __forceinline__ float pow4(float x) { float y = x * x; return y * y; }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// common.wgsl:45 rhoOf: ideal-gas density at temperature T (floor 200 K)
__device__ __forceinline__ float rhoOf(const Params *__restrict__ P, float T) {
  return P->rho0 * P->t0 / fmax_(T, 200.0f);
}

// exp(x): Cody-Waite reduction x = k*ln2 + r with ln2 split into a 16-bit
// head (0.693145751953125) and a tail, |r| <= ln2/2, then the degree-7 Taylor
// polynomial of e^r by Horner's rule with fused multiply-adds, scaled by 2^k
// through the exponent field (two steps for subnormal results).
__device__ __forceinline__ float pexp(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return __builtin_inff();
  if (x < -103.97208f) return 0.0f;
  float kf = rintf(x * 1.44269504088896341f);
  float r = fmaf(-kf, 0.693145751953125f, x);
  r = fmaf(-kf, 1.428606765330187e-06f, r);
  float p = 1.9841270e-4f;
  p = fmaf(p, r, 1.3888889e-3f);
  p = fmaf(p, r, 8.3333333e-3f);
  p = fmaf(p, r, 4.1666667e-2f);
  p = fmaf(p, r, 1.6666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  int k = (int)kf;
  if (k < -125) { p = p * 5.42101086242752217e-20f; k += 64; }
  if (k > 127) { p = p * 2.0f; k -= 1; }
  return p * __uint_as_float((uint32_t)(k + 127) << 23);
}

// PCG hash (common.wgsl:54-58).
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

// Grid geometry passed by value to every kernel.
struct Geo {
  int nx, ny, nz;          // cell dims
  int vx, vy;              // velocity padded dims x, y (nx+1, ny+1)
  uint32_t n, nv;          // cell / padded-face counts
  FastDiv dnx, dny, dvx, dvy;
  // brick grids (BR_X x BR_Y x BR_Z) over the cells and the padded faces
  uint32_t cbricks, fbricks;
  FastDiv dcbx, dcby, dfbx, dfby;
};

// Brick order for the gather-heavy advection kernels: block b (after
// xcd_tile) takes brick b of the brick grid (x fastest, then y, then z) and
// thread t its cell (t % 16, (t / 16) % 4, t / 64).  A block's trilinear
// samples touch the rows and planes around its brick, which the blocks in
// flight on the same XCD share: consecutive blocks walk x, then y.  The
// kernels that stage the brick's velocity in LDS make the halo's share of
// the tile the cost to minimise; 16x4x4 measured fastest of 16x2x8, 16x4x4,
// 32x2x4 and 8x4x8.
constexpr int BR_X = 16, BR_Y = 4, BR_Z = 4;
__device__ __forceinline__ void brick_cell(uint32_t b, uint32_t t, const FastDiv &dbx, const FastDiv &dby, int &x,
                                           int &y, int &z) {
  const uint32_t r = fdiv(b, dbx), bx = b - r * dbx.d;
  const uint32_t bz = fdiv(r, dby), by = r - bz * dby.d;
  x = (int)(bx * BR_X + (t & (BR_X - 1)));
  y = (int)(by * BR_Y + ((t / BR_X) & (BR_Y - 1)));
  z = (int)(bz * BR_Z + t / (BR_X * BR_Y));
}

__device__ __forceinline__ bool inb(const Geo &g, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz;
}
__device__ __forceinline__ uint32_t sidx(const Geo &g, int x, int y, int z) {
  return (uint32_t)x + (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
}
__device__ __forceinline__ uint32_t vidx(const Geo &g, int x, int y, int z) {
  return (uint32_t)x + (uint32_t)g.vx * ((uint32_t)y + (uint32_t)g.vy * (uint32_t)z);
}
__device__ __forceinline__ void decode_cell(const Geo &g, uint32_t i, int &x, int &y, int &z) {
  uint32_t r = fdiv(i, g.dnx);
  x = (int)(i - r * (uint32_t)g.nx);
  uint32_t q = fdiv(r, g.dny);
  y = (int)(r - q * (uint32_t)g.ny);
  z = (int)q;
}
__device__ __forceinline__ void decode_face(const Geo &g, uint32_t i, int &x, int &y, int &z) {
  uint32_t r = fdiv(i, g.dvx);
  x = (int)(i - r * (uint32_t)g.vx);
  uint32_t q = fdiv(r, g.dvy);
  y = (int)(r - q * (uint32_t)g.vy);
  z = (int)q;
}

// Workgroups are dispatched to the 8 XCDs round-robin (XCD = id % 8), each
// XCD with its own L2.  Remap so that every XCD gets a contiguous range of
// tiles: neighbouring tiles then share halos in the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q = nb >> 3;
  if (bid >= (q << 3)) return bid;
  return (bid & 7) * q + (bid >> 3);
}

// Deterministic block reductions (fixed pairing order) for FP64 stats.
// v[first] + v[first + step] + ... (count terms), added in that order from
// 0 with eight loads in flight: the bits of the plain loop, without a load
// round trip per term.
__device__ __forceinline__ double seq_sum(const double *__restrict__ v, int count, size_t first, size_t step) {
  double a = 0.0;
  for (int k0 = 0; k0 < count; k0 += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = k0 + u < count ? v[first + (size_t)(k0 + u) * step] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k0 + u < count) a += x[u];
  }
  return a;
}
// The same for a running maximum from 0 (compare-select, in order).
__device__ __forceinline__ float seq_max(const float *__restrict__ v, int count, size_t first, size_t step) {
  float m = 0.0f;
  for (int k0 = 0; k0 < count; k0 += 8) {
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = k0 + u < count ? v[first + (size_t)(k0 + u) * step] : 0.0f;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k0 + u < count) m = fmax_(m, x[u]);
  }
  return m;
}
// Terms of thread t in a loop over b = t, t + nth, ... < n.
__device__ __forceinline__ int strided_count(int t, int n, int nth) { return t < n ? (n - 1 - t) / nth + 1 : 0; }

template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < NT / 64; q++) r += sh[q];
  return r;  // valid in thread 0
}
template <int NT>
__device__ __forceinline__ float block_max(float v, float *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax_(v, __shfl_xor(v, o, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  float r = 0.0f;
  if (threadIdx.x == 0)
    for (int q = 0; q < NT / 64; q++) r = fmax_(r, sh[q]);
  return r;
}

}  // namespace pyro
