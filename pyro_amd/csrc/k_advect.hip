// k_advect.hip -- MacCormack advection on the staggered MAC grid (gfx950).
//
// Reference passes: advect.wgsl adv_vel1 (:4-20), adv_vel2 (:22-70),
// adv_s1 (:72-91), adv_s2 (:93-125), with the samplers of tmpl_comp.wgsl
// (:2-25), tmpl_scalar.wgsl (:2-33) and the RK2 traces of vel.wgsl (:2-19).
// Velocity samples are trilinear gathers from the padded component arrays;
// the gathers hit L1/L2 (each face's stencil overlaps its neighbours'), the
// compulsory HBM traffic is one read of each input and one write per output.
#include "kernels.hpp"

namespace pyro {

struct VelPtrs {
  const float *__restrict__ u;
  const float *__restrict__ v;
  const float *__restrict__ w;
};

static_assert(BR_X * BR_Y * BR_Z == NT, "one brick per block");

// Read-only arrays are gathered through buffer resources (RB, common.hpp):
// a sample's eight loads share one 32-bit byte offset (+4 and the row/plane
// strides), instead of eight 64-bit address computations.
__device__ __forceinline__ float ldf(const RB &b, uint32_t idx) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, idx * 4u, 0, 0));
}
// Elements idx and idx + 1 in one 8-byte load (4-byte alignment is all a
// multi-dword buffer load needs): the x-pair of a trilinear sample.
__device__ __forceinline__ float2 ld2(const RB &b, uint32_t idx) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.r, idx * 4u, 0, 0));
}
__device__ __forceinline__ float4 ld4(const RB &b, uint32_t idx) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(b.r, idx * 16u, 0, 0));
}
struct VelR {
  RB u, v, w;
};

// Sample coordinates clamped to [0, hi] (tmpl_comp.wgsl:5, tmpl_scalar.wgsl:5
// clamp them with compare-select).  A coordinate here is a non-negative grid
// position minus a trace displacement and staggering offsets, never -0, so
// fmaxf / fminf (one instruction each) give the same value as the
// compare-select pair: the only inputs on which they differ are +-0 ties
// with the 0 bound, and for a NaN coordinate both return the bound.  After
// the clamp the coordinate is >= 0, so truncation to int is its floor.
__device__ __forceinline__ float pclamp(float x, float hi) { return __builtin_fminf(__builtin_fmaxf(x, 0.0f), hi); }
__device__ __forceinline__ VelR velr_of(const Geo &g, const VelPtrs &V) {
  const uint32_t bytes = g.nv * 4u;
  return VelR{rb_of(V.u, bytes), rb_of(V.v, bytes), rb_of(V.w, bytes)};
}

// Trilinear sample of a staggered component whose nodes are integer points
// of g (tmpl_comp.wgsl:2-15); hi = largest valid node index.
__device__ __forceinline__ float triC(const Geo &g, const RB &X, float gx0, float gy0, float gz0,
                                      int hx, int hy, int hz) {
  float gx = pclamp(gx0, (float)hx), gy = pclamp(gy0, (float)hy), gz = pclamp(gz0, (float)hz);
  int bx = (int)gx, by = (int)gy, bz = (int)gz;
  bx = bx > hx - 1 ? hx - 1 : bx;
  by = by > hy - 1 ? hy - 1 : by;
  bz = bz > hz - 1 ? hz - 1 : bz;
  float fx = gx - (float)bx, fy = gy - (float)by, fz = gz - (float)bz;
  uint32_t sy = (uint32_t)g.vx, sz = (uint32_t)g.vx * (uint32_t)g.vy;
  uint32_t i = vidx(g, bx, by, bz);
  const float2 p00 = ld2(X, i), p10 = ld2(X, i + sy), p01 = ld2(X, i + sz), p11 = ld2(X, i + sy + sz);
  float c00 = mixf_(p00.x, p00.y, fx);
  float c10 = mixf_(p10.x, p10.y, fx);
  float c01 = mixf_(p01.x, p01.y, fx);
  float c11 = mixf_(p11.x, p11.y, fx);
  return mixf_(mixf_(c00, c10, fy), mixf_(c01, c11, fy), fz);
}
// ---------------------------------------------------------------- velocity tile
// Each block stages the three padded velocity arrays over its brick plus a
// halo of VT_H nodes in LDS.  A cell brick and a face brick index the same
// padded grid (face (x, y, z) is the low face of cell (x, y, z)), so one
// tile shape serves every kernel, all four of which use it (adv_s1 with it
// 278 -> 239 us at pool 2 cm: its velocity gathers kept the texture
// addressers busy).  Two nodes (15 KB per 16x4x4 block) measured best
// against three (29.6 KB per 16x2x8 block: fewer blocks in flight, longer
// staging).
//
// Staging also finds the largest |velocity| M of the tile.  Every velocity
// sample is a convex combination of tile nodes, so no trace point moves more
// than k M cells (k = dt/dx) from its start, and with k M < 1.4 every corner
// of every sample lies in the tile (the halo needs k M <= 1.5).  Such a block
// ("fast") reads all corners from LDS with no per-sample test; any other
// block tests each sample and gathers the corners outside the tile from
// global memory.  The arithmetic is the same on both paths.
constexpr int VT_H = 2;
constexpr int VT_X = BR_X + 2 * VT_H, VT_Y = BR_Y + 2 * VT_H, VT_Z = BR_Z + 2 * VT_H;
constexpr int VT_N = VT_X * VT_Y * VT_Z;
constexpr float FAST_REACH = 1.4f;
static_assert(VT_X * VT_Y + VT_X + 1 < 256, "corner offsets fit a ds_read2 offset");

struct VTile {
  const float *s;  // component c at s + c * VT_N, x fastest, then y, then z
  int x0, y0, z0;  // global node index of the tile's first entry
  int k0;          // x0 + VT_X y0 + VT_X VT_Y z0
};
// Tile entry of global node (x, y, z) (any node of the tile; x, y, z >= 0):
// full-rate 24-bit multiplies (node indices are far below 2^24).
__device__ __forceinline__ int tli(const VTile &T, int x, int y, int z) {
  return (int)(__umul24((uint32_t)z, (uint32_t)(VT_X * VT_Y)) + __umul24((uint32_t)y, (uint32_t)VT_X) + (uint32_t)x) -
         T.k0;
}
// Float bounds of the staggered samples, converted once per thread.
struct Bnd {
  float nx, ny, nz, nx1, ny1, nz1;
};
__device__ __forceinline__ Bnd bnd_of(const Geo &g) {
  return Bnd{(float)g.nx, (float)g.ny, (float)g.nz, (float)(g.nx - 1), (float)(g.ny - 1), (float)(g.nz - 1)};
}

__device__ __forceinline__ void brick_origin(uint32_t b, const FastDiv &dbx, const FastDiv &dby, int &x0, int &y0,
                                             int &z0) {
  const uint32_t r = fdiv(b, dbx), bx = b - r * dbx.d;
  const uint32_t bz = fdiv(r, dby), by = r - bz * dby.d;
  x0 = (int)(bx * BR_X);
  y0 = (int)(by * BR_Y);
  z0 = (int)(bz * BR_Z);
}

// Fill the tile (entries outside the padded arrays read 0 through the
// buffer resource's range check; no sample reads them), wait for it, and
// return whether the block is fast.  M is reduced on the bit patterns of
// |v| (unsigned order = float order for non-negative floats, and a NaN or an
// infinity exceeds every finite value, so such a tile is never fast).
__device__ __forceinline__ bool stage_vtile(const Geo &g, const VelR &V, float *sm, uint32_t *smax, float k,
                                            int bx0, int by0, int bz0, VTile &T) {
  T = VTile{sm, bx0 - VT_H, by0 - VT_H, bz0 - VT_H, 0};
  T.k0 = T.x0 + VT_X * T.y0 + VT_X * VT_Y * T.z0;
  if (threadIdx.x == 0) *smax = 0u;
  // entry e = t + NT j of every component: the three padded arrays share
  // the layout, so each entry's offset is computed once; all loads are
  // issued before the first store
  constexpr int NE = VT_N / NT;
  static_assert(VT_N % NT == 0, "whole staging rounds");
  uint32_t bits[NE][3];
#pragma unroll
  for (int j = 0; j < NE; j++) {
    const int e = (int)threadIdx.x + NT * j;
    const int lz = e / (VT_X * VT_Y), r = e - lz * (VT_X * VT_Y);
    const int ly = r / VT_X, lx = r - ly * VT_X;
    const int gx = T.x0 + lx, gy = T.y0 + ly, gz = T.z0 + lz;
    const bool in = (unsigned)gx < (unsigned)g.vx && (unsigned)gy < (unsigned)g.vy && (unsigned)gz <= (unsigned)g.nz;
    const uint32_t off = in ? vidx(g, gx, gy, gz) * 4u : RB_OOB;
    bits[j][0] = __builtin_amdgcn_raw_buffer_load_b32(V.u.r, off, 0, 0);
    bits[j][1] = __builtin_amdgcn_raw_buffer_load_b32(V.v.r, off, 0, 0);
    bits[j][2] = __builtin_amdgcn_raw_buffer_load_b32(V.w.r, off, 0, 0);
  }
  uint32_t mb = 0u;
#pragma unroll
  for (int j = 0; j < NE; j++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sm[c * VT_N + (int)threadIdx.x + NT * j] = __builtin_bit_cast(float, bits[j][c]);
      const uint32_t ab = bits[j][c] & 0x7fffffffu;
      mb = ab > mb ? ab : mb;
    }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = (uint32_t)__shfl_xor((int)mb, o, 64);
    mb = t > mb ? t : mb;
  }
  __syncthreads();  // the reset of *smax is visible
  if ((threadIdx.x & 63) == 0) atomicMax(smax, mb);
  __syncthreads();
  const uint32_t m = (uint32_t)__builtin_amdgcn_readfirstlane((int)*smax);
  return m < 0x7f800000u && k * __builtin_bit_cast(float, m) < FAST_REACH;
}

// One axis of a staggered trilinear sample (tmpl_comp.wgsl:2-15): the
// coordinate clamped to [0, h], its base node min(floor, h - 1) and weight.
// A coordinate is a non-negative grid position minus a trace displacement
// and staggering offsets, never -0, so the clamp's fmaxf/fminf pair equals
// the reference's compare-select clamp (for a NaN coordinate both give the
// bound); after it the truncation to int is the floor.  A fast block's
// coordinates are finite, and there one v_med3 does the clamp.
struct Ax {
  int b;
  float f;
};
template <bool FAST>
__device__ __forceinline__ Ax axis(float g0, int h, float hf) {
  const float g = FAST ? __builtin_amdgcn_fmed3f(g0, 0.0f, hf) : pclamp(g0, hf);
  int b = (int)g;
  b = b > h - 1 ? h - 1 : b;
  return Ax{b, g - (float)b};
}
template <bool FAST>
__device__ __forceinline__ Ax axis(float g0, int h) {
  return axis<FAST>(g0, h, (float)h);
}
// The four x-pairs of a sample's corners: from the tile, or gathered.
template <bool FAST>
__device__ __forceinline__ void corners(const Geo &g, const VTile &T, int c, const RB &X, int bx, int by, int bz,
                                        float2 &p00, float2 &p10, float2 &p01, float2 &p11) {
  const int lx = bx - T.x0, ly = by - T.y0, lz = bz - T.z0;
  if (FAST || ((unsigned)lx < (unsigned)(VT_X - 1) && (unsigned)ly < (unsigned)(VT_Y - 1) &&
               (unsigned)lz < (unsigned)(VT_Z - 1))) {
    const float *q = T.s + c * VT_N + tli(T, bx, by, bz);
    p00 = make_float2(q[0], q[1]);
    p10 = make_float2(q[VT_X], q[VT_X + 1]);
    p01 = make_float2(q[VT_X * VT_Y], q[VT_X * VT_Y + 1]);
    p11 = make_float2(q[VT_X * VT_Y + VT_X], q[VT_X * VT_Y + VT_X + 1]);
  } else {
    const uint32_t sy = (uint32_t)g.vx, sz = (uint32_t)g.vx * (uint32_t)g.vy, i = vidx(g, bx, by, bz);
    p00 = ld2(X, i);
    p10 = ld2(X, i + sy);
    p01 = ld2(X, i + sz);
    p11 = ld2(X, i + sy + sz);
  }
}
// triC (tmpl_comp.wgsl:2-15) of velocity component c through the tile.
template <bool FAST>
__device__ __forceinline__ float triT(const Geo &g, const VTile &T, int c, const RB &X, Ax ax, Ax ay, Ax az) {
  float2 p00, p10, p01, p11;
  corners<FAST>(g, T, c, X, ax.b, ay.b, az.b, p00, p10, p01, p11);
  const float c00 = mixf_(p00.x, p00.y, ax.f);
  const float c10 = mixf_(p10.x, p10.y, ax.f);
  const float c01 = mixf_(p01.x, p01.y, ax.f);
  const float c11 = mixf_(p11.x, p11.y, ax.f);
  return mixf_(mixf_(c00, c10, ay.f), mixf_(c01, c11, ay.f), az.f);
}
// triC and mmC (tmpl_comp.wgsl:2-25) of one point from the same corners:
// the sample is returned, (min, max) of the eight corners stored in *mm.
template <bool FAST>
__device__ __forceinline__ float triMMCT(const Geo &g, const VTile &T, int c, const RB &X, float gx, float gy,
                                         float gz, int hx, int hy, int hz, float2 *mm) {
  const Ax ax = axis<FAST>(gx, hx), ay = axis<FAST>(gy, hy), az = axis<FAST>(gz, hz);
  float2 p[4];
  corners<FAST>(g, T, c, X, ax.b, ay.b, az.b, p[0], p[1], p[2], p[3]);
  float l = 1e30f, h = -1e30f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    l = fmin_(l, p[j].x);
    h = fmax_(h, p[j].x);
    l = fmin_(l, p[j].y);
    h = fmax_(h, p[j].y);
  }
  *mm = make_float2(l, h);
  const float c00 = mixf_(p[0].x, p[0].y, ax.f);
  const float c10 = mixf_(p[1].x, p[1].y, ax.f);
  const float c01 = mixf_(p[2].x, p[2].y, ax.f);
  const float c11 = mixf_(p[3].x, p[3].y, ax.f);
  return mixf_(mixf_(c00, c10, ay.f), mixf_(c01, c11, ay.f), az.f);
}
// velAt (vel.wgsl:2-13): the three components' samples share their axes
// pairwise -- v and w the x axis (px - 1/2, bound nx - 1), u and w the y axis
// (py - 1/2, ny - 1), u and v the z axis (pz - 1/2, nz - 1).
template <bool FAST>
__device__ __forceinline__ void velAtT(const Geo &g, const Bnd &B, const VTile &T, const VelR &V, float px,
                                       float py, float pz, float &ux, float &uy, float &uz) {
  const Ax x0 = axis<FAST>(px, g.nx, B.nx), x1 = axis<FAST>(px - 0.5f, g.nx - 1, B.nx1);
  const Ax y0 = axis<FAST>(py, g.ny, B.ny), y1 = axis<FAST>(py - 0.5f, g.ny - 1, B.ny1);
  const Ax z0 = axis<FAST>(pz, g.nz, B.nz), z1 = axis<FAST>(pz - 0.5f, g.nz - 1, B.nz1);
  ux = triT<FAST>(g, T, 0, V.u, x0, y1, z1);
  uy = triT<FAST>(g, T, 1, V.v, x1, y0, z1);
  uz = triT<FAST>(g, T, 2, V.w, x1, y1, z0);
}

// Tile node (x, y, z) of component c.
__device__ __forceinline__ float tn(const VTile &T, int c, int x, int y, int z) {
  return T.s[c * VT_N + tli(T, x, y, z)];
}
// The velocity at a trace's start point, where every sample weight is 0 or
// 1/2.  Valid on a fast block (finite velocities) for the interior points
// named below, where no clamp acts and every base node is the natural one.
// A weight-0 lerp a + (b - a) * 0 is a up to the sign of a zero result, and
// + - * keep "equal up to the sign of zero" for finite operands, so these
// sums differ from the general sample at most in the sign of a zero.  The
// start velocity only moves the start point p (p - hk u, p + hk u), and
// every start coordinate is >= 1/2, so the trace's points are bit-identical.
//   u face (x, y, z) at (x, y+1/2, z+1/2): 1 <= x <= nx-1, y <= ny-2, z <= nz-2
//   v face at (x+1/2, y, z+1/2):           x <= nx-2, 1 <= y <= ny-1, z <= nz-2
//   w face at (x+1/2, y+1/2, z):           x <= nx-2, y <= ny-2, 1 <= z <= nz-1
//   cell centre (x+1/2, y+1/2, z+1/2):     x <= nx-2, y <= ny-2, z <= nz-2
__device__ __forceinline__ float half_(float a, float b) { return mixf_(a, b, 0.5f); }
__device__ __forceinline__ void start_u(const VTile &T, int x, int y, int z, float &ux, float &uy, float &uz) {
  ux = tn(T, 0, x, y, z);
  uy = half_(half_(tn(T, 1, x - 1, y, z), tn(T, 1, x, y, z)), half_(tn(T, 1, x - 1, y + 1, z), tn(T, 1, x, y + 1, z)));
  uz = half_(half_(tn(T, 2, x - 1, y, z), tn(T, 2, x, y, z)), half_(tn(T, 2, x - 1, y, z + 1), tn(T, 2, x, y, z + 1)));
}
__device__ __forceinline__ void start_v(const VTile &T, int x, int y, int z, float &ux, float &uy, float &uz) {
  ux = half_(half_(tn(T, 0, x, y - 1, z), tn(T, 0, x + 1, y - 1, z)), half_(tn(T, 0, x, y, z), tn(T, 0, x + 1, y, z)));
  uy = tn(T, 1, x, y, z);
  uz = half_(half_(tn(T, 2, x, y - 1, z), tn(T, 2, x, y, z)), half_(tn(T, 2, x, y - 1, z + 1), tn(T, 2, x, y, z + 1)));
}
__device__ __forceinline__ void start_w(const VTile &T, int x, int y, int z, float &ux, float &uy, float &uz) {
  ux = half_(half_(tn(T, 0, x, y, z - 1), tn(T, 0, x + 1, y, z - 1)), half_(tn(T, 0, x, y, z), tn(T, 0, x + 1, y, z)));
  uy = half_(half_(tn(T, 1, x, y, z - 1), tn(T, 1, x, y + 1, z - 1)), half_(tn(T, 1, x, y, z), tn(T, 1, x, y + 1, z)));
  uz = tn(T, 2, x, y, z);
}
__device__ __forceinline__ void start_c(const VTile &T, int x, int y, int z, float &ux, float &uy, float &uz) {
  ux = half_(tn(T, 0, x, y, z), tn(T, 0, x + 1, y, z));
  uy = half_(tn(T, 1, x, y, z), tn(T, 1, x, y + 1, z));
  uz = half_(tn(T, 2, x, y, z), tn(T, 2, x, y, z + 1));
}
// Start type of a trace: face component 0..2, or 3 for a cell centre.
__device__ __forceinline__ bool start_interior(const Geo &g, int c, int x, int y, int z) {
  switch (c) {
    case 0: return x >= 1 && x <= g.nx - 1 && y <= g.ny - 2 && z <= g.nz - 2;
    case 1: return x <= g.nx - 2 && y >= 1 && y <= g.ny - 1 && z <= g.nz - 2;
    case 2: return x <= g.nx - 2 && y <= g.ny - 2 && z >= 1 && z <= g.nz - 1;
    default: return x <= g.nx - 2 && y <= g.ny - 2 && z <= g.nz - 2;
  }
}
template <bool FAST>
__device__ __forceinline__ void start_vel(const Geo &g, const Bnd &B, const VTile &T, const VelR &V, int c, int x,
                                          int y, int z, float px, float py, float pz, float &ux, float &uy,
                                          float &uz) {
  if (FAST && start_interior(g, c, x, y, z)) {
    if (c == 0) start_u(T, x, y, z, ux, uy, uz);
    else if (c == 1) start_v(T, x, y, z, ux, uy, uz);
    else if (c == 2) start_w(T, x, y, z, ux, uy, uz);
    else start_c(T, x, y, z, ux, uy, uz);
  } else {
    velAtT<FAST>(g, B, T, V, px, py, pz, ux, uy, uz);
  }
}
// traceBack (vel.wgsl:2-13) from the start point of type c when BACK, the
// forward trace (vel.wgsl:14-19) when FWD; both start from the same sample.
template <bool FAST, bool BACK, bool FWD>
__device__ __forceinline__ void traceT(const Geo &g, const Bnd &B, const VTile &T, const VelR &V, float k, int c,
                                       int x, int y, int z, float px, float py, float pz, float &bx, float &by,
                                       float &bz, float &fx, float &fy, float &fz) {
  float u1x, u1y, u1z, u2x, u2y, u2z;
  start_vel<FAST>(g, B, T, V, c, x, y, z, px, py, pz, u1x, u1y, u1z);
  const float hk = 0.5f * k;
  if (BACK) {
    velAtT<FAST>(g, B, T, V, px - hk * u1x, py - hk * u1y, pz - hk * u1z, u2x, u2y, u2z);
    bx = px - k * u2x;
    by = py - k * u2y;
    bz = pz - k * u2z;
  }
  if (FWD) {
    velAtT<FAST>(g, B, T, V, px + hk * u1x, py + hk * u1y, pz + hk * u1z, u2x, u2y, u2z);
    fx = px + k * u2x;
    fy = py + k * u2y;
    fz = pz + k * u2z;
  }
}

// ---------------------------------------------------------------- adv_vel1
// Besides the sample at the back-trace foot, each face stores (min, max) of
// that sample's eight corners: adv_vel2's limiter (advect.wgsl:22-70, mmC at
// the same foot of the same field) then reads them instead of tracing back
// again.
template <bool FAST>
__device__ __forceinline__ void adv_vel1_face(const Params *__restrict__ P, const Geo &g, const VTile &T,
                                              const VelR &V, float k, int x, int y, int z, float *__restrict__ hu,
                                              float *__restrict__ hv, float *__restrict__ hw,
                                              float2 *__restrict__ mu, float2 *__restrict__ mv,
                                              float2 *__restrict__ mw) {
  const uint32_t i = vidx(g, x, y, z);
  const Bnd B = bnd_of(g);
  float bx, by, bz, fx, fy, fz;
  float2 mm;
  if (x <= g.nx && y <= g.ny - 1 && z <= g.nz - 1) {
    traceT<FAST, true, false>(g, B, T, V, k, 0, x, y, z, (float)x, (float)y + 0.5f, (float)z + 0.5f, bx, by, bz, fx,
                              fy, fz);
    hu[i] = triMMCT<FAST>(g, T, 0, V.u, bx, by - 0.5f, bz - 0.5f, g.nx, g.ny - 1, g.nz - 1, &mm);
    mu[i] = mm;
  }
  if (x <= g.nx - 1 && y <= g.ny && z <= g.nz - 1) {
    traceT<FAST, true, false>(g, B, T, V, k, 1, x, y, z, (float)x + 0.5f, (float)y, (float)z + 0.5f, bx, by, bz, fx,
                              fy, fz);
    hv[i] = triMMCT<FAST>(g, T, 1, V.v, bx - 0.5f, by, bz - 0.5f, g.nx - 1, g.ny, g.nz - 1, &mm);
    mv[i] = mm;
  }
  if (x <= g.nx - 1 && y <= g.ny - 1 && z <= g.nz) {
    traceT<FAST, true, false>(g, B, T, V, k, 2, x, y, z, (float)x + 0.5f, (float)y + 0.5f, (float)z, bx, by, bz, fx,
                              fy, fz);
    hw[i] = triMMCT<FAST>(g, T, 2, V.w, bx - 0.5f, by - 0.5f, bz, g.nx - 1, g.ny - 1, g.nz, &mm);
    mw[i] = mm;
  }
}
__global__ __launch_bounds__(NT) void k_adv_vel1(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                  float *__restrict__ hu, float *__restrict__ hv,
                                                  float *__restrict__ hw, float2 *__restrict__ mu,
                                                  float2 *__restrict__ mv, float2 *__restrict__ mw) {
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dfbx, g.dfby, x, y, z);
  brick_origin(blk, g.dfbx, g.dfby, x0, y0, z0);
  const VelR V = velr_of(g, VP);
  const float k = P->dt / P->dx;
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.vx || y >= g.vy || z > g.nz) return;
  if (fast) adv_vel1_face<true>(P, g, T, V, k, x, y, z, hu, hv, hw, mu, mv, mw);
  else adv_vel1_face<false>(P, g, T, V, k, x, y, z, hu, hv, hw, mu, mv, mw);
}
void launch_adv_vel1(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_vel1<<<f.g.fbricks, NT, 0, s>>>(f.P, f.g, V, f.hat[0], f.hat[1], f.hat[2], f.mm[0], f.mm[1], f.mm[2]);
}

// ---------------------------------------------------------------- adv_vel2
__device__ __forceinline__ bool solidC(const Geo &g, const uint8_t *__restrict__ fl, int x, int y, int z) {
  return inb(g, x, y, z) && fl[sidx(g, x, y, z)] == 1;
}

template <bool FAST>
__device__ __forceinline__ void adv_vel2_face(const Geo &g, const VTile &T, const VelR &V, float k, int x, int y,
                                              int z, const RB &Hu, const RB &Hv, const RB &Hw,
                                              const float2 *__restrict__ mu, const float2 *__restrict__ mv,
                                              const float2 *__restrict__ mw, float *__restrict__ ou,
                                              float *__restrict__ ov, float *__restrict__ ow,
                                              const uint8_t *__restrict__ fl) {
  const uint32_t i = vidx(g, x, y, z);
  const Bnd B = bnd_of(g);
  const bool sc = solidC(g, fl, x, y, z);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int hx = g.nx - 1 + (c == 0), hy = g.ny - 1 + (c == 1), hz = g.nz - 1 + (c == 2);
    if (x > hx || y > hy || z > hz) continue;
    float *O = c == 0 ? ou : (c == 1 ? ov : ow);
    if ((c == 1 && y == 0) || sc || solidC(g, fl, x - (c == 0), y - (c == 1), z - (c == 2))) {
      O[i] = 0.0f;
      continue;
    }
    const float offx = c == 0 ? 0.0f : 0.5f, offy = c == 1 ? 0.0f : 0.5f, offz = c == 2 ? 0.0f : 0.5f;
    const RB &Vc = c == 0 ? V.u : (c == 1 ? V.v : V.w);
    const RB &Hc = c == 0 ? Hu : (c == 1 ? Hv : Hw);
    // limiter bounds at the back-trace foot, from adv_vel1
    const float2 mm = (c == 0 ? mu : (c == 1 ? mv : mw))[i];
    const float px = (float)x + offx, py = (float)y + offy, pz = (float)z + offz;
    float bx, by, bz, fx, fy, fz;
    traceT<FAST, false, true>(g, B, T, V, k, c, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
    fx = fx - offx; fy = fy - offy; fz = fz - offz;
    const float t = triC(g, Hc, fx, fy, fz, hx, hy, hz);
    O[i] = clampf_(ldf(Hc, i) + 0.5f * (ldf(Vc, i) - t), mm.x, mm.y);
  }
}
// The component loop is unrolled (each start type and offset known at
// compile time) and the kernel held to at least 5 waves per SIMD (at most 102
// VGPRs; with the back trace still in it, unbounded it took 105 and ran 4
// waves): 466 -> 347 us at pool 2 cm, and 289 us without the back trace.
__global__ __launch_bounds__(NT, 5) void k_adv_vel2(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                  const float *__restrict__ hu, const float *__restrict__ hv,
                                                  const float *__restrict__ hw, const float2 *__restrict__ mu,
                                                  const float2 *__restrict__ mv, const float2 *__restrict__ mw,
                                                  float *__restrict__ ou, float *__restrict__ ov,
                                                  float *__restrict__ ow, const uint8_t *__restrict__ fl) {
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dfbx, g.dfby, x, y, z);
  brick_origin(blk, g.dfbx, g.dfby, x0, y0, z0);
  const VelR V = velr_of(g, VP);
  const float k = P->dt / P->dx;
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.vx || y >= g.vy || z > g.nz) return;
  const RB Hu = rb_of(hu, g.nv * 4u), Hv = rb_of(hv, g.nv * 4u), Hw = rb_of(hw, g.nv * 4u);
  if (fast) adv_vel2_face<true>(g, T, V, k, x, y, z, Hu, Hv, Hw, mu, mv, mw, ou, ov, ow, fl);
  else adv_vel2_face<false>(g, T, V, k, x, y, z, Hu, Hv, Hw, mu, mv, mw, ou, ov, ow, fl);
}
void launch_adv_vel2(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_vel2<<<f.g.fbricks, NT, 0, s>>>(f.P, f.g, V, f.hat[0], f.hat[1], f.hat[2], f.mm[0], f.mm[1], f.mm[2],
                                            f.vb[0], f.vb[1], f.vb[2], f.flags);
}

// ---------------------------------------------------------------- scalars
// Base cell and weights of a cell-centred trilinear sample (tmpl_scalar.wgsl:2-15).
struct SSample {
  uint32_t i;
  float fx, fy, fz;
};
__device__ __forceinline__ SSample sbase(const Geo &g, float gx0, float gy0, float gz0) {
  float gx = pclamp(gx0, (float)(g.nx - 1)), gy = pclamp(gy0, (float)(g.ny - 1)), gz = pclamp(gz0, (float)(g.nz - 1));
  int bx = (int)gx, by = (int)gy, bz = (int)gz;
  bx = bx > g.nx - 2 ? g.nx - 2 : bx;
  by = by > g.ny - 2 ? g.ny - 2 : by;
  bz = bz > g.nz - 2 ? g.nz - 2 : bz;
  SSample s;
  s.fx = gx - (float)bx;
  s.fy = gy - (float)by;
  s.fz = gz - (float)bz;
  s.i = sidx(g, bx, by, bz);
  return s;
}
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
  return make_float4(mixf_(a.x, b.x, t), mixf_(a.y, b.y, t), mixf_(a.z, b.z, t), mixf_(a.w, b.w, t));
}
__device__ __forceinline__ float4 triS(const Geo &g, const RB &X, const SSample &s) {
  uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny, i = s.i;
  float4 c00 = mix4(ld4(X, i), ld4(X, i + 1), s.fx);
  float4 c10 = mix4(ld4(X, i + sy), ld4(X, i + sy + 1), s.fx);
  float4 c01 = mix4(ld4(X, i + sz), ld4(X, i + sz + 1), s.fx);
  float4 c11 = mix4(ld4(X, i + sy + sz), ld4(X, i + sy + sz + 1), s.fx);
  return mix4(mix4(c00, c10, s.fy), mix4(c01, c11, s.fy), s.fz);
}
__device__ __forceinline__ void mmS(const Geo &g, const RB &X, const SSample &s, float4 &lo, float4 &hi) {
  uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny, i = s.i;
  const uint32_t o[8] = {0, 1, sy, sy + 1, sz, sz + 1, sy + sz, sy + sz + 1};
  float4 l = make_float4(1e30f, 1e30f, 1e30f, 1e30f), h = make_float4(-1e30f, -1e30f, -1e30f, -1e30f);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    float4 v = ld4(X, i + o[j]);
    l.x = fmin_(l.x, v.x); l.y = fmin_(l.y, v.y); l.z = fmin_(l.z, v.z); l.w = fmin_(l.w, v.w);
    h.x = fmax_(h.x, v.x); h.y = fmax_(h.y, v.y); h.z = fmax_(h.z, v.z); h.w = fmax_(h.w, v.w);
  }
  lo = l;
  hi = h;
}
// Distance beyond the open side/top boundaries (common.wgsl:79-82).
__device__ __forceinline__ float outside(const Geo &g, float x, float y, float z) {
  return fmax_(fmax_(-0.5f - x, x - ((float)g.nx - 0.5f)),
               fmax_(fmax_(-0.5f - z, z - ((float)g.nz - 0.5f)), y - ((float)g.ny - 0.5f)));
}

// adv_s1 (advect.wgsl:72-91): the back trace from the cell centre (through
// the velocity tile) and the sample of both fields at its foot, ambient
// beyond the open boundaries.
__device__ __forceinline__ void adv_s1_out(const Params *__restrict__ P, const Geo &g, uint32_t i, float bx,
                                           float by, float bz, const RB &s0, const RB &s1, float4 *__restrict__ h0,
                                           float4 *__restrict__ h1) {
  bx = bx - 0.5f; by = by - 0.5f; bz = bz - 0.5f;
  SSample s = sbase(g, bx, by, bz);
  float w = clampf_(outside(g, bx, by, bz), 0.0f, 1.0f);
  float4 a = triS(g, s0, s);
  float4 c = triS(g, s1, s);
  h0[i] = make_float4(mixf_(a.x, P->t0, w), mixf_(a.y, 0.0f, w), mixf_(a.z, P->yo2, w), mixf_(a.w, 0.0f, w));
  h1[i] = make_float4(mixf_(c.x, 0.0f, w), mixf_(c.y, 0.0f, w), mixf_(c.z, 0.0f, w), mixf_(c.w, 0.0f, w));
}
__global__ __launch_bounds__(NT) void k_adv_s1(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                const float4 *__restrict__ s0p, const float4 *__restrict__ s1p,
                                                float4 *__restrict__ h0, float4 *__restrict__ h1) {
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z;
  brick_cell(blk, threadIdx.x, g.dcbx, g.dcby, x, y, z);
  const VelR V = velr_of(g, VP);
  const RB s0 = rb_of(s0p, g.n * 16u), s1 = rb_of(s1p, g.n * 16u);
  const float k = P->dt / P->dx;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
  float bx, by, bz, fx, fy, fz;
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  int x0, y0, z0;
  brick_origin(blk, g.dcbx, g.dcby, x0, y0, z0);
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  if (fast) traceT<true, true, false>(g, bnd_of(g), T, V, k, 3, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
  else traceT<false, true, false>(g, bnd_of(g), T, V, k, 3, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
  adv_s1_out(P, g, sidx(g, x, y, z), bx, by, bz, s0, s1, h0, h1);
}
void launch_adv_s1(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_s1<<<f.g.cbricks, NT, 0, s>>>(f.P, f.g, V, f.s0, f.s1, f.h0, f.h1);
}

__device__ __forceinline__ float mcl(float r, float s, float t, float lo, float hi) {
  return clampf_(r + 0.5f * (s - t), lo, hi);
}

// adv_s2 (advect.wgsl:93-125).
template <bool FAST>
__device__ __forceinline__ void adv_s2_cell(const Geo &g, const VTile &T, const VelR &V, float k, int x, int y,
                                            int z, const RB &s0, const RB &s1, const RB &h0, const RB &h1,
                                            float4 *__restrict__ o0, float4 *__restrict__ o1) {
  const uint32_t i = sidx(g, x, y, z);
  const Bnd B = bnd_of(g);
  const float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
  float bx, by, bz, fx, fy, fz;
  traceT<FAST, true, true>(g, B, T, V, k, 3, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
  bx = bx - 0.5f; by = by - 0.5f; bz = bz - 0.5f;
  fx = fx - 0.5f; fy = fy - 0.5f; fz = fz - 0.5f;
  float4 r0 = ld4(h0, i), r1 = ld4(h1, i);
  if (outside(g, bx, by, bz) <= 0.0f && outside(g, fx, fy, fz) <= 0.0f) {
    SSample sb = sbase(g, bx, by, bz), sf = sbase(g, fx, fy, fz);
    float4 l0, u0, l1, u1;
    mmS(g, s0, sb, l0, u0);
    mmS(g, s1, sb, l1, u1);
    float4 t0 = triS(g, h0, sf), t1 = triS(g, h1, sf);
    float4 a = ld4(s0, i), c = ld4(s1, i);
    r0 = make_float4(mcl(r0.x, a.x, t0.x, l0.x, u0.x), mcl(r0.y, a.y, t0.y, l0.y, u0.y),
                     mcl(r0.z, a.z, t0.z, l0.z, u0.z), mcl(r0.w, a.w, t0.w, l0.w, u0.w));
    r1 = make_float4(mcl(r1.x, c.x, t1.x, l1.x, u1.x), mcl(r1.y, c.y, t1.y, l1.y, u1.y),
                     mcl(r1.z, c.z, t1.z, l1.z, u1.z), mcl(r1.w, c.w, t1.w, l1.w, u1.w));
  }
  o0[i] = r0;
  o1[i] = r1;
}
__global__ __launch_bounds__(NT) void k_adv_s2(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                const float4 *__restrict__ s0p, const float4 *__restrict__ s1p,
                                                const float4 *__restrict__ h0p, const float4 *__restrict__ h1p,
                                                float4 *__restrict__ o0, float4 *__restrict__ o1) {
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dcbx, g.dcby, x, y, z);
  brick_origin(blk, g.dcbx, g.dcby, x0, y0, z0);
  const VelR V = velr_of(g, VP);
  const float k = P->dt / P->dx;
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  const RB s0 = rb_of(s0p, g.n * 16u), s1 = rb_of(s1p, g.n * 16u), h0 = rb_of(h0p, g.n * 16u),
           h1 = rb_of(h1p, g.n * 16u);
  if (fast) adv_s2_cell<true>(g, T, V, k, x, y, z, s0, s1, h0, h1, o0, o1);
  else adv_s2_cell<false>(g, T, V, k, x, y, z, s0, s1, h0, h1, o0, o1);
}
void launch_adv_s2(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_s2<<<f.g.cbricks, NT, 0, s>>>(f.P, f.g, V, f.s0, f.s1, f.h0, f.h1, f.s0b, f.s1b);
}

// ---------------------------------------------------------------- fused passes
// The step runs adv_vel1 with adv_s1, then adv_vel2 with adv_s2: each pair
// needs the same velocity tile and neither output feeds the other, so one
// block per face brick stages the tile once and computes its face triples'
// velocity samples, then its cells' scalar samples (the cells of a brick are
// its faces inside the grid).  The arithmetic of each output is that of the
// single kernels above; adv_vel1's velocity-bound (VALU) work and adv_s1's
// gather-bound work then share the CU.
__global__ __launch_bounds__(NT) void k_adv_pass1(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                   float *__restrict__ hu, float *__restrict__ hv,
                                                   float *__restrict__ hw, float2 *__restrict__ mu,
                                                   float2 *__restrict__ mv, float2 *__restrict__ mw,
                                                   const float4 *__restrict__ s0p, const float4 *__restrict__ s1p,
                                                   float4 *__restrict__ h0, float4 *__restrict__ h1) {
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dfbx, g.dfby, x, y, z);
  brick_origin(blk, g.dfbx, g.dfby, x0, y0, z0);
  const VelR V = velr_of(g, VP);
  const float k = P->dt / P->dx;
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.vx || y >= g.vy || z > g.nz) return;
  if (fast) adv_vel1_face<true>(P, g, T, V, k, x, y, z, hu, hv, hw, mu, mv, mw);
  else adv_vel1_face<false>(P, g, T, V, k, x, y, z, hu, hv, hw, mu, mv, mw);
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  const RB s0 = rb_of(s0p, g.n * 16u), s1 = rb_of(s1p, g.n * 16u);
  const float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
  float bx, by, bz, fx, fy, fz;
  if (fast) traceT<true, true, false>(g, bnd_of(g), T, V, k, 3, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
  else traceT<false, true, false>(g, bnd_of(g), T, V, k, 3, x, y, z, px, py, pz, bx, by, bz, fx, fy, fz);
  adv_s1_out(P, g, sidx(g, x, y, z), bx, by, bz, s0, s1, h0, h1);
}
void launch_adv_pass1(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_pass1<<<f.g.fbricks, NT, 0, s>>>(f.P, f.g, V, f.hat[0], f.hat[1], f.hat[2], f.mm[0], f.mm[1], f.mm[2],
                                           f.s0, f.s1, f.h0, f.h1);
}

__global__ __launch_bounds__(NT) void k_adv_pass2(const Params *__restrict__ P, Geo g, VelPtrs VP,
                                                   const float *__restrict__ hu, const float *__restrict__ hv,
                                                   const float *__restrict__ hw, const float2 *__restrict__ mu,
                                                   const float2 *__restrict__ mv, const float2 *__restrict__ mw,
                                                   float *__restrict__ ou, float *__restrict__ ov,
                                                   float *__restrict__ ow, const uint8_t *__restrict__ fl,
                                                   const float4 *__restrict__ s0p, const float4 *__restrict__ s1p,
                                                   const float4 *__restrict__ h0p, const float4 *__restrict__ h1p,
                                                   float4 *__restrict__ o0, float4 *__restrict__ o1) {
  __shared__ float sm[3 * VT_N];
  __shared__ uint32_t smax;
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dfbx, g.dfby, x, y, z);
  brick_origin(blk, g.dfbx, g.dfby, x0, y0, z0);
  const VelR V = velr_of(g, VP);
  const float k = P->dt / P->dx;
  VTile T;
  const bool fast = stage_vtile(g, V, sm, &smax, k, x0, y0, z0, T);
  if (x >= g.vx || y >= g.vy || z > g.nz) return;
  {
    const RB Hu = rb_of(hu, g.nv * 4u), Hv = rb_of(hv, g.nv * 4u), Hw = rb_of(hw, g.nv * 4u);
    if (fast) adv_vel2_face<true>(g, T, V, k, x, y, z, Hu, Hv, Hw, mu, mv, mw, ou, ov, ow, fl);
    else adv_vel2_face<false>(g, T, V, k, x, y, z, Hu, Hv, Hw, mu, mv, mw, ou, ov, ow, fl);
  }
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  const RB s0 = rb_of(s0p, g.n * 16u), s1 = rb_of(s1p, g.n * 16u), h0 = rb_of(h0p, g.n * 16u),
           h1 = rb_of(h1p, g.n * 16u);
  if (fast) adv_s2_cell<true>(g, T, V, k, x, y, z, s0, s1, h0, h1, o0, o1);
  else adv_s2_cell<false>(g, T, V, k, x, y, z, s0, s1, h0, h1, o0, o1);
}
void launch_adv_pass2(const Fields &f, hipStream_t s) {
  VelPtrs V{f.va[0], f.va[1], f.va[2]};
  k_adv_pass2<<<f.g.fbricks, NT, 0, s>>>(f.P, f.g, V, f.hat[0], f.hat[1], f.hat[2], f.mm[0], f.mm[1], f.mm[2],
                                           f.vb[0], f.vb[1], f.vb[2], f.flags, f.s0, f.s1, f.h0, f.h1, f.s0b, f.s1b);
}

}  // namespace pyro
