// k_fluid.hip -- vorticity, body forces, divergence and projection on gfx950.
//
// Reference passes: fluid.wgsl curl (:3-29), forces (:31-109), divergence
// (:111-128), project (:130-152).  One thread per cell (curl, divergence) or
// per padded face triple (forces, project); the velocity components live in
// the reference's padded (nx+1)(ny+1)(nz+1) layout, x fastest.
#include "kernels.hpp"
#include "curl.hpp"

namespace pyro {

// ---------------------------------------------------------------- curl
template <bool STATS>
__global__ __launch_bounds__(NT) void k_curl(const Params *__restrict__ P, Geo g, const float *__restrict__ U,
                                              const float *__restrict__ V, const float *__restrict__ W,
                                              float4 *__restrict__ curl, float *__restrict__ part) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  float spd = 0.0f;
  if (i < g.n) spd = curl_cell<STATS>(P, g, U, V, W, curl, i);
  if (STATS) {
    __shared__ float sh[NT / 64];
    float m = block_max<NT>(spd, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
  }
}

void launch_curl(const Fields &f, StatParts *st, hipStream_t s) {
  int nb = nblocks(f.g.n);
  if (st)
    k_curl<true><<<nb, NT, 0, s>>>(f.P, f.g, f.va[0], f.va[1], f.va[2], f.curl, st->curl);
  else
    k_curl<false><<<nb, NT, 0, s>>>(f.P, f.g, f.va[0], f.va[1], f.va[2], f.curl, nullptr);
}

// ---------------------------------------------------------------- forces
__device__ __forceinline__ float spongeSide(const Params *__restrict__ P, const Geo &g, int x, int z) {
  int e = x < g.nx - 1 - x ? x : g.nx - 1 - x;
  int e2 = z < g.nz - 1 - z ? z : g.nz - 1 - z;
  if (e2 < e) e = e2;
  return 1.0f - smoothstep_(0.0f, P->sponge, (float)e);
}
__device__ __forceinline__ float spongeTop(const Params *__restrict__ P, const Geo &g, int y) {
  float h = (float)g.ny;
  return smoothstep_(h - 1.5f * P->sponge, h - 1.0f, (float)y);
}

// A block takes a 16 x 4 x 4 brick of face triples (brick_cell over the
// padded faces).  Every face reads its own cell b = (x, y, z) and, for
// component k, the cell a one step down along k, so the block's cells are
// the brick grown by one on the low side of each axis: C_X x C_Y x C_Z =
// 17 x 5 x 5.  Their confinement needs |curl| one cell further out on both
// sides (19 x 7 x 7, coordinates clamped to the grid as fluid.wgsl's
// curl lookups are).  Both tiles are staged in LDS (the |curl| tile only by
// a block where some cell's confinement gate is open), the confinement of
// each tile cell is evaluated once (a face triple used to evaluate four),
// and the faces then read T, smoke, solid flags and forces from LDS.
constexpr int FC_X = BR_X + 1, FC_Y = BR_Y + 1, FC_Z = BR_Z + 1, FC_N = FC_X * FC_Y * FC_Z;
constexpr int FW_X = BR_X + 3, FW_Y = BR_Y + 3, FW_Z = BR_Z + 3, FW_N = FW_X * FW_Y * FW_Z;
constexpr int FC_R = (FC_N + NT - 1) / NT, FW_R = (FW_N + NT - 1) / NT;

// Vorticity confinement at a cell centre (fluid.wgsl:167-181) from its T,
// smoke and curl w; m points at the cell's |curl| in the |curl| tile, whose
// entries m[+-1], m[+-FW_X], m[+-FW_X FW_Y] are the clamped neighbours.
__device__ __forceinline__ float confine_gate(const Params *__restrict__ P, float T, float smoke) {
  return fmax_(smoothstep_(P->vort_t, P->vort_t + 500.0f, T), 0.35f * smoothstep_(0.0f, 2e-4f, smoke));
}
__device__ __forceinline__ void confine_tile(const Params *__restrict__ P, float gate, float4 w, const float *m,
                                             float &fx, float &fy, float &fz) {
  fx = fy = fz = 0.0f;
  if (gate <= 0.0f) return;
  float gx = m[1] - m[-1];
  float gy = m[FW_X] - m[-FW_X];
  float gz = m[FW_X * FW_Y] - m[-FW_X * FW_Y];
  float gl = len3(gx, gy, gz);
  if (gl < 1e-6f) return;
  float nx = gx / gl, ny = gy / gl, nz = gz / gl;
  float cx = ny * w.z - nz * w.y, cy = nz * w.x - nx * w.z, cz = nx * w.y - ny * w.x;
  float vd = P->vort * P->dx;
  fx = vd * cx * gate;
  fy = vd * cy * gate;
  fz = vd * cz * gate;
  float fl = len3(fx, fy, fz);
  if (fl > 60.0f) {
    float sc = 60.0f / fl;
    fx *= sc; fy *= sc; fz *= sc;
  }
}

__global__ __launch_bounds__(NT) void k_forces(const Params *__restrict__ P, Geo g, float *__restrict__ U,
                                                float *__restrict__ V, float *__restrict__ W,
                                                const float4 *__restrict__ s0, const float4 *__restrict__ s1,
                                                const float4 *__restrict__ curl, const uint8_t *__restrict__ fl) {
  __shared__ float sw[FW_N];                           // |curl|
  __shared__ float sf[3][FC_N], st[FC_N], sm[FC_N];    // confinement force, T, smoke
  __shared__ uint8_t sc[FC_N];                         // bit 0 inside the grid, bit 1 solid
  const uint32_t blk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
  int x, y, z, x0, y0, z0;
  brick_cell(blk, threadIdx.x, g.dfbx, g.dfby, x, y, z);
  {
    const uint32_t r = fdiv(blk, g.dfbx), bz = fdiv(r, g.dfby);
    x0 = (int)((blk - r * g.dfbx.d) * BR_X);
    y0 = (int)((r - bz * g.dfby.d) * BR_Y);
    z0 = (int)(bz * BR_Z);
  }
  // the tile cells' state at clamped coordinates (a cell outside the grid is
  // flagged and never read) and the face's velocities, all loads in flight
  // together
  float4 ca[FC_R];
  float cs1[FC_R], gate[FC_R];
  uint8_t cf[FC_R];
  bool cin[FC_R];
  uint32_t cidx[FC_R];
  bool open = false;
#pragma unroll
  for (int j = 0; j < FC_R; j++) {
    const int e = (int)threadIdx.x + NT * j;
    cin[j] = false;
    ca[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    cs1[j] = 0.0f;
    cf[j] = 0;
    cidx[j] = 0u;
    if (e < FC_N) {
      const int lz = e / (FC_X * FC_Y), r = e - lz * (FC_X * FC_Y), ly = r / FC_X, lx = r - ly * FC_X;
      const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = z0 - 1 + lz;
      cin[j] = inb(g, gx, gy, gz);
      cidx[j] = sidx(g, clampi(gx, 0, g.nx - 1), clampi(gy, 0, g.ny - 1), clampi(gz, 0, g.nz - 1));
      ca[j] = s0[cidx[j]];
      cs1[j] = s1[cidx[j]].x;
      cf[j] = fl[cidx[j]];
    }
  }
  const bool face = x < g.vx && y < g.vy && z <= g.nz;
  const uint32_t i = face ? vidx(g, x, y, z) : 0u;
  float uvw[3];
  uvw[0] = U[i];
  uvw[1] = V[i];
  uvw[2] = W[i];
#pragma unroll
  for (int j = 0; j < FC_R; j++) {
    gate[j] = confine_gate(P, ca[j].x, ca[j].w + cs1[j]);
    open |= cin[j] && gate[j] > 0.0f;
  }
  // confinement acts only where the gate is open (fluid.wgsl:167-181 returns
  // 0 otherwise): a block none of whose cells has it open stages no |curl|
  float4 cw[FC_R];
#pragma unroll
  for (int j = 0; j < FC_R; j++) cw[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (__syncthreads_or(open)) {
    float wl[FW_R];
#pragma unroll
    for (int j = 0; j < FW_R; j++) {
      const int e = (int)threadIdx.x + NT * j;
      wl[j] = 0.0f;
      if (e < FW_N) {
        const int lz = e / (FW_X * FW_Y), r = e - lz * (FW_X * FW_Y), ly = r / FW_X, lx = r - ly * FW_X;
        const int gx = clampi(x0 - 2 + lx, 0, g.nx - 1), gy = clampi(y0 - 2 + ly, 0, g.ny - 1),
                  gz = clampi(z0 - 2 + lz, 0, g.nz - 1);
        wl[j] = curl[sidx(g, gx, gy, gz)].w;
      }
    }
#pragma unroll
    for (int j = 0; j < FC_R; j++)
      if (cin[j] && gate[j] > 0.0f) cw[j] = curl[cidx[j]];
#pragma unroll
    for (int j = 0; j < FW_R; j++) {
      const int e = (int)threadIdx.x + NT * j;
      if (e < FW_N) sw[e] = wl[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < FC_R; j++) {
    const int e = (int)threadIdx.x + NT * j;
    if (e >= FC_N) continue;
    const int lz = e / (FC_X * FC_Y), r = e - lz * (FC_X * FC_Y), ly = r / FC_X, lx = r - ly * FC_X;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (cin[j]) confine_tile(P, gate[j], cw[j], sw + (lx + 1) + FW_X * ((ly + 1) + FW_Y * (lz + 1)), fx, fy, fz);
    sf[0][e] = fx;
    sf[1][e] = fy;
    sf[2][e] = fz;
    st[e] = ca[j].x;
    sm[e] = ca[j].w + cs1[j];
    sc[e] = (uint8_t)((cin[j] ? 1u : 0u) | (cin[j] && cf[j] == 1 ? 2u : 0u));
  }
  __syncthreads();
  if (!face) return;
  const float dt = P->dt;
  const int cb = (x - x0 + 1) + FC_X * ((y - y0 + 1) + FC_Y * (z - z0 + 1));
  const bool ib = sc[cb] & 1u;
  // the face's own cell (x, y, z) is the b side of all three components
  const float cbx = sf[0][cb], cby = sf[1][cb], cbz = sf[2][cb];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int hx = g.nx - 1 + (k == 0), hy = g.ny - 1 + (k == 1), hz = g.nz - 1 + (k == 2);
    if (x > hx || y > hy || z > hz) continue;
    const int ca_ = cb - (k == 0 ? 1 : (k == 1 ? FC_X : FC_X * FC_Y));
    if ((k == 1 && y == 0) || (sc[cb] & 2u) || (sc[ca_] & 2u)) continue;
    const bool ia = sc[ca_] & 1u;
    float f = 0.0f;
    if (k == 1) {
      float Ta = P->t0, Tb = P->t0, ma = 0.0f, mb = 0.0f;
      if (ia) { Ta = st[ca_]; ma = sm[ca_]; }
      if (ib) { Tb = st[cb]; mb = sm[cb]; }
      if (!ia) { Ta = Tb; ma = mb; }
      if (!ib) { Tb = Ta; mb = ma; }
      float Tf = 0.5f * (Ta + Tb);
      f += P->g * P->buoy * (Tf - P->t0) / Tf - P->g * P->smoke_w * 0.5f * (ma + mb);
    }
    float fcx = 0.0f, fcy = 0.0f, fcz = 0.0f, n = 0.0f;
    if (ia) {
      fcx += sf[0][ca_]; fcy += sf[1][ca_]; fcz += sf[2][ca_]; n += 1.0f;
    }
    if (ib) {
      fcx += cbx; fcy += cby; fcz += cbz; n += 1.0f;
    }
    if (n > 0.0f) f += (k == 0 ? fcx : (k == 1 ? fcy : fcz)) / n;
    float *C = k == 0 ? U : (k == 1 ? V : W);
    float u = uvw[k];
    u += f * dt;
    int cx = clampi(x, 0, g.nx - 1), cy = clampi(y, 0, g.ny - 1), cz = clampi(z, 0, g.nz - 1);
    float ss = spongeSide(P, g, cx, cz);
    if (k != 1) {
      float Tl = P->t0;
      if (ib) Tl = st[cb];
      float hot = smoothstep_(350.0f, 900.0f, Tl);
      float kr = P->wind_relax * (1.0f - 0.85f * hot) + 8.0f * ss;
      float wv = k == 0 ? P->wind_x : P->wind_z;
      u = mixf_(u, wv, 1.0f - pexp(-kr * dt));
    } else {
      float stp = spongeTop(P, g, cy);
      if (stp > 0.0f) u = mixf_(u, fmax_(u, 0.0f), 1.0f - pexp(-8.0f * stp * dt));
    }
    u = clampf_(u, -40.0f, 40.0f);
    C[i] = u;
  }
}

void launch_forces(const Fields &f, hipStream_t s) {
  k_forces<<<f.g.fbricks, NT, 0, s>>>(f.P, f.g, f.vb[0], f.vb[1], f.vb[2], f.s0, f.s1, f.curl, f.flags);
}

// ---------------------------------------------------------------- divergence
__global__ __launch_bounds__(NT) void k_divergence(const Params *__restrict__ P, Geo g,
                                                    const float *__restrict__ U, const float *__restrict__ V,
                                                    const float *__restrict__ W, const float4 *__restrict__ s1,
                                                    const uint8_t *__restrict__ fl, float *__restrict__ b) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= g.n) return;
  int x, y, z;
  decode_cell(g, i, x, y, z);
  uint32_t j = vidx(g, x, y, z);
  // the faces are loaded with the flag (not behind it): one round trip
  const uint8_t f = fl[i];
  const float u1 = U[j + 1], u0 = U[j], v1 = V[j + g.vx], v0 = V[j];
  const float w1 = W[j + (uint32_t)g.vx * (uint32_t)g.vy], w0 = W[j], sz = s1[i].z;
  if (f == 1) { b[i] = 0.0f; return; }
  float d = (u1 - u0 + v1 - v0 + w1 - w0) / P->dx;
  b[i] = d - sz;
}

void launch_divergence(const Fields &f, hipStream_t s) {
  k_divergence<<<nblocks(f.g.n), NT, 0, s>>>(f.P, f.g, f.vb[0], f.vb[1], f.vb[2], f.s1, f.flags, f.b0);
}

// ---------------------------------------------------------------- project
__global__ __launch_bounds__(NT) void k_project(const Params *__restrict__ P, Geo g, float *__restrict__ U,
                                                 float *__restrict__ V, float *__restrict__ W,
                                                 const float *__restrict__ phi, const uint8_t *__restrict__ fl) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= g.nv) return;
  int x, y, z;
  decode_face(g, i, x, y, z);
  const float dx = P->dx;
  // every load is issued first (a cell outside the grid reads cell 0 and is
  // selected away), then the three components are formed
  const bool ib = inb(g, x, y, z);
  const uint32_t ic = ib ? sidx(g, x, y, z) : 0u;
  const float pcl = phi[ic];
  const uint8_t fcl = fl[ic];
  float cv[3], pav[3];
  uint8_t fav[3];
  bool ia[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int ax = x - (k == 0), ay = y - (k == 1), az = z - (k == 2);
    ia[k] = inb(g, ax, ay, az);
    const uint32_t j = ia[k] ? sidx(g, ax, ay, az) : 0u;
    pav[k] = phi[j];
    fav[k] = fl[j];
    cv[k] = (k == 0 ? U : (k == 1 ? V : W))[i];
  }
  const float pc = ib ? pcl : 0.0f;
  const bool sc = ib && fcl == 1;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int hx = g.nx - 1 + (k == 0), hy = g.ny - 1 + (k == 1), hz = g.nz - 1 + (k == 2);
    if (x > hx || y > hy || z > hz) continue;
    float *C = k == 0 ? U : (k == 1 ? V : W);
    float u = 0.0f;
    if (!((k == 1 && y == 0) || sc || (ia[k] && fav[k] == 1))) {
      u = cv[k];
      const float pa = ia[k] ? pav[k] : 0.0f;
      u -= (pc - pa) / dx;
    }
    C[i] = u;
  }
}

void launch_project(const Fields &f, hipStream_t s) {
  k_project<<<nblocks(f.g.nv), NT, 0, s>>>(f.P, f.g, f.vb[0], f.vb[1], f.vb[2], f.phi0, f.flags);
}

// ---------------------------------------------------------------- step helpers
__global__ void k_set_step(Params *P, float dt, float time, float ign_on, uint32_t frame) {
  if (threadIdx.x == 0) {
    P->dt = dt;
    P->time = time;
    P->ign_on = ign_on;
    P->frame = frame;
  }
}
void launch_set_step(Params *P, float dt, float time, float ign_on, uint32_t frame, hipStream_t s) {
  k_set_step<<<1, 64, 0, s>>>(P, dt, time, ign_on, frame);
}

// Harness noise: U[-1,1] from the PCG hash of (component, face index, seed).
__global__ __launch_bounds__(NT) void k_noise(Geo g, float *U, float *V, float *W, float amp, uint32_t seed) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= g.nv) return;
  int x, y, z;
  decode_face(g, i, x, y, z);
  for (int k = 0; k < 3; k++) {
    int hx = g.nx - 1 + (k == 0), hy = g.ny - 1 + (k == 1), hz = g.nz - 1 + (k == 2);
    if (x > hx || y > hy || z > hz) continue;
    uint32_t h = pcg(pcg(i * 3u + (uint32_t)k) ^ (seed * 0x9e3779b9u));
    float u01 = (float)(h >> 8) * (1.0f / 16777216.0f);
    float nz = 2.0f * u01 - 1.0f;
    float *C = k == 0 ? U : (k == 1 ? V : W);
    C[i] += amp * nz;
  }
}
void launch_velocity_noise(const Fields &f, float amp, uint32_t seed, hipStream_t s) {
  k_noise<<<nblocks(f.g.nv), NT, 0, s>>>(f.g, f.va[0], f.va[1], f.va[2], amp, seed);
}

// ---------------------------------------------------------------- stats
// Deterministic combination of the per-block partials into the 1040 slots of
// stats.wgsl:2-3 (FP64; maxima are exact in any order).  The reductions are
// independent, so each runs in a block of its own, concurrently: blocks 0-4
// the react sums, 5 the react maximum, 6-8 the solid sums, 9 the curl
// maximum, 10-11 the residual sums, 12 the per-layer sums (and the unused
// slots 12-15).  Each keeps its one-block order: every thread sums its
// strided partials in order, then a fixed tree combines the threads.
// `ran` has a bit per stats-producing stage that ran in this step
// (STATS_RAN_*); the slots of a stage that did not run are 0, as they are in a
// step whose accumulators start at 0.
constexpr int STATS_JOBS = 13;
__global__ __launch_bounds__(1024) void k_stats_finalize(Geo g, StatParts st, uint32_t ran) {
  __shared__ double sh[1024];
  __shared__ float shf[1024];
  const int t = threadIdx.x, job = blockIdx.x;
  double *out = st.out;
  const bool react = ran & STATS_RAN_REACT, solid = ran & STATS_RAN_SOLID;
  const bool curl = ran & STATS_RAN_CURL, resid = ran & STATS_RAN_RESID;
  if (job == 12) {
    // layer sums 16+y: one thread per layer, blocks in order
    for (int y = t; y < 1024; y += 1024)
      out[16 + y] = y < g.ny && react ? seq_sum(st.react, st.react_gx, (size_t)y * st.react_gx * 6 + 5, 6) : 0.0;
    if (t >= 12 && t < 16) out[t] = 0.0;
    return;
  }
  auto tree = [&](double v) -> double {
    sh[t] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if (t < s) sh[t] += sh[t + s];
      __syncthreads();
    }
    return sh[0];
  };
  auto treemax = [&](float v) -> float {
    shf[t] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if (t < s) shf[t] = fmax_(shf[t], shf[t + s]);
      __syncthreads();
    }
    return shf[0];
  };
  const int nreact = g.ny * st.react_gx;
  // the block-uniform `ran` tests select what a block sums; every thread still
  // reaches each barrier of the tree
  if (job < 5) {
    const int slot_r[5] = {0, 1, 3, 6, 8};
    const double a =
        react ? seq_sum(st.react, strided_count(t, nreact, 1024), (size_t)t * 6 + job, 6 * 1024) : 0.0;
    const double r = tree(a);
    if (t == 0) out[slot_r[job]] = r;
  } else if (job == 5) {
    const float a = react ? seq_max(st.react_t, strided_count(t, nreact, 1024), (size_t)t, 1024) : 0.0f;
    const float r = treemax(a);
    if (t == 0) out[4] = (double)r;
  } else if (job < 9) {
    const int q = job - 6, slot_s[3] = {2, 7, 9};
    const double a =
        solid ? seq_sum(st.solid, strided_count(t, st.n_solid, 1024), (size_t)t * 3 + q, 3 * 1024) : 0.0;
    const double r = tree(a);
    if (t == 0) out[slot_s[q]] = r;
  } else if (job == 9) {
    const float a = curl ? seq_max(st.curl, strided_count(t, st.n_curl, 1024), (size_t)t, 1024) : 0.0f;
    const float r = treemax(a);
    if (t == 0) out[5] = (double)r;
  } else {
    const int q = job - 10;
    const double a =
        resid ? seq_sum(st.resid, strided_count(t, st.n_resid, 1024), (size_t)t * 2 + q, 2 * 1024) : 0.0;
    const double r = tree(a);
    if (t == 0) out[10 + q] = r;
  }
}
void launch_stats_finalize(const Geo &g, const StatParts &st, uint32_t ran, hipStream_t s) {
  k_stats_finalize<<<STATS_JOBS, 1024, 0, s>>>(g, st, ran);
}

// ---------------------------------------------------------------- copy
// Plain streaming copy (16 B per lane): the calibration kernel for HBM byte
// counters (FETCH_SIZE / WRITE_SIZE) and achievable bandwidth.
__global__ __launch_bounds__(NT) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, uint32_t n4) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}
void launch_copy(const float4 *src, float4 *dst, uint32_t n4, hipStream_t s) {
  k_copy<<<nblocks(n4), NT, 0, s>>>(src, dst, n4);
}

// ---------------------------------------------------------------- debug math
__global__ void k_debug_math(int fn, const float *in, const float *in2, float *out, uint32_t n) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  float x = in[i];
  out[i] = fn == 0 ? pexp(x) : (fn == 1 ? sqrtf(x) : x / in2[i]);
}
void launch_debug_math(int fn, const float *in, const float *in2, float *out, uint32_t n, hipStream_t s) {
  k_debug_math<<<nblocks(n), NT, 0, s>>>(fn, in, in2, out, n);
}

}  // namespace pyro
