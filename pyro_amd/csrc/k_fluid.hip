// k_fluid.hip -- vorticity, body forces, divergence and projection on gfx950.
//
// Reference passes: fluid.wgsl curl (:3-29), forces (:31-109), divergence
// (:111-128), project (:130-152).  One thread per cell (curl, divergence) or
// per padded face triple (forces, project); the velocity components live in
// the reference's padded (nx+1)(ny+1)(nz+1) layout, x fastest.
#include "kernels.hpp"

namespace pyro {

// ---------------------------------------------------------------- curl
// Face-averaged velocity at the clamped cell c (fluid.wgsl:134-138).
__device__ __forceinline__ void uc(const Geo &g, const float *__restrict__ U, const float *__restrict__ V,
                                   const float *__restrict__ W, int x, int y, int z, float &ox, float &oy,
                                   float &oz) {
  x = clampi(x, 0, g.nx - 1);
  y = clampi(y, 0, g.ny - 1);
  z = clampi(z, 0, g.nz - 1);
  uint32_t i = vidx(g, x, y, z);
  ox = 0.5f * (U[i] + U[i + 1]);
  oy = 0.5f * (V[i] + V[i + g.vx]);
  oz = 0.5f * (W[i] + W[i + (uint32_t)g.vx * (uint32_t)g.vy]);
}

template <bool STATS>
__global__ __launch_bounds__(NT) void k_curl(const Params *__restrict__ P, Geo g, const float *__restrict__ U,
                                              const float *__restrict__ V, const float *__restrict__ W,
                                              float4 *__restrict__ curl, float *__restrict__ part) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  float spd = 0.0f;
  if (i < g.n) {
    int x, y, z;
    decode_cell(g, i, x, y, z);
    float tdx = 2.0f * P->dx;
    float pyx, pyy, pyz, myx, myy, myz, pzx, pzy, pzz, mzx, mzy, mzz, pxx, pxy, pxz, mxx, mxy, mxz;
    uc(g, U, V, W, x, y + 1, z, pyx, pyy, pyz);
    uc(g, U, V, W, x, y - 1, z, myx, myy, myz);
    uc(g, U, V, W, x, y, z + 1, pzx, pzy, pzz);
    uc(g, U, V, W, x, y, z - 1, mzx, mzy, mzz);
    uc(g, U, V, W, x + 1, y, z, pxx, pxy, pxz);
    uc(g, U, V, W, x - 1, y, z, mxx, mxy, mxz);
    float wx = ((pyz - myz) - (pzy - mzy)) / tdx;
    float wy = ((pzx - mzx) - (pxz - mxz)) / tdx;
    float wz = ((pxy - mxy) - (pyx - myx)) / tdx;
    curl[i] = make_float4(wx, wy, wz, len3(wx, wy, wz));
    if (STATS) {
      float cx, cy, cz;
      uc(g, U, V, W, x, y, z, cx, cy, cz);
      spd = len3(cx, cy, cz);
    }
  }
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
__device__ __forceinline__ bool solidAt(const Geo &g, const uint8_t *__restrict__ fl, int x, int y, int z) {
  return inb(g, x, y, z) && fl[sidx(g, x, y, z)] == 1;
}
__device__ __forceinline__ float cmag(const Geo &g, const float4 *__restrict__ curl, int x, int y, int z) {
  return curl[sidx(g, clampi(x, 0, g.nx - 1), clampi(y, 0, g.ny - 1), clampi(z, 0, g.nz - 1))].w;
}
// Vorticity confinement at a cell centre (fluid.wgsl:167-181).
__device__ void confine(const Params *__restrict__ P, const Geo &g, const float4 *__restrict__ s0,
                        const float4 *__restrict__ s1, const float4 *__restrict__ curl, int x, int y, int z,
                        float &fx, float &fy, float &fz) {
  fx = fy = fz = 0.0f;
  uint32_t i = sidx(g, x, y, z);
  float4 a = s0[i];
  float T = a.x;
  float smoke = a.w + s1[i].x;
  float gate = fmax_(smoothstep_(P->vort_t, P->vort_t + 500.0f, T), 0.35f * smoothstep_(0.0f, 2e-4f, smoke));
  if (gate <= 0.0f) return;
  float gx = cmag(g, curl, x + 1, y, z) - cmag(g, curl, x - 1, y, z);
  float gy = cmag(g, curl, x, y + 1, z) - cmag(g, curl, x, y - 1, z);
  float gz = cmag(g, curl, x, y, z + 1) - cmag(g, curl, x, y, z - 1);
  float gl = len3(gx, gy, gz);
  if (gl < 1e-6f) return;
  float nx = gx / gl, ny = gy / gl, nz = gz / gl;
  float4 w = curl[i];
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

__global__ __launch_bounds__(NT) void k_forces(const Params *__restrict__ P, Geo g, float *__restrict__ U,
                                                float *__restrict__ V, float *__restrict__ W,
                                                const float4 *__restrict__ s0, const float4 *__restrict__ s1,
                                                const float4 *__restrict__ curl, const uint8_t *__restrict__ fl) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= g.nv) return;
  int x, y, z;
  decode_face(g, i, x, y, z);
  const float dt = P->dt;
  // the face's own cell (x, y, z) is the b side of all three components:
  // its confinement force is evaluated once
  float cbx = 0.0f, cby = 0.0f, cbz = 0.0f;
  if (inb(g, x, y, z)) confine(P, g, s0, s1, curl, x, y, z, cbx, cby, cbz);
  for (int k = 0; k < 3; k++) {
    int hx = g.nx - 1 + (k == 0), hy = g.ny - 1 + (k == 1), hz = g.nz - 1 + (k == 2);
    if (x > hx || y > hy || z > hz) continue;
    int ax = x - (k == 0), ay = y - (k == 1), az = z - (k == 2);
    if ((k == 1 && y == 0) || solidAt(g, fl, x, y, z) || solidAt(g, fl, ax, ay, az)) continue;
    bool ia = inb(g, ax, ay, az), ib = inb(g, x, y, z);
    float f = 0.0f;
    if (k == 1) {
      float Ta = P->t0, Tb = P->t0, ma = 0.0f, mb = 0.0f;
      if (ia) { uint32_t j = sidx(g, ax, ay, az); float4 q = s0[j]; Ta = q.x; ma = q.w + s1[j].x; }
      if (ib) { uint32_t j = sidx(g, x, y, z); float4 q = s0[j]; Tb = q.x; mb = q.w + s1[j].x; }
      if (!ia) { Ta = Tb; ma = mb; }
      if (!ib) { Tb = Ta; mb = ma; }
      float Tf = 0.5f * (Ta + Tb);
      f += P->g * P->buoy * (Tf - P->t0) / Tf - P->g * P->smoke_w * 0.5f * (ma + mb);
    }
    float fcx = 0.0f, fcy = 0.0f, fcz = 0.0f, n = 0.0f;
    if (ia) {
      float cx, cy, cz;
      confine(P, g, s0, s1, curl, ax, ay, az, cx, cy, cz);
      fcx += cx; fcy += cy; fcz += cz; n += 1.0f;
    }
    if (ib) {
      fcx += cbx; fcy += cby; fcz += cbz; n += 1.0f;
    }
    if (n > 0.0f) f += (k == 0 ? fcx : (k == 1 ? fcy : fcz)) / n;
    float *C = k == 0 ? U : (k == 1 ? V : W);
    float u = C[i];
    u += f * dt;
    int cx = clampi(x, 0, g.nx - 1), cy = clampi(y, 0, g.ny - 1), cz = clampi(z, 0, g.nz - 1);
    float ss = spongeSide(P, g, cx, cz);
    if (k != 1) {
      float Tl = P->t0;
      if (ib) Tl = s0[sidx(g, x, y, z)].x;
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
  k_forces<<<nblocks(f.g.nv), NT, 0, s>>>(f.P, f.g, f.vb[0], f.vb[1], f.vb[2], f.s0, f.s1, f.curl, f.flags);
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
