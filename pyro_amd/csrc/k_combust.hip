// k_combust.hip -- pyrolysis (solid) and gas-phase reaction (react) on gfx950.
//
// Reference passes: combust.wgsl solid (:3-137) with rayIrr (:18-39) and
// faceIrr (:41-48); react (:139-273).  Statistics (stats.wgsl slot map) are
// reduced per block in FP64 in a fixed order and combined by
// k_stats_finalize, so every stats step is bit-reproducible.
#include "kernels.hpp"
#include "curl.hpp"

namespace pyro {

__device__ __forceinline__ float kapP(const Params *__restrict__ P, float T, float yo, float ys) {
  float fv = rhoOf(P, T) * fmax_(ys, 0.0f) / RHO_SOOT;
  return 1862.0f * fv * T + P->kap_gas * clampf_(1.0f - yo / P->yo2, 0.0f, 1.0f);
}
__device__ __forceinline__ float cpOf(float T) { return clampf_(1000.0f + 0.24f * (T - 300.0f), 1000.0f, 1450.0f); }
// Material slot of a SOL cell; slots beyond the 8-entry table are clamped to
// the last entry so that malformed input cannot read past Params.
__device__ __forceinline__ uint32_t slotOf(float4 s) {
  uint32_t k = (uint32_t)s.x;
  return k > 7u ? 7u : k;
}
__device__ __forceinline__ uint32_t kindOf(const Params *__restrict__ P, float4 s) {
  return (uint32_t)(P->mats[slotOf(s)].a[3] + 0.5f);
}
__device__ __forceinline__ bool isSolidS(const Params *__restrict__ P, float4 s) {
  if (kindOf(P, s) == 2u) return s.y > 0.5f;
  return s.y + s.z > P->burnout;
}

// Radiative transfer along one ray of 10 strided steps (combust.wgsl:18-39).
__device__ __forceinline__ float rayIrr(const Params *__restrict__ P, const Geo &g, const float4 *__restrict__ solIn,
                        const float4 *__restrict__ s0, int px, int py, int pz, int dx_, int dy_, int dz_) {
  // strides 1,1,1,2,2,3,4,5,7,9 packed one per nibble (no private array)
  const uint64_t strides = 0x9754322111ull;
  // The sample points do not depend on the marching, so each half of the ray
  // (five samples) is loaded before it is marched (a point outside the grid
  // reads cell 0, and the march stops before it): two round trips per ray
  // instead of ten.
  float tr = 1.0f, acc = 0.0f;
  const float dl = len3((float)dx_, (float)dy_, (float)dz_) * P->dx;
  bool done = false;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    float4 so[5], q[5];
    bool in[5];
#pragma unroll
    for (int u = 0; u < 5; u++) {
      const int st = 5 * h + u;
      in[u] = inb(g, px, py, pz);
      const uint32_t j = in[u] ? sidx(g, px, py, pz) : 0u;
      so[u] = solIn[j];
      q[u] = s0[j];
      const int sd = (int)((strides >> (4 * st)) & 15u);
      px += dx_ * sd;
      py += dy_ * sd;
      pz += dz_ * sd;
    }
#pragma unroll
    for (int u = 0; u < 5; u++) {
      const int st = 5 * h + u;
      if (done || !in[u]) { done = true; break; }
      if (isSolidS(P, so[u])) {
        acc += tr * pow4(so[u].w);
        tr = 0.0f;
        done = true;
        break;
      }
      const float T = q[u].x;
      const int sd = (int)((strides >> (4 * st)) & 15u);
      const float len = (float)sd * dl;
      const float e = 1.0f - pexp(-kapP(P, T, q[u].z, q[u].w) * len);
      acc += tr * e * pow4(T);
      tr *= 1.0f - e;
    }
    if (done) break;
  }
  acc += tr * pow4(P->t0);
  return SIGMA * acc;
}
__device__ __forceinline__ void dirOf(int k, int &x, int &y, int &z) {
  x = k == 0 ? 1 : (k == 1 ? -1 : 0);
  y = k == 2 ? 1 : (k == 3 ? -1 : 0);
  z = k == 4 ? 1 : (k == 5 ? -1 : 0);
}
// Direction of ray r of face k (combust.wgsl:41-48): the normal (r = 0),
// then the normal plus and minus the first tangent (r = 1, 2) and plus and
// minus the second (r = 3, 4).
__device__ __forceinline__ void rayDir(int k, int r, int &dx_, int &dy_, int &dz_) {
  dirOf(k, dx_, dy_, dz_);
  int t1x = 0, t1y = 1, t1z = 0, t2x = 0, t2y = 0, t2z = 1;
  if (dy_ != 0) { t1x = 1; t1y = 0; t1z = 0; }
  if (dz_ != 0) { t2x = 1; t2y = 0; t2z = 0; }
  const int s = (r == 1 || r == 3) ? 1 : ((r == 2 || r == 4) ? -1 : 0);
  const bool second = r >= 3;
  dx_ += s * (second ? t2x : t1x);
  dy_ += s * (second ? t2y : t1y);
  dz_ += s * (second ? t2z : t1z);
}
// Faces of a solid cell whose irradiance the cell needs (combust.wgsl:60-75):
// bit k set for a neighbour in direction k (dirOf) that is inside the grid,
// not below the ground and not solid.  Inert cells (kind 2) need none.
__device__ __forceinline__ uint32_t ray_faces(const Params *__restrict__ P, const Geo &g,
                                              const float4 *__restrict__ solIn, int x, int y, int z, float4 so) {
  if (kindOf(P, so) == 2u) return 0u;
  uint32_t mask = 0u;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int dx_, dy_, dz_;
    dirOf(k, dx_, dy_, dz_);
    const int nx_ = x + dx_, ny_ = y + dy_, nz_ = z + dz_;
    if (ny_ < 0 || !inb(g, nx_, ny_, nz_)) continue;
    if (!isSolidS(P, solIn[sidx(g, nx_, ny_, nz_)])) mask |= 1u << k;
  }
  return mask;
}

// Surface energy balance and pyrolysis of one solid cell (combust.wgsl:60-128).
// The irradiance of the cell's ray faces (ray_faces) is given: face k at
// irr[popc(mask & ((1 << k) - 1))].
__device__ __forceinline__ void solid_cell(const Params *__restrict__ P, const Geo &g, const float4 *__restrict__ solIn,
                                        const float4 *__restrict__ s0, int x, int y, int z, float4 so,
                                        const float *irr, uint32_t mask, float4 &o, float4 &po, uint32_t &fl,
                                        float &stPyro, float &stMass, float &stArea) {
  fl = 1;
  const Mat &m = P->mats[slotOf(so)];
  uint32_t kind = kindOf(P, so);
  float fuel = so.y, chr = so.z, Ts = so.w;
  if (kind != 2u) {
    float nF = 0.0f, Tg = 0.0f, Yo = 0.0f, qin = 0.0f, qcond = 0.0f;
    for (int k = 0; k < 6; k++) {
      int dx_, dy_, dz_;
      dirOf(k, dx_, dy_, dz_);
      int nx_ = x + dx_, ny_ = y + dy_, nz_ = z + dz_;
      if (ny_ < 0) continue;
      bool nin = inb(g, nx_, ny_, nz_);
      if (nin) {
        float4 sn = solIn[sidx(g, nx_, ny_, nz_)];
        if (isSolidS(P, sn)) {
          qcond += m.b[3] / P->dx * (sn.w - Ts);
          continue;
        }
      }
      nF += 1.0f;
      if (nin) {
        float4 q = s0[sidx(g, nx_, ny_, nz_)];
        Tg += q.x;
        Yo += q.z;
        qin += irr[__popc(mask & ((1u << k) - 1u))];
      } else {
        Tg += P->t0;
        Yo += P->yo2;
        qin += SIGMA * pow4(P->t0);
      }
    }
    float layer = kind == 1u ? 0.001f : fmin_(P->layer, P->dx);
    float cap = m.a[0] * m.b[2] * layer;
    if (nF > 0.0f) {
      Tg /= nF;
      Yo /= nF;
      qin /= nF;
      float qnet = m.b[1] * (qin - SIGMA * pow4(Ts)) + P->h_conv * (Tg - Ts) + qcond / nF;
      if (P->ign_on > 0.5f) {
        float dI = len3(((float)x + 0.5f) * P->dx - P->ign_x, ((float)y + 0.5f) * P->dx - P->ign_y,
                        ((float)z + 0.5f) * P->dx - P->ign_z);
        qnet += P->ign_flux * clampf_(1.0f - dI / (1.5f * P->ign_r), 0.0f, 1.0f);
      }
      float mv = 0.0f, mc = 0.0f;
      if (kind == 1u) {
        float Tb = m.a[2];
        mv = smoothstep_(Tb - 60.0f, Tb, Ts) * fmax_(qnet, 0.0f) / m.a[1] * m.c[1];
        Ts = fmin_(Ts + P->dt * fmax_(qnet - mv * m.a[1], fmin_(qnet, 0.0f)) / cap, Tb);
        fuel -= mv * nF / P->dx * P->dt;
      } else {
        if (fuel > 0.0f) mv = smoothstep_(m.a[2] - 20.0f, m.a[2] + 30.0f, Ts) * fmax_(qnet, 0.0f) / m.a[1] * m.c[1];
        float chi = kind == 0u ? m.b[0] : 0.0f;
        float kin = P->char_a * pexp(-P->char_e / (RGAS * Ts)) * chr * layer;
        float dif = P->h_mass * rhoOf(P, Tg) * Yo / 1.33f;
        mc = fmin_(kin, dif);
        Ts += P->dt * (qnet - mv * m.a[1] + 0.5f * mc * P->dh_char) / cap;
        float virgin = fmin_(mv / (1.0f - chi) * nF / P->dx * P->dt, fuel);
        fuel -= virgin;
        chr += virgin * chi - mc * nF / P->dx * P->dt;
      }
      stPyro = mv * nF * P->dx * P->dx;
      stArea = mv > 1e-4f ? nF * P->dx * P->dx : 0.0f;
      po = make_float4(mv, mc, nF, 0.0f);
    } else {
      Ts += P->dt * qcond / (m.a[0] * m.b[2] * P->dx);
    }
    Ts = clampf_(Ts, 250.0f, 2000.0f);
    fuel = fmax_(fuel, 0.0f);
    chr = fmax_(chr, 0.0f);
    o = make_float4(so.x, fuel, chr, Ts);
    if (fuel + chr <= P->burnout) {
      o = make_float4(0.0f, 0.0f, 0.0f, P->t0);
      fl = 0;
      po = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  stMass = (o.y + o.z) * P->dx * P->dx * P->dx * (kind == 2u ? 0.0f : 1.0f);
}

// Sparse form of the per-cell pass of combust.wgsl:50-132.  Two facts hold at
// the start of every step (set up by launch_init_flags and kept by this
// kernel): flags[i] == 1 exactly when sol_in[i] is solid, and a cell that is
// not solid holds the same SOL value in both ping-pong buffers (solids only
// ever disappear).  So a cell with flag 0 needs no work: the reference would
// copy SOL unchanged, write pyr = 0 and flag 0, all already in place.  A cell
// that burns out gets flag FLAG_BURNT (read as "not solid" by every
// consumer, which test == 1); on the next step its SOL value is copied into
// the other buffer and the flag becomes 0.
constexpr uint8_t FLAG_BURNT = 2;

// The ray marches of a block's solid cells are dealt out one ray per
// thread: each thread lists its cell's ray faces in LDS, the block's threads
// take the listed faces' rays in turn (5 rays of 10 samples per face), each
// face sums its rays as faceIrr does (combust.wgsl:41-48), and each cell then
// reads its faces' irradiance back in its own order.  A face's value goes to
// its own slot, so the listing order does not matter.  (One thread per cell
// marched up to 6 faces in turn, and its wave every face any of its cells
// needed.)
template <bool STATS>
__device__ __forceinline__ void solid_block(const Params *__restrict__ P, const Geo &g,
                                            const float4 *__restrict__ solIn, float4 *__restrict__ solOut,
                                            float4 *__restrict__ pyr, uint8_t *__restrict__ flags,
                                            const float4 *__restrict__ s0, double *__restrict__ part) {
  __shared__ uint16_t sfm[6 * NT];  // listed face: thread << 3 | k
  __shared__ float sfi[6 * NT];     // its irradiance
  __shared__ float sray[5 * NT];    // the rays of a batch of NT faces
  __shared__ int snf;
  const uint32_t c0 = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT;
  const uint32_t i = c0 + threadIdx.x;
  float stPyro = 0.0f, stMass = 0.0f, stArea = 0.0f;
  if (threadIdx.x == 0) snf = 0;
  const uint8_t fprev = i < g.n ? flags[i] : 0;
  int x = 0, y = 0, z = 0;
  float4 so = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t mask = 0u;
  if (fprev == 1) {
    so = solIn[i];
    decode_cell(g, i, x, y, z);
    mask = ray_faces(P, g, solIn, x, y, z, so);
  }
  // snf = 0 is visible after this barrier; a block without solid or burnt
  // cells has nothing to do (its statistics partials are +0)
  if (!__syncthreads_or(fprev != 0)) {
    if (STATS && threadIdx.x == 0) {
      part[blockIdx.x * 3 + 0] = 0.0;
      part[blockIdx.x * 3 + 1] = 0.0;
      part[blockIdx.x * 3 + 2] = 0.0;
    }
    return;
  }
  int base = 0;
  if (mask) {
    base = atomicAdd(&snf, __popc(mask));
    int s = base;
#pragma unroll
    for (int k = 0; k < 6; k++)
      if (mask & (1u << k)) sfm[s++] = (uint16_t)((threadIdx.x << 3) | k);
  }
  __syncthreads();
  // the listed faces in batches of NT: the batch's rays one per thread, then
  // each face's five rays summed in faceIrr's order
  const int nf = snf;
  for (int f0 = 0; f0 < nf; f0 += NT) {
    const int nr = 5 * (nf - f0 < NT ? nf - f0 : NT);
    for (int j = threadIdx.x; j < nr; j += NT) {
      const int f = f0 + j / 5, r = j - 5 * (j / 5);
      const int owner = sfm[f] >> 3, k = sfm[f] & 7;
      int ox, oy, oz, nx_, ny_, nz_, dx_, dy_, dz_;
      decode_cell(g, c0 + (uint32_t)owner, ox, oy, oz);
      dirOf(k, nx_, ny_, nz_);
      rayDir(k, r, dx_, dy_, dz_);
      sray[j] = rayIrr(P, g, solIn, s0, ox + nx_, oy + ny_, oz + nz_, dx_, dy_, dz_);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nr / 5; j += NT) {
      const float *q = sray + 5 * j;
      sfi[f0 + j] = 0.2f * ((((q[0] + q[1]) + q[2]) + q[3]) + q[4]);
    }
    __syncthreads();
  }
  if (i < g.n) {
    if (fprev == FLAG_BURNT) {
      solOut[i] = solIn[i];
      pyr[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      flags[i] = 0;
    } else if (fprev == 1) {
      float4 o = so, po = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      uint32_t fl = 0;
      solid_cell(P, g, solIn, s0, x, y, z, so, sfi + base, mask, o, po, fl, stPyro, stMass, stArea);
      solOut[i] = o;
      pyr[i] = po;
      flags[i] = fl ? 1 : FLAG_BURNT;
    }
  }
  if (STATS) {
    __shared__ double sh[NT / 64];
    double a = block_sum<NT>((double)stPyro, sh);
    double b = block_sum<NT>((double)stMass, sh);
    double c = block_sum<NT>((double)stArea, sh);
    if (threadIdx.x == 0) {
      part[blockIdx.x * 3 + 0] = a;
      part[blockIdx.x * 3 + 1] = b;
      part[blockIdx.x * 3 + 2] = c;
    }
  }
}

template <bool STATS>
__global__ __launch_bounds__(NT) void k_solid(const Params *__restrict__ P, Geo g, const float4 *__restrict__ solIn,
                                               float4 *__restrict__ solOut, float4 *__restrict__ pyr,
                                               uint8_t *__restrict__ flags, const float4 *__restrict__ s0,
                                               double *__restrict__ part) {
  solid_block<STATS>(P, g, solIn, solOut, pyr, flags, s0, part);
}

// The step's solid and curl stages in one launch (the same block and cell
// mapping): curl reads only the velocity and writes only curl, which solid
// neither reads nor writes, so each block computes its cells' vorticity and
// then runs the solid pass; the dense, memory-bound curl fills the CUs that
// wait on the few blocks with ray marches.
template <bool STATS>
__global__ __launch_bounds__(NT) void k_solid_curl(const Params *__restrict__ P, Geo g,
                                                    const float4 *__restrict__ solIn, float4 *__restrict__ solOut,
                                                    float4 *__restrict__ pyr, uint8_t *__restrict__ flags,
                                                    const float4 *__restrict__ s0, double *__restrict__ part,
                                                    const float *__restrict__ U, const float *__restrict__ V,
                                                    const float *__restrict__ W, float4 *__restrict__ curl,
                                                    float *__restrict__ cpart) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  float spd = 0.0f;
  if (i < g.n) spd = curl_cell<STATS>(P, g, U, V, W, curl, i);
  if (STATS) {
    __shared__ float sh[NT / 64];
    const float m = block_max<NT>(spd, sh);
    if (threadIdx.x == 0) cpart[blockIdx.x] = m;
  }
  solid_block<STATS>(P, g, solIn, solOut, pyr, flags, s0, part);
}
void launch_solid_curl(const Fields &f, StatParts *st, hipStream_t s) {
  const int nb = nblocks(f.g.n);
  if (st)
    k_solid_curl<true><<<nb, NT, 0, s>>>(f.P, f.g, f.sol_in, f.sol_out, f.pyr, f.flags, f.s0, st->solid, f.va[0],
                                         f.va[1], f.va[2], f.curl, st->curl);
  else
    k_solid_curl<false><<<nb, NT, 0, s>>>(f.P, f.g, f.sol_in, f.sol_out, f.pyr, f.flags, f.s0, nullptr, f.va[0],
                                          f.va[1], f.va[2], f.curl, nullptr);
}

void launch_solid(const Fields &f, StatParts *st, hipStream_t s) {
  int nb = nblocks(f.g.n);
  if (st)
    k_solid<true><<<nb, NT, 0, s>>>(f.P, f.g, f.sol_in, f.sol_out, f.pyr, f.flags, f.s0, st->solid);
  else
    k_solid<false><<<nb, NT, 0, s>>>(f.P, f.g, f.sol_in, f.sol_out, f.pyr, f.flags, f.s0, nullptr);
}

// flags = isSolidS(sol) (combust.wgsl:11-14), establishing the invariant above.
__global__ __launch_bounds__(NT) void k_init_flags(const Params *__restrict__ P, Geo g,
                                                    const float4 *__restrict__ sol, uint8_t *__restrict__ flags) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < g.n) flags[i] = isSolidS(P, sol[i]) ? 1 : 0;
}
void launch_init_flags(const Params *P, const Geo &g, const float4 *sol, uint8_t *flags, hipStream_t s) {
  k_init_flags<<<nblocks(g.n), NT, 0, s>>>(P, g, sol, flags);
}

// ---------------------------------------------------------------- react
__device__ __forceinline__ float spongeSideR(const Params *__restrict__ P, const Geo &g, int x, int z) {
  int e = x < g.nx - 1 - x ? x : g.nx - 1 - x;
  int e2 = z < g.nz - 1 - z ? z : g.nz - 1 - z;
  if (e2 < e) e = e2;
  return 1.0f - smoothstep_(0.0f, P->sponge, (float)e);
}
__device__ __forceinline__ float spongeTopR(const Params *__restrict__ P, const Geo &g, int y) {
  float h = (float)g.ny;
  return smoothstep_(h - 1.5f * P->sponge, h - 1.0f, (float)y);
}

// Blocks tile one y-layer each: grid = (ceil(nx*nz / NT), ny), so the
// per-layer heat release (stats slots 16+y) is a sum of whole blocks.
template <bool STATS>
__global__ __launch_bounds__(NT) void k_react(const Params *__restrict__ P, Geo g, const float4 *__restrict__ s0in,
                                               const float4 *__restrict__ s1in, float4 *__restrict__ s0out,
                                               float4 *__restrict__ s1out, const uint8_t *__restrict__ flags,
                                               const float4 *__restrict__ sol, const float4 *__restrict__ pyr,
                                               const float4 *__restrict__ curl, double *__restrict__ part,
                                               float *__restrict__ part_t) {
  const int lb = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int y = lb / (int)gridDim.x, bxl = lb - y * (int)gridDim.x;
  const uint32_t i2 = (uint32_t)bxl * NT + threadIdx.x;
  float stH = 0.0f, stR = 0.0f, stS = 0.0f, stF = 0.0f, stM = 0.0f, stT = 0.0f, stL = 0.0f;
  if (i2 < (uint32_t)g.nx * (uint32_t)g.nz) {
    uint32_t zq = fdiv(i2, g.dnx);
    const int x = (int)(i2 - zq * (uint32_t)g.nx), z = (int)zq;
    const uint32_t i = sidx(g, x, y, z);
    const float dx = P->dx, dt = P->dt;
    if (flags[i] == 1) {
      s0out[i] = make_float4(sol[i].w, 0.0f, P->yo2, 0.0f);
      s1out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      float4 a = s0in[i];
      float4 b = s1in[i];
      float nut = (P->cs * dx) * (P->cs * dx) * curl[i].w;
      if (P->cs > 0.0f) {
        float kd = fmin_(nut / 0.5f * dt / (dx * dx), 0.15f);
        // every neighbour is loaded (one outside the grid at the cell itself,
        // always in range) so the loads are in flight together; a skipped
        // neighbour (solid, or below the ground) adds +0, which leaves these
        // sums unchanged: a running sum started at +0 is never -0
        float lx = 0.0f, ly = 0.0f, lz = 0.0f, lw = 0.0f, lb = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; k++) {
          int ddx, ddy, ddz;
          dirOf(k, ddx, ddy, ddz);
          const int nx_ = x + ddx, ny_ = y + ddy, nz_ = z + ddz;
          const bool in = inb(g, nx_, ny_, nz_);
          const uint32_t j = in ? sidx(g, nx_, ny_, nz_) : i;
          const float4 q = s0in[j];
          const float qb = s1in[j].x;
          const bool fluid = in && flags[j] != 1;
          const bool amb = !in && ny_ >= 0;  // open side or top: ambient air
          lx += fluid ? q.x - a.x : (amb ? P->t0 - a.x : 0.0f);
          ly += fluid ? q.y - a.y : (amb ? 0.0f - a.y : 0.0f);
          lz += fluid ? q.z - a.z : (amb ? P->yo2 - a.z : 0.0f);
          lw += fluid ? q.w - a.w : (amb ? 0.0f - a.w : 0.0f);
          lb += fluid ? qb - b.x : (amb ? -b.x : 0.0f);
        }
        a.x += kd * lx; a.y += kd * ly; a.z += kd * lz; a.w += kd * lw;
        b.x += kd * lb;
      }
      float T = a.x;
      float yf = fmax_(a.y, 0.0f);
      float yo = clampf_(a.z, 0.0f, P->yo2);
      float ys = fmax_(a.w, 0.0f);
      float ym = fmax_(b.x, 0.0f);
      float S = 0.0f;
      for (int k = 0; k < 6; k++) {
        int ddx, ddy, ddz;
        dirOf(k, ddx, ddy, ddz);
        int nx_ = x + ddx, ny_ = y + ddy, nz_ = z + ddz;
        if (!inb(g, nx_, ny_, nz_)) continue;
        uint32_t j = sidx(g, nx_, ny_, nz_);
        if (flags[j] != 1) continue;
        float4 pr = pyr[j];
        float Ts = sol[j].w;
        float rho = rhoOf(P, T);
        if (pr.x > 0.0f) {
          float dm = pr.x * dt / (rho * dx);
          float inv = 1.0f / (1.0f + dm);
          yf = (yf + dm) * inv;
          yo *= inv;
          ys *= inv;
          ym *= inv;
          T = (T + dm * Ts) * inv;
          S += pr.x / (rho * dx);
        }
        T += P->h_conv * (Ts - T) * dt / (rho * cpOf(T) * dx);
        if (pr.y > 0.0f) {
          yo = fmax_(yo - 1.33f * pr.y * dt / (rho * dx), 0.0f);
          T += 0.5f * P->dh_char * pr.y * dt / (rho * cpOf(T) * dx);
        }
      }
      if (P->ign_on > 0.5f) {
        float dI = len3(((float)x + 0.5f) * dx - P->ign_x, ((float)y + 0.5f) * dx - P->ign_y,
                        ((float)z + 0.5f) * dx - P->ign_z);
        float w = clampf_(1.0f - dI / P->ign_r, 0.0f, 1.0f);
        if (w > 0.0f) {
          T = fmax_(T, mixf_(P->t0, P->ign_temp, smoothstep_(0.0f, 0.5f, w)));
          yf += P->ign_fuel * w * dt;
        }
      }
      float rho = rhoOf(P, T);
      float tauU = 0.049f * dx * dx / (nut + 1e-6f);
      float tauG = sqrtf(2.0f * dx / P->g);
      float tau = fmax_(fmin_(tauU, tauG), P->tau_min);
      float gate = smoothstep_(P->t_ign - 100.0f, P->t_ign + 100.0f, T);
      float lim = fmin_(yf, yo / P->s_o2);
      float dY = lim * (1.0f - pexp(-gate * dt / tau));
      yf -= dY;
      yo = fmax_(yo - P->s_o2 * dY, 0.0f);
      T += dY * P->dhc / cpOf(T + 0.5f * dY * P->dhc / cpOf(T));
      float q = rho * dY * P->dhc / dt;
      float rich = smoothstep_(0.0f, 0.03f, yf);
      ys += P->soot_form * dY * (0.25f + 0.75f * rich);
      float oxg = smoothstep_(P->soot_ox_t, P->soot_ox_t + 300.0f, T) * (yo / P->yo2);
      ys *= pexp(-P->soot_ox * oxg * dt);
      float cold = 1.0f - smoothstep_(450.0f, 650.0f, T);
      float cv = yf * fmin_(P->smoke_conv * cold * dt, 1.0f);
      yf -= cv;
      ym += cv;
      rho = rhoOf(P, T);
      float kap = kapP(P, T, yo, ys);
      float tl = kap * P->rad_len;
      float esc = tl > 1e-3f ? (1.0f - pexp(-tl)) / tl : 1.0f;
      float prad = 4.0f * kap * esc * SIGMA * (pow4(T) - pow4(P->t0));
      float dTr = fmin_(prad * dt / (rho * cpOf(T)), 0.5f * fmax_(T - P->t0, 0.0f));
      T -= dTr;
      float prr = dTr * rho * cpOf(T) / dt;
      S += (q - prr) / (rho * cpOf(T) * T);
      float sp = fmax_(spongeSideR(P, g, x, z), spongeTopR(P, g, y));
      float dec = 1.0f - pexp(-dt * 4.0f * sp);
      T = mixf_(T, P->t0, dec);
      yf *= 1.0f - dec;
      yo = mixf_(yo, P->yo2, dec);
      ys *= 1.0f - dec;
      ym *= 1.0f - dec;
      T = clampf_(T, 250.0f, 2600.0f);
      s0out[i] = make_float4(T, yf, yo, ys);
      s1out[i] = make_float4(ym, q, S, 0.0f);
      if (STATS) {
        const float dV = dx * dx * dx;
        rho = rhoOf(P, T);
        stH = q * dV;
        stR = prr * dV;
        stS = rho * ys * dV;
        stF = rho * yf * dV;
        stM = rho * ym * dV;
        stT = T;
        stL = stH > 1.0f ? stH : 0.0f;
      }
    }
  }
  if (STATS) {
    __shared__ double sh[NT / 64];
    __shared__ float shf[NT / 64];
    double vH = block_sum<NT>((double)stH, sh);
    double vR = block_sum<NT>((double)stR, sh);
    double vS = block_sum<NT>((double)stS, sh);
    double vF = block_sum<NT>((double)stF, sh);
    double vM = block_sum<NT>((double)stM, sh);
    double vL = block_sum<NT>((double)stL, sh);
    float vT = block_max<NT>(stT, shf);
    if (threadIdx.x == 0) {
      size_t b = (size_t)y * gridDim.x + bxl;
      part[b * 6 + 0] = vH;
      part[b * 6 + 1] = vR;
      part[b * 6 + 2] = vS;
      part[b * 6 + 3] = vF;
      part[b * 6 + 4] = vM;
      part[b * 6 + 5] = vL;
      part_t[b] = vT;
    }
  }
}

void launch_react(const Fields &f, StatParts *st, hipStream_t s) {
  dim3 grid((unsigned)((f.g.nx * f.g.nz + NT - 1) / NT), (unsigned)f.g.ny);
  if (st)
    k_react<true><<<grid, NT, 0, s>>>(f.P, f.g, f.s0b, f.s1b, f.s0, f.s1, f.flags, f.sol_out, f.pyr, f.curl,
                                      st->react, st->react_t);
  else
    k_react<false><<<grid, NT, 0, s>>>(f.P, f.g, f.s0b, f.s1b, f.s0, f.s1, f.flags, f.sol_out, f.pyr, f.curl,
                                       nullptr, nullptr);
}

}  // namespace pyro
