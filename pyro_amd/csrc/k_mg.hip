// k_mg.hip -- geometric multigrid pressure solve on gfx950.
//
// Reference: mg.wgsl mg_smooth (:5-36), mg_restrict (:38-83), mg_prolong
// (:85-120), mg_resid (:122-158); V-cycle schedule sim.rs:451-491.
// Operator: (sum of fluid neighbours - cnt*phi)/h^2; ground (y<0) Neumann,
// side/top neighbours outside the box count with phi = 0, solid neighbours
// are skipped (Neumann).
//
// The reference dispatches the coarsest level's 80 red-black SOR iterations
// as 160 separate passes; here they run in one workgroup with phi, b and the
// solid mask resident in LDS (160 KB per CU on gfx950), one barrier per
// colour pass.
#include "kernels.hpp"

namespace pyro {

// Neighbour sum and count of cell (x,y,z) (mg.wgsl:25-33 / :56-64).
__device__ __forceinline__ void nsum(const int d0, const int d1, const int d2, const float *__restrict__ p,
                                     const uint8_t *__restrict__ f, int x, int y, int z, uint32_t i, float &sum,
                                     float &cnt) {
  const uint32_t sy = (uint32_t)d0, sz = (uint32_t)d0 * (uint32_t)d1;
  sum = 0.0f;
  cnt = 0.0f;
  // +x
  if (x + 1 < d0) { if (f[i + 1] != 1) { sum += p[i + 1]; cnt += 1.0f; } } else cnt += 1.0f;
  // -x
  if (x - 1 >= 0) { if (f[i - 1] != 1) { sum += p[i - 1]; cnt += 1.0f; } } else cnt += 1.0f;
  // +y
  if (y + 1 < d1) { if (f[i + sy] != 1) { sum += p[i + sy]; cnt += 1.0f; } } else cnt += 1.0f;
  // -y (ground: skipped)
  if (y - 1 >= 0) { if (f[i - sy] != 1) { sum += p[i - sy]; cnt += 1.0f; } }
  // +z
  if (z + 1 < d2) { if (f[i + sz] != 1) { sum += p[i + sz]; cnt += 1.0f; } } else cnt += 1.0f;
  // -z
  if (z - 1 >= 0) { if (f[i - sz] != 1) { sum += p[i - sz]; cnt += 1.0f; } } else cnt += 1.0f;
}

struct LvArgs {
  int d0, d1, d2;
  uint32_t n;
  float hh;
  FastDiv dhx, dny;  // ceil(d0/2), d1
  uint32_t nhalf;    // ceil(d0/2)*d1*d2
};
static LvArgs lv_args(const MgLevel &L) {
  LvArgs a;
  a.d0 = L.d[0]; a.d1 = L.d[1]; a.d2 = L.d[2];
  a.n = L.n;
  a.hh = L.hh;
  uint32_t hx = (uint32_t)(L.d[0] + 1) / 2;
  a.dhx = make_fastdiv(hx);
  a.dny = make_fastdiv((uint32_t)L.d[1]);
  a.nhalf = hx * (uint32_t)L.d[1] * (uint32_t)L.d[2];
  return a;
}

// One colour of the red-black sweep: cells with (x+y+z)&1 == color.
__global__ __launch_bounds__(NT) void k_mg_smooth(LvArgs a, const float *__restrict__ b,
                                                   const uint8_t *__restrict__ f, float *__restrict__ p, int color,
                                                   float omega) {
  uint32_t t = blockIdx.x * NT + threadIdx.x;
  if (t >= a.nhalf) return;
  uint32_t r = fdiv(t, a.dhx);
  int gx = (int)(t - r * a.dhx.d);
  uint32_t zq = fdiv(r, a.dny);
  int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
  int x = 2 * gx + ((y + z + color) & 1);
  if (x >= a.d0) return;
  uint32_t i = (uint32_t)x + (uint32_t)a.d0 * ((uint32_t)y + (uint32_t)a.d1 * (uint32_t)z);
  if (f[i] == 1) return;
  float sum, cnt;
  nsum(a.d0, a.d1, a.d2, p, f, x, y, z, i, sum, cnt);
  if (cnt < 0.5f) return;
  p[i] = mixf_(p[i], (sum - a.hh * b[i]) / cnt, omega);
}
void launch_mg_smooth(const MgLevel &L, int color, float omega, hipStream_t s) {
  LvArgs a = lv_args(L);
  k_mg_smooth<<<nblocks(a.nhalf), NT, 0, s>>>(a, L.b, L.f, L.phi, color, omega);
}

// Fine residual at (x,y,z) (mg.wgsl:52-66).
__device__ __forceinline__ float resid_at(const LvArgs &F, const float *__restrict__ p, const float *__restrict__ b,
                                          const uint8_t *__restrict__ f, int x, int y, int z) {
  uint32_t i = (uint32_t)x + (uint32_t)F.d0 * ((uint32_t)y + (uint32_t)F.d1 * (uint32_t)z);
  if (f[i] == 1) return 0.0f;
  float pcen = p[i];
  float sum, cnt;
  nsum(F.d0, F.d1, F.d2, p, f, x, y, z, i, sum, cnt);
  return b[i] - (sum - cnt * pcen) / F.hh;
}

// Residual + average over fluid children; coarse phi = 0; coarse cell solid
// iff every child is solid (mg.wgsl:67-83).
__global__ __launch_bounds__(NT) void k_mg_restrict(LvArgs F, const float *__restrict__ pf,
                                                     const float *__restrict__ bf, const uint8_t *__restrict__ ff,
                                                     int c0, int c1, int c2, FastDiv dc0, FastDiv dc1,
                                                     float *__restrict__ bc, float *__restrict__ pc,
                                                     uint8_t *__restrict__ fc) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= (uint32_t)c0 * (uint32_t)c1 * (uint32_t)c2) return;
  uint32_t r = fdiv(i, dc0);
  int x = (int)(i - r * (uint32_t)c0);
  uint32_t zq = fdiv(r, dc1);
  int y = (int)(r - zq * (uint32_t)c1), z = (int)zq;
  float acc = 0.0f, nfl = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    int a = 2 * x + (k & 1), bb = 2 * y + ((k >> 1) & 1), c = 2 * z + ((k >> 2) & 1);
    if (a >= F.d0 || bb >= F.d1 || c >= F.d2) continue;
    if (ff[(uint32_t)a + (uint32_t)F.d0 * ((uint32_t)bb + (uint32_t)F.d1 * (uint32_t)c)] != 1) {
      nfl += 1.0f;
      acc += resid_at(F, pf, bf, ff, a, bb, c);
    }
  }
  bc[i] = nfl > 0.0f ? acc / fmax_(nfl, 1.0f) : 0.0f;
  pc[i] = 0.0f;
  fc[i] = nfl > 0.0f ? 0 : 1;
}
void launch_mg_restrict(const MgLevel &Fl, const MgLevel &C, hipStream_t s) {
  LvArgs F = lv_args(Fl);
  k_mg_restrict<<<nblocks(C.n), NT, 0, s>>>(F, Fl.phi, Fl.b, Fl.f, C.d[0], C.d[1], C.d[2],
                                            make_fastdiv((uint32_t)C.d[0]), make_fastdiv((uint32_t)C.d[1]), C.b,
                                            C.phi, C.f);
}

// Masked, renormalised trilinear correction (mg.wgsl:97-120).
__global__ __launch_bounds__(NT) void k_mg_prolong(int f0, int f1, int f2, FastDiv df0, FastDiv df1,
                                                    const uint8_t *__restrict__ ff, float *__restrict__ pf, int c0,
                                                    int c1, int c2, const float *__restrict__ pc,
                                                    const uint8_t *__restrict__ fc) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= (uint32_t)f0 * (uint32_t)f1 * (uint32_t)f2) return;
  if (ff[i] == 1) return;
  uint32_t r = fdiv(i, df0);
  int x = (int)(i - r * (uint32_t)f0);
  uint32_t zq = fdiv(r, df1);
  int y = (int)(r - zq * (uint32_t)f1), z = (int)zq;
  float px = ((float)x + 0.5f) * 0.5f - 0.5f, py = ((float)y + 0.5f) * 0.5f - 0.5f,
        pz = ((float)z + 0.5f) * 0.5f - 0.5f;
  int i0x = (int)floorf(px), i0y = (int)floorf(py), i0z = (int)floorf(pz);
  float frx = px - (float)i0x, fry = py - (float)i0y, frz = pz - (float)i0z;
  float acc = 0.0f, ws = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
    int cx = clampi(i0x + ox, 0, c0 - 1), cy = clampi(i0y + oy, 0, c1 - 1), cz = clampi(i0z + oz, 0, c2 - 1);
    uint32_t j = (uint32_t)cx + (uint32_t)c0 * ((uint32_t)cy + (uint32_t)c1 * (uint32_t)cz);
    if (fc[j] == 1) continue;
    float wx = ox == 1 ? frx : 1.0f - frx, wy = oy == 1 ? fry : 1.0f - fry, wz = oz == 1 ? frz : 1.0f - frz;
    float w = wx * wy * wz;
    acc += w * pc[j];
    ws += w;
  }
  if (ws > 1e-4f) pf[i] += acc / ws;
}
void launch_mg_prolong(const MgLevel &C, const MgLevel &F, hipStream_t s) {
  k_mg_prolong<<<nblocks(F.n), NT, 0, s>>>(F.d[0], F.d[1], F.d[2], make_fastdiv((uint32_t)F.d[0]),
                                           make_fastdiv((uint32_t)F.d[1]), F.f, F.phi, C.d[0], C.d[1], C.d[2],
                                           C.phi, C.f);
}

// ---------------------------------------------------------------- coarse
constexpr int COARSE_NT = 1024;
constexpr int COARSE_MAX = 16384;  // cells; phi + b (f32) + mask (u8) = 144 KB of LDS

__global__ __launch_bounds__(COARSE_NT) void k_mg_coarse_lds(LvArgs a, const float *__restrict__ b,
                                                              const uint8_t *__restrict__ f, float *__restrict__ p,
                                                              int iters, float omega) {
  __shared__ float sp[COARSE_MAX];
  __shared__ float sb[COARSE_MAX];
  __shared__ uint8_t sf[COARSE_MAX];
  for (uint32_t i = threadIdx.x; i < a.n; i += COARSE_NT) {
    sp[i] = p[i];
    sb[i] = b[i];
    sf[i] = f[i];
  }
  __syncthreads();
  for (int it = 0; it < iters; it++) {
    for (int color = 0; color < 2; color++) {
      for (uint32_t t = threadIdx.x; t < a.nhalf; t += COARSE_NT) {
        uint32_t r = fdiv(t, a.dhx);
        int gx = (int)(t - r * a.dhx.d);
        uint32_t zq = fdiv(r, a.dny);
        int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
        int x = 2 * gx + ((y + z + color) & 1);
        if (x >= a.d0) continue;
        uint32_t i = (uint32_t)x + (uint32_t)a.d0 * ((uint32_t)y + (uint32_t)a.d1 * (uint32_t)z);
        if (sf[i] == 1) continue;
        float sum, cnt;
        nsum(a.d0, a.d1, a.d2, sp, sf, x, y, z, i, sum, cnt);
        if (cnt < 0.5f) continue;
        sp[i] = mixf_(sp[i], (sum - a.hh * sb[i]) / cnt, omega);
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < a.n; i += COARSE_NT) p[i] = sp[i];
}

// Same schedule for a coarsest level larger than LDS: one workgroup, phi in
// global memory (L2-resident), a workgroup barrier between colour passes.
__global__ __launch_bounds__(COARSE_NT) void k_mg_coarse_global(LvArgs a, const float *__restrict__ b,
                                                                 const uint8_t *__restrict__ f, float *p, int iters,
                                                                 float omega) {
  for (int it = 0; it < iters; it++) {
    for (int color = 0; color < 2; color++) {
      for (uint32_t t = threadIdx.x; t < a.nhalf; t += COARSE_NT) {
        uint32_t r = fdiv(t, a.dhx);
        int gx = (int)(t - r * a.dhx.d);
        uint32_t zq = fdiv(r, a.dny);
        int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
        int x = 2 * gx + ((y + z + color) & 1);
        if (x >= a.d0) continue;
        uint32_t i = (uint32_t)x + (uint32_t)a.d0 * ((uint32_t)y + (uint32_t)a.d1 * (uint32_t)z);
        if (f[i] == 1) continue;
        float sum, cnt;
        nsum(a.d0, a.d1, a.d2, p, f, x, y, z, i, sum, cnt);
        if (cnt < 0.5f) continue;
        p[i] = mixf_(p[i], (sum - a.hh * b[i]) / cnt, omega);
      }
      __syncthreads();
    }
  }
}

bool mg_coarse_fits_lds(const MgLevel &L) { return L.n <= (uint32_t)COARSE_MAX; }

void launch_mg_coarse(const MgLevel &L, int iters, float omega, hipStream_t s) {
  LvArgs a = lv_args(L);
  if (mg_coarse_fits_lds(L))
    k_mg_coarse_lds<<<1, COARSE_NT, 0, s>>>(a, L.b, L.f, L.phi, iters, omega);
  else
    k_mg_coarse_global<<<1, COARSE_NT, 0, s>>>(a, L.b, L.f, L.phi, iters, omega);
}

// ---------------------------------------------------------------- resid
// Per-block FP64 sums of r^2 and b^2 over fluid cells (mg.wgsl:134-158).
__global__ __launch_bounds__(NT) void k_mg_resid(LvArgs a, FastDiv d0, const float *__restrict__ p,
                                                  const float *__restrict__ b, const uint8_t *__restrict__ f,
                                                  double *__restrict__ part) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  double r2 = 0.0, b2 = 0.0;
  if (i < a.n && f[i] != 1) {
    uint32_t r = fdiv(i, d0);
    int x = (int)(i - r * (uint32_t)a.d0);
    uint32_t zq = fdiv(r, a.dny);
    int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
    float rr = resid_at(a, p, b, f, x, y, z);
    float bb = b[i];
    r2 = (double)rr * (double)rr;
    b2 = (double)bb * (double)bb;
  }
  __shared__ double sh[NT / 64];
  double s1 = block_sum<NT>(r2, sh);
  double s2 = block_sum<NT>(b2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = s1;
    part[blockIdx.x * 2 + 1] = s2;
  }
}
int mg_resid_blocks(const MgLevel &L) { return nblocks(L.n); }
void launch_mg_resid(const MgLevel &L, StatParts *st, hipStream_t s) {
  LvArgs a = lv_args(L);
  k_mg_resid<<<nblocks(L.n), NT, 0, s>>>(a, make_fastdiv((uint32_t)L.d[0]), L.phi, L.b, L.f, st->resid);
}

}  // namespace pyro
