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

// Neighbour sum and count of cell (x,y,z) (mg.wgsl:25-33 / :56-64): fluid
// neighbours add phi and count 1, neighbours outside the box on a side or the
// top count wdir (1 in mg.wgsl), the ground and solids are skipped.  The
// count is formed as fluid + wdir * outside, which for wdir = 1 is the same
// small integer as the reference's running sum of 1.0.
__device__ __forceinline__ void nsum(const int d0, const int d1, const int d2, const float *__restrict__ p,
                                     const uint8_t *__restrict__ f, int x, int y, int z, uint32_t i, float wdir,
                                     float &sum, float &cnt) {
  const uint32_t sy = (uint32_t)d0, sz = (uint32_t)d0 * (uint32_t)d1;
  sum = 0.0f;
  int fl = 0, out = 0;
  if (x + 1 < d0) { if (f[i + 1] != 1) { sum += p[i + 1]; fl++; } } else out++;
  if (x - 1 >= 0) { if (f[i - 1] != 1) { sum += p[i - 1]; fl++; } } else out++;
  if (y + 1 < d1) { if (f[i + sy] != 1) { sum += p[i + sy]; fl++; } } else out++;
  if (y - 1 >= 0) { if (f[i - sy] != 1) { sum += p[i - sy]; fl++; } }  // ground: skipped
  if (z + 1 < d2) { if (f[i + sz] != 1) { sum += p[i + sz]; fl++; } } else out++;
  if (z - 1 >= 0) { if (f[i - sz] != 1) { sum += p[i - sz]; fl++; } } else out++;
  cnt = (float)fl + wdir * (float)out;
}

struct LvArgs {
  int d0, d1, d2;
  uint32_t n;
  float hh;
  float wdir;        // weight of a neighbour outside the box (1 = mg.wgsl)
  FastDiv dhx, dny;  // ceil(d0/2), d1
  uint32_t nhalf;    // ceil(d0/2)*d1*d2
};
static LvArgs lv_args(const MgLevel &L) {
  LvArgs a;
  a.d0 = L.d[0]; a.d1 = L.d[1]; a.d2 = L.d[2];
  a.n = L.n;
  a.hh = L.hh;
  a.wdir = L.wdir;
  uint32_t hx = (uint32_t)(L.d[0] + 1) / 2;
  a.dhx = make_fastdiv(hx);
  a.dny = make_fastdiv((uint32_t)L.d[1]);
  a.nhalf = hx * (uint32_t)L.d[1] * (uint32_t)L.d[2];
  return a;
}

// Fine residual at (x,y,z) (mg.wgsl:52-66).
__device__ __forceinline__ float resid_at(const LvArgs &F, const float *__restrict__ p, const float *__restrict__ b,
                                          const uint8_t *__restrict__ f, int x, int y, int z) {
  uint32_t i = (uint32_t)x + (uint32_t)F.d0 * ((uint32_t)y + (uint32_t)F.d1 * (uint32_t)z);
  if (f[i] == 1) return 0.0f;
  float pcen = p[i];
  float sum, cnt;
  nsum(F.d0, F.d1, F.d2, p, f, x, y, z, i, F.wdir, sum, cnt);
  return b[i] - (sum - cnt * pcen) / F.hh;
}

// Neighbour codes (setup pass below): bit 7 solid, bit 6 no counted
// neighbour, bits 0-2 fluid neighbours, bits 3-5 neighbours outside the box.
constexpr uint8_t NC_SOLID = 0x80, NC_NONE = 0x40;
__device__ __forceinline__ float nc_count(uint32_t code, float wdir) {
  return (float)(code & 7u) + wdir * (float)((code >> 3) & 7u);
}

// x / c for an integer count c in 1..6 given r = RN(1/c) and x != -0:
// q0 = x r, the remainder e = x - c q0 (exact through the FMA), q = q0 + e r
// (Markstein's correction).  Checked exhaustively over every finite float
// x != -0 for each c: q equals the IEEE quotient whenever q is not subnormal
// (the only mismatches are c = 6 with 0 < |x| < 2^-125), and a subnormal q
// takes the IEEE division.  In the sweep, x = (neighbour sum from +0) - h^2 b
// is never -0: a sum that starts at +0 never becomes -0, and s - t is -0
// only for s = -0.  For x = -0 the result is +0 (IEEE: -0), and with c = r =
// 0 it is +-0; the coarse solve uses q only in phi + (q - phi) omega, which
// is the same for q = +0 and q = -0 (+0 for phi = +-0, else q - phi = -phi).
__device__ __forceinline__ float div_count(float x, float c, float r) {
  const float q0 = x * r;
  const float e = __builtin_fmaf(-q0, c, x);
  float q = __builtin_fmaf(e, r, q0);
  if (__builtin_amdgcn_classf(q, 0x90)) {  // -subnormal | +subnormal
    float xs = x;
    asm volatile("" : "+v"(xs));  // a branch, not a select: the division runs only for these x
    q = xs / c;
  }
  return q;
}

// Residual + average over fluid children, coarse phi = 0 (mg.wgsl:67-81).
// The coarse flags (mg.wgsl:82) are produced once per step by the setup pass.
// A block takes 16 x 4 x 4 coarse cells (one per thread); the fine φ and
// neighbour codes of their children plus a ring of 1 are staged in LDS, φ
// zeroed on solid and outside cells, so each child's neighbour sum adds the
// six LDS values in the order of nsum (x+1, x-1, y+1, y-1, z+1, z-1): a
// skipped neighbour contributes +0.0, which never changes the running sum.
constexpr int RS_TX = 16, RS_TY = 4, RS_TZ = 4;
constexpr int RS_FX = 2 * RS_TX + 2, RS_FY = 2 * RS_TY + 2, RS_FZ = 2 * RS_TZ + 2;
constexpr int RS_FN = RS_FX * RS_FY * RS_FZ;
// VEC (fine x-size a multiple of 4, arrays below 2 GiB): the footprint's x
// range is widened to 16-byte alignment, [2 x0 - 4, 2 x0 + 36), and staged
// as whole float4 / uchar4 groups: thread t < 250 loads group t % 10 of the
// rows t / 10 + 25 it, it = 0..3 (a group lies entirely inside or outside
// the level; outside reads 0 and codes 0x40).
constexpr int RS_VX = 2 * RS_TX + 8, RS_VQ = RS_VX / 4, RS_VN = RS_VX * RS_FY * RS_FZ;
static_assert(RS_VQ * 25 == 250 && RS_FY * RS_FZ == 100, "restriction staging map");
template <bool VEC>
__global__ __launch_bounds__(NT) void k_mg_restrict(LvArgs F, int ntx, int nty, const float *__restrict__ pf,
                                                     const float *__restrict__ bf, const uint8_t *__restrict__ ncf,
                                                     int c0, int c1, int c2, float *__restrict__ bc,
                                                     float *__restrict__ pc) {
  constexpr int FX = VEC ? RS_VX : RS_FX;
  __shared__ __align__(16) float sp[VEC ? RS_VN : RS_FN];
  __shared__ __align__(16) uint8_t sc[VEC ? RS_VN : RS_FN];
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = bid % ntx, rr = bid / ntx, ty = rr % nty, tz = rr / nty;
  // footprint origin
  const int fx0 = 2 * RS_TX * tx - (VEC ? 4 : 1), fy0 = 2 * RS_TY * ty - 1, fz0 = 2 * RS_TZ * tz - 1;
  const int sy = F.d0, sz = F.d0 * F.d1;
  const int lcx = threadIdx.x % RS_TX, lr = threadIdx.x / RS_TX, lcy = lr % RS_TY, lcz = lr / RS_TY;
  const int x = RS_TX * tx + lcx, y = RS_TY * ty + lcy, z = RS_TZ * tz + lcz;
  // the eight children's h^2 b, issued with the staging loads (a child
  // outside the level, also of a coarse cell outside it, reads element 0
  // and is skipped)
  float bv[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int a = 2 * x + (k & 1), bb = 2 * y + ((k >> 1) & 1), c = 2 * z + ((k >> 2) & 1);
    const bool in = a < F.d0 && bb < F.d1 && c < F.d2;
    bv[k] = bf[in ? a + sy * bb + sz * c : 0];
  }
  if (VEC) {
    const RB rp = rb_of(pf, F.n * 4u), rcd = rb_of(ncf, F.n);
    const int t = threadIdx.x, q = t % RS_VQ, r0 = t / RS_VQ;
    const int gx = fx0 + 4 * q;
    float4 pv[4];
    uint32_t cv[4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int r = r0 + 25 * it, lz = r / RS_FY, ly = r - RS_FY * lz;
      const int gy = fy0 + ly, gz = fz0 + lz;
      const bool ok = t < 250 && (uint32_t)gx < (uint32_t)F.d0 && (uint32_t)gy < (uint32_t)F.d1 &&
                      (uint32_t)gz < (uint32_t)F.d2;
      const uint32_t g = (uint32_t)gx + (uint32_t)sy * (uint32_t)gy + (uint32_t)sz * (uint32_t)gz;
      pv[it] = rb_ld4(rp, ok ? 4u * g : RB_OOB);
      const uint32_t c = rb_ld32(rcd, ok ? g : RB_OOB);
      cv[it] = ok ? c : 0x40404040u;
    }
    if (t < 250) {
#pragma unroll
      for (int it = 0; it < 4; it++) {
        float4 v = pv[it];
        const uint32_t c = cv[it];
        if (c & 0x80808080u) {  // solids read 0 in the neighbour sums
          if (c & 0x80u) v.x = 0.0f;
          if (c & 0x8000u) v.y = 0.0f;
          if (c & 0x800000u) v.z = 0.0f;
          if (c & 0x80000000u) v.w = 0.0f;
        }
        const int e = 4 * (q + RS_VQ * (r0 + 25 * it));
        *reinterpret_cast<float4 *>(sp + e) = v;
        *reinterpret_cast<uint32_t *>(sc + e) = c;
      }
    }
  } else {
    // every load is unconditional (outside the level at element 0) and
    // selected afterwards, so all of a thread's loads are in flight at once
#pragma unroll
    for (int q = 0; q < (RS_FN + NT - 1) / NT; q++) {
      const int s = threadIdx.x + q * NT;
      const int se = s < RS_FN ? s : RS_FN - 1;
      const int lx = se % RS_FX, r2 = se / RS_FX, ly = r2 % RS_FY, lz = r2 / RS_FY;
      const int gx = fx0 + lx, gy = fy0 + ly, gz = fz0 + lz;
      const bool in = gx >= 0 && gx < F.d0 && gy >= 0 && gy < F.d1 && gz >= 0 && gz < F.d2;
      const int g = in ? gx + sy * gy + sz * gz : 0;
      const uint8_t c = ncf[g];
      const float p = pf[g];
      if (s < RS_FN) {
        sp[s] = (in && !(c & NC_SOLID)) ? p : 0.0f;
        sc[s] = in ? c : NC_NONE;
      }
    }
  }
  __syncthreads();
  if (x >= c0 || y >= c1 || z >= c2) return;
  float acc = 0.0f, nfl = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int a = 2 * x + (k & 1), bb = 2 * y + ((k >> 1) & 1), c = 2 * z + ((k >> 2) & 1);
    if (a >= F.d0 || bb >= F.d1 || c >= F.d2) continue;
    const int pi = ((c - fz0) * RS_FY + (bb - fy0)) * FX + (a - fx0);
    const uint8_t code = sc[pi];
    if (code & NC_SOLID) continue;
    nfl += 1.0f;
    float sum = 0.0f;
    sum += sp[pi + 1];
    sum += sp[pi - 1];
    sum += sp[pi + FX];
    sum += sp[pi - FX];
    sum += sp[pi + FX * RS_FY];
    sum += sp[pi - FX * RS_FY];
    const float cnt = nc_count(code, F.wdir);
    acc += bv[k] - (sum - cnt * sp[pi]) / F.hh;
  }
  const uint32_t i = (uint32_t)x + (uint32_t)c0 * ((uint32_t)y + (uint32_t)c1 * (uint32_t)z);
  bc[i] = nfl > 0.0f ? acc / fmax_(nfl, 1.0f) : 0.0f;
  pc[i] = 0.0f;
}
void launch_mg_restrict(const MgLevel &Fl, const MgLevel &C, hipStream_t s) {
  LvArgs F = lv_args(Fl);
  const int ntx = (C.d[0] + RS_TX - 1) / RS_TX, nty = (C.d[1] + RS_TY - 1) / RS_TY, ntz = (C.d[2] + RS_TZ - 1) / RS_TZ;
  const bool vec = (Fl.d[0] % 4) == 0 && (uint64_t)Fl.n * 4u < (1ull << 31);
  if (vec)
    k_mg_restrict<true><<<ntx * nty * ntz, NT, 0, s>>>(F, ntx, nty, Fl.phi, Fl.b, Fl.nc, C.d[0], C.d[1], C.d[2], C.b,
                                                       C.phi);
  else
    k_mg_restrict<false><<<ntx * nty * ntz, NT, 0, s>>>(F, ntx, nty, Fl.phi, Fl.b, Fl.nc, C.d[0], C.d[1], C.d[2],
                                                        C.b, C.phi);
}

// Masked, renormalised trilinear correction (mg.wgsl:97-120).  A block takes
// 32 x 8 x 4 fine cells (4 per thread along y); the coarse φ and flags they
// interpolate from (18 x 6 x 4 cells, with the clamping of mg.wgsl applied
// while staging) sit in LDS.  Weights and the order of the sums are those of
// the per-cell formula.
constexpr int PL_TX = 32, PL_TY = 8, PL_TZ = 4;
constexpr int PL_CX = PL_TX / 2 + 2, PL_CY = PL_TY / 2 + 2, PL_CZ = PL_TZ / 2 + 2;
constexpr int PL_CN = PL_CX * PL_CY * PL_CZ;
__global__ __launch_bounds__(NT) void k_mg_prolong(int f0, int f1, int f2, int ntx, int nty,
                                                    const uint8_t *__restrict__ ff, float *__restrict__ pf, int c0,
                                                    int c1, int c2, const float *__restrict__ pc,
                                                    const uint8_t *__restrict__ fc) {
  __shared__ float sp[PL_CN];
  __shared__ uint8_t sf[PL_CN];
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = bid % ntx, rr = bid / ntx, ty = rr % nty, tz = rr / nty;
  const int x0 = PL_TX * tx, y0 = PL_TY * ty, z0 = PL_TZ * tz;
  const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1, cz0 = z0 / 2 - 1;  // coarse footprint origin (unclamped)
  // a thread's cells: x = x0 + lx, y = y0 + ly, z = z0 + q for q < PL_TZ
  static_assert(NT / PL_TX == PL_TY && PL_TZ % 2 == 0, "one (x, y) column of PL_TZ cells per thread");
  constexpr int NQ = PL_TZ;
  const int lx = threadIdx.x % PL_TX, ly = threadIdx.x / PL_TX;
  const int x = x0 + lx, y = y0 + ly;
  // the thread's cells' flags and phi, issued with the staging loads (a
  // cell outside the level reads element 0 and is skipped)
  uint8_t fv[NQ];
  float pv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int z = z0 + q;
    const uint32_t i =
        (x < f0 && y < f1 && z < f2) ? (uint32_t)x + (uint32_t)f0 * ((uint32_t)y + (uint32_t)f1 * (uint32_t)z) : 0u;
    fv[q] = ff[i];
    pv[q] = pf[i];
  }
  bool solid = false;
  for (int s = threadIdx.x; s < PL_CN; s += NT) {
    const int u = s % PL_CX, r2 = s / PL_CX, v = r2 % PL_CY, w = r2 / PL_CY;
    const int cx = clampi(cx0 + u, 0, c0 - 1), cy = clampi(cy0 + v, 0, c1 - 1), cz = clampi(cz0 + w, 0, c2 - 1);
    const uint32_t j = (uint32_t)cx + (uint32_t)c0 * ((uint32_t)cy + (uint32_t)c1 * (uint32_t)cz);
    sp[s] = pc[j];
    const uint8_t fl = fc[j];
    sf[s] = fl;
    solid |= fl == 1;
  }
  // whether any coarse cell of the footprint is solid (block-uniform)
  solid = __syncthreads_or(solid);
  if (x >= f0 || y >= f1) return;
  // A fine index's fraction is 0.75 (even) or 0.25 (odd), exactly, so the
  // weights are exact: wx wy is formed once per thread, and z = z0 + q (z0 a
  // multiple of 4) has the parity of q, so frz is a constant; the product
  // keeps the per-cell formula's order, (wx wy) wz.
  const float px = ((float)x + 0.5f) * 0.5f - 0.5f, py = ((float)y + 0.5f) * 0.5f - 0.5f;
  const int i0x = (int)floorf(px), i0y = (int)floorf(py);
  const float frx = px - (float)i0x, fry = py - (float)i0y;
  float wxy[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float wx = (k & 1) ? frx : 1.0f - frx, wy = (k >> 1) ? fry : 1.0f - fry;
    wxy[k] = wx * wy;
  }
  const int sxy = (i0y - cy0) * PL_CX + (i0x - cx0);
  if (!solid) {
    // No solid corner anywhere in the block: every weight is the plain
    // product and they sum to exactly 1 (multiples of 1/64 below 2, added
    // exactly), so the correction is acc itself; acc is formed with the
    // weights and in the order of the general loop below.
    float w8[2][8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      w8[0][k] = wxy[k & 3] * ((k >> 2) ? 0.75f : 1.0f - 0.75f);
      w8[1][k] = wxy[k & 3] * ((k >> 2) ? 0.25f : 1.0f - 0.25f);
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int z = z0 + q;
      if (z >= f2 || fv[q] == 1) continue;
      const uint32_t i = (uint32_t)x + (uint32_t)f0 * ((uint32_t)y + (uint32_t)f1 * (uint32_t)z);
      const int i0z = (int)floorf(((float)z + 0.5f) * 0.5f - 0.5f);
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const int si = (i0z + oz - cz0) * (PL_CY * PL_CX) + sxy + oy * PL_CX + ox;
        acc += w8[q & 1][k] * sp[si];
      }
      pf[i] = pv[q] + acc;
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int z = z0 + q;
    if (z >= f2) continue;
    const uint32_t i = (uint32_t)x + (uint32_t)f0 * ((uint32_t)y + (uint32_t)f1 * (uint32_t)z);
    if (fv[q] == 1) continue;
    const float pz = ((float)z + 0.5f) * 0.5f - 0.5f;
    const int i0z = (int)floorf(pz);
    const float frz = (q & 1) ? 0.25f : 0.75f;  // = pz - i0z
    // a solid corner adds weight 0 (its coarse phi is 0, never updated):
    // +-0 added to these sums, which start at +0 and are never -0, leaves
    // them unchanged, so this is the per-cell loop that skips solid corners
    float acc = 0.0f, ws = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
      const int si = (i0z + oz - cz0) * (PL_CY * PL_CX) + sxy + oy * PL_CX + ox;
      const float wgt = sf[si] == 1 ? 0.0f : wxy[k & 3] * (oz == 1 ? frz : 1.0f - frz);
      acc += wgt * sp[si];
      ws += wgt;
    }
    // with no solid corner the weights sum to exactly 1 and acc / 1 = acc:
    // the division runs (a branch, not a select) only for the other cells
    float corr = acc;
    if (ws != 1.0f) {
      float a = acc;
      asm volatile("" : "+v"(a));
      corr = a / ws;
    }
    if (ws > 1e-4f) pf[i] = pv[q] + corr;
  }
}
void launch_mg_prolong(const MgLevel &C, const MgLevel &F, hipStream_t s) {
  const int ntx = (F.d[0] + PL_TX - 1) / PL_TX, nty = (F.d[1] + PL_TY - 1) / PL_TY, ntz = (F.d[2] + PL_TZ - 1) / PL_TZ;
  k_mg_prolong<<<ntx * nty * ntz, NT, 0, s>>>(F.d[0], F.d[1], F.d[2], ntx, nty, F.f, F.phi, C.d[0], C.d[1], C.d[2],
                                              C.phi, C.f);
}

// ---------------------------------------------------------------- setup
// Per-level neighbour codes, computed once per step (flags are fixed during
// a step and restriction derives the same coarse flags in every cycle):
//   bit 7: solid (never updated, phi is Neumann for its neighbours)
//   bit 6: fluid cell without a counted neighbour (mg.wgsl:34 skips it)
//   bits 0-2: fluid neighbours; bits 3-5: neighbours outside the box on a side
//   or the top.  cnt of mg.wgsl:25-33 = fluid + wdir * outside; with
//   wdir = 1 it is the same small integer, exactly.

__global__ __launch_bounds__(NT) void k_mg_counts(LvArgs a, FastDiv d0, const uint8_t *__restrict__ f,
                                                   uint8_t *__restrict__ nc, uint8_t *__restrict__ nm) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= a.n) return;
  uint32_t r = fdiv(i, d0);
  int x = (int)(i - r * (uint32_t)a.d0);
  uint32_t zq = fdiv(r, a.dny);
  int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
  const uint32_t sy = (uint32_t)a.d0, sz = (uint32_t)a.d0 * (uint32_t)a.d1;
  const bool hxp = x + 1 < a.d0, hxm = x > 0, hyp = y + 1 < a.d1, hym = y > 0, hzp = z + 1 < a.d2, hzm = z > 0;
  // the cell's and its six neighbours' flags are loaded together (a
  // neighbour outside the level reads the cell itself and is not used)
  const uint8_t fc = f[i];
  const uint8_t fxp = f[hxp ? i + 1 : i], fxm = f[hxm ? i - 1 : i];
  const uint8_t fyp = f[hyp ? i + sy : i], fym = f[hym ? i - sy : i];
  const uint8_t fzp = f[hzp ? i + sz : i], fzm = f[hzm ? i - sz : i];
  if (fc == 1) {
    nc[i] = NC_SOLID;
    if (nm) nm[i] = 0;
    return;
  }
  const uint32_t m = (hxp && fxp != 1 ? 1u : 0u) | (hxm && fxm != 1 ? 2u : 0u) | (hyp && fyp != 1 ? 4u : 0u) |
                     (hym && fym != 1 ? 8u : 0u) | (hzp && fzp != 1 ? 16u : 0u) | (hzm && fzm != 1 ? 32u : 0u);
  // neighbours outside the box on a side or the top; the ground (y - 1) is not counted
  const int out = (hxp ? 0 : 1) + (hxm ? 0 : 1) + (hyp ? 0 : 1) + (hzp ? 0 : 1) + (hzm ? 0 : 1);
  const int fl = __builtin_popcount(m);
  nc[i] = (fl + out) == 0 ? NC_NONE : (uint8_t)(fl | (out << 3));
  if (nm) nm[i] = (uint8_t)m;
}

// The same codes four cells at a time along x (level x-size a multiple of
// 4): a thread reads the flag words of its cells and of their y and z
// neighbours and the two bytes beside them in x, and writes one code word.
__device__ __forceinline__ uint32_t byte_of(uint32_t w, int j) { return (w >> (8 * j)) & 0xffu; }
__global__ __launch_bounds__(NT) void k_mg_counts4(LvArgs a, FastDiv d0, const uint8_t *__restrict__ f,
                                                    uint8_t *__restrict__ nc, uint8_t *__restrict__ nm) {
  const uint32_t w = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;  // word index
  if (w >= a.n / 4u) return;
  const uint32_t i = 4u * w;
  uint32_t r = fdiv(i, d0);
  const int x0 = (int)(i - r * (uint32_t)a.d0);
  uint32_t zq = fdiv(r, a.dny);
  const int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
  const uint32_t sy = (uint32_t)a.d0, sz = (uint32_t)a.d0 * (uint32_t)a.d1;
  const bool hyp = y + 1 < a.d1, hym = y > 0, hzp = z + 1 < a.d2, hzm = z > 0;
  const bool hx0 = x0 > 0, hx3 = x0 + 4 < a.d0;
  const uint32_t *f4 = reinterpret_cast<const uint32_t *>(f);
  // a neighbour outside the level reads the cell's own word or byte and is not used
  const uint32_t C = f4[w];
  const uint32_t YP = f4[hyp ? w + sy / 4u : w], YM = f4[hym ? w - sy / 4u : w];
  const uint32_t ZP = f4[hzp ? w + sz / 4u : w], ZM = f4[hzm ? w - sz / 4u : w];
  const uint32_t XM = f[hx0 ? i - 1u : i], XP = f[hx3 ? i + 4u : i];
  uint32_t codes = 0u, masks = 0u;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int x = x0 + j;
    const bool hxp = x + 1 < a.d0, hxm = x > 0;
    const uint32_t fc = byte_of(C, j);
    const uint32_t fxp = j < 3 ? byte_of(C, j + 1) : XP, fxm = j > 0 ? byte_of(C, j - 1) : XM;
    uint32_t code, m = 0u;
    if (fc == 1u) {
      code = NC_SOLID;
    } else {
      m = (hxp && fxp != 1u ? 1u : 0u) | (hxm && fxm != 1u ? 2u : 0u) | (hyp && byte_of(YP, j) != 1u ? 4u : 0u) |
          (hym && byte_of(YM, j) != 1u ? 8u : 0u) | (hzp && byte_of(ZP, j) != 1u ? 16u : 0u) |
          (hzm && byte_of(ZM, j) != 1u ? 32u : 0u);
      // neighbours outside the box on a side or the top; the ground (y - 1) is not counted
      const int out = (hxp ? 0 : 1) + (hxm ? 0 : 1) + (hyp ? 0 : 1) + (hzp ? 0 : 1) + (hzm ? 0 : 1);
      const int fl = __builtin_popcount(m);
      code = (fl + out) == 0 ? NC_NONE : (uint32_t)(fl | (out << 3));
    }
    codes |= code << (8 * j);
    masks |= m << (8 * j);
  }
  reinterpret_cast<uint32_t *>(nc)[w] = codes;
  if (nm) reinterpret_cast<uint32_t *>(nm)[w] = masks;
}

// Coarse cell is solid iff every child inside the fine box is solid (mg.wgsl:74-82).
__global__ __launch_bounds__(NT) void k_mg_coarse_flags(LvArgs F, const uint8_t *__restrict__ ff, int c0, int c1,
                                                         int c2, FastDiv dc0, FastDiv dc1, uint8_t *__restrict__ fc) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= (uint32_t)c0 * (uint32_t)c1 * (uint32_t)c2) return;
  uint32_t r = fdiv(i, dc0);
  int x = (int)(i - r * (uint32_t)c0);
  uint32_t zq = fdiv(r, dc1);
  int y = (int)(r - zq * (uint32_t)c1), z = (int)zq;
  int nfl = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    int a = 2 * x + (k & 1), bb = 2 * y + ((k >> 1) & 1), c = 2 * z + ((k >> 2) & 1);
    if (a >= F.d0 || bb >= F.d1 || c >= F.d2) continue;
    nfl += ff[(uint32_t)a + (uint32_t)F.d0 * ((uint32_t)bb + (uint32_t)F.d1 * (uint32_t)c)] != 1;
  }
  fc[i] = nfl > 0 ? 0 : 1;
}

void launch_mg_setup(const MgLevel *L, int nl, hipStream_t s) {
  for (int l = 0; l < nl; l++) {
    if (l > 0) {
      const MgLevel &F = L[l - 1], &C = L[l];
      k_mg_coarse_flags<<<nblocks(C.n), NT, 0, s>>>(lv_args(F), F.f, C.d[0], C.d[1], C.d[2],
                                                   make_fastdiv((uint32_t)C.d[0]), make_fastdiv((uint32_t)C.d[1]), C.f);
    }
    if (L[l].d[0] % 4 == 0)
      k_mg_counts4<<<nblocks(L[l].n / 4u), NT, 0, s>>>(lv_args(L[l]), make_fastdiv((uint32_t)L[l].d[0]), L[l].f,
                                                        L[l].nc, L[l].nm);
    else
      k_mg_counts<<<nblocks(L[l].n), NT, 0, s>>>(lv_args(L[l]), make_fastdiv((uint32_t)L[l].d[0]), L[l].f, L[l].nc,
                                                 L[l].nm);
  }
}

// ---------------------------------------------------------------- sweep
// One full red-black sweep (colour 0 then colour 1, sim.rs:444-448) of a
// 32x8x4 tile per workgroup, staged in LDS.  The LDS footprint is x in
// [x0-4, x0+36) (16-byte aligned rows of 10 float4) by y, z with a halo of 2:
// phi and neighbour codes on the whole footprint, h^2 b on the ring-1 box.
// Colour-0 cells are updated on the tile plus a ring of 1 (the values the
// tile's colour-1 cells read), then colour-1 cells on the tile.  phi is read
// from `pin` and written to `pout` (ping-pong), so a tile never sees a
// neighbour's already-swept values.
// Solid and outside cells hold 0 in the LDS copy of phi, so the neighbour
// sum needs no mask: a skipped neighbour adds +0.0, which cannot change the
// running sum (it starts at +0.0 and never becomes -0.0).  The per-cell
// arithmetic is that of mg.wgsl:25-35, so the result equals two colour passes.
// VEC (level x-size a multiple of 4): whole float4/uchar4 rows; otherwise
// one element at a time.
constexpr int SW_TX = 32, SW_TY = 8, SW_TZ = 4, SW_NT = 256;
constexpr int SW_PX = SW_TX + 8, SW_PY = SW_TY + 4, SW_PZ = SW_TZ + 4;  // footprint
constexpr int SW_BY = SW_TY + 2, SW_BZ = SW_TZ + 2;                      // ring-1 rows for b
constexpr int SW_PN = SW_PX * SW_PY * SW_PZ;
constexpr int SW_Q = SW_PX / 4;  // float4 per footprint row

// The N cells of one colour that a thread updates in a pass (mg.wgsl:25-35
// per cell).  ICNT (outside weight 1): the count is the integer fluid +
// outside and ct[code & 63] holds (count, RN(1/count)) for div_count (code
// 0x40 / 0x80 reads (0, 0): a finite, discarded result).  The cells of a
// colour are independent (their neighbours are of the other colour, which
// the pass does not write), so all of the pass's LDS reads are issued
// first, the quotients are formed branch-free with one branch per pass for a
// subnormal quotient (div_count's IEEE fallback), and every cell is stored:
// one that is not updated (solid, or no counted neighbour) gets its own
// value back.  act[q] = false: slot q is not this thread's (no read, no
// store).
template <bool ICNT, int N>
__device__ __forceinline__ void sweep_cells(float *sp, const uint8_t *sc, const float *hb, const float2 *ct,
                                            const int *pi, const bool *act, float omega, float wdir) {
  uint32_t code[N];
  float nb[N][6], cen[N];
#pragma unroll
  for (int q = 0; q < N; q++) {
    const int p = act[q] ? pi[q] : SW_PX * SW_PY + SW_PX + 1;  // any interior entry
    code[q] = sc[p];
    nb[q][0] = sp[p + 1];
    nb[q][1] = sp[p - 1];
    nb[q][2] = sp[p + SW_PX];
    nb[q][3] = sp[p - SW_PX];
    nb[q][4] = sp[p + SW_PX * SW_PY];
    nb[q][5] = sp[p - SW_PX * SW_PY];
    cen[q] = sp[p];
  }
  float xs[N], qs[N];
  bool redo = false;
#pragma unroll
  for (int q = 0; q < N; q++) {
    // the sum starts with its first term: it then differs from +0 + ... only
    // in the sign of a zero sum (all six terms -0), and the update maps a
    // quotient of either sign of zero to the same value (cen + (q - cen) w)
    float sum = nb[q][0];
#pragma unroll
    for (int k = 1; k < 6; k++) sum += nb[q][k];
    xs[q] = sum - hb[q];
    if (ICNT) {
      const float2 cr = ct[code[q] & 63u];
      const float q0 = xs[q] * cr.y;
      const float e = __builtin_fmaf(-q0, cr.x, xs[q]);
      qs[q] = __builtin_fmaf(e, cr.y, q0);
      redo |= __builtin_amdgcn_classf(qs[q], 0x90);  // -subnormal | +subnormal
    } else {
      qs[q] = xs[q] / nc_count(code[q], wdir);
    }
  }
  if (ICNT && redo) {
#pragma unroll
    for (int q = 0; q < N; q++)
      if (__builtin_amdgcn_classf(qs[q], 0x90)) {
        float x = xs[q];
        asm volatile("" : "+v"(x));
        qs[q] = x / ct[code[q] & 63u].x;
      }
  }
#pragma unroll
  for (int q = 0; q < N; q++) {
    const float nv = mixf_(cen[q], qs[q], omega);
    if (act[q]) sp[pi[q]] = (code[q] & (NC_SOLID | NC_NONE)) ? cen[q] : nv;
  }
}

// h^2 b of the cell at grid (gx, gy, gz), read from global memory (0 outside
// the grid: the cell is then never updated).  Not staged in LDS, which
// keeps a block at 19.7 KB (8 blocks per CU).  VEC: through the buffer
// resource (arrays below 2 GiB); otherwise a bounds-checked load.
template <bool VEC>
__device__ __forceinline__ float sweep_hb(const LvArgs &a, const RB &rbb, const float *__restrict__ b, int gx, int gy,
                                          int gz) {
  const bool ok = (uint32_t)gx < (uint32_t)a.d0 && (uint32_t)gy < (uint32_t)a.d1 && (uint32_t)gz < (uint32_t)a.d2;
  if (VEC) {
    // 24-bit multiplies (full rate): d0, d1, z and the row index y + d1 z of
    // a cell inside the level are below 2^24 (launch_mg_sweep takes VEC only
    // for d1 d2 <= 2^24); the offset of a cell outside is replaced by the
    // out-of-range one
    const uint32_t g = (uint32_t)gx + __umul24((uint32_t)a.d0, (uint32_t)gy + __umul24((uint32_t)a.d1, (uint32_t)gz));
    return a.hh * __builtin_bit_cast(float, rb_ld32(rbb, ok ? 4u * g : RB_OOB));
  }
  const uint32_t g = (uint32_t)gx + (uint32_t)a.d0 * ((uint32_t)gy + (uint32_t)a.d1 * (uint32_t)gz);
  return ok ? a.hh * b[g] : 0.0f;
}

// Staging of the VEC path: thread t < 240 loads x-group q = t % 10 of
// footprint row ly = (t / 10) % 12 in planes (t / 120) + 2 it, it = 0..3
// (LDS float4 index t + 240 it).  Loads outside the grid take the
// out-of-range offset and read 0 (codes: 0x40, never updated).  The VEC path
// needs the level's arrays below 2 GiB (launch_mg_sweep checks).
constexpr int SW_PT = SW_Q * SW_PY * 2;
static_assert(SW_PZ == 8 && SW_PT <= SW_NT, "staging map");

template <bool FINE, bool VEC, bool ICNT>
__global__ __launch_bounds__(SW_NT) void k_mg_sweep(LvArgs a, int ntx, int nty, const float *__restrict__ b,
                                                     const uint8_t *__restrict__ nc, const float *__restrict__ pin,
                                                     float *__restrict__ pout, float omega) {
  __shared__ __align__(16) float sp[SW_PN];
  __shared__ __align__(16) uint8_t sc[SW_PN];
  __shared__ float2 ct[64];
  const int t = threadIdx.x;
  if (ICNT && t < 64) {
    const uint32_t cnt = (t & 7u) + ((t >> 3) & 7u);
    ct[t] = make_float2((float)cnt, cnt ? 1.0f / (float)cnt : 0.0f);
  }
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = bid % ntx, rr = bid / ntx, ty = rr % nty, tz = rr / nty;
  const int x0 = tx * SW_TX - 4, y0 = ty * SW_TY - 2, z0 = tz * SW_TZ - 2;  // footprint origin, x0 % 4 == 0
  const int sy = a.d0, sz = a.d0 * a.d1;
  const RB rbb = rb_of(b, VEC ? a.n * 4u : 0u);
  if (VEC) {
    const RB rp = rb_of(pin, a.n * 4u), rcd = rb_of(nc, a.n);
    float4 pv[4];
    uint32_t cv[4];
    {
      const int q = t % SW_Q, r0 = t / SW_Q, ly = r0 % SW_PY, lz0 = r0 / SW_PY;
      const int gx = x0 + 4 * q, gy = y0 + ly;
      const bool okxy = t < SW_PT && (uint32_t)gx < (uint32_t)a.d0 && (uint32_t)gy < (uint32_t)a.d1;
      // offsets by 24-bit multiplies (the VEC bounds of sweep_hb)
#pragma unroll
      for (int it = 0; it < 4; it++) {
        const int gz = z0 + lz0 + 2 * it;
        const bool ok = okxy && (uint32_t)gz < (uint32_t)a.d2;
        const uint32_t g =
            (uint32_t)gx + __umul24((uint32_t)sy, (uint32_t)gy + __umul24((uint32_t)a.d1, (uint32_t)gz));
        pv[it] = rb_ld4(rp, ok ? 4u * g : RB_OOB);
        const uint32_t c = rb_ld32(rcd, ok ? g : RB_OOB);
        cv[it] = ok ? c : 0x40404040u;
      }
    }
    if (t < SW_PT) {
#pragma unroll
      for (int it = 0; it < 4; it++) {
        float4 v = pv[it];
        const uint32_t c = cv[it];
        if (c & 0x80808080u) {  // solids read 0 in the neighbour sums
          if (c & 0x80u) v.x = 0.0f;
          if (c & 0x8000u) v.y = 0.0f;
          if (c & 0x800000u) v.z = 0.0f;
          if (c & 0x80000000u) v.w = 0.0f;
        }
        *reinterpret_cast<float4 *>(sp + 4 * (t + SW_PT * it)) = v;
        *reinterpret_cast<uint32_t *>(sc + 4 * (t + SW_PT * it)) = c;
      }
    }
  } else {
    // every load unconditional (outside the level at element 0) and
    // selected, so a thread's loads are all in flight together
#pragma unroll
    for (int q = 0; q < (SW_PN + SW_NT - 1) / SW_NT; q++) {
      const int i = t + q * SW_NT;
      const int ie = i < SW_PN ? i : SW_PN - 1;
      const int lx = ie % SW_PX, row = ie / SW_PX, ly = row % SW_PY, lz = row / SW_PY;
      const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
      const bool in = gx >= 0 && gx < a.d0 && gy >= 0 && gy < a.d1 && gz >= 0 && gz < a.d2;
      const int g = in ? gx + sy * gy + sz * gz : 0;
      const uint8_t c = nc[g];
      const float p = pin[g];
      if (i < SW_PN) {
        sp[i] = (in && !(c & NC_SOLID)) ? p : 0.0f;
        sc[i] = in ? c : NC_NONE;
      }
    }
  }
  // Colour of (gx, gy, gz) = (gx + gy + gz) & 1 with gx = x0 + lx etc.; x0 is
  // a multiple of 4, so the colour-c cell of an x-pair starting at an odd
  // lx is lx + ((K + ly + lz + c) & 1) with K = y0 + z0 + 1 (ly, lz counted
  // from the box the pass walks).
  const int K = y0 + z0 + 1;
  // h^2 b of the six cells this thread updates (maps below), loaded while
  // the tile is staged rather than inside the colour passes
  float hb0[4], hb1[2];
  {
    const int px = t % 17, r0 = t / 17;
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int r = r0 + 15 * it, lz = r / SW_BY, ly = r - SW_BY * lz;
      const int lx = 3 + 2 * px + ((K + ly + lz) & 1);
      hb0[it] = sweep_hb<VEC>(a, rbb, b, x0 + lx, y0 + 1 + ly, z0 + 1 + lz);
    }
  }
  {
    const int px = t % 16, r0 = t / 16;
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int r = r0 + 16 * it, ly = r % SW_TY, lz = r / SW_TY;
      const int lx = 4 + 2 * px + ((K + ly + lz) & 1);
      hb1[it] = sweep_hb<VEC>(a, rbb, b, x0 + lx, y0 + 2 + ly, z0 + 2 + lz);
    }
  }
  __syncthreads();
  // colour 0 on the tile + ring 1: x-pairs from lx = 3 (17 of them), rows
  // (ly, lz) of the 10 x 6 ring-1 box (LDS offset 1 in y and z); thread
  // t < 255 takes x-pair t % 17 of rows r0 + 15 it, r0 = t / 17, it = 0..3
  static_assert((SW_TX + 2) / 2 == 17 && SW_BY * SW_BZ == 60, "colour-0 map");
  {
    const int px = t % 17, r0 = t / 17;
    int pi[4];
    bool act[4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int r = r0 + 15 * it, lz = r / SW_BY, ly = r - SW_BY * lz;
      const int lx = 3 + 2 * px + ((K + ly + lz) & 1);
      pi[it] = (int)__umul24((uint32_t)((lz + 1) * SW_PY + (ly + 1)), (uint32_t)SW_PX) + lx;
      act[it] = t < 255;
    }
    sweep_cells<ICNT, 4>(sp, sc, hb0, ct, pi, act, omega, a.wdir);
  }
  __syncthreads();
  // colour 1 on the tile: x-pairs from lx = 4 (16), rows of the 8 x 4 tile
  // (LDS offset 2); thread t takes x-pair t % 16 of rows t / 16 and t / 16 + 16
  static_assert(SW_TX / 2 == 16 && SW_TY == 8 && SW_TZ == 4, "colour-1 map");
  {
    const int px = t % 16, r0 = t / 16;
    int pi[2];
    const bool act[2] = {true, true};
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int r = r0 + 16 * it, ly = r % SW_TY, lz = r / SW_TY;
      const int lx = 4 + 2 * px + ((K + ly + lz) & 1);
      pi[it] = (int)__umul24((uint32_t)((lz + 2) * SW_PY + (ly + 2)), (uint32_t)SW_PX) + lx;
    }
    sweep_cells<ICNT, 2>(sp, sc, hb1, ct, pi, act, omega, a.wdir);
  }
  __syncthreads();
  if (VEC) {
    constexpr int WN = (SW_TX / 4) * SW_TY * SW_TZ;
    static_assert(WN == SW_NT, "one float4 of the tile per thread");
    const int q = t % (SW_TX / 4), row = t / (SW_TX / 4), ly = row % SW_TY, lz = row / SW_TY;
    const int gx = x0 + 4 + 4 * q, gy = y0 + 2 + ly, gz = z0 + 2 + lz;
    if (gx < a.d0 && gy < a.d1 && gz < a.d2) {
      const uint32_t g = (uint32_t)gx + __umul24((uint32_t)sy, (uint32_t)gy + __umul24((uint32_t)a.d1, (uint32_t)gz));
      const int pi = ((lz + 2) * SW_PY + (ly + 2)) * SW_PX + 4 + 4 * q;
      float4 v = *reinterpret_cast<const float4 *>(sp + pi);
      const uint32_t cv = *reinterpret_cast<const uint32_t *>(sc + pi);
      if (cv & 0x80808080u) {  // solid cells keep their phi (zeroed only in LDS)
        const float4 o = *reinterpret_cast<const float4 *>(pin + g);
        if (cv & 0x80u) v.x = o.x;
        if (cv & 0x8000u) v.y = o.y;
        if (cv & 0x800000u) v.z = o.z;
        if (cv & 0x80000000u) v.w = o.w;
      }
      *reinterpret_cast<float4 *>(pout + g) = v;
    }
  } else {
    constexpr int TN = SW_TX * SW_TY * SW_TZ;
    for (int i = t; i < TN; i += SW_NT) {
      const int lx = i % SW_TX, row = i / SW_TX, ly = row % SW_TY, lz = row / SW_TY;
      const int gx = x0 + 4 + lx, gy = y0 + 2 + ly, gz = z0 + 2 + lz;
      if (gx >= a.d0 || gy >= a.d1 || gz >= a.d2) continue;
      const int g = gx + sy * gy + sz * gz;
      const int pi = ((lz + 2) * SW_PY + (ly + 2)) * SW_PX + 4 + lx;
      pout[g] = (sc[pi] & NC_SOLID) ? pin[g] : sp[pi];
    }
  }
}

void launch_mg_sweep(const MgLevel &L, const float *pin, float *pout, float omega, hipStream_t s, bool fine) {
  LvArgs a = lv_args(L);
  int ntx = (L.d[0] + SW_TX - 1) / SW_TX, nty = (L.d[1] + SW_TY - 1) / SW_TY, ntz = (L.d[2] + SW_TZ - 1) / SW_TZ;
  int nb = ntx * nty * ntz;
  // VEC: whole float4 rows, buffer loads with 32-bit byte offsets below 2 GiB,
  // row indices y + d1 z below 2^24 (sweep_hb's 24-bit multiplies)
  bool vec = (L.d[0] % 4) == 0 && (uint64_t)L.n * 4u < (1ull << 31) && (uint64_t)L.d[1] * L.d[2] <= (1ull << 24);
  if (L.wdir == 1.0f) {
    if (fine && vec)
      k_mg_sweep<true, true, true><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
    else if (vec)
      k_mg_sweep<false, true, true><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
    else
      k_mg_sweep<false, false, true><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
  } else {
    if (fine && vec)
      k_mg_sweep<true, true, false><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
    else if (vec)
      k_mg_sweep<false, true, false><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
    else
      k_mg_sweep<false, false, false><<<nb, SW_NT, 0, s>>>(a, ntx, nty, L.b, L.nc, pin, pout, omega);
  }
}

// ---------------------------------------------------------------- coarse
// All red-black SOR iterations of the coarsest level in one workgroup
// (sim.rs:472-473: 80 iterations, omega = 1.6).  phi lives in a zero-padded
// LDS grid (the ghost ring is the Dirichlet phi = 0 of mg.wgsl:29, solids are
// zeroed as in the sweep); each thread owns up to CO_K cells per colour and
// keeps their LDS index, count and h^2 b in registers.  Within one colour
// pass no cell reads a cell of its own colour, so the loads of a chunk of
// CO_C cells are issued together before any store.
constexpr int CO_NT = 1024, CO_K = 8, CO_C = 2;
constexpr int CO_PAD_MAX = 36000;  // padded cells: 144 KB of LDS (< 65536: 16-bit index)

__global__ __launch_bounds__(CO_NT) void k_mg_coarse(LvArgs a, const float *__restrict__ b,
                                                      const uint8_t *__restrict__ f, const uint8_t *__restrict__ nc,
                                                      float *__restrict__ p, int iters, float omega) {
  __shared__ float sp[CO_PAD_MAX];
  const int P0 = a.d0 + 2, P1 = a.d1 + 2, P2 = a.d2 + 2;
  const int PN = P0 * P1 * P2;
  const uint32_t sy = (uint32_t)a.d0, sz = (uint32_t)a.d0 * (uint32_t)a.d1;
  const int nt = (int)blockDim.x;  // whole waves covering a colour's cells
  for (int i = threadIdx.x; i < PN; i += nt) {
    int lx = i % P0, t = i / P0, ly = t % P1, lz = t / P1;
    int gx = lx - 1, gy = ly - 1, gz = lz - 1;
    float v = 0.0f;
    if (gx >= 0 && gy >= 0 && gz >= 0 && gx < a.d0 && gy < a.d1 && gz < a.d2) {
      uint32_t g = (uint32_t)gx + sy * (uint32_t)gy + sz * (uint32_t)gz;
      v = f[g] == 1 ? 0.0f : p[g];
    }
    sp[i] = v;
  }
  // per owned cell: LDS index (bits 0-15) | fluid (16-18) | outside (19-21) | active (31);
  // an inactive slot points at grid cell (0, 0, 0), whose six neighbours
  // are inside the padded array, and is never stored
  const uint32_t idle = (uint32_t)(1 + P0 + P0 * P1);
  uint32_t pk[2][CO_K];
  float hb[2][CO_K];
#pragma unroll
  for (int c = 0; c < 2; c++) {
#pragma unroll
    for (int j = 0; j < CO_K; j++) {
      pk[c][j] = idle;
      hb[c][j] = 0.0f;
      uint32_t t = threadIdx.x + (uint32_t)j * (uint32_t)nt;
      if (t >= a.nhalf) continue;
      uint32_t r = fdiv(t, a.dhx);
      int gxh = (int)(t - r * a.dhx.d);
      uint32_t zq = fdiv(r, a.dny);
      int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
      int x = 2 * gxh + ((y + z + c) & 1);
      if (x >= a.d0) continue;
      uint32_t g = (uint32_t)x + sy * (uint32_t)y + sz * (uint32_t)z;
      uint8_t code = nc[g];
      if (code & (NC_SOLID | NC_NONE)) continue;
      uint32_t li = (uint32_t)(((z + 1) * P1 + (y + 1)) * P0 + (x + 1));
      pk[c][j] = li | ((uint32_t)(code & 63u) << 16) | 0x80000000u;
      hb[c][j] = a.hh * b[g];
    }
  }
  __syncthreads();
  const int o1 = 1, o2 = P0, o3 = P0 * P1;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
      for (int j0 = 0; j0 < CO_K; j0 += CO_C) {
        if ((uint32_t)j0 * (uint32_t)nt >= a.nhalf) break;  // no thread owns a cell here (uniform)
        float nb[CO_C][7];
#pragma unroll
        for (int q = 0; q < CO_C; q++) {
          const int i = (int)(pk[c][j0 + q] & 0xffffu);
          nb[q][0] = sp[i + o1]; nb[q][1] = sp[i - o1]; nb[q][2] = sp[i + o2];
          nb[q][3] = sp[i - o2]; nb[q][4] = sp[i + o3]; nb[q][5] = sp[i - o3];
          nb[q][6] = sp[i];
        }
#pragma unroll
        for (int q = 0; q < CO_C; q++) {
          const uint32_t w = pk[c][j0 + q];
          float sum = 0.0f;
          sum += nb[q][0];
          sum += nb[q][1];
          sum += nb[q][2];
          sum += nb[q][3];
          sum += nb[q][4];
          sum += nb[q][5];
          const float cnt = nc_count(w >> 16, a.wdir);
          const float nv = mixf_(nb[q][6], (sum - hb[c][j0 + q]) / cnt, omega);
          if (w & 0x80000000u) sp[(int)(w & 0xffffu)] = nv;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < PN; i += nt) {
    int lx = i % P0, t = i / P0, ly = t % P1, lz = t / P1;
    int gx = lx - 1, gy = ly - 1, gz = lz - 1;
    if (gx >= 0 && gy >= 0 && gz >= 0 && gx < a.d0 && gy < a.d1 && gz < a.d2) {
      uint32_t g = (uint32_t)gx + sy * (uint32_t)gy + sz * (uint32_t)gz;
      if (f[g] != 1) p[g] = sp[i];
    }
  }
}

// The same solve with one thread per (x, y) column and the column's phi in
// registers, for an outside weight of 1 (integer counts 1..6, div_count).
// LDS holds phi plane-major: plane z + 1 (z = -1 .. ZJ - 1) of CZ_PLANE
// floats, the cell of column c = x + d0 * y at offset c.  CZ_PLANE is a
// multiple of 64, so a read or store of one slot by the 64 lanes of a wave
// (consecutive columns) touches 64 distinct banks.  Columns c >= d0 * d1 are
// zero; a lane on a side of the domain reads its missing x or y neighbour
// from the zero block at CZ_ZERO + (the neighbour's column mod 64), the bank
// the neighbour would have used, which gives the Dirichlet phi = 0 beyond the
// sides.  A cell reads its four x/y neighbours in LDS and its two z
// neighbours from registers.  Register slot j of a column holds
// z = j - ((x + y) & 1), so the cells of colour c are the slots j = c, c + 2,
// ... on every lane; slots outside the column hold phi = 0 and count 0 (each
// pass computes them like the others and they keep +0), which gives the
// Dirichlet phi = 0 above the top and below the bottom.
//
// EX (whole waves): the block has blockDim.x = 64 k < d0 d1 threads, one per
// column c < blockDim.x; the cells of the remaining columns (each SIMD issues
// its waves' VALU work in turn, so a wave of columns beyond a multiple of
// four adds a whole wave's slots to one SIMD) are taken one per thread and colour: extra
// index e = thread of colour c is leftover column blockDim.x + e % NL, cell
// z = ((c + parity) & 1) + 2 (e / NL), NL = d0 d1 - blockDim.x leftover
// columns (consecutive lanes, consecutive columns: distinct banks).  An extra cell
// reads all six neighbours in LDS (z +- 1: planes; the planes z = -1 and
// z = d2 hold +0) and keeps its phi in a register.
constexpr int CZ_ZERO = CO_NT, CZ_PLANE = CO_NT + 64, CZ_EX_NT = 768;

template <int ZJ, bool EX = false>
__global__ __launch_bounds__(EX ? CZ_EX_NT : CO_NT) void k_mg_coarse_z(LvArgs a, const float *__restrict__ b,
                                                                        const uint8_t *__restrict__ f,
                                                                        const uint8_t *__restrict__ nc,
                                                                        float *__restrict__ p, int iters,
                                                                        float omega) {
  constexpr int PN = (ZJ + 1) * CZ_PLANE;  // planes z + 1 = j + 1 - par for slots j < ZJ
  static_assert(PN <= CO_PAD_MAX, "planes exceed the LDS budget");
  static_assert(!EX || ZJ > 12, "extra cells are taken by the long-column passes");
  __shared__ float sp[PN];
  const uint32_t sy = (uint32_t)a.d0, sz = (uint32_t)a.d0 * (uint32_t)a.d1;
  for (int i = threadIdx.x; i < PN; i += (int)blockDim.x) sp[i] = 0.0f;
  const int ncol = a.d0 * a.d1;  // <= CO_NT (host)
  const int col = threadIdx.x;
  const bool has = col < ncol;
  const int x = has ? col % a.d0 : 0, y = has ? col / a.d0 : 0;
  const int par = (x + y) & 1;
  // LDS index of slot j is B + j * CZ_PLANE.  A lane without a column
  // (ncol <= col < CO_NT) reads columns col +- 1 and the zero block, all
  // inside a plane, and never stores.
  const int B = (1 - par) * CZ_PLANE + col;
  const int cxp = !has || x < a.d0 - 1 ? col + 1 : CZ_ZERO + ((col + 1) & 63);
  const int cxm = !has || x > 0 ? col - 1 : CZ_ZERO + ((col - 1) & 63);
  const int cyp = has && y < a.d1 - 1 ? col + a.d0 : CZ_ZERO + ((col + a.d0) & 63);
  const int cym = has && y > 0 ? col - a.d0 : CZ_ZERO + ((col - a.d0) & 63);
  float ph[ZJ + 2], hb[ZJ], rc[ZJ];
  uint32_t cp[(ZJ + 3) / 4];
  ph[0] = 0.0f;
  ph[ZJ + 1] = 0.0f;
#pragma unroll
  for (int k = 0; k < (ZJ + 3) / 4; k++) cp[k] = 0u;
#pragma unroll
  for (int j = 0; j < ZJ; j++) {
    const int z = j - par;
    // a slot that is never updated (outside the column, solid, or a cell
    // with no fluid neighbour and no open face) holds phi = 0, count 0 and
    // reciprocal 0: its quotient is then +-0 and the update keeps +0, so the
    // passes need no per-slot select.  Such a cell's phi is read only by
    // solid neighbours, which are not updated, and is not written back.
    float v = 0.0f, h = 0.0f, r = 0.0f;
    uint32_t cnt = 0u;
    if (has && z >= 0 && z < a.d2) {
      const uint32_t g = (uint32_t)x + sy * (uint32_t)y + sz * (uint32_t)z;
      const uint8_t code = nc[g];
      if (!(code & (NC_SOLID | NC_NONE))) {
        v = p[g];
        cnt = (code & 7u) + ((code >> 3) & 7u);
        h = a.hh * b[g];
        r = 1.0f / (float)cnt;
      }
    }
    ph[j + 1] = v;
    hb[j] = h;
    rc[j] = r;
    cp[j >> 2] |= cnt << (8 * (j & 3));
  }
  // extra cell of each colour (EX): LDS index, neighbour indices, count,
  // reciprocal, h^2 b, phi; a cell that is never updated holds 0 as above
  bool ev[2] = {false, false};
  int ei[2] = {0, 0}, en[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  float ec[2] = {0.0f, 0.0f}, er[2] = {0.0f, 0.0f}, eh[2] = {0.0f, 0.0f}, ep[2] = {0.0f, 0.0f};
  if constexpr (EX) {
    const int nl = ncol - (int)blockDim.x;  // 1 .. 63 leftover columns (the host launches EX only then)
    const int k = col / nl;
    const int cl = (int)blockDim.x + (col - k * nl);  // leftover column
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int xl = cl % a.d0, yl = cl / a.d0;
      const int z = ((c + xl + yl) & 1) + 2 * k;
      if (z >= a.d2) continue;
      const int pl = (z + 1) * CZ_PLANE;
      ev[c] = true;
      ei[c] = pl + cl;
      en[c][0] = pl + (xl < a.d0 - 1 ? cl + 1 : CZ_ZERO + ((cl + 1) & 63));
      en[c][1] = pl + (xl > 0 ? cl - 1 : CZ_ZERO + ((cl - 1) & 63));
      en[c][2] = pl + (yl < a.d1 - 1 ? cl + a.d0 : CZ_ZERO + ((cl + a.d0) & 63));
      en[c][3] = pl + (yl > 0 ? cl - a.d0 : CZ_ZERO + ((cl - a.d0) & 63));
      const uint32_t g = (uint32_t)xl + sy * (uint32_t)yl + sz * (uint32_t)z;
      const uint8_t code = nc[g];
      if (!(code & (NC_SOLID | NC_NONE))) {
        const uint32_t cnt = (code & 7u) + ((code >> 3) & 7u);
        ep[c] = p[g];
        ec[c] = (float)cnt;
        er[c] = 1.0f / (float)cnt;
        eh[c] = a.hh * b[g];
      }
    }
  }
  __syncthreads();  // the zero fill is complete
  if (has) {
#pragma unroll
    for (int j = 0; j < ZJ; j++) {
      const int z = j - par;
      if (z >= 0 && z < a.d2) sp[B + j * CZ_PLANE] = ph[j + 1];
    }
  }
  if constexpr (EX) {
#pragma unroll
    for (int c = 0; c < 2; c++)
      if (ev[c]) sp[ei[c]] = ep[c];
  }
  __syncthreads();
  const int pb = (1 - par) * CZ_PLANE;
  const float *sxp = sp + pb + cxp, *sxm = sp + pb + cxm, *syp = sp + pb + cyp, *sym = sp + pb + cym;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
      for (int k = 0; k < (ZJ + 3) / 4; k++) asm volatile("" : "+v"(cp[k]));  // conversions stay in the loop
      // every slot of the colour is computed and stored (reads and stores
      // stay inside sp: planes j + 1 - par <= ZJ, columns < CZ_PLANE; a slot
      // that is never updated stores its +0); the stores follow all reads of
      // the pass, so no read waits on them.  Each sum starts with its first
      // term rather than +0 + term: the two differ only in the sign of a
      // zero sum, and the update maps either to the same value.
      if constexpr (ZJ <= 12) {
        // Short columns (one wave per SIMD on small levels): the slots of a
        // colour are independent (red-black), so all of the pass's LDS reads
        // are issued first, the fast quotients of every slot are formed
        // branch-free, and one branch per pass redoes by IEEE division any
        // quotient that came out subnormal (div_count); the slots' latencies
        // overlap instead of queueing behind one another.
        constexpr int NS = ZJ / 2;
        float xs[NS], qs[NS], nb[4][NS];
        bool redo = false;
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const int j = c + 2 * t;
          nb[0][t] = sxp[j * CZ_PLANE];
          nb[1][t] = sxm[j * CZ_PLANE];
          nb[2][t] = syp[j * CZ_PLANE];
          nb[3][t] = sym[j * CZ_PLANE];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const int j = c + 2 * t;
          const float cn = (float)((cp[j >> 2] >> (8 * (j & 3))) & 0xffu);
          float sum = nb[0][t];
          sum += nb[1][t];
          sum += nb[2][t];
          sum += nb[3][t];
          sum += ph[j + 2];
          sum += ph[j];
          const float x = sum - hb[j];
          const float q0 = x * rc[j];
          const float e = __builtin_fmaf(-q0, cn, x);
          xs[t] = x;
          qs[t] = __builtin_fmaf(e, rc[j], q0);
          redo |= __builtin_amdgcn_classf(qs[t], 0x90);
        }
        if (redo) {
#pragma unroll
          for (int t = 0; t < NS; t++) {
            const int j = c + 2 * t;
            if (__builtin_amdgcn_classf(qs[t], 0x90)) {
              float x = xs[t];
              asm volatile("" : "+v"(x));
              qs[t] = x / (float)((cp[j >> 2] >> (8 * (j & 3))) & 0xffu);
            }
          }
        }
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const int j = c + 2 * t;
          ph[j + 1] = mixf_(ph[j + 1], qs[t], omega);
        }
      } else {
        // Long columns: many waves share each SIMD and hide the latency;
        // registers are the limit (no per-pass arrays).  The next slot's
        // four reads are issued before this slot's arithmetic.
        float nb[4] = {sxp[c * CZ_PLANE], sxm[c * CZ_PLANE], syp[c * CZ_PLANE], sym[c * CZ_PLANE]};
#pragma unroll
        for (int j = c; j < ZJ; j += 2) {
          float nx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
          if (j + 2 < ZJ) {
            nx[0] = sxp[(j + 2) * CZ_PLANE];
            nx[1] = sxm[(j + 2) * CZ_PLANE];
            nx[2] = syp[(j + 2) * CZ_PLANE];
            nx[3] = sym[(j + 2) * CZ_PLANE];
          }
          const float cn = (float)((cp[j >> 2] >> (8 * (j & 3))) & 0xffu);
          float sum = nb[0];
          sum += nb[1];
          sum += nb[2];
          sum += nb[3];
          sum += ph[j + 2];
          sum += ph[j];
          ph[j + 1] = mixf_(ph[j + 1], div_count(sum - hb[j], cn, rc[j]), omega);
#pragma unroll
          for (int k = 0; k < 4; k++) nb[k] = nx[k];
        }
      }
#pragma unroll
      for (int j = c; j < ZJ; j += 2) sp[B + j * CZ_PLANE] = ph[j + 1];
      if constexpr (EX) {
        if (ev[c]) {
          const int i = ei[c];
          float sum = sp[en[c][0]];
          sum += sp[en[c][1]];
          sum += sp[en[c][2]];
          sum += sp[en[c][3]];
          sum += sp[i + CZ_PLANE];
          sum += sp[i - CZ_PLANE];
          ep[c] = mixf_(ep[c], div_count(sum - eh[c], ec[c], er[c]), omega);
          sp[i] = ep[c];
        }
      }
      __syncthreads();
    }
  }
  if (has) {
#pragma unroll
    for (int j = 0; j < ZJ; j++) {
      const int z = j - par;
      if (z >= 0 && z < a.d2) {
        const uint32_t g = (uint32_t)x + sy * (uint32_t)y + sz * (uint32_t)z;
        if (!(nc[g] & (NC_SOLID | NC_NONE))) p[g] = ph[j + 1];
      }
    }
  }
  if constexpr (EX) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (!ev[c]) continue;
      const int z = ei[c] / CZ_PLANE - 1, cl = ei[c] - (z + 1) * CZ_PLANE;
      const uint32_t g = (uint32_t)cl + sz * (uint32_t)z;  // column index cl = x + d0 y
      if (!(nc[g] & (NC_SOLID | NC_NONE))) p[g] = ep[c];
    }
  }
}

// Same schedule for a coarsest level larger than LDS: one workgroup, phi in
// global memory (L2-resident), a workgroup barrier between colour passes.
__global__ __launch_bounds__(1024) void k_mg_coarse_global(LvArgs a, const float *__restrict__ b,
                                                            const uint8_t *__restrict__ f, float *p, int iters,
                                                            float omega) {
  for (int it = 0; it < iters; it++) {
    for (int color = 0; color < 2; color++) {
      for (uint32_t t = threadIdx.x; t < a.nhalf; t += 1024) {
        uint32_t r = fdiv(t, a.dhx);
        int gx = (int)(t - r * a.dhx.d);
        uint32_t zq = fdiv(r, a.dny);
        int y = (int)(r - zq * (uint32_t)a.d1), z = (int)zq;
        int x = 2 * gx + ((y + z + color) & 1);
        if (x >= a.d0) continue;
        uint32_t i = (uint32_t)x + (uint32_t)a.d0 * ((uint32_t)y + (uint32_t)a.d1 * (uint32_t)z);
        if (f[i] == 1) continue;
        float sum, cnt;
        nsum(a.d0, a.d1, a.d2, p, f, x, y, z, i, a.wdir, sum, cnt);
        if (cnt < 0.5f) continue;
        p[i] = mixf_(p[i], (sum - a.hh * b[i]) / cnt, omega);
      }
      __syncthreads();
    }
  }
}

bool mg_coarse_fits_lds(const MgLevel &L) {
  uint32_t pad = (uint32_t)(L.d[0] + 2) * (uint32_t)(L.d[1] + 2) * (uint32_t)(L.d[2] + 2);
  uint32_t half = (uint32_t)((L.d[0] + 1) / 2) * (uint32_t)L.d[1] * (uint32_t)L.d[2];
  return pad <= (uint32_t)CO_PAD_MAX && half <= (uint32_t)(CO_NT * CO_K);
}

void launch_mg_coarse(const MgLevel &L, int iters, float omega, hipStream_t s) {
  LvArgs a = lv_args(L);
  const int ncol = L.d[0] * L.d[1];
  const bool colz = L.wdir == 1.0f && ncol <= CO_NT;
  // one thread per column, whole waves only: a wave without columns would
  // still run every colour pass's arithmetic on the CU
  const int colt = ((ncol + 63) / 64) * 64;
  if (colz && L.d[2] + 1 <= 12) {
    k_mg_coarse_z<12><<<1, colt, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
    return;
  }
  if (colz && L.d[2] + 1 <= 22) {
    // waves run on SIMD w % 4.  Per iteration (both colours) a SIMD issues
    // ZJ slots for each of its column waves plus two for each of its waves
    // holding extra cells: take the extra-cell layout when its busiest SIMD
    // issues fewer slots.
    const int zj = L.d[2] + 1 <= 21 ? 21 : 22;
    const int full = (ncol / 64) * 64, zh = (L.d[2] + 1) / 2;
    const int wc = colt / 64, wf = full / 64, we = ((ncol - full) * zh + 63) / 64;
    const int load_col = zj * ((wc + 3) / 4);
    int load_ex = 0;
    for (int sm = 0; sm < 4; sm++) {
      int l = 0;
      for (int w = sm; w < wf; w += 4) l += zj + (w < we ? 2 : 0);
      load_ex = l > load_ex ? l : load_ex;
    }
    const bool ex = full >= 64 && full <= CZ_EX_NT && ncol > full && (ncol - full) * zh <= full && load_ex < load_col;
    if (zj == 21) {
      if (ex)
        k_mg_coarse_z<21, true><<<1, full, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
      else
        k_mg_coarse_z<21><<<1, colt, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
    } else {
      if (ex)
        k_mg_coarse_z<22, true><<<1, full, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
      else
        k_mg_coarse_z<22><<<1, colt, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
    }
    return;
  }
  if (mg_coarse_fits_lds(L)) {
    // one cell of a colour per thread where that fits, whole waves only
    // (the level fits: nhalf <= CO_NT * CO_K)
    const uint32_t nt = a.nhalf >= (uint32_t)CO_NT ? (uint32_t)CO_NT : ((a.nhalf + 63u) / 64u) * 64u;
    k_mg_coarse<<<1, nt, 0, s>>>(a, L.b, L.f, L.nc, L.phi, iters, omega);
  }
  else
    k_mg_coarse_global<<<1, 1024, 0, s>>>(a, L.b, L.f, L.phi, iters, omega);
}

// ---------------------------------------------------------------- resid
// Per-block FP64 sums of r^2 and b^2 over fluid cells (mg.wgsl:134-158).
__global__ __launch_bounds__(NT) void k_mg_resid(LvArgs a, FastDiv d0, const float *__restrict__ p,
                                                  const float *__restrict__ b, const uint8_t *__restrict__ f,
                                                  double *__restrict__ part) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
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

// ---------------------------------------------------------------- V-cycle
void enqueue_vcycle(const MgLevel *lv, int nl, int coarse_iters, hipStream_t s) {
  auto sweeps = [&](const MgLevel &L, bool fine) {
    launch_mg_sweep(L, L.phi, L.tmp, 1.0f, s, fine);
    launch_mg_sweep(L, L.tmp, L.phi, 1.0f, s, fine);
  };
  for (int l = 0; l < nl - 1; l++) {
    sweeps(lv[l], l == 0);
    launch_mg_restrict(lv[l], lv[l + 1], s);
  }
  launch_mg_coarse(lv[nl - 1], coarse_iters, 1.6f, s);
  for (int l = nl - 2; l >= 0; l--) {
    launch_mg_prolong(lv[l + 1], lv[l], s);
    sweeps(lv[l], l == 0);
  }
}

}  // namespace pyro
