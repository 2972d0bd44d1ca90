/*
 * pyro_oracle.c -- CPU restatement of pyro's per-timestep solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * pyro_amd/csrc; it is loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.
 *
 * Every function follows the WGSL text it cites (paths relative to the
 * reference repository).  Arithmetic is FP32 in the WGSL operation order and
 * must be compiled with -ffp-contract=off.  Builtins whose bits the WGSL spec
 * leaves to the backend are fixed as follows (the reference was validated on
 * Metal, README.md:25):
 *   mix(a,b,t)        = a + (b - a) * t          (Metal's documented form)
 *   smoothstep(e0,e1,x)= t*t*(3-2t), t = clamp((x-e0)/(e1-e0),0,1)
 *   length(v)         = sqrt((x*x + y*y) + z*z)   (IEEE sqrt)
 *   exp(x)            = oracle_expf: Cody-Waite reduction by ln2 and a
 *                       degree-7 Taylor polynomial, <= 2 ulp (tested).
 * Stats are FP64 sums in a fixed order.
 */
#include "pyro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define STATS_N 1040
#define MG_MAX_LEVELS 8

typedef struct { float a[4], b[4], c[4]; } OMat;
typedef struct {
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
  OMat mats[8];
} OParams; /* scene.rs:426-477 == common.wgsl:14-28, 576 bytes */

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } v4;
typedef struct { int x, y, z; } i3;

typedef struct {
  uint32_t d[3];
  float h;
  float *phi, *b;
  uint32_t *f;
  int own_f;
} OLevel;

struct OSim {
  OParams P;
  size_t n, nv;
  float *vel[2][3], *hat[3];
  v4 *s0, *s1, *s0b, *s1b, *h0, *h1, *sol[2], *pyr, *curl;
  uint32_t *flags;
  OLevel lv[MG_MAX_LEVELS];
  int nl;
  int cur, sol_cur;
  uint32_t cycles;
  float ign_until;
  double time;
  int want_stats;
  double stats[STATS_N];
};

struct OMg {
  OLevel lv[MG_MAX_LEVELS];
  int nl;
};

static const float SIGMA = 5.670374e-8f;
static const float RGAS = 8.314f;
static const float RHO_SOOT = 1800.0f;
static const i3 DIRS[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

void osim_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Builtins                                                            */
/* ------------------------------------------------------------------ */
static inline float fmin_(float a, float b) { return a < b ? a : b; }
static inline float fmax_(float a, float b) { return a > b ? a : b; }
static inline float clampf_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
static inline float mixf_(float a, float b, float t) { return a + (b - a) * t; }
static inline float smoothstep_(float e0, float e1, float x) {
  float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
static inline float len3(v3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline float pow4(float x) { float y = x * x; return y * y; }

/* exp by Cody-Waite range reduction (x = k ln2 + r, |r| <= ln2/2, ln2 split
 * into a 16-bit head and a tail) and the degree-7 Taylor polynomial of e^r
 * evaluated by Horner's rule with fused multiply-adds; 2^k applied by exponent
 * construction (two steps for subnormal results). */
float oracle_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.97208f) return 0.0f;
  float kf = rintf(x * 1.44269504088896341f);
  float r = fmaf(-kf, 0.693145751953125f, x);
  r = fmaf(-kf, 1.428606765330187e-06f, r);
  float p = 1.9841270e-4f;          /* 1/5040 */
  p = fmaf(p, r, 1.3888889e-3f);    /* 1/720 */
  p = fmaf(p, r, 8.3333333e-3f);    /* 1/120 */
  p = fmaf(p, r, 4.1666667e-2f);    /* 1/24 */
  p = fmaf(p, r, 1.6666667e-1f);    /* 1/6 */
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  int k = (int)kf;
  if (k < -125) { p = p * 5.42101086242752217e-20f; k += 64; } /* 2^-64 */
  if (k > 127) { p = p * 2.0f; k -= 1; }
  union { uint32_t u; float f; } sc;
  sc.u = (uint32_t)(k + 127) << 23;
  return p * sc.f;
}
#define EXPF oracle_expf

/* common.wgsl:54-58 */
static inline uint32_t pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

/* Harness noise (no reference counterpart): U[-1,1] from the PCG hash of
 * (component, face index, seed); 24 random bits -> exact float. */
float oracle_noise(uint32_t comp, uint32_t idx, uint32_t seed) {
  uint32_t h = pcg(pcg(idx * 3u + comp) ^ (seed * 0x9e3779b9u));
  float u01 = (float)(h >> 8) * (1.0f / 16777216.0f);
  return 2.0f * u01 - 1.0f;
}

/* ------------------------------------------------------------------ */
/* Indexing and helpers: common.wgsl:38-89                             */
/* ------------------------------------------------------------------ */
#define NX ((int)s->P.nx)
#define NY ((int)s->P.ny)
#define NZ ((int)s->P.nz)
static inline int inb(const OSim *s, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < NX && y < NY && z < NZ;
}
static inline size_t sidx(const OSim *s, int x, int y, int z) {
  return (size_t)x + (size_t)NX * ((size_t)y + (size_t)NY * (size_t)z);
}
static inline size_t vidx(const OSim *s, int x, int y, int z) {
  return (size_t)x + (size_t)(NX + 1) * ((size_t)y + (size_t)(NY + 1) * (size_t)z);
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float rhoOf(const OSim *s, float T) { return s->P.rho0 * s->P.t0 / fmax_(T, 200.0f); }
static inline float kapP(const OSim *s, float T, float yo, float ys) {
  float fv = rhoOf(s, T) * fmax_(ys, 0.0f) / RHO_SOOT;
  return 1862.0f * fv * T + s->P.kap_gas * clampf_(1.0f - yo / s->P.yo2, 0.0f, 1.0f);
}
static inline float cpOf(float T) { return clampf_(1000.0f + 0.24f * (T - 300.0f), 1000.0f, 1450.0f); }
static inline v3 compOff(int k) {
  v3 r = {0.5f, 0.5f, 0.5f};
  if (k == 0) r.x = 0.0f; else if (k == 1) r.y = 0.0f; else r.z = 0.0f;
  return r;
}
static inline i3 compHi(const OSim *s, int k) {
  i3 d = {NX - 1, NY - 1, NZ - 1};
  if (k == 0) d.x += 1; else if (k == 1) d.y += 1; else d.z += 1;
  return d;
}
static inline int compValid(const OSim *s, int k, int x, int y, int z) {
  i3 h = compHi(s, k);
  return x >= 0 && y >= 0 && z >= 0 && x <= h.x && y <= h.y && z <= h.z;
}
static inline float outside(const OSim *s, v3 gc) {
  float dxf = (float)NX, dyf = (float)NY, dzf = (float)NZ;
  return fmax_(fmax_(-0.5f - gc.x, gc.x - (dxf - 0.5f)),
               fmax_(fmax_(-0.5f - gc.z, gc.z - (dzf - 0.5f)), gc.y - (dyf - 0.5f)));
}
static inline float spongeSide(const OSim *s, int x, int z) {
  int e = x < NX - 1 - x ? x : NX - 1 - x;
  int e2 = z < NZ - 1 - z ? z : NZ - 1 - z;
  if (e2 < e) e = e2;
  return 1.0f - smoothstep_(0.0f, s->P.sponge, (float)e);
}
static inline float spongeTop(const OSim *s, int y) {
  float h = (float)NY;
  return smoothstep_(h - 1.5f * s->P.sponge, h - 1.0f, (float)y);
}

/* tmpl_comp.wgsl:2-15 */
static float triC(const OSim *s, const float *X, v3 g0, i3 hi) {
  float gx = clampf_(g0.x, 0.0f, (float)hi.x), gy = clampf_(g0.y, 0.0f, (float)hi.y),
        gz = clampf_(g0.z, 0.0f, (float)hi.z);
  int bx = (int)floorf(gx), by = (int)floorf(gy), bz = (int)floorf(gz);
  if (bx > hi.x - 1) bx = hi.x - 1;
  if (by > hi.y - 1) by = hi.y - 1;
  if (bz > hi.z - 1) bz = hi.z - 1;
  float fx = gx - (float)bx, fy = gy - (float)by, fz = gz - (float)bz;
  size_t sy = (size_t)(NX + 1), sz = (size_t)(NX + 1) * (size_t)(NY + 1);
  size_t i = vidx(s, bx, by, bz);
  float c00 = mixf_(X[i], X[i + 1], fx);
  float c10 = mixf_(X[i + sy], X[i + sy + 1], fx);
  float c01 = mixf_(X[i + sz], X[i + sz + 1], fx);
  float c11 = mixf_(X[i + sy + sz], X[i + sy + sz + 1], fx);
  return mixf_(mixf_(c00, c10, fy), mixf_(c01, c11, fy), fz);
}
/* tmpl_comp.wgsl:16-25 */
static void mmC(const OSim *s, const float *X, v3 g0, i3 hi, float *lo, float *hi_out) {
  float gx = clampf_(g0.x, 0.0f, (float)hi.x), gy = clampf_(g0.y, 0.0f, (float)hi.y),
        gz = clampf_(g0.z, 0.0f, (float)hi.z);
  int bx = (int)floorf(gx), by = (int)floorf(gy), bz = (int)floorf(gz);
  if (bx > hi.x - 1) bx = hi.x - 1;
  if (by > hi.y - 1) by = hi.y - 1;
  if (bz > hi.z - 1) bz = hi.z - 1;
  float rl = 1e30f, rh = -1e30f;
  for (int j = 0; j < 8; j++) {
    float v = X[vidx(s, bx + (j & 1), by + ((j >> 1) & 1), bz + ((j >> 2) & 1))];
    rl = fmin_(rl, v);
    rh = fmax_(rh, v);
  }
  *lo = rl;
  *hi_out = rh;
}
/* tmpl_scalar.wgsl:2-15 */
static v4 triS(const OSim *s, const v4 *X, v3 gc0) {
  float gx = clampf_(gc0.x, 0.0f, (float)(NX - 1)), gy = clampf_(gc0.y, 0.0f, (float)(NY - 1)),
        gz = clampf_(gc0.z, 0.0f, (float)(NZ - 1));
  int bx = (int)floorf(gx), by = (int)floorf(gy), bz = (int)floorf(gz);
  if (bx > NX - 2) bx = NX - 2;
  if (by > NY - 2) by = NY - 2;
  if (bz > NZ - 2) bz = NZ - 2;
  float fx = gx - (float)bx, fy = gy - (float)by, fz = gz - (float)bz;
  size_t sy = (size_t)NX, sz = (size_t)NX * (size_t)NY;
  size_t i = sidx(s, bx, by, bz);
  const float *a000 = &X[i].x, *a100 = &X[i + 1].x, *a010 = &X[i + sy].x, *a110 = &X[i + sy + 1].x;
  const float *a001 = &X[i + sz].x, *a101 = &X[i + sz + 1].x, *a011 = &X[i + sy + sz].x,
              *a111 = &X[i + sy + sz + 1].x;
  float r[4];
  for (int q = 0; q < 4; q++) {
    float c00 = mixf_(a000[q], a100[q], fx);
    float c10 = mixf_(a010[q], a110[q], fx);
    float c01 = mixf_(a001[q], a101[q], fx);
    float c11 = mixf_(a011[q], a111[q], fx);
    r[q] = mixf_(mixf_(c00, c10, fy), mixf_(c01, c11, fy), fz);
  }
  v4 o = {r[0], r[1], r[2], r[3]};
  return o;
}
/* tmpl_scalar.wgsl:17-19 */
static v4 advS(const OSim *s, const v4 *X, v3 gc, v4 amb) {
  v4 t = triS(s, X, gc);
  float w = clampf_(outside(s, gc), 0.0f, 1.0f);
  v4 o = {mixf_(t.x, amb.x, w), mixf_(t.y, amb.y, w), mixf_(t.z, amb.z, w), mixf_(t.w, amb.w, w)};
  return o;
}
/* tmpl_scalar.wgsl:20-33 */
static void mmS(const OSim *s, const v4 *X, v3 gc0, v4 *lo, v4 *hi) {
  float gx = clampf_(gc0.x, 0.0f, (float)(NX - 1)), gy = clampf_(gc0.y, 0.0f, (float)(NY - 1)),
        gz = clampf_(gc0.z, 0.0f, (float)(NZ - 1));
  int bx = (int)floorf(gx), by = (int)floorf(gy), bz = (int)floorf(gz);
  if (bx > NX - 2) bx = NX - 2;
  if (by > NY - 2) by = NY - 2;
  if (bz > NZ - 2) bz = NZ - 2;
  v4 l = {1e30f, 1e30f, 1e30f, 1e30f}, h = {-1e30f, -1e30f, -1e30f, -1e30f};
  for (int j = 0; j < 8; j++) {
    v4 v = X[sidx(s, bx + (j & 1), by + ((j >> 1) & 1), bz + ((j >> 2) & 1))];
    l.x = fmin_(l.x, v.x); l.y = fmin_(l.y, v.y); l.z = fmin_(l.z, v.z); l.w = fmin_(l.w, v.w);
    h.x = fmax_(h.x, v.x); h.y = fmax_(h.y, v.y); h.z = fmax_(h.z, v.z); h.w = fmax_(h.w, v.w);
  }
  *lo = l;
  *hi = h;
}

/* vel.wgsl:2-19 (velocity field = vel[cur]) */
static v3 velAt(const OSim *s, const float *const *V, v3 p) {
  v3 r;
  v3 o0 = compOff(0), o1 = compOff(1), o2 = compOff(2);
  v3 q0 = {p.x - o0.x, p.y - o0.y, p.z - o0.z};
  v3 q1 = {p.x - o1.x, p.y - o1.y, p.z - o1.z};
  v3 q2 = {p.x - o2.x, p.y - o2.y, p.z - o2.z};
  r.x = triC(s, V[0], q0, compHi(s, 0));
  r.y = triC(s, V[1], q1, compHi(s, 1));
  r.z = triC(s, V[2], q2, compHi(s, 2));
  return r;
}
static v3 traceBack(const OSim *s, const float *const *V, v3 x) {
  float k = s->P.dt / s->P.dx;
  v3 u1 = velAt(s, V, x);
  float hk = 0.5f * k;
  v3 m = {x.x - hk * u1.x, x.y - hk * u1.y, x.z - hk * u1.z};
  v3 u2 = velAt(s, V, m);
  v3 r = {x.x - k * u2.x, x.y - k * u2.y, x.z - k * u2.z};
  return r;
}
static v3 traceFwd(const OSim *s, const float *const *V, v3 x) {
  float k = s->P.dt / s->P.dx;
  v3 u1 = velAt(s, V, x);
  float hk = 0.5f * k;
  v3 m = {x.x + hk * u1.x, x.y + hk * u1.y, x.z + hk * u1.z};
  v3 u2 = velAt(s, V, m);
  v3 r = {x.x + k * u2.x, x.y + k * u2.y, x.z + k * u2.z};
  return r;
}

/* ------------------------------------------------------------------ */
/* solid: combust.wgsl:3-137                                           */
/* ------------------------------------------------------------------ */
/* material slot, clamped to the 8-entry table (malformed input guard) */
static inline uint32_t slotOf(v4 so) {
  uint32_t k = (uint32_t)so.x;
  return k > 7u ? 7u : k;
}
static inline uint32_t kindOf(const OSim *s, v4 so) {
  return (uint32_t)(s->P.mats[slotOf(so)].a[3] + 0.5f);
}
static inline int isSolidS(const OSim *s, v4 so) {
  if (kindOf(s, so) == 2u) return so.y > 0.5f;
  return so.y + so.z > s->P.burnout;
}
static inline int solidN(const OSim *s, const v4 *solIn, int x, int y, int z) {
  if (!inb(s, x, y, z)) return 0;
  return isSolidS(s, solIn[sidx(s, x, y, z)]);
}
/* combust.wgsl:18-39 */
static float rayIrr(const OSim *s, const v4 *solIn, int px, int py, int pz, i3 d) {
  static const int strides[10] = {1, 1, 1, 2, 2, 3, 4, 5, 7, 9};
  float tr = 1.0f, acc = 0.0f;
  v3 df = {(float)d.x, (float)d.y, (float)d.z};
  float dl = len3(df) * s->P.dx;
  for (int st = 0; st < 10; st++) {
    if (!inb(s, px, py, pz)) break;
    size_t j = sidx(s, px, py, pz);
    v4 so = solIn[j];
    if (isSolidS(s, so)) { acc += tr * pow4(so.w); tr = 0.0f; break; }
    v4 g = s->s0[j];
    float T = g.x;
    float len = (float)strides[st] * dl;
    float e = 1.0f - EXPF(-kapP(s, T, g.z, g.w) * len);
    acc += tr * e * pow4(T);
    tr *= 1.0f - e;
    px += d.x * strides[st];
    py += d.y * strides[st];
    pz += d.z * strides[st];
  }
  acc += tr * pow4(s->P.t0);
  return SIGMA * acc;
}
/* combust.wgsl:41-48 */
static float faceIrr(const OSim *s, const v4 *solIn, int x, int y, int z, int k) {
  i3 d = DIRS[k];
  i3 t1 = {0, 1, 0}, t2 = {0, 0, 1};
  if (d.y != 0) { t1.x = 1; t1.y = 0; t1.z = 0; }
  if (d.z != 0) { t2.x = 1; t2.y = 0; t2.z = 0; }
  i3 a = {d.x + t1.x, d.y + t1.y, d.z + t1.z}, b = {d.x - t1.x, d.y - t1.y, d.z - t1.z};
  i3 c = {d.x + t2.x, d.y + t2.y, d.z + t2.z}, e = {d.x - t2.x, d.y - t2.y, d.z - t2.z};
  float r0 = rayIrr(s, solIn, x, y, z, d);
  float r1 = rayIrr(s, solIn, x, y, z, a);
  float r2 = rayIrr(s, solIn, x, y, z, b);
  float r3 = rayIrr(s, solIn, x, y, z, c);
  float r4 = rayIrr(s, solIn, x, y, z, e);
  return 0.2f * ((((r0 + r1) + r2) + r3) + r4);
}

static void stage_solid(OSim *s) {
  const v4 *solIn = s->sol[s->sol_cur];
  v4 *solOut = s->sol[1 - s->sol_cur];
  const OParams *P = &s->P;
  double *zp = calloc((size_t)NZ * 3, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++) {
    double aP = 0, aM = 0, aA = 0;
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t i = sidx(s, x, y, z);
        v4 so = solIn[i];
        v4 o = so;
        v4 po = {0, 0, 0, 0};
        uint32_t fl = 0;
        float stPyro = 0.0f, stMass = 0.0f, stArea = 0.0f;
        if (isSolidS(s, so)) {
          fl = 1;
          const OMat *m = &P->mats[slotOf(so)];
          uint32_t kind = kindOf(s, so);
          float fuel = so.y, chr = so.z, Ts = so.w;
          if (kind != 2u) {
            float nF = 0, Tg = 0, Yo = 0, qin = 0, qcond = 0;
            for (int k = 0; k < 6; k++) {
              int nx_ = x + DIRS[k].x, ny_ = y + DIRS[k].y, nz_ = z + DIRS[k].z;
              if (ny_ < 0) continue;
              if (solidN(s, solIn, nx_, ny_, nz_)) {
                qcond += m->b[3] / P->dx * (solIn[sidx(s, nx_, ny_, nz_)].w - Ts);
                continue;
              }
              nF += 1.0f;
              if (inb(s, nx_, ny_, nz_)) {
                v4 g = s->s0[sidx(s, nx_, ny_, nz_)];
                Tg += g.x;
                Yo += g.z;
                qin += faceIrr(s, solIn, nx_, ny_, nz_, k);
              } else {
                Tg += P->t0;
                Yo += P->yo2;
                qin += SIGMA * pow4(P->t0);
              }
            }
            float layer = kind == 1u ? 0.001f : fmin_(P->layer, P->dx);
            float cap = m->a[0] * m->b[2] * layer;
            if (nF > 0.0f) {
              Tg /= nF;
              Yo /= nF;
              qin /= nF;
              float qnet = m->b[1] * (qin - SIGMA * pow4(Ts)) + P->h_conv * (Tg - Ts) + qcond / nF;
              if (P->ign_on > 0.5f) {
                v3 dv = {((float)x + 0.5f) * P->dx - P->ign_x, ((float)y + 0.5f) * P->dx - P->ign_y,
                         ((float)z + 0.5f) * P->dx - P->ign_z};
                float dI = len3(dv);
                qnet += P->ign_flux * clampf_(1.0f - dI / (1.5f * P->ign_r), 0.0f, 1.0f);
              }
              float mv = 0.0f, mc = 0.0f;
              if (kind == 1u) {
                float Tb = m->a[2];
                mv = smoothstep_(Tb - 60.0f, Tb, Ts) * fmax_(qnet, 0.0f) / m->a[1] * m->c[1];
                Ts = fmin_(Ts + P->dt * fmax_(qnet - mv * m->a[1], fmin_(qnet, 0.0f)) / cap, Tb);
                fuel -= mv * nF / P->dx * P->dt;
              } else {
                if (fuel > 0.0f)
                  mv = smoothstep_(m->a[2] - 20.0f, m->a[2] + 30.0f, Ts) * fmax_(qnet, 0.0f) / m->a[1] * m->c[1];
                float chi = kind == 0u ? m->b[0] : 0.0f;
                float kin = P->char_a * EXPF(-P->char_e / (RGAS * Ts)) * chr * layer;
                float dif = P->h_mass * rhoOf(s, Tg) * Yo / 1.33f;
                mc = fmin_(kin, dif);
                Ts += P->dt * (qnet - mv * m->a[1] + 0.5f * mc * P->dh_char) / cap;
                float virgin = fmin_(mv / (1.0f - chi) * nF / P->dx * P->dt, fuel);
                fuel -= virgin;
                chr += virgin * chi - mc * nF / P->dx * P->dt;
              }
              stPyro = mv * nF * P->dx * P->dx;
              stArea = mv > 1e-4f ? nF * P->dx * P->dx : 0.0f;
              po.x = mv; po.y = mc; po.z = nF; po.w = 0.0f;
            } else {
              Ts += P->dt * qcond / (m->a[0] * m->b[2] * P->dx);
            }
            Ts = clampf_(Ts, 250.0f, 2000.0f);
            fuel = fmax_(fuel, 0.0f);
            chr = fmax_(chr, 0.0f);
            o.x = so.x; o.y = fuel; o.z = chr; o.w = Ts;
            if (fuel + chr <= P->burnout) {
              o.x = 0; o.y = 0; o.z = 0; o.w = P->t0;
              fl = 0;
              po.x = po.y = po.z = po.w = 0.0f;
            }
          }
          stMass = (o.y + o.z) * P->dx * P->dx * P->dx * (kind == 2u ? 0.0f : 1.0f);
        }
        solOut[i] = o;
        s->pyr[i] = po;
        s->flags[i] = fl;
        aP += stPyro;
        aM += stMass;
        aA += stArea;
      }
    zp[3 * z] = aP; zp[3 * z + 1] = aM; zp[3 * z + 2] = aA;
  }
  if (s->want_stats)
    for (int z = 0; z < NZ; z++) {
      s->stats[2] += zp[3 * z]; s->stats[7] += zp[3 * z + 1]; s->stats[9] += zp[3 * z + 2];
    }
  free(zp);
}

/* ------------------------------------------------------------------ */
/* curl: fluid.wgsl:3-29                                               */
/* ------------------------------------------------------------------ */
static v3 uc(const OSim *s, const float *const *V, int x, int y, int z) {
  x = clampi(x, 0, NX - 1); y = clampi(y, 0, NY - 1); z = clampi(z, 0, NZ - 1);
  v3 r = {0.5f * (V[0][vidx(s, x, y, z)] + V[0][vidx(s, x + 1, y, z)]),
          0.5f * (V[1][vidx(s, x, y, z)] + V[1][vidx(s, x, y + 1, z)]),
          0.5f * (V[2][vidx(s, x, y, z)] + V[2][vidx(s, x, y, z + 1)])};
  return r;
}
static void stage_curl(OSim *s) {
  const float *const *V = (const float *const *)s->vel[s->cur];
  double *zm = calloc((size_t)NZ, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++) {
    float mx = 0.0f;
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        float tdx = 2.0f * s->P.dx;
        v3 w = {((uc(s, V, x, y + 1, z).z - uc(s, V, x, y - 1, z).z) - (uc(s, V, x, y, z + 1).y - uc(s, V, x, y, z - 1).y)) / tdx,
                ((uc(s, V, x, y, z + 1).x - uc(s, V, x, y, z - 1).x) - (uc(s, V, x + 1, y, z).z - uc(s, V, x - 1, y, z).z)) / tdx,
                ((uc(s, V, x + 1, y, z).y - uc(s, V, x - 1, y, z).y) - (uc(s, V, x, y + 1, z).x - uc(s, V, x, y - 1, z).x)) / tdx};
        v4 o = {w.x, w.y, w.z, len3(w)};
        s->curl[sidx(s, x, y, z)] = o;
        float spd = len3(uc(s, V, x, y, z));
        mx = fmax_(mx, spd);
      }
    zm[z] = mx;
  }
  if (s->want_stats)
    for (int z = 0; z < NZ; z++) s->stats[5] = fmax(s->stats[5], zm[z]);
  free(zm);
}

/* ------------------------------------------------------------------ */
/* Advection: advect.wgsl                                              */
/* ------------------------------------------------------------------ */
static void stage_adv_vel1(OSim *s) { /* advect.wgsl:4-20 */
  const float *const *V = (const float *const *)s->vel[s->cur];
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= NZ; z++)
    for (int y = 0; y <= NY; y++)
      for (int x = 0; x <= NX; x++) {
        size_t i = vidx(s, x, y, z);
        for (int k = 0; k < 3; k++) {
          if (!compValid(s, k, x, y, z)) continue;
          v3 off = compOff(k);
          v3 p = {(float)x + off.x, (float)y + off.y, (float)z + off.z};
          v3 b = traceBack(s, V, p);
          v3 q = {b.x - off.x, b.y - off.y, b.z - off.z};
          s->hat[k][i] = triC(s, V[k], q, compHi(s, k));
        }
      }
}
static inline int solidAt(const OSim *s, int x, int y, int z) {
  return inb(s, x, y, z) && s->flags[sidx(s, x, y, z)] == 1u;
}
static void stage_adv_vel2(OSim *s) { /* advect.wgsl:22-70 */
  const float *const *V = (const float *const *)s->vel[s->cur];
  float **O = s->vel[1 - s->cur];
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= NZ; z++)
    for (int y = 0; y <= NY; y++)
      for (int x = 0; x <= NX; x++) {
        size_t i = vidx(s, x, y, z);
        for (int k = 0; k < 3; k++) {
          if (!compValid(s, k, x, y, z)) continue;
          int ax = x - (k == 0), ay = y - (k == 1), az = z - (k == 2);
          if ((k == 1 && y == 0) || solidAt(s, x, y, z) || solidAt(s, ax, ay, az)) {
            O[k][i] = 0.0f;
            continue;
          }
          v3 off = compOff(k);
          i3 hi = compHi(s, k);
          v3 p = {(float)x + off.x, (float)y + off.y, (float)z + off.z};
          v3 bb = traceBack(s, V, p);
          v3 ff = traceFwd(s, V, p);
          v3 b = {bb.x - off.x, bb.y - off.y, bb.z - off.z};
          v3 f = {ff.x - off.x, ff.y - off.y, ff.z - off.z};
          float lo, hh;
          mmC(s, V[k], b, hi, &lo, &hh);
          float r = clampf_(s->hat[k][i] + 0.5f * (V[k][i] - triC(s, s->hat[k], f, hi)), lo, hh);
          O[k][i] = r;
        }
      }
}
static void stage_adv_s1(OSim *s) { /* advect.wgsl:72-91 */
  const float *const *V = (const float *const *)s->vel[s->cur];
  v4 amb0 = {s->P.t0, 0.0f, s->P.yo2, 0.0f}, amb1 = {0, 0, 0, 0};
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++)
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t i = sidx(s, x, y, z);
        v3 p = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
        v3 bb = traceBack(s, V, p);
        v3 b = {bb.x - 0.5f, bb.y - 0.5f, bb.z - 0.5f};
        s->h0[i] = advS(s, s->s0, b, amb0);
        s->h1[i] = advS(s, s->s1, b, amb1);
      }
}
static inline float mc4(float r, float s0, float t, float lo, float hi) {
  return clampf_(r + 0.5f * (s0 - t), lo, hi);
}
static void stage_adv_s2(OSim *s) { /* advect.wgsl:93-125 */
  const float *const *V = (const float *const *)s->vel[s->cur];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++)
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t i = sidx(s, x, y, z);
        v3 p = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
        v3 bb = traceBack(s, V, p);
        v3 ff = traceFwd(s, V, p);
        v3 b = {bb.x - 0.5f, bb.y - 0.5f, bb.z - 0.5f};
        v3 f = {ff.x - 0.5f, ff.y - 0.5f, ff.z - 0.5f};
        v4 r0 = s->h0[i], r1 = s->h1[i];
        if (outside(s, b) <= 0.0f && outside(s, f) <= 0.0f) {
          v4 l0, u0, l1, u1;
          mmS(s, s->s0, b, &l0, &u0);
          mmS(s, s->s1, b, &l1, &u1);
          v4 t0 = triS(s, s->h0, f), t1 = triS(s, s->h1, f);
          v4 a = s->s0[i], c = s->s1[i];
          r0.x = mc4(r0.x, a.x, t0.x, l0.x, u0.x); r0.y = mc4(r0.y, a.y, t0.y, l0.y, u0.y);
          r0.z = mc4(r0.z, a.z, t0.z, l0.z, u0.z); r0.w = mc4(r0.w, a.w, t0.w, l0.w, u0.w);
          r1.x = mc4(r1.x, c.x, t1.x, l1.x, u1.x); r1.y = mc4(r1.y, c.y, t1.y, l1.y, u1.y);
          r1.z = mc4(r1.z, c.z, t1.z, l1.z, u1.z); r1.w = mc4(r1.w, c.w, t1.w, l1.w, u1.w);
        }
        s->s0b[i] = r0;
        s->s1b[i] = r1;
      }
}

/* ------------------------------------------------------------------ */
/* react: combust.wgsl:139-273                                         */
/* ------------------------------------------------------------------ */
static void stage_react(OSim *s) {
  const OParams *P = &s->P;
  const v4 *s0in = s->s0b, *s1in = s->s1b;
  const v4 *sol = s->sol[1 - s->sol_cur];
  const float dV = P->dx * P->dx * P->dx;
  double *zs = calloc((size_t)NZ * 5, sizeof(double));
  double *zt = calloc((size_t)NZ, sizeof(double));
  double *zl = calloc((size_t)NZ * (size_t)NY, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++) {
    double aH = 0, aR = 0, aS = 0, aF = 0, aM = 0;
    float tmax = 0.0f;
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t i = sidx(s, x, y, z);
        float stH = 0, stR = 0, stS = 0, stF = 0, stM = 0, stT = 0;
        if (s->flags[i] == 1u) {
          v4 a0 = {sol[i].w, 0.0f, P->yo2, 0.0f}, a1 = {0, 0, 0, 0};
          s->s0[i] = a0;
          s->s1[i] = a1;
        } else {
          v4 a = s0in[i];
          v4 b = s1in[i];
          float nut = (P->cs * P->dx) * (P->cs * P->dx) * s->curl[i].w;
          if (P->cs > 0.0f) {
            float kd = fmin_(nut / 0.5f * P->dt / (P->dx * P->dx), 0.15f);
            v4 lapA = {0, 0, 0, 0};
            float lapB = 0.0f;
            for (int k = 0; k < 6; k++) {
              int nx_ = x + DIRS[k].x, ny_ = y + DIRS[k].y, nz_ = z + DIRS[k].z;
              if (inb(s, nx_, ny_, nz_)) {
                size_t j = sidx(s, nx_, ny_, nz_);
                if (s->flags[j] != 1u) {
                  lapA.x += s0in[j].x - a.x; lapA.y += s0in[j].y - a.y;
                  lapA.z += s0in[j].z - a.z; lapA.w += s0in[j].w - a.w;
                  lapB += s1in[j].x - b.x;
                }
              } else if (ny_ >= 0) {
                lapA.x += P->t0 - a.x; lapA.y += 0.0f - a.y;
                lapA.z += P->yo2 - a.z; lapA.w += 0.0f - a.w;
                lapB -= b.x;
              }
            }
            a.x += kd * lapA.x; a.y += kd * lapA.y; a.z += kd * lapA.z; a.w += kd * lapA.w;
            b.x += kd * lapB;
          }
          float T = a.x;
          float yf = fmax_(a.y, 0.0f);
          float yo = clampf_(a.z, 0.0f, P->yo2);
          float ys = fmax_(a.w, 0.0f);
          float ym = fmax_(b.x, 0.0f);
          float S = 0.0f;
          for (int k = 0; k < 6; k++) {
            int nx_ = x + DIRS[k].x, ny_ = y + DIRS[k].y, nz_ = z + DIRS[k].z;
            if (!inb(s, nx_, ny_, nz_)) continue;
            size_t j = sidx(s, nx_, ny_, nz_);
            if (s->flags[j] != 1u) continue;
            v4 pr = s->pyr[j];
            float Ts = sol[j].w;
            float rho = rhoOf(s, T);
            if (pr.x > 0.0f) {
              float dm = pr.x * P->dt / (rho * P->dx);
              float inv = 1.0f / (1.0f + dm);
              yf = (yf + dm) * inv;
              yo *= inv;
              ys *= inv;
              ym *= inv;
              T = (T + dm * Ts) * inv;
              S += pr.x / (rho * P->dx);
            }
            T += P->h_conv * (Ts - T) * P->dt / (rho * cpOf(T) * P->dx);
            if (pr.y > 0.0f) {
              yo = fmax_(yo - 1.33f * pr.y * P->dt / (rho * P->dx), 0.0f);
              T += 0.5f * P->dh_char * pr.y * P->dt / (rho * cpOf(T) * P->dx);
            }
          }
          if (P->ign_on > 0.5f) {
            v3 dv = {((float)x + 0.5f) * P->dx - P->ign_x, ((float)y + 0.5f) * P->dx - P->ign_y,
                     ((float)z + 0.5f) * P->dx - P->ign_z};
            float dI = len3(dv);
            float w = clampf_(1.0f - dI / P->ign_r, 0.0f, 1.0f);
            if (w > 0.0f) {
              T = fmax_(T, mixf_(P->t0, P->ign_temp, smoothstep_(0.0f, 0.5f, w)));
              yf += P->ign_fuel * w * P->dt;
            }
          }
          float rho = rhoOf(s, T);
          float tauU = 0.049f * P->dx * P->dx / (nut + 1e-6f);
          float tauG = sqrtf(2.0f * P->dx / P->g);
          float tau = fmax_(fmin_(tauU, tauG), P->tau_min);
          float gate = smoothstep_(P->t_ign - 100.0f, P->t_ign + 100.0f, T);
          float lim = fmin_(yf, yo / P->s_o2);
          float dY = lim * (1.0f - EXPF(-gate * P->dt / tau));
          yf -= dY;
          yo = fmax_(yo - P->s_o2 * dY, 0.0f);
          T += dY * P->dhc / cpOf(T + 0.5f * dY * P->dhc / cpOf(T));
          float q = rho * dY * P->dhc / P->dt;
          float rich = smoothstep_(0.0f, 0.03f, yf);
          ys += P->soot_form * dY * (0.25f + 0.75f * rich);
          float oxg = smoothstep_(P->soot_ox_t, P->soot_ox_t + 300.0f, T) * (yo / P->yo2);
          ys *= EXPF(-P->soot_ox * oxg * P->dt);
          float cold = 1.0f - smoothstep_(450.0f, 650.0f, T);
          float cv = yf * fmin_(P->smoke_conv * cold * P->dt, 1.0f);
          yf -= cv;
          ym += cv;
          rho = rhoOf(s, T);
          float kap = kapP(s, T, yo, ys);
          float tl = kap * P->rad_len;
          float esc = tl > 1e-3f ? (1.0f - EXPF(-tl)) / tl : 1.0f;
          float prad = 4.0f * kap * esc * SIGMA * (pow4(T) - pow4(P->t0));
          float dTr = fmin_(prad * P->dt / (rho * cpOf(T)), 0.5f * fmax_(T - P->t0, 0.0f));
          T -= dTr;
          float prr = dTr * rho * cpOf(T) / P->dt;
          S += (q - prr) / (rho * cpOf(T) * T);
          float sp = fmax_(spongeSide(s, x, z), spongeTop(s, y));
          float dec = 1.0f - EXPF(-P->dt * 4.0f * sp);
          T = mixf_(T, P->t0, dec);
          yf *= 1.0f - dec;
          yo = mixf_(yo, P->yo2, dec);
          ys *= 1.0f - dec;
          ym *= 1.0f - dec;
          T = clampf_(T, 250.0f, 2600.0f);
          v4 o0 = {T, yf, yo, ys}, o1 = {ym, q, S, 0.0f};
          s->s0[i] = o0;
          s->s1[i] = o1;
          rho = rhoOf(s, T);
          stH = q * dV; stR = prr * dV; stS = rho * ys * dV; stF = rho * yf * dV; stM = rho * ym * dV; stT = T;
          if (stH > 1.0f && y < 1024) zl[(size_t)z * NY + y] += stH;
        }
        aH += stH; aR += stR; aS += stS; aF += stF; aM += stM;
        tmax = fmax_(tmax, stT);
      }
    zs[5 * z] = aH; zs[5 * z + 1] = aR; zs[5 * z + 2] = aS; zs[5 * z + 3] = aF; zs[5 * z + 4] = aM;
    zt[z] = tmax;
  }
  if (s->want_stats) {
    for (int z = 0; z < NZ; z++) {
      s->stats[0] += zs[5 * z]; s->stats[1] += zs[5 * z + 1]; s->stats[3] += zs[5 * z + 2];
      s->stats[6] += zs[5 * z + 3]; s->stats[8] += zs[5 * z + 4];
      s->stats[4] = fmax(s->stats[4], zt[z]);
    }
    for (int y = 0; y < NY && y < 1024; y++)
      for (int z = 0; z < NZ; z++) s->stats[16 + y] += zl[(size_t)z * NY + y];
  }
  free(zs); free(zt); free(zl);
}

/* ------------------------------------------------------------------ */
/* forces: fluid.wgsl:31-109                                           */
/* ------------------------------------------------------------------ */
static inline float cm(const OSim *s, int x, int y, int z) {
  return s->curl[sidx(s, clampi(x, 0, NX - 1), clampi(y, 0, NY - 1), clampi(z, 0, NZ - 1))].w;
}
static v3 confine(const OSim *s, int x, int y, int z) {
  const OParams *P = &s->P;
  v3 zero = {0, 0, 0};
  size_t i = sidx(s, x, y, z);
  float T = s->s0[i].x;
  float smoke = s->s0[i].w + s->s1[i].x;
  float gate = fmax_(smoothstep_(P->vort_t, P->vort_t + 500.0f, T), 0.35f * smoothstep_(0.0f, 2e-4f, smoke));
  if (gate <= 0.0f) return zero;
  v3 gr = {cm(s, x + 1, y, z) - cm(s, x - 1, y, z), cm(s, x, y + 1, z) - cm(s, x, y - 1, z),
           cm(s, x, y, z + 1) - cm(s, x, y, z - 1)};
  float gl = len3(gr);
  if (gl < 1e-6f) return zero;
  v3 n = {gr.x / gl, gr.y / gl, gr.z / gl};
  v4 w = s->curl[i];
  v3 cr = {n.y * w.z - n.z * w.y, n.z * w.x - n.x * w.z, n.x * w.y - n.y * w.x};
  float vd = P->vort * P->dx;
  v3 f = {vd * cr.x * gate, vd * cr.y * gate, vd * cr.z * gate};
  float fl = len3(f);
  if (fl > 60.0f) {
    float sc = 60.0f / fl;
    f.x *= sc; f.y *= sc; f.z *= sc;
  }
  return f;
}
static void stage_forces(OSim *s) {
  const OParams *P = &s->P;
  float **V = s->vel[1 - s->cur];
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= NZ; z++)
    for (int y = 0; y <= NY; y++)
      for (int x = 0; x <= NX; x++) {
        size_t i = vidx(s, x, y, z);
        for (int k = 0; k < 3; k++) {
          if (!compValid(s, k, x, y, z)) continue;
          int ax = x - (k == 0), ay = y - (k == 1), az = z - (k == 2);
          if ((k == 1 && y == 0) || solidAt(s, x, y, z) || solidAt(s, ax, ay, az)) continue;
          int ia = inb(s, ax, ay, az), ib = inb(s, x, y, z);
          float f = 0.0f;
          if (k == 1) {
            float Ta = P->t0, Tb = P->t0, ma = 0.0f, mb = 0.0f;
            if (ia) { v4 q = s->s0[sidx(s, ax, ay, az)]; Ta = q.x; ma = q.w + s->s1[sidx(s, ax, ay, az)].x; }
            if (ib) { v4 q = s->s0[sidx(s, x, y, z)]; Tb = q.x; mb = q.w + s->s1[sidx(s, x, y, z)].x; }
            if (!ia) { Ta = Tb; ma = mb; }
            if (!ib) { Tb = Ta; mb = ma; }
            float Tf = 0.5f * (Ta + Tb);
            f += P->g * P->buoy * (Tf - P->t0) / Tf - P->g * P->smoke_w * 0.5f * (ma + mb);
          }
          v3 fc = {0, 0, 0};
          float n = 0.0f;
          if (ia) { v3 c = confine(s, ax, ay, az); fc.x += c.x; fc.y += c.y; fc.z += c.z; n += 1.0f; }
          if (ib) { v3 c = confine(s, x, y, z); fc.x += c.x; fc.y += c.y; fc.z += c.z; n += 1.0f; }
          if (n > 0.0f) f += (k == 0 ? fc.x : (k == 1 ? fc.y : fc.z)) / n;
          float u = V[k][i];
          u += f * P->dt;
          int cx = clampi(x, 0, NX - 1), cy = clampi(y, 0, NY - 1), cz = clampi(z, 0, NZ - 1);
          float ss = spongeSide(s, cx, cz);
          if (k != 1) {
            float Tl = P->t0;
            if (ib) Tl = s->s0[sidx(s, x, y, z)].x;
            float hot = smoothstep_(350.0f, 900.0f, Tl);
            float kr = P->wind_relax * (1.0f - 0.85f * hot) + 8.0f * ss;
            float wv = k == 0 ? P->wind_x : P->wind_z;
            u = mixf_(u, wv, 1.0f - EXPF(-kr * P->dt));
          } else {
            float st = spongeTop(s, cy);
            if (st > 0.0f) u = mixf_(u, fmax_(u, 0.0f), 1.0f - EXPF(-8.0f * st * P->dt));
          }
          u = clampf_(u, -40.0f, 40.0f);
          V[k][i] = u;
        }
      }
}

/* divergence: fluid.wgsl:111-128 */
static void stage_divergence(OSim *s) {
  float **V = s->vel[1 - s->cur];
  float *b = s->lv[0].b;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++)
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t i = sidx(s, x, y, z);
        if (s->flags[i] == 1u) { b[i] = 0.0f; continue; }
        float d = (V[0][vidx(s, x + 1, y, z)] - V[0][vidx(s, x, y, z)] + V[1][vidx(s, x, y + 1, z)] -
                   V[1][vidx(s, x, y, z)] + V[2][vidx(s, x, y, z + 1)] - V[2][vidx(s, x, y, z)]) / s->P.dx;
        b[i] = d - s->s1[i].z;
      }
}

/* project: fluid.wgsl:130-152 */
static void stage_project(OSim *s) {
  float **V = s->vel[1 - s->cur];
  const float *phi = s->lv[0].phi;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= NZ; z++)
    for (int y = 0; y <= NY; y++)
      for (int x = 0; x <= NX; x++) {
        size_t i = vidx(s, x, y, z);
        for (int k = 0; k < 3; k++) {
          if (!compValid(s, k, x, y, z)) continue;
          int ax = x - (k == 0), ay = y - (k == 1), az = z - (k == 2);
          float u = 0.0f;
          if (!((k == 1 && y == 0) || solidAt(s, x, y, z) || solidAt(s, ax, ay, az))) {
            u = V[k][i];
            float pc = inb(s, x, y, z) ? phi[sidx(s, x, y, z)] : 0.0f;
            float pa = inb(s, ax, ay, az) ? phi[sidx(s, ax, ay, az)] : 0.0f;
            u -= (pc - pa) / s->P.dx;
          }
          V[k][i] = u;
        }
      }
}

/* ------------------------------------------------------------------ */
/* Multigrid: mg.wgsl, sim.rs:425-491                                  */
/* ------------------------------------------------------------------ */
static inline int linb(const OLevel *L, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < (int)L->d[0] && y < (int)L->d[1] && z < (int)L->d[2];
}
static inline size_t lidx(const OLevel *L, int x, int y, int z) {
  return (size_t)x + (size_t)L->d[0] * ((size_t)y + (size_t)L->d[1] * (size_t)z);
}
/* mg.wgsl:5-36, one colour */
static void mg_smooth_color(OLevel *L, int color, float omega) {
  const int nx = (int)L->d[0], ny = (int)L->d[1], nz = (int)L->d[2];
  const float hh = L->h * L->h;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = ((y + z + color) & 1); x < nx; x += 2) {
        size_t i = lidx(L, x, y, z);
        if (L->f[i] == 1u) continue;
        float sum = 0.0f, cnt = 0.0f;
        for (int k = 0; k < 6; k++) {
          int a = x + DIRS[k].x, b = y + DIRS[k].y, c = z + DIRS[k].z;
          if (b < 0) continue;
          if (!linb(L, a, b, c)) { cnt += 1.0f; continue; }
          size_t j = lidx(L, a, b, c);
          if (L->f[j] == 1u) continue;
          sum += L->phi[j];
          cnt += 1.0f;
        }
        if (cnt < 0.5f) continue;
        L->phi[i] = mixf_(L->phi[i], (sum - hh * L->b[i]) / cnt, omega);
      }
}
/* sim.rs:425-449: iters x (colour 0, colour 1) */
static void mg_smooth(OLevel *L, uint32_t iters, float omega) {
  for (uint32_t it = 0; it < iters; it++) {
    mg_smooth_color(L, 0, omega);
    mg_smooth_color(L, 1, omega);
  }
}
/* mg.wgsl:52-66 */
static float mg_resid_at(const OLevel *F, int x, int y, int z) {
  size_t i = lidx(F, x, y, z);
  if (F->f[i] == 1u) return 0.0f;
  float pcen = F->phi[i];
  float sum = 0.0f, cnt = 0.0f;
  for (int k = 0; k < 6; k++) {
    int a = x + DIRS[k].x, b = y + DIRS[k].y, c = z + DIRS[k].z;
    if (b < 0) continue;
    if (!linb(F, a, b, c)) { cnt += 1.0f; continue; }
    size_t j = lidx(F, a, b, c);
    if (F->f[j] == 1u) continue;
    sum += F->phi[j];
    cnt += 1.0f;
  }
  return F->b[i] - (sum - cnt * pcen) / (F->h * F->h);
}
/* mg.wgsl:38-83 */
static void mg_restrict(const OLevel *F, OLevel *C) {
  const int nx = (int)C->d[0], ny = (int)C->d[1], nz = (int)C->d[2];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        float r = 0.0f, nfl = 0.0f;
        for (int k = 0; k < 8; k++) {
          int a = 2 * x + (k & 1), b = 2 * y + ((k >> 1) & 1), c = 2 * z + ((k >> 2) & 1);
          if (!linb(F, a, b, c)) continue;
          if (F->f[lidx(F, a, b, c)] != 1u) { nfl += 1.0f; r += mg_resid_at(F, a, b, c); }
        }
        size_t i = lidx(C, x, y, z);
        C->b[i] = nfl > 0.0f ? r / fmax_(nfl, 1.0f) : 0.0f;
        C->phi[i] = 0.0f;
        C->f[i] = nfl > 0.0f ? 0u : 1u;
      }
}
/* mg.wgsl:85-120 */
static void mg_prolong(const OLevel *C, OLevel *F) {
  const int nx = (int)F->d[0], ny = (int)F->d[1], nz = (int)F->d[2];
  const int cx = (int)C->d[0], cy = (int)C->d[1], cz = (int)C->d[2];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = lidx(F, x, y, z);
        if (F->f[i] == 1u) continue;
        float px = ((float)x + 0.5f) * 0.5f - 0.5f, py = ((float)y + 0.5f) * 0.5f - 0.5f,
              pz = ((float)z + 0.5f) * 0.5f - 0.5f;
        int i0x = (int)floorf(px), i0y = (int)floorf(py), i0z = (int)floorf(pz);
        float frx = px - (float)i0x, fry = py - (float)i0y, frz = pz - (float)i0z;
        float acc = 0.0f, ws = 0.0f;
        for (int k = 0; k < 8; k++) {
          int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
          int ccx = clampi(i0x + ox, 0, cx - 1), ccy = clampi(i0y + oy, 0, cy - 1), ccz = clampi(i0z + oz, 0, cz - 1);
          size_t j = lidx(C, ccx, ccy, ccz);
          if (C->f[j] == 1u) continue;
          float wx = ox == 1 ? frx : 1.0f - frx, wy = oy == 1 ? fry : 1.0f - fry, wz = oz == 1 ? frz : 1.0f - frz;
          float w = wx * wy * wz;
          acc += w * C->phi[j];
          ws += w;
        }
        if (ws > 1e-4f) F->phi[i] += acc / ws;
      }
}
/* sim.rs:451-491 */
static void mg_vcycles(OLevel *lv, int nl, uint32_t cycles, uint32_t coarse_iters) {
  for (uint32_t c = 0; c < cycles; c++) {
    for (int l = 0; l < nl - 1; l++) {
      mg_smooth(&lv[l], 2, 1.0f);
      mg_restrict(&lv[l], &lv[l + 1]);
    }
    mg_smooth(&lv[nl - 1], coarse_iters, 1.6f);
    for (int l = nl - 2; l >= 0; l--) {
      mg_prolong(&lv[l + 1], &lv[l]);
      mg_smooth(&lv[l], 2, 1.0f);
    }
  }
}
/* mg.wgsl:122-158, FP64 sums (row order) */
static void mg_resid_norms(const OLevel *L, double *r2o, double *b2o) {
  const int nx = (int)L->d[0], ny = (int)L->d[1], nz = (int)L->d[2];
  double *zr = calloc((size_t)nz * 2, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++) {
    double r2 = 0, b2 = 0;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        size_t i = lidx(L, x, y, z);
        if (L->f[i] == 1u) continue;
        float r = mg_resid_at(L, x, y, z);
        r2 += (double)r * (double)r;
        b2 += (double)L->b[i] * (double)L->b[i];
      }
    zr[2 * z] = r2;
    zr[2 * z + 1] = b2;
  }
  double r2 = 0, b2 = 0;
  for (int z = 0; z < nz; z++) { r2 += zr[2 * z]; b2 += zr[2 * z + 1]; }
  free(zr);
  *r2o = r2;
  *b2o = b2;
}

/* sim.rs:142-168: halve (div_ceil) until 4 levels or min dim < 16 */
static int build_levels(OLevel *lv, uint32_t nx, uint32_t ny, uint32_t nz, float h,
                        uint32_t max_levels, uint32_t *flags0) {
  uint32_t d[3] = {nx, ny, nz};
  int nl = 0;
  for (uint32_t l = 0; l < max_levels && l < MG_MAX_LEVELS; l++) {
    size_t cnt = (size_t)d[0] * d[1] * d[2];
    OLevel *L = &lv[nl++];
    L->d[0] = d[0]; L->d[1] = d[1]; L->d[2] = d[2];
    L->h = h;
    L->phi = calloc(cnt, sizeof(float));
    L->b = calloc(cnt, sizeof(float));
    if (l == 0 && flags0) { L->f = flags0; L->own_f = 0; }
    else { L->f = calloc(cnt, sizeof(uint32_t)); L->own_f = 1; }
    uint32_t mn = d[0] < d[1] ? d[0] : d[1];
    if (d[2] < mn) mn = d[2];
    if (mn < 16) break;
    d[0] = (d[0] + 1) / 2; d[1] = (d[1] + 1) / 2; d[2] = (d[2] + 1) / 2;
    h *= 2.0f;
  }
  return nl;
}
static void free_levels(OLevel *lv, int nl) {
  for (int l = 0; l < nl; l++) {
    free(lv[l].phi);
    free(lv[l].b);
    if (lv[l].own_f) free(lv[l].f);
  }
}

/* ------------------------------------------------------------------ */
/* Sim lifecycle: sim.rs:108-232, 246-423                              */
/* ------------------------------------------------------------------ */
OSim *osim_create(const void *params, size_t params_bytes, const float *sol_init,
                  const float *s0_init, uint32_t cycles, float ign_until) {
  if (params_bytes != sizeof(OParams) || !params || !sol_init) return NULL;
  OSim *s = calloc(1, sizeof(OSim));
  memcpy(&s->P, params, sizeof(OParams));
  s->n = (size_t)s->P.nx * s->P.ny * s->P.nz;
  s->nv = (size_t)(s->P.nx + 1) * (s->P.ny + 1) * (s->P.nz + 1);
  for (int c = 0; c < 2; c++)
    for (int k = 0; k < 3; k++) s->vel[c][k] = calloc(s->nv, sizeof(float));
  for (int k = 0; k < 3; k++) s->hat[k] = calloc(s->nv, sizeof(float));
  v4 **f4[] = {&s->s0, &s->s1, &s->s0b, &s->s1b, &s->h0, &s->h1, &s->sol[0], &s->sol[1], &s->pyr, &s->curl};
  for (size_t q = 0; q < sizeof(f4) / sizeof(f4[0]); q++) *f4[q] = calloc(s->n, sizeof(v4));
  s->flags = calloc(s->n, sizeof(uint32_t));
  for (size_t i = 0; i < s->n; i++) {
    if (s0_init) memcpy(&s->s0[i], &s0_init[4 * i], 16);
    else { s->s0[i].x = s->P.t0; s->s0[i].y = 0; s->s0[i].z = s->P.yo2; s->s0[i].w = 0; }
  }
  memcpy(s->sol[0], sol_init, s->n * 16);
  memcpy(s->sol[1], sol_init, s->n * 16);
  s->nl = build_levels(s->lv, s->P.nx, s->P.ny, s->P.nz, s->P.dx, 4, s->flags);
  s->cycles = cycles;
  s->ign_until = ign_until;
  return s;
}

void osim_destroy(OSim *s) {
  if (!s) return;
  for (int c = 0; c < 2; c++)
    for (int k = 0; k < 3; k++) free(s->vel[c][k]);
  for (int k = 0; k < 3; k++) free(s->hat[k]);
  free(s->s0); free(s->s1); free(s->s0b); free(s->s1b); free(s->h0); free(s->h1);
  free(s->sol[0]); free(s->sol[1]); free(s->pyr); free(s->curl); free(s->flags);
  free_levels(s->lv, s->nl);
  free(s);
}

int osim_levels(const OSim *s, uint32_t *d) {
  if (d)
    for (int l = 0; l < s->nl; l++) { d[3 * l] = s->lv[l].d[0]; d[3 * l + 1] = s->lv[l].d[1]; d[3 * l + 2] = s->lv[l].d[2]; }
  return s->nl;
}

void osim_begin_step(OSim *s, float dt, int want_stats) {
  s->P.dt = dt;
  s->P.time = (float)s->time;
  s->P.ign_on = ((float)s->time < s->ign_until) ? 1.0f : 0.0f;
  s->want_stats = want_stats;
  if (want_stats) memset(s->stats, 0, sizeof(s->stats));
}

int osim_run_stage(OSim *s, int stage) {
  switch (stage) {
    case O_STAGE_SOLID: stage_solid(s); break;
    case O_STAGE_CURL: stage_curl(s); break;
    case O_STAGE_ADV_VEL1: stage_adv_vel1(s); break;
    case O_STAGE_ADV_VEL2: stage_adv_vel2(s); break;
    case O_STAGE_ADV_S1: stage_adv_s1(s); break;
    case O_STAGE_ADV_S2: stage_adv_s2(s); break;
    case O_STAGE_REACT: stage_react(s); break;
    case O_STAGE_FORCES: stage_forces(s); break;
    case O_STAGE_DIVERGENCE: stage_divergence(s); break;
    case O_STAGE_VCYCLES: mg_vcycles(s->lv, s->nl, s->cycles, 80); break;
    case O_STAGE_RESID: {
      double r2, b2;
      mg_resid_norms(&s->lv[0], &r2, &b2);
      if (s->want_stats) { s->stats[10] = r2; s->stats[11] = b2; }
      break;
    }
    case O_STAGE_PROJECT: stage_project(s); break;
    default: return -1;
  }
  return 0;
}

void osim_end_step(OSim *s) {
  s->cur = 1 - s->cur;
  s->sol_cur = 1 - s->sol_cur;
  s->time = (double)((float)s->time + s->P.dt); /* sim.rs keeps time in f32 */
  s->P.frame += 1;
}

void osim_step(OSim *s, float dt, int want_stats) {
  osim_begin_step(s, dt, want_stats);
  for (int st = O_STAGE_SOLID; st <= O_STAGE_DIVERGENCE; st++) osim_run_stage(s, st);
  osim_run_stage(s, O_STAGE_VCYCLES);
  if (want_stats) osim_run_stage(s, O_STAGE_RESID);
  osim_run_stage(s, O_STAGE_PROJECT);
  osim_end_step(s);
}

double osim_time(const OSim *s) { return s->time; }

/* The simulation's Params blob (576 bytes) as the next step starts from it. */
void osim_get_params(const OSim *s, void *out) { memcpy(out, &s->P, sizeof(OParams)); }

/* Clock of a simulation continued from a checkpoint: time, frame counter
   (Params.frame) and the ignition end time. */
void osim_set_clock(OSim *s, double time, uint32_t frame, float ign_until) {
  s->time = time;
  s->P.frame = frame;
  s->ign_until = ign_until;
}

void osim_read_stats(const OSim *s, double *out) { memcpy(out, s->stats, sizeof(s->stats)); }

static void *field_ptr(const OSim *s, int field, size_t *bytes) {
  size_t n4 = s->n * 16, nv = s->nv * 4;
  switch (field) {
    case O_F_S0: *bytes = n4; return s->s0;
    case O_F_S1: *bytes = n4; return s->s1;
    case O_F_S0B: *bytes = n4; return s->s0b;
    case O_F_S1B: *bytes = n4; return s->s1b;
    case O_F_H0: *bytes = n4; return s->h0;
    case O_F_H1: *bytes = n4; return s->h1;
    case O_F_SOL: *bytes = n4; return s->sol[s->sol_cur];
    case O_F_SOL_OUT: *bytes = n4; return s->sol[1 - s->sol_cur];
    case O_F_PYR: *bytes = n4; return s->pyr;
    case O_F_FLAGS: *bytes = s->n * 4; return s->flags;
    case O_F_CURL: *bytes = n4; return s->curl;
    case O_F_U: case O_F_V: case O_F_W: *bytes = nv; return s->vel[s->cur][field - O_F_U];
    case O_F_UB: case O_F_VB: case O_F_WB: *bytes = nv; return s->vel[1 - s->cur][field - O_F_UB];
    case O_F_HU: case O_F_HV: case O_F_HW: *bytes = nv; return s->hat[field - O_F_HU];
    default: break;
  }
  if (field >= O_F_LEVEL_BASE) {
    int l = (field - O_F_LEVEL_BASE) / 4, q = (field - O_F_LEVEL_BASE) % 4;
    if (l < s->nl && q < 3) {
      const OLevel *L = &s->lv[l];
      size_t cnt = (size_t)L->d[0] * L->d[1] * L->d[2];
      *bytes = cnt * 4;
      return q == 0 ? (void *)L->phi : (q == 1 ? (void *)L->b : (void *)L->f);
    }
  }
  *bytes = 0;
  return NULL;
}

int osim_read_field(const OSim *s, int field, void *dst, size_t bytes) {
  size_t nb;
  void *p = field_ptr(s, field, &nb);
  if (!p || nb != bytes) return -1;
  memcpy(dst, p, nb);
  return 0;
}

int osim_write_field(OSim *s, int field, const void *src, size_t bytes) {
  size_t nb;
  void *p = field_ptr(s, field, &nb);
  if (!p || nb != bytes) return -1;
  memcpy(p, src, nb);
  return 0;
}

/* render.wgsl:16-37 (pack): the renderer's three volumes, as the f32
 * values handed to textureStore before the rgba16float conversion.
 * out = vol | sol | vel, each N x vec4, x fastest:
 *   vol = (soot ppm, T, smoke g/m3, heat release kW/m3), (0, T, 0, 0) in solids
 *   sol = (occupancy, Ts*occ, char fraction*occ, slot*occ)
 *   vel = (cell-centre u, v, w, 0) */
void osim_render_pack(const OSim *s, float *out) {
  const size_t n = s->n;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; z++)
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        const size_t i = sidx(s, x, y, z);
        const v4 a = s->s0[i], b = s->s1[i], so = s->sol[s->sol_cur][i];
        const float rho = rhoOf(s, a.x);
        const int solid = s->flags[i] == 1u;
        float *vol = out + 4 * i, *sl = out + 4 * (n + i), *ve = out + 4 * (2 * n + i);
        if (solid) {
          vol[0] = 0.0f; vol[1] = a.x; vol[2] = 0.0f; vol[3] = 0.0f;
        } else {
          vol[0] = rho * a.w / RHO_SOOT * 1e6f;
          vol[1] = a.x;
          vol[2] = rho * b.x * 1e3f;
          vol[3] = b.y * 1e-3f;
        }
        const float tot = so.y + so.z;
        const float oc = solid ? 1.0f : 0.0f;
        sl[0] = oc;
        sl[1] = so.w * oc;
        sl[2] = so.z / fmax_(tot, 1e-6f) * oc;
        sl[3] = so.x * oc;
        const float *U = s->vel[s->cur][0], *V = s->vel[s->cur][1], *W = s->vel[s->cur][2];
        ve[0] = 0.5f * (U[vidx(s, x, y, z)] + U[vidx(s, x + 1, y, z)]);
        ve[1] = 0.5f * (V[vidx(s, x, y, z)] + V[vidx(s, x, y + 1, z)]);
        ve[2] = 0.5f * (W[vidx(s, x, y, z)] + W[vidx(s, x, y, z + 1)]);
        ve[3] = 0.0f;
      }
}

void osim_add_velocity_noise(OSim *s, float amp, uint32_t seed) {
  for (int k = 0; k < 3; k++) {
    float *V = s->vel[s->cur][k];
#pragma omp parallel for schedule(static)
    for (int z = 0; z <= NZ; z++)
      for (int y = 0; y <= NY; y++)
        for (int x = 0; x <= NX; x++) {
          if (!compValid(s, k, x, y, z)) continue;
          size_t i = vidx(s, x, y, z);
          V[i] += amp * oracle_noise((uint32_t)k, (uint32_t)i, seed);
        }
  }
}

/* ------------------------------------------------------------------ */
/* Stand-alone MG handle                                               */
/* ------------------------------------------------------------------ */
OMg *omg_create(uint32_t nx, uint32_t ny, uint32_t nz, float h, uint32_t levels,
                const uint32_t *flags0) {
  OMg *m = calloc(1, sizeof(OMg));
  m->nl = build_levels(m->lv, nx, ny, nz, h, levels, NULL);
  if (flags0) memcpy(m->lv[0].f, flags0, (size_t)nx * ny * nz * 4);
  return m;
}
void omg_destroy(OMg *m) {
  if (!m) return;
  free_levels(m->lv, m->nl);
  free(m);
}
int omg_level_dims(const OMg *m, int l, uint32_t *d3) {
  if (l < 0 || l >= m->nl) return m->nl;
  d3[0] = m->lv[l].d[0]; d3[1] = m->lv[l].d[1]; d3[2] = m->lv[l].d[2];
  return m->nl;
}
void omg_set(OMg *m, int l, const float *phi, const float *b) {
  size_t cnt = (size_t)m->lv[l].d[0] * m->lv[l].d[1] * m->lv[l].d[2];
  if (phi) memcpy(m->lv[l].phi, phi, cnt * 4);
  if (b) memcpy(m->lv[l].b, b, cnt * 4);
}
void omg_get(const OMg *m, int l, float *phi, float *b, uint32_t *flags) {
  size_t cnt = (size_t)m->lv[l].d[0] * m->lv[l].d[1] * m->lv[l].d[2];
  if (phi) memcpy(phi, m->lv[l].phi, cnt * 4);
  if (b) memcpy(b, m->lv[l].b, cnt * 4);
  if (flags) memcpy(flags, m->lv[l].f, cnt * 4);
}
void omg_smooth(OMg *m, int l, uint32_t iters, float omega) { mg_smooth(&m->lv[l], iters, omega); }
void omg_restrict(OMg *m, int l) { mg_restrict(&m->lv[l], &m->lv[l + 1]); }
void omg_prolong(OMg *m, int l) { mg_prolong(&m->lv[l + 1], &m->lv[l]); }
void omg_vcycles(OMg *m, uint32_t cycles, uint32_t coarse_iters) { mg_vcycles(m->lv, m->nl, cycles, coarse_iters); }
void omg_resid_norms(const OMg *m, double *r2, double *b2) { mg_resid_norms(&m->lv[0], r2, b2); }

/* ------------------------------------------------------------------ */
/* FP64 Poisson (mg.wgsl operator) for the pressure oracle             */
/* ------------------------------------------------------------------ */
/* y = -A x over fluid cells (SPD form); solids -> 0 */
static void negA64(uint32_t nx, uint32_t ny, uint32_t nz, double ih2, const uint32_t *f,
                   const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < (int)nz; z++)
    for (int yy = 0; yy < (int)ny; yy++)
      for (int xx = 0; xx < (int)nx; xx++) {
        size_t i = (size_t)xx + nx * ((size_t)yy + ny * (size_t)z);
        if (f[i] == 1u) { y[i] = 0.0; continue; }
        double sum = 0.0, cnt = 0.0;
        for (int k = 0; k < 6; k++) {
          int a = xx + DIRS[k].x, b = yy + DIRS[k].y, c = z + DIRS[k].z;
          if (b < 0) continue;
          if (a < 0 || c < 0 || a >= (int)nx || b >= (int)ny || c >= (int)nz) { cnt += 1.0; continue; }
          size_t j = (size_t)a + nx * ((size_t)b + ny * (size_t)c);
          if (f[j] == 1u) continue;
          sum += x[j];
          cnt += 1.0;
        }
        y[i] = -(sum - cnt * x[i]) * ih2;
      }
}
static double dot64(size_t n, const double *a, const double *b) {
  /* fixed-order blocked sum */
  size_t nb = (n + 65535) / 65536;
  double *part = calloc(nb, sizeof(double));
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)nb; q++) {
    double acc = 0.0;
    size_t e = ((size_t)q + 1) * 65536 < n ? ((size_t)q + 1) * 65536 : n;
    for (size_t i = (size_t)q * 65536; i < e; i++) acc += a[i] * b[i];
    part[q] = acc;
  }
  double t = 0.0;
  for (size_t q = 0; q < nb; q++) t += part[q];
  free(part);
  return t;
}
void oracle_residual64(uint32_t nx, uint32_t ny, uint32_t nz, double h, const uint32_t *flags,
                       const double *b, const double *phi, double *r) {
  size_t n = (size_t)nx * ny * nz;
  negA64(nx, ny, nz, 1.0 / (h * h), flags, phi, r);
  for (size_t i = 0; i < n; i++) r[i] = flags[i] == 1u ? 0.0 : b[i] + r[i];
}
int oracle_poisson_cg64(uint32_t nx, uint32_t ny, uint32_t nz, double h, const uint32_t *flags,
                        const double *b, double *phi, double tol, int max_iter, double *rel) {
  size_t n = (size_t)nx * ny * nz;
  double ih2 = 1.0 / (h * h);
  double *r = malloc(n * 8), *p = malloc(n * 8), *q = malloc(n * 8), *zv = malloc(n * 8), *dg = malloc(n * 8);
  /* Jacobi preconditioner: diagonal of -A = cnt/h^2 */
  for (int z = 0; z < (int)nz; z++)
    for (int yy = 0; yy < (int)ny; yy++)
      for (int xx = 0; xx < (int)nx; xx++) {
        size_t i = (size_t)xx + nx * ((size_t)yy + ny * (size_t)z);
        double cnt = 0.0;
        if (flags[i] != 1u)
          for (int k = 0; k < 6; k++) {
            int a = xx + DIRS[k].x, bb = yy + DIRS[k].y, c = z + DIRS[k].z;
            if (bb < 0) continue;
            if (a < 0 || c < 0 || a >= (int)nx || bb >= (int)ny || c >= (int)nz) { cnt += 1.0; continue; }
            if (flags[(size_t)a + nx * ((size_t)bb + ny * (size_t)c)] == 1u) continue;
            cnt += 1.0;
          }
        dg[i] = cnt > 0.5 ? 1.0 / (cnt * ih2) : 0.0;
      }
  /* residual of -A phi = -b  ->  r = -b - (-A phi) */
  negA64(nx, ny, nz, ih2, flags, phi, q);
  for (size_t i = 0; i < n; i++) r[i] = flags[i] == 1u ? 0.0 : (-b[i] - q[i]);
  double bn = 0.0;
  for (size_t i = 0; i < n; i++) if (flags[i] != 1u) bn += b[i] * b[i];
  bn = sqrt(bn);
  if (bn == 0.0) bn = 1.0;
  for (size_t i = 0; i < n; i++) { zv[i] = dg[i] * r[i]; p[i] = zv[i]; }
  double rz = dot64(n, r, zv);
  int it = 0;
  double rr = sqrt(dot64(n, r, r)) / bn;
  while (it < max_iter && rr > tol) {
    negA64(nx, ny, nz, ih2, flags, p, q);
    double pq = dot64(n, p, q);
    if (pq == 0.0) break;
    double alpha = rz / pq;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) { phi[i] += alpha * p[i]; r[i] -= alpha * q[i]; zv[i] = dg[i] * r[i]; }
    double rz2 = dot64(n, r, zv);
    double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) p[i] = zv[i] + beta * p[i];
    rr = sqrt(dot64(n, r, r)) / bn;
    it++;
  }
  if (rel) *rel = rr;
  free(r); free(p); free(q); free(zv); free(dg);
  return it;
}
