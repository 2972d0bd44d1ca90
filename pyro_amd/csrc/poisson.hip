// poisson.hip -- stand-alone pressure solves on one grid (C ABI
// pyro_poisson_*): the reference V-cycles, FP64-residual iterative
// refinement around FP32 V-cycles, and flexible MGPCG with FP64 vectors and
// dot products.
//
// Operator: level 0 of mg.wgsl (ground Neumann, phi = 0 outside the sides
// and top, solids Neumann).  Two hierarchies share the buffers:
//  - `ref`: the reference's (sim.rs:142-168: halve until 4 levels or a
//    dimension below 16), outside weight 1 -- the plain method, bit-equal to
//    Sim::vcycles;
//  - `pc`: the preconditioner's, halved down to a few hundred cells, with the
//    outside weight of level l set to 2*2^l/(2^l+1) so that the Dirichlet
//    point phi = 0 sits at the same distance beyond the wall on every level
//    (on level 0 it is half a fine cell beyond the wall).  The reference
//    weight 1 moves that point outward on each coarser level, which is what
//    slows its V-cycle to ~0.55 per cycle at 32^3 and stalls it at 64^3.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "net.hpp"
#include "pyro_hip.h"
#include "solve_guard.hpp"

using namespace pyro;

namespace {
thread_local std::string g_perr;
int pfail(int code, const std::string &m) {
  g_perr = m;
  return code;
}
#define PCHK(x)                                                                                 \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) return pfail(PYRO_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
}  // namespace

struct PyroPoisson {
  int d[3] = {0, 0, 0};
  float h = 0.0f;
  uint32_t n = 0;
  MgLevel pc[12];
  int nl_pc = 0;
  MgLevel ref[4];
  int nl_ref = 0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;
  double *w = nullptr, *t = nullptr;  // learned preconditioner: N(r) and the residual after it
  double *part = nullptr, *dots = nullptr;
  double *cgpart = nullptr, *dev = nullptr;  // fused-CG partial sums and the iteration's scalars
  double nfluid = 0.0;
  unsigned int *count = nullptr;  // device fluid-cell counter
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  int coarse_iters_pc = 80;
  PyroNet *net = nullptr;
  // bit 0: the network in the first preconditioner application only (else in
  // every one); bit 1: the network on level 1 of the preconditioner hierarchy
  int learned_mode = 0;
  double nfluid1 = -1.0;  // fluid cells of level 1, counted when first needed (-1: not yet)
  // sealed fluid components (pure-Neumann pockets): b is made consistent by
  // subtracting its mean over each (segment k of seal_idx: cells
  // seal_off[k] .. seal_off[k+1]-1), once per right-hand side
  uint8_t *fprev = nullptr;          // level-0 flags at the last labelling
  uint8_t *cc_open = nullptr;        // open[root]
  unsigned int *cc_flag = nullptr;   // [0] sealed-cell counter, [1] flags changed
  uint32_t *seal_idx = nullptr, *seal_off = nullptr;
  uint32_t seal_cap = 0, seal_comp = 0, seal_cells = 0;
  int seal_mode = 1;                 // 0 never, 1 all methods but VCYCLE, 2 all
  bool labelled = false, b_fresh = true;
};

namespace {
template <class T>
int palloc(PyroPoisson *P, T **ptr, size_t count) {
  void *v = nullptr;
  size_t bytes = count * sizeof(T);
  if (hipMalloc(&v, bytes < 16 ? 16 : bytes) != hipSuccess) return pfail(PYRO_E_OOM, "hipMalloc failed");
  P->allocs.push_back(v);
  if (hipMemset(v, 0, bytes < 16 ? 16 : bytes) != hipSuccess) return pfail(PYRO_E_HIP, "hipMemset failed");
  *ptr = static_cast<T *>(v);
  return 0;
}

int sync(PyroPoisson *P) {
  PCHK(hipStreamSynchronize(P->stream));
  PCHK(hipGetLastError());
  return 0;
}

// dots[0..k) of pairs (a_i, b_i), read back to the host
int dots(PyroPoisson *P, std::initializer_list<std::pair<const double *, const double *>> pairs, double *out) {
  const double *a[4], *b[4];
  int k = 0;
  for (auto &pr : pairs) { a[k] = pr.first; b[k] = pr.second; k++; }
  launch_dots(a, b, k, P->n, P->part, P->dots, P->stream);
  PCHK(hipMemcpyAsync(out, P->dots, k * sizeof(double), hipMemcpyDeviceToHost, P->stream));
  return sync(P);
}

// z = M(r): one preconditioner V-cycle on FP32 copies, from a zero guess.
void precond(PyroPoisson *P, const double *r, double *z) {
  MgLevel &L0 = P->pc[0];
  launch_to_f32(P->n, r, L0.b, P->stream);
  launch_zero32(P->n, L0.phi, P->stream);
  enqueue_vcycle(P->pc, P->nl_pc, P->coarse_iters_pc, P->stream);
  launch_to_f64(P->n, L0.phi, z, P->stream);
}

int rel_resid(PyroPoisson *P, double *rel, double *bb_out) {
  launch_resid64(P->pc[0], P->x, P->b, P->r, P->stream);
  double d[2];
  int rc = dots(P, {{P->r, P->r}, {P->b, P->b}}, d);
  if (rc) return rc;
  *rel = d[1] > 0.0 ? std::sqrt(d[0] / d[1]) : std::sqrt(d[0]);
  if (bb_out) *bb_out = d[1];
  return 0;
}
// solid flags from a simulation's level-0 flags (1 solid; 0 and 2, burnt out, are fluid)
__global__ __launch_bounds__(256) void k_flags_from_sim(uint32_t n, const uint8_t *__restrict__ src,
                                                        uint8_t *__restrict__ f) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) f[i] = src[i] == 1 ? 1 : 0;
}

// number of fluid cells: per-wave ballots summed in LDS, one integer atomic
// per block (exact and order-independent)
constexpr int COUNT_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_count_fluid(uint32_t n, const uint8_t *__restrict__ f,
                                                     unsigned int *__restrict__ cnt) {
  unsigned int c = 0;
  for (uint32_t base = blockIdx.x * 256u; base < n; base += COUNT_BLOCKS * 256u) {
    const uint32_t i = base + threadIdx.x;
    const unsigned long long m = __ballot(i < n && f[i] != 1);
    c += (unsigned int)__popcll(m);
  }
  __shared__ unsigned int sh[4];
  if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(cnt, sh[0] + sh[1] + sh[2] + sh[3]);
}

// s * N(r) + V(r - s A N(r)) with s = (N(r).r) / (N(r).A N(r)), the step that
// minimises the energy norm of the error along N(r) (A is negative definite,
// so both dots are negative for a useful direction and s > 0).  ok = false,
// z untouched, if the direction is unusable (s not positive and finite).
// With bit 1 of the mode the direction is Pr N(R r) instead of N(r): R the
// restriction to level 1 of the preconditioner hierarchy (the mean of r over
// each coarse cell's fluid children), N applied there (spacing 2h, 1/8 of the
// voxels), Pr the V-cycle's masked trilinear prolongation.
int learned_precond(PyroPoisson *P, const double *r, double rr, double *z, bool *ok) {
  *ok = false;
  int rc = 0;
  if ((P->learned_mode & 2) && P->nl_pc > 1) {
    MgLevel &L0 = P->pc[0], &L1 = P->pc[1];
    const G3 g1{L1.d[0], L1.d[1], L1.d[2]};
    if (P->nfluid1 < 0.0) {
      unsigned int c = 0;
      PCHK(hipMemsetAsync(P->count, 0, sizeof(unsigned int), P->stream));
      k_count_fluid<<<COUNT_BLOCKS, 256, 0, P->stream>>>(L1.n, L1.f, P->count);
      PCHK(hipMemcpyAsync(&c, P->count, sizeof(c), hipMemcpyDeviceToHost, P->stream));
      rc = sync(P);
      if (rc) return rc;
      P->nfluid1 = (double)c;
    }
    // L1.b = R r (restriction of the residual of phi = 0, b = r), into t as FP64
    launch_to_f32_zero(P->n, r, L0.b, L0.phi, P->stream);
    launch_mg_restrict(L0, L1, P->stream);
    launch_to_f64(L1.n, L1.b, P->t, P->stream);
    double rr1 = 0.0;
    {
      const double *a[1] = {P->t}, *b[1] = {P->t};
      launch_dots(a, b, 1, L1.n, P->part, P->dots, P->stream);
      PCHK(hipMemcpyAsync(&rr1, P->dots, sizeof(double), hipMemcpyDeviceToHost, P->stream));
      rc = sync(P);
      if (rc) return rc;
    }
    if (!(rr1 > 0.0) || !(P->nfluid1 > 0.0)) return 0;
    rc = net_apply64(P->net, g1, P->t, L1.f, rr1, P->nfluid1, (double)L1.h, P->w, P->stream);
    if (rc) return pfail(rc, std::string("learned preconditioner: ") + net_error());
    // w = Pr N(R r): the level-1 direction through FP32 phi and the prolongation
    launch_to_f32(L1.n, P->w, L1.phi, P->stream);
    launch_zero32(P->n, L0.phi, P->stream);
    launch_mg_prolong(L1, L0, P->stream);
    launch_to_f64(P->n, L0.phi, P->w, P->stream);
  } else {
    const G3 g{P->d[0], P->d[1], P->d[2]};
    rc = net_apply64(P->net, g, r, P->pc[0].f, rr, P->nfluid, P->h, P->w, P->stream);
    if (rc) return pfail(rc, std::string("learned preconditioner: ") + net_error());
  }
  launch_apply64(P->pc[0], P->w, P->t, P->stream);
  double d[2];
  rc = dots(P, {{P->w, r}, {P->w, P->t}}, d);
  if (rc) return rc;
  const double sc = d[1] != 0.0 ? d[0] / d[1] : 0.0;
  if (!(sc > 0.0) || !std::isfinite(sc)) return 0;
  // t = r - s A w, then z = V(t) + s w
  launch_xpby64(P->n, r, -sc, P->t, P->stream);
  precond(P, P->t, z);
  launch_axpy64(P->n, sc, P->w, z, P->stream);
  *ok = true;
  return 0;
}


// Sealed components of the level-0 fluid cells (launch_cc_label), sorted on
// the host into per-component segments of ascending cell index.  The union-
// find parents and the (root, cell) pairs use the learned preconditioner's
// buffers w and t, which hold nothing between solves.
int relabel(PyroPoisson *P) {
  MgLevel &L0 = P->pc[0];
  uint32_t *parent = reinterpret_cast<uint32_t *>(P->w);
  uint2 *pairs = reinterpret_cast<uint2 *>(P->t);
  launch_cc_label(L0, parent, P->cc_open, P->stream);
  PCHK(hipMemsetAsync(P->cc_flag, 0, sizeof(unsigned int), P->stream));
  launch_cc_sealed(L0, parent, P->cc_open, pairs, P->cc_flag, P->stream);
  unsigned int cnt = 0;
  PCHK(hipMemcpyAsync(&cnt, P->cc_flag, sizeof(cnt), hipMemcpyDeviceToHost, P->stream));
  int rc = sync(P);
  if (rc) return rc;
  std::vector<uint2> h(cnt);
  if (cnt) {
    PCHK(hipMemcpyAsync(h.data(), pairs, cnt * sizeof(uint2), hipMemcpyDeviceToHost, P->stream));
    rc = sync(P);
    if (rc) return rc;
  }
  std::sort(h.begin(), h.end(), [](const uint2 &a, const uint2 &b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
  std::vector<uint32_t> idx(cnt), off;
  for (uint32_t k = 0; k < cnt; k++) {
    if (k == 0 || h[k].x != h[k - 1].x) off.push_back(k);
    idx[k] = h[k].y;
  }
  off.push_back(cnt);
  if (cnt + 1 > P->seal_cap) {
    if (P->seal_idx) (void)hipFree(P->seal_idx);
    if (P->seal_off) (void)hipFree(P->seal_off);
    P->seal_idx = P->seal_off = nullptr;
    P->seal_cap = 0;
    const uint32_t cap = cnt + 1 > 1024 ? cnt + 1 : 1024;
    if (hipMalloc(&P->seal_idx, cap * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&P->seal_off, cap * sizeof(uint32_t)) != hipSuccess)
      return pfail(PYRO_E_OOM, "hipMalloc failed (sealed components)");
    P->seal_cap = cap;
  }
  if (cnt) PCHK(hipMemcpyAsync(P->seal_idx, idx.data(), cnt * sizeof(uint32_t), hipMemcpyHostToDevice, P->stream));
  PCHK(hipMemcpyAsync(P->seal_off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, P->stream));
  rc = sync(P);
  if (rc) return rc;
  P->seal_comp = (uint32_t)off.size() - 1;
  P->seal_cells = cnt;
  P->labelled = true;
  P->b_fresh = true;
  return 0;
}

int set_flags_device(PyroPoisson *P, const uint8_t *sim_flags) {
  k_flags_from_sim<<<nblocks(P->n), 256, 0, P->stream>>>(P->n, sim_flags, P->pc[0].f);
  launch_mg_setup(P->pc, P->nl_pc, P->stream);
  PCHK(hipMemsetAsync(P->count, 0, 2 * sizeof(unsigned int), P->stream));
  k_count_fluid<<<COUNT_BLOCKS, 256, 0, P->stream>>>(P->n, P->pc[0].f, P->count);
  launch_flags_changed(P->n, P->pc[0].f, P->fprev, P->count + 1, P->stream);
  unsigned int c[2] = {0, 0};
  PCHK(hipMemcpyAsync(c, P->count, sizeof(c), hipMemcpyDeviceToHost, P->stream));
  int rc = sync(P);
  if (rc) return rc;
  P->nfluid = (double)c[0];
  P->nfluid1 = -1.0;
  if (c[1] || !P->labelled) return relabel(P);
  return 0;
}

bool projects(const PyroPoisson *P, int method) {
  return P->seal_mode == 2 || (P->seal_mode == 1 && method != PYRO_SOLVER_VCYCLE);
}
}  // namespace

extern "C" {

const char *pyro_poisson_last_error(void) { return g_perr.c_str(); }

int pyro_poisson_create(uint32_t nx, uint32_t ny, uint32_t nz, float h, const uint32_t *flags, PyroPoisson **out) {
  if (!flags || !out || nx < 2 || ny < 2 || nz < 2 || !(h > 0.0f)) return pfail(PYRO_E_INVALID, "bad argument");
  if ((uint64_t)nx * ny * nz >= (1ull << 31)) return pfail(PYRO_E_INVALID, "grid too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return pfail(PYRO_E_NODEVICE, "no HIP device available");
  PyroPoisson *P = new PyroPoisson();
  auto bail = [&](int rc) { pyro_poisson_destroy(P); return rc; };
  P->d[0] = (int)nx; P->d[1] = (int)ny; P->d[2] = (int)nz;
  P->h = h;
  P->n = nx * ny * nz;
  if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(pfail(PYRO_E_HIP, "hipStreamCreate failed"));
  int d[3] = {(int)nx, (int)ny, (int)nz};
  float hl = h;
  int rc = 0;
  bool ref_open = true;
  for (int l = 0; l < 12 && !rc; l++) {
    MgLevel &L = P->pc[P->nl_pc++];
    L.d[0] = d[0]; L.d[1] = d[1]; L.d[2] = d[2];
    L.h = hl;
    L.hh = hl * hl;
    L.n = (uint32_t)d[0] * (uint32_t)d[1] * (uint32_t)d[2];
    L.wdir = l == 0 ? 1.0f : (float)(2.0 * std::ldexp(1.0, l) / (std::ldexp(1.0, l) + 1.0));
    rc |= palloc(P, &L.phi, L.n);
    if (!rc) rc |= palloc(P, &L.b, L.n);
    if (!rc) rc |= palloc(P, &L.tmp, L.n);
    if (!rc) rc |= palloc(P, &L.f, L.n);
    if (!rc) rc |= palloc(P, &L.nc, L.n);
    if (!rc && l == 0) rc |= palloc(P, &L.nm, L.n);  // the FP64 operator's neighbour mask
    // reference hierarchy: the first levels of the same halving, a level is
    // kept and the halving stops after 4 levels or once min(dims) < 16
    // (sim.rs:146-168).  It never outlasts `pc`, which stops only below
    // 512 cells or min(dims) < 4.
    int mn = d[0] < d[1] ? d[0] : d[1];
    mn = d[2] < mn ? d[2] : mn;
    if (ref_open) {
      P->ref[P->nl_ref] = L;
      P->ref[P->nl_ref].wdir = 1.0f;
      P->nl_ref++;
      ref_open = P->nl_ref < 4 && mn >= 16;
    }
    if (L.n <= 512 || mn < 4) break;
    for (int qd = 0; qd < 3; qd++) d[qd] = (d[qd] + 1) / 2;
    hl *= 2.0f;
  }
  if (!rc) rc |= palloc(P, &P->x, P->n);
  if (!rc) rc |= palloc(P, &P->b, P->n);
  if (!rc) rc |= palloc(P, &P->r, P->n);
  if (!rc) rc |= palloc(P, &P->p, P->n);
  if (!rc) rc |= palloc(P, &P->q, P->n);
  if (!rc) rc |= palloc(P, &P->z, P->n);
  if (!rc) rc |= palloc(P, &P->w, P->n);
  if (!rc) rc |= palloc(P, &P->t, P->n);
  if (!rc) rc |= palloc(P, &P->count, 8);
  if (!rc) rc |= palloc(P, &P->cc_flag, 8);
  if (!rc) rc |= palloc(P, &P->fprev, P->n);
  if (!rc) rc |= palloc(P, &P->cc_open, P->n);
  if (!rc) rc |= palloc(P, &P->cgpart, 2 * (size_t)cg_part_blocks(P->n));
  if (!rc) rc |= palloc(P, &P->dev, 4);
  if (!rc) rc |= palloc(P, &P->part, DOT_PART);
  if (!rc) rc |= palloc(P, &P->dots, 4);
  if (rc) return bail(rc < 0 ? rc : PYRO_E_OOM);
  std::vector<uint8_t> f8(P->n);
  for (uint32_t i = 0; i < P->n; i++) {
    f8[i] = flags[i] == 1u ? 1 : 0;
    P->nfluid += f8[i] ? 0.0 : 1.0;
  }
  if (hipMemcpy(P->pc[0].f, f8.data(), P->n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P->fprev, f8.data(), P->n, hipMemcpyHostToDevice) != hipSuccess)
    return bail(pfail(PYRO_E_HIP, "upload failed"));
  launch_mg_setup(P->pc, P->nl_pc, P->stream);
  if (sync(P)) return bail(PYRO_E_HIP);
  if ((rc = relabel(P))) return bail(rc);
  *out = P;
  return 0;
}

int pyro_poisson_destroy(PyroPoisson *P) {
  if (!P) return 0;
  if (P->stream) (void)hipStreamSynchronize(P->stream);
  for (void *v : P->allocs) (void)hipFree(v);
  if (P->seal_idx) (void)hipFree(P->seal_idx);
  if (P->seal_off) (void)hipFree(P->seal_off);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
  return 0;
}

int pyro_poisson_levels(const PyroPoisson *P, int *ref_levels, int *pc_levels) {
  if (!P) return pfail(PYRO_E_INVALID, "null handle");
  if (ref_levels) *ref_levels = P->nl_ref;
  if (pc_levels) *pc_levels = P->nl_pc;
  return 0;
}

int pyro_poisson_set(PyroPoisson *P, const double *b, const double *phi) {
  if (!P) return pfail(PYRO_E_INVALID, "null handle");
  if (b) PCHK(hipMemcpy(P->b, b, P->n * sizeof(double), hipMemcpyHostToDevice));
  if (b) P->b_fresh = true;
  if (phi) PCHK(hipMemcpy(P->x, phi, P->n * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int pyro_poisson_get(PyroPoisson *P, double *phi) {
  if (!P || !phi) return pfail(PYRO_E_INVALID, "null argument");
  int rc = sync(P);
  if (rc) return rc;
  PCHK(hipMemcpy(phi, P->x, P->n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pyro_poisson_residual(PyroPoisson *P, double *rel) {
  if (!P || !rel) return pfail(PYRO_E_INVALID, "null argument");
  return rel_resid(P, rel, nullptr);
}

int pyro_poisson_solve(PyroPoisson *P, int method, double tol, int max_iters, PyroSolveInfo *info) {
  if (!P || max_iters < 1 || !(tol >= 0.0)) return pfail(PYRO_E_INVALID, "bad argument");
  if (method < PYRO_SOLVER_VCYCLE || method > PYRO_SOLVER_LEARNED) return pfail(PYRO_E_INVALID, "unknown method");
  if (method == PYRO_SOLVER_LEARNED && !P->net) return pfail(PYRO_E_STATE, "no network attached");
  hipEvent_t e0, e1;
  PCHK(hipEventCreate(&e0));
  PCHK(hipEventCreate(&e1));
  PCHK(hipEventRecord(e0, P->stream));
  if (P->b_fresh && P->seal_comp && projects(P, method)) {
    launch_seal_project(P->seal_comp, P->seal_off, P->seal_idx, P->b, P->stream);
    P->b_fresh = false;
  }
  int it = 0, cycles = 0, rc = 0, fallbacks = 0, diverged = 0;
  double rel = 0.0;
  const uint32_t n = P->n;
  // Every residual read is checked (solve_guard.hpp): NaN/Inf fails the
  // solve, growth past kDivergeFactor times the best residual stops it with
  // info->diverged set.  Returns true when the loop must stop.
  ResidWatch rw;
  auto watch = [&](int k) -> bool {
    switch (rw.step(rel, k)) {
      case Watch::NonFinite:
        rc = pfail(PYRO_E_NONFINITE, "non-finite residual after " + std::to_string(k) + " iterations");
        return true;
      case Watch::Diverged: diverged = 1; return true;
      default: return false;
    }
  };
  if (method == PYRO_SOLVER_VCYCLE) {
    // the reference's V-cycles on the FP32 level-0 state, FP64 residual after each
    MgLevel &L0 = P->ref[0];
    launch_to_f32(n, P->x, L0.phi, P->stream);
    launch_to_f32(n, P->b, L0.b, P->stream);
    double bb = 0.0;
    rc = rel_resid(P, &rel, &bb);
    if (bb <= 0.0) bb = 1.0;
    if (!rc) watch(0);
    while (!rc && !diverged && rel > tol && it < max_iters) {
      enqueue_vcycle(P->ref, P->nl_ref, 80, P->stream);
      launch_to_f64(n, L0.phi, P->x, P->stream);
      cycles++;
      it++;
      launch_resid_rr(P->pc[0], P->x, P->b, P->r, P->cgpart, P->dev, P->stream);
      double rr = 0.0;
      PCHK(hipMemcpyAsync(&rr, P->dev + 2, sizeof(double), hipMemcpyDeviceToHost, P->stream));
      rc = sync(P);
      if (rc) break;
      rel = std::sqrt(rr / bb);
      if (watch(it)) break;
    }
  } else if (method == PYRO_SOLVER_REFINE) {
    // x += V(b - A x) with the residual in FP64
    double bb = 0.0;
    rc = rel_resid(P, &rel, &bb);
    if (bb <= 0.0) bb = 1.0;
    if (!rc) watch(0);
    while (!rc && !diverged && rel > tol && it < max_iters) {
      MgLevel &L0 = P->pc[0];
      launch_to_f32_zero(n, P->r, L0.b, L0.phi, P->stream);
      enqueue_vcycle(P->pc, P->nl_pc, P->coarse_iters_pc, P->stream);
      launch_add_f32_to_f64(n, L0.phi, P->x, P->stream);
      cycles++;
      it++;
      launch_resid_rr(L0, P->x, P->b, P->r, P->cgpart, P->dev, P->stream);
      double rr = 0.0;
      PCHK(hipMemcpyAsync(&rr, P->dev + 2, sizeof(double), hipMemcpyDeviceToHost, P->stream));
      rc = sync(P);
      if (rc) break;
      rel = std::sqrt(rr / bb);
      if (watch(it)) break;
    }
  } else {
    // flexible CG (Notay's FCG(1)): p_{k+1} = z_{k+1} - (z_{k+1}.q_k / p_k.q_k) p_k.
    // LEARNED: the preconditioner is s N(r) + V(r - s A N(r)) (the network's
    // step, optimally scaled, then a V-cycle on what it leaves); the plain
    // V-cycle takes over for good if the network's direction is unusable.
    // The iteration's scalars (p.q, p.r, r.r, z.q) stay on the device; the
    // host reads only r.r for the convergence test.
    bool use_net = method == PYRO_SOLVER_LEARNED;
    MgLevel &L0 = P->pc[0];
    auto apply_m = [&](double rr_now, bool with_dot) -> int {
      if (use_net) {
        bool ok = false;
        int e = learned_precond(P, P->r, rr_now, P->z, &ok);
        if (e) return e;
        if (ok) {
          if (P->learned_mode & 1) use_net = false;
          if (with_dot) {
            const double *a[1] = {P->z}, *b[1] = {P->q};
            launch_dots(a, b, 1, n, P->part, P->dev + 3, P->stream);
          }
          return 0;
        }
        use_net = false;
        fallbacks++;
      }
      launch_to_f32_zero(n, P->r, L0.b, L0.phi, P->stream);
      enqueue_vcycle(P->pc, P->nl_pc, P->coarse_iters_pc, P->stream);
      if (with_dot)
        launch_to_f64_dot(n, L0.phi, P->q, P->z, P->cgpart, P->dev, P->stream);
      else
        launch_to_f64(n, L0.phi, P->z, P->stream);
      return 0;
    };
    double bb = 0.0;
    rc = rel_resid(P, &rel, &bb);  // r = b - A x
    if (bb <= 0.0) bb = 1.0;
    if (!rc) watch(0);
    if (!rc && !diverged && rel > tol) {
      rc = apply_m(rel * rel * bb, false);
      cycles++;
      if (!rc) PCHK(hipMemcpyAsync(P->p, P->z, n * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
      while (!rc && it < max_iters) {
        launch_apply_dots(L0, P->p, P->r, P->q, P->cgpart, P->dev, P->stream);  // dev[0] p.q, dev[1] p.r
        launch_update_rr(n, P->dev, P->p, P->q, P->x, P->r, P->cgpart, P->dev, P->stream);  // dev[2] r.r
        it++;
        double rr = 0.0;
        PCHK(hipMemcpyAsync(&rr, P->dev + 2, sizeof(double), hipMemcpyDeviceToHost, P->stream));
        rc = sync(P);
        if (rc) break;
        rel = std::sqrt(rr / bb);
        if (watch(it) || rel <= tol) break;
        rc = apply_m(rr, true);  // z, dev[3] z.q
        if (rc) break;
        cycles++;
        launch_xpby_dev(n, P->z, P->dev, P->p, P->stream);
      }
      if (!rc) rc = rel_resid(P, &rel, nullptr);  // true residual of the final x
      if (!rc && !std::isfinite(rel)) rc = pfail(PYRO_E_NONFINITE, "non-finite residual of the final iterate");
    }
  }
  PCHK(hipEventRecord(e1, P->stream));
  PCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  PCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (info) {
    info->iterations = it;
    info->vcycles = cycles;
    info->rel_resid = rel;
    info->ms = ms;
    info->fallbacks = fallbacks;
    info->diverged = diverged;
  }
  return 0;
}

int pyro_poisson_set_sealed_projection(PyroPoisson *P, int mode) {
  if (!P || mode < 0 || mode > 2) return pfail(PYRO_E_INVALID, "bad argument");
  P->seal_mode = mode;
  return 0;
}

int pyro_poisson_sealed_info(const PyroPoisson *P, uint32_t *components, uint32_t *cells) {
  if (!P) return pfail(PYRO_E_INVALID, "null handle");
  if (components) *components = P->seal_comp;
  if (cells) *cells = P->seal_cells;
  return 0;
}

int pyro_poisson_get_rhs(PyroPoisson *P, double *b) {
  if (!P || !b) return pfail(PYRO_E_INVALID, "null argument");
  PCHK(hipStreamSynchronize(P->stream));
  PCHK(hipMemcpy(b, P->b, P->n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pyro_poisson_attach_net(PyroPoisson *P, PyroNet *net, int mode) {
  if (!P || mode < 0 || mode > 3) return pfail(PYRO_E_INVALID, "bad argument");
  P->net = net;
  P->learned_mode = mode;
  return 0;
}

int pyro_poisson_set_flags_from_sim(PyroPoisson *P, PyroSim *sim) {
  if (!P || !sim) return pfail(PYRO_E_INVALID, "null argument");
  void *fp = nullptr;
  size_t bytes = 0;
  if (pyro_sim_sync(sim) || pyro_sim_device_ptr(sim, PYRO_FIELD_LEVEL_BASE + 2, &fp, &bytes))
    return pfail(PYRO_E_INVALID, std::string("simulation: ") + pyro_last_error());
  if (bytes != P->n) return pfail(PYRO_E_INVALID, "grid mismatch");
  return set_flags_device(P, static_cast<const uint8_t *>(fp));
}

// b = the simulation's current level-0 right-hand side, phi = its level-0 phi
// (warm start) or 0.
int pyro_poisson_set_from_sim(PyroPoisson *P, PyroSim *sim, int warm) {
  if (!P || !sim) return pfail(PYRO_E_INVALID, "null argument");
  void *bp = nullptr, *pp = nullptr;
  size_t nb = 0, npb = 0;
  if (pyro_sim_sync(sim) || pyro_sim_device_ptr(sim, PYRO_FIELD_LEVEL_BASE + 1, &bp, &nb) ||
      pyro_sim_device_ptr(sim, PYRO_FIELD_LEVEL_BASE + 0, &pp, &npb))
    return pfail(PYRO_E_INVALID, std::string("simulation: ") + pyro_last_error());
  if (nb != (size_t)P->n * 4) return pfail(PYRO_E_INVALID, "grid mismatch");
  launch_to_f64(P->n, static_cast<const float *>(bp), P->b, P->stream);
  P->b_fresh = true;
  if (warm)
    launch_to_f64(P->n, static_cast<const float *>(pp), P->x, P->stream);
  else
    launch_zero64(P->n, P->x, P->stream);
  return sync(P);
}

// One simulation step whose pressure solve runs `method` to `tol` (at most
// max_iters outer iterations) from the previous step's phi, instead of the
// fixed V-cycle count: stages up to the divergence (sim.rs:246-403), the
// solve, phi back to the simulation in FP32, then the residual statistics
// (when requested) and the projection (sim.rs:405-423).
int pyro_sim_step_solve(PyroSim *sim, PyroPoisson *P, int method, double tol, int max_iters, float dt,
                        int want_stats, PyroSolveInfo *info) {
  if (!sim || !P) return pfail(PYRO_E_INVALID, "null argument");
  // every argument is checked before the step begins
  if (max_iters < 1 || !(tol >= 0.0) || !std::isfinite(tol)) return pfail(PYRO_E_INVALID, "bad argument");
  if (method < PYRO_SOLVER_VCYCLE || method > PYRO_SOLVER_LEARNED) return pfail(PYRO_E_INVALID, "unknown method");
  if (method == PYRO_SOLVER_LEARNED && !P->net) return pfail(PYRO_E_STATE, "no network attached");
  uint32_t sd[3] = {0, 0, 0};
  if (pyro_sim_info(sim, sd, nullptr, nullptr, nullptr, nullptr))
    return pfail(PYRO_E_INVALID, std::string("simulation: ") + pyro_last_error());
  if ((int)sd[0] != P->d[0] || (int)sd[1] != P->d[1] || (int)sd[2] != P->d[2])
    return pfail(PYRO_E_INVALID, "grid mismatch between the simulation and the pressure problem");
  if (pyro_sim_begin_step(sim, dt, want_stats))
    return pfail(PYRO_E_STATE, std::string("simulation: ") + pyro_last_error());
  // From here a failure aborts the step: vel[cur], sol[sol_cur], time and
  // frame stay those before it; s0/s1 and phi hold the partial step.
  auto abort_with = [&](int rc) {
    (void)pyro_sim_abort_step(sim);
    return rc;
  };
  auto simfail = [&]() {
    return abort_with(pfail(PYRO_E_STATE, std::string("simulation: ") + pyro_last_error()));
  };
  for (int st = PYRO_STAGE_SOLID; st <= PYRO_STAGE_DIVERGENCE; st++)
    if (pyro_sim_run_stage(sim, st)) return simfail();
  int rc = pyro_poisson_set_flags_from_sim(P, sim);
  if (!rc) rc = pyro_poisson_set_from_sim(P, sim, 1);
  PyroSolveInfo si{};
  if (!rc) rc = pyro_poisson_solve(P, method, tol, max_iters, &si);
  if (rc) return abort_with(rc);
  if (si.diverged) {
    if (info) *info = si;
    return abort_with(pfail(PYRO_E_DIVERGED, "pressure solve diverged: relative residual " +
                                                 std::to_string(si.rel_resid) + " after " +
                                                 std::to_string(si.iterations) + " iterations"));
  }
  void *pp = nullptr;
  size_t nb = 0;
  if (pyro_sim_device_ptr(sim, PYRO_FIELD_LEVEL_BASE + 0, &pp, &nb)) return simfail();
  launch_to_f32(P->n, P->x, static_cast<float *>(pp), P->stream);
  rc = sync(P);
  if (rc) return abort_with(rc);
  if (want_stats && pyro_sim_run_stage(sim, PYRO_STAGE_RESID)) return simfail();
  if (pyro_sim_run_stage(sim, PYRO_STAGE_PROJECT)) return simfail();
  if (pyro_sim_end_step(sim)) return pfail(PYRO_E_STATE, std::string("simulation: ") + pyro_last_error());
  if (info) *info = si;
  return 0;
}

// Training pair from the simulation's current pressure problem: b as it
// stands, phi = the MGPCG solution from zero to `tol`.
int pyro_net_collect(PyroNet *net, PyroSim *sim, PyroPoisson *P, double tol, PyroSolveInfo *info) {
  if (!net || !sim || !P) return pfail(PYRO_E_INVALID, "null argument");
  int rc = pyro_poisson_set_flags_from_sim(P, sim);
  if (!rc) rc = pyro_poisson_set_from_sim(P, sim, 0);
  if (rc) return rc;
  launch_zero64(P->n, P->x, P->stream);
  PyroSolveInfo si;
  rc = pyro_poisson_solve(P, PYRO_SOLVER_MGPCG, tol, 400, &si);
  if (rc) return rc;
  if (info) *info = si;
  double bb[1];
  rc = dots(P, {{P->b, P->b}}, bb);
  if (rc) return rc;
  const G3 g{P->d[0], P->d[1], P->d[2]};
  rc = net_add_pair64(net, g, P->b, P->x, P->pc[0].f, bb[0], P->nfluid, P->h, P->stream);
  if (rc) return pfail(rc, net_error());
  return sync(P);
}

}  // extern "C"
