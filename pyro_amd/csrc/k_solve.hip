// k_solve.hip -- FP64 vector kernels of the pressure solvers (iterative
// refinement and flexible MGPCG).
//
// The operator is the level-0 operator of mg.wgsl:25-35, evaluated in FP64:
//   (A x)_i = (sum of fluid-neighbour x - cnt_i x_i) / h^2   on fluid cells,
// with cnt from the level's neighbour codes (sides/top outside count 1, the
// ground and solids are skipped).  Solid cells carry 0.  Dot products are
// reduced per block in a fixed order and then by one block, so every run is
// bit-reproducible.
#include "kernels.hpp"

namespace pyro {

struct Op64 {
  int d0, d1, d2;
  uint32_t n;
  double ih2;  // 1 / h^2
  FastDiv dd0, dd1;
};

Op64 op64_of(const MgLevel &L) {
  Op64 o;
  o.d0 = L.d[0]; o.d1 = L.d[1]; o.d2 = L.d[2];
  o.n = L.n;
  o.ih2 = 1.0 / ((double)L.h * (double)L.h);
  o.dd0 = make_fastdiv((uint32_t)L.d[0]);
  o.dd1 = make_fastdiv((uint32_t)L.d[1]);
  return o;
}

// Neighbours from the level's 6-bit mask (nm: in the level and not solid),
// summed in the order +x, -x, +y, -y, +z, -z.  All six are loaded at once
// (a masked one at i itself, always in range) and a masked term adds +0,
// which leaves the sum unchanged: a running sum started at +0 is never -0.
__device__ __forceinline__ double apply_at(const Op64 &o, const uint8_t *__restrict__ nm, const uint8_t *__restrict__ nc,
                                           const double *__restrict__ x, uint32_t i) {
  const uint8_t code = nc[i];
  if (code & 0x80) return 0.0;
  const uint32_t m = nm[i];
  const uint32_t sy = (uint32_t)o.d0, sz = (uint32_t)o.d0 * (uint32_t)o.d1;
  const double xc = x[i];
  const double v0 = x[(m & 1u) ? i + 1 : i], v1 = x[(m & 2u) ? i - 1 : i];
  const double v2 = x[(m & 4u) ? i + sy : i], v3 = x[(m & 8u) ? i - sy : i];
  const double v4 = x[(m & 16u) ? i + sz : i], v5 = x[(m & 32u) ? i - sz : i];
  double sum = 0.0;
  sum += (m & 1u) ? v0 : 0.0;
  sum += (m & 2u) ? v1 : 0.0;
  sum += (m & 4u) ? v2 : 0.0;
  sum += (m & 8u) ? v3 : 0.0;
  sum += (m & 16u) ? v4 : 0.0;
  sum += (m & 32u) ? v5 : 0.0;
  const double cnt = (double)((code & 7u) + ((code >> 3) & 7u));
  return (sum - cnt * xc) * o.ih2;
}

// r = b - A x  (fluid cells; 0 on solids)
__global__ __launch_bounds__(NT) void k_resid64(Op64 o, const uint8_t *__restrict__ nm, const uint8_t *__restrict__ nc,
                                                 const double *__restrict__ x, const double *__restrict__ b,
                                                 double *__restrict__ r) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= o.n) return;
  r[i] = (nc[i] & 0x80) ? 0.0 : b[i] - apply_at(o, nm, nc, x, i);
}
// y = A x
__global__ __launch_bounds__(NT) void k_apply64(Op64 o, const uint8_t *__restrict__ nm, const uint8_t *__restrict__ nc,
                                                 const double *__restrict__ x, double *__restrict__ y) {
  uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= o.n) return;
  y[i] = apply_at(o, nm, nc, x, i);
}

void launch_resid64(const MgLevel &L, const double *x, const double *b, double *r, hipStream_t s) {
  k_resid64<<<nblocks(L.n), NT, 0, s>>>(op64_of(L), L.nm, L.nc, x, b, r);
}
void launch_apply64(const MgLevel &L, const double *x, double *y, hipStream_t s) {
  k_apply64<<<nblocks(L.n), NT, 0, s>>>(op64_of(L), L.nm, L.nc, x, y);
}

// Up to 4 dot products at once: out[k] = sum_i a_k[i] * b_k[i].
struct DotArgs {
  const double *a[4];
  const double *b[4];
  int k;
};
constexpr int DOT_BLOCKS = 1024;
template <int K>
__global__ __launch_bounds__(NT) void k_dot_partial(DotArgs d, uint32_t n, double *__restrict__ part) {
  double acc[K];
#pragma unroll
  for (int q = 0; q < K; q++) acc[q] = 0.0;
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += DOT_BLOCKS * NT) {
#pragma unroll
    for (int q = 0; q < K; q++) acc[q] += d.a[q][i] * d.b[q][i];
  }
  __shared__ double sh[NT / 64];
#pragma unroll
  for (int q = 0; q < K; q++) {
    double v = block_sum<NT>(acc[q], sh);
    if (threadIdx.x == 0) part[q * DOT_BLOCKS + blockIdx.x] = v;
  }
}
__global__ __launch_bounds__(1024) void k_dot_final(const double *__restrict__ part, int k, double *__restrict__ out) {
  __shared__ double sh[1024];
  for (int q = 0; q < k; q++) {
    sh[threadIdx.x] = part[q * DOT_BLOCKS + threadIdx.x];
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
    __syncthreads();
  }
}
void launch_dots(const double *const *a, const double *const *b, int k, uint32_t n, double *part, double *out,
                 hipStream_t s) {
  DotArgs d;
  for (int q = 0; q < 4; q++) {
    d.a[q] = q < k ? a[q] : nullptr;
    d.b[q] = q < k ? b[q] : nullptr;
  }
  d.k = k;
  switch (k) {
    case 1: k_dot_partial<1><<<DOT_BLOCKS, NT, 0, s>>>(d, n, part); break;
    case 2: k_dot_partial<2><<<DOT_BLOCKS, NT, 0, s>>>(d, n, part); break;
    case 3: k_dot_partial<3><<<DOT_BLOCKS, NT, 0, s>>>(d, n, part); break;
    default: k_dot_partial<4><<<DOT_BLOCKS, NT, 0, s>>>(d, n, part); break;
  }
  k_dot_final<<<1, 1024, 0, s>>>(part, k, out);
}

// y += alpha * x
__global__ __launch_bounds__(NT) void k_axpy64(uint32_t n, double alpha, const double *__restrict__ x,
                                                double *__restrict__ y) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) y[i] += alpha * x[i];
}
void launch_axpy64(uint32_t n, double alpha, const double *x, double *y, hipStream_t s) {
  k_axpy64<<<nblocks(n), NT, 0, s>>>(n, alpha, x, y);
}
// p = z + beta * p
__global__ __launch_bounds__(NT) void k_xpby64(uint32_t n, const double *__restrict__ z, double beta,
                                                double *__restrict__ p) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}
void launch_xpby64(uint32_t n, const double *z, double beta, double *p, hipStream_t s) {
  k_xpby64<<<nblocks(n), NT, 0, s>>>(n, z, beta, p);
}
__global__ __launch_bounds__(NT) void k_to_f32(uint32_t n, const double *__restrict__ a, float *__restrict__ b) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ __launch_bounds__(NT) void k_to_f64(uint32_t n, const float *__restrict__ a, double *__restrict__ b) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ __launch_bounds__(NT) void k_add_f32_to_f64(uint32_t n, const float *__restrict__ a, double *__restrict__ b) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] += (double)a[i];
}
void launch_to_f32(uint32_t n, const double *a, float *b, hipStream_t s) { k_to_f32<<<nblocks(n), NT, 0, s>>>(n, a, b); }
void launch_to_f64(uint32_t n, const float *a, double *b, hipStream_t s) { k_to_f64<<<nblocks(n), NT, 0, s>>>(n, a, b); }
void launch_add_f32_to_f64(uint32_t n, const float *a, double *b, hipStream_t s) {
  k_add_f32_to_f64<<<nblocks(n), NT, 0, s>>>(n, a, b);
}
// zero fill (a kernel rather than a memset node, so it is graph-safe)
__global__ __launch_bounds__(NT) void k_zero32(uint32_t n, float *__restrict__ a) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) a[i] = 0.0f;
}
void launch_zero32(uint32_t n, float *a, hipStream_t s) { k_zero32<<<nblocks(n), NT, 0, s>>>(n, a); }
__global__ __launch_bounds__(NT) void k_zero64(uint32_t n, double *__restrict__ a) {
  uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) a[i] = 0.0;
}
void launch_zero64(uint32_t n, double *a, hipStream_t s) { k_zero64<<<nblocks(n), NT, 0, s>>>(n, a); }
__global__ __launch_bounds__(NT) void k_zero_u8(uint32_t n, uint8_t *a) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ---------------------------------------------------------------- fused CG steps
// Per-block FP64 partial sums go to part[k * nb + block] (nb = blocks of the
// launch); k_sum_final adds them in a fixed order into dev[slot + k], so the
// scalars of the iteration stay on the device and every run is bit-reproducible.
__global__ __launch_bounds__(1024) void k_sum_final(const double *__restrict__ part, int nb, int k, double *dev,
                                                    int slot) {
  __shared__ double sh[1024];
  for (int q = 0; q < k; q++) {
    const int t = (int)threadIdx.x;
    sh[t] = seq_sum(part, strided_count(t, nb, 1024), (size_t)q * nb + t, 1024);
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dev[slot + q] = sh[0];
    __syncthreads();
  }
}

// q = A p with partials of p.q and p.r
__global__ __launch_bounds__(NT) void k_apply_dots(Op64 o, const uint8_t *__restrict__ nm, const uint8_t *__restrict__ nc,
                                                    const double *__restrict__ p, const double *__restrict__ r,
                                                    double *__restrict__ q, double *__restrict__ part) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  double pq = 0.0, pr = 0.0;
  if (i < o.n) {
    const double qi = apply_at(o, nm, nc, p, i);
    q[i] = qi;
    pq = p[i] * qi;
    pr = p[i] * r[i];
  }
  __shared__ double sh[NT / 64];
  const double a = block_sum<NT>(pq, sh);
  const double b = block_sum<NT>(pr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}
// alpha = dev[1] / dev[0]; x += alpha p; r -= alpha q; partials of r.r
__global__ __launch_bounds__(NT) void k_update_rr(uint32_t n, const double *__restrict__ dev,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ part) {
  const double alpha = dev[0] != 0.0 ? dev[1] / dev[0] : 0.0;
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rr = ri * ri;
  }
  __shared__ double sh[NT / 64];
  const double a = block_sum<NT>(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// preconditioner input: b32 = (float)r, phi = 0
__global__ __launch_bounds__(NT) void k_to_f32_zero(uint32_t n, const double *__restrict__ r, float *__restrict__ b,
                                                     float *__restrict__ phi) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) {
    b[i] = (float)r[i];
    phi[i] = 0.0f;
  }
}
// z = (double)phi with partials of z.q
__global__ __launch_bounds__(NT) void k_to_f64_dot(uint32_t n, const float *__restrict__ phi,
                                                    const double *__restrict__ q, double *__restrict__ z,
                                                    double *__restrict__ part) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  double zq = 0.0;
  if (i < n) {
    const double zi = (double)phi[i];
    z[i] = zi;
    zq = zi * q[i];
  }
  __shared__ double sh[NT / 64];
  const double a = block_sum<NT>(zq, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// p = z + beta p with beta = -dev[3] / dev[0]
__global__ __launch_bounds__(NT) void k_xpby_dev(uint32_t n, const double *__restrict__ z,
                                                  const double *__restrict__ dev, double *__restrict__ p) {
  const double beta = dev[0] != 0.0 ? -dev[3] / dev[0] : 0.0;
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}

// r = b - A x with partials of r.r
__global__ __launch_bounds__(NT) void k_resid_rr(Op64 o, const uint8_t *__restrict__ nm, const uint8_t *__restrict__ nc,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r, double *__restrict__ part) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  double rr = 0.0;
  if (i < o.n) {
    const double ri = (nc[i] & 0x80) ? 0.0 : b[i] - apply_at(o, nm, nc, x, i);
    r[i] = ri;
    rr = ri * ri;
  }
  __shared__ double sh[NT / 64];
  const double a = block_sum<NT>(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

uint32_t cg_part_blocks(uint32_t n) { return nblocks(n); }
void launch_resid_rr(const MgLevel &L, const double *x, const double *b, double *r, double *part, double *dev,
                     hipStream_t s) {
  const unsigned nb = nblocks(L.n);
  k_resid_rr<<<nb, NT, 0, s>>>(op64_of(L), L.nm, L.nc, x, b, r, part);
  k_sum_final<<<1, 1024, 0, s>>>(part, (int)nb, 1, dev, 2);
}
void launch_apply_dots(const MgLevel &L, const double *p, const double *r, double *q, double *part, double *dev,
                       hipStream_t s) {
  const unsigned nb = nblocks(L.n);
  k_apply_dots<<<nb, NT, 0, s>>>(op64_of(L), L.nm, L.nc, p, r, q, part);
  k_sum_final<<<1, 1024, 0, s>>>(part, (int)nb, 2, dev, 0);
}
void launch_update_rr(uint32_t n, const double *dev, const double *p, const double *q, double *x, double *r,
                      double *part, double *dev_out, hipStream_t s) {
  const unsigned nb = nblocks(n);
  k_update_rr<<<nb, NT, 0, s>>>(n, dev, p, q, x, r, part);
  k_sum_final<<<1, 1024, 0, s>>>(part, (int)nb, 1, dev_out, 2);
}
void launch_to_f32_zero(uint32_t n, const double *r, float *b, float *phi, hipStream_t s) {
  k_to_f32_zero<<<nblocks(n), NT, 0, s>>>(n, r, b, phi);
}
void launch_to_f64_dot(uint32_t n, const float *phi, const double *q, double *z, double *part, double *dev,
                       hipStream_t s) {
  const unsigned nb = nblocks(n);
  k_to_f64_dot<<<nb, NT, 0, s>>>(n, phi, q, z, part);
  k_sum_final<<<1, 1024, 0, s>>>(part, (int)nb, 1, dev, 3);
}
void launch_xpby_dev(uint32_t n, const double *z, const double *dev, double *p, hipStream_t s) {
  k_xpby_dev<<<nblocks(n), NT, 0, s>>>(n, z, dev, p);
}


// ---------------------------------------------------------------- sealed fluid components
// Connected components of the fluid cells (6-neighbour) by union-find:
// parent[i] starts at i; every fluid cell unites itself with its fluid +x,
// +y, +z neighbours, always hanging the larger root under the smaller by a
// compare-and-swap on the root.  A link only ever points a cell at a smaller
// index, so a find walks strictly decreasing indices and a failed
// compare-and-swap restarts from a strictly smaller root: both loops end.
// Afterwards every path is compressed; the root of a component is then its
// smallest cell index, whatever the order of the links.  A component is
// open when one of its cells lies on a side (x = 0 or d0-1, z = 0 or d2-1)
// or at the top (y = d1-1), where phi = 0 lies beyond the face, and sealed
// otherwise: a pure-Neumann pocket, consistent only when b sums to zero.
__device__ __forceinline__ uint32_t cc_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t cc_find(uint32_t *parent, uint32_t x) {
  uint32_t p = cc_load(parent + x);
  while (p != x) {  // path halving: parent[x] is replaced by an ancestor
    const uint32_t gp = cc_load(parent + p);
    if (gp == p) return p;
    __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
    p = cc_load(parent + x);
  }
  return x;
}
// without path halving: in the compression pass the only store to
// parent[i] must be thread i's own (the root), or a halving store of another
// thread passing through i could leave an intermediate ancestor there
__device__ __forceinline__ uint32_t cc_find_ro(const uint32_t *parent, uint32_t x) {
  uint32_t p = cc_load(parent + x);
  while (p != x) {
    x = p;
    p = cc_load(parent + x);
  }
  return x;
}
__device__ __forceinline__ void cc_unite(uint32_t *parent, uint32_t a, uint32_t b) {
  uint32_t ra = cc_find(parent, a), rb = cc_find(parent, b);
  while (ra != rb) {
    if (ra < rb) {
      const uint32_t t = ra;
      ra = rb;
      rb = t;
    }
    const uint32_t old = atomicCAS(parent + ra, ra, rb);
    if (old == ra) return;
    ra = cc_find(parent, old);  // ra was linked meanwhile: go on from its root
    rb = cc_find(parent, rb);
  }
}
struct CcDims {
  int d0, d1, d2;
  uint32_t n;
  FastDiv dd0, dd1;
};
CcDims cc_dims(const MgLevel &L) {
  CcDims c;
  c.d0 = L.d[0]; c.d1 = L.d[1]; c.d2 = L.d[2];
  c.n = L.n;
  c.dd0 = make_fastdiv((uint32_t)L.d[0]);
  c.dd1 = make_fastdiv((uint32_t)L.d[1]);
  return c;
}
__global__ __launch_bounds__(NT) void k_cc_init(uint32_t n, uint32_t *parent) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i < n) parent[i] = i;
}
__global__ __launch_bounds__(NT) void k_cc_link(CcDims c, const uint8_t *__restrict__ f, uint32_t *parent) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= c.n || f[i] == 1) return;
  const uint32_t r = fdiv(i, c.dd0), x = i - r * c.dd0.d, z = fdiv(r, c.dd1), y = r - z * c.dd1.d;
  const uint32_t sy = (uint32_t)c.d0, sz = (uint32_t)c.d0 * (uint32_t)c.d1;
  if (x + 1 < (uint32_t)c.d0 && f[i + 1] != 1) cc_unite(parent, i, i + 1);
  if (y + 1 < (uint32_t)c.d1 && f[i + sy] != 1) cc_unite(parent, i, i + sy);
  if (z + 1 < (uint32_t)c.d2 && f[i + sz] != 1) cc_unite(parent, i, i + sz);
}
// compress to the roots; mark the roots of components with an open face
__global__ __launch_bounds__(NT) void k_cc_root_open(CcDims c, const uint8_t *__restrict__ f, uint32_t *parent,
                                                     uint8_t *__restrict__ open) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= c.n || f[i] == 1) return;
  const uint32_t root = cc_find_ro(parent, i);
  __hip_atomic_store(parent + i, root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t r = fdiv(i, c.dd0), x = i - r * c.dd0.d, z = fdiv(r, c.dd1), y = r - z * c.dd1.d;
  if (x == 0 || x + 1 == (uint32_t)c.d0 || z == 0 || z + 1 == (uint32_t)c.d2 || y + 1 == (uint32_t)c.d1)
    open[root] = 1;
}
// (root, cell) of every fluid cell in a sealed component, in any order
__global__ __launch_bounds__(NT) void k_cc_sealed(uint32_t n, const uint8_t *__restrict__ f,
                                                  const uint32_t *__restrict__ parent,
                                                  const uint8_t *__restrict__ open, uint2 *out,
                                                  unsigned int *count) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= n || f[i] == 1) return;
  const uint32_t root = parent[i];
  if (open[root]) return;
  const unsigned int k = atomicAdd(count, 1u);
  out[k] = make_uint2(root, i);
}
void launch_cc_label(const MgLevel &L, uint32_t *parent, uint8_t *open, hipStream_t s) {
  const CcDims c = cc_dims(L);
  k_cc_init<<<nblocks(L.n), NT, 0, s>>>(L.n, parent);
  k_zero_u8<<<nblocks(L.n), NT, 0, s>>>(L.n, open);
  k_cc_link<<<nblocks(L.n), NT, 0, s>>>(c, L.f, parent);
  k_cc_root_open<<<nblocks(L.n), NT, 0, s>>>(c, L.f, parent, open);
}
void launch_cc_sealed(const MgLevel &L, const uint32_t *parent, const uint8_t *open, uint2 *out,
                      unsigned int *count, hipStream_t s) {
  k_cc_sealed<<<nblocks(L.n), NT, 0, s>>>(L.n, L.f, parent, open, out, count);
}

// b -= mean of b over each sealed component (segment k of idx: cells
// off[k] .. off[k+1]-1, ascending).  One block per component; the sum is
// formed in a fixed order, so the result is bit-reproducible.
__global__ __launch_bounds__(NT) void k_seal_project(const uint32_t *__restrict__ off,
                                                     const uint32_t *__restrict__ idx, double *b) {
  __shared__ double sh[NT / 64];
  __shared__ double mean_s;
  const uint32_t a = off[blockIdx.x], e = off[blockIdx.x + 1];
  double acc = 0.0;
  for (uint32_t j = a + threadIdx.x; j < e; j += NT) acc += b[idx[j]];
  const double sum = block_sum<NT>(acc, sh);  // valid in thread 0
  if (threadIdx.x == 0) mean_s = sum / (double)(e - a);
  __syncthreads();
  const double mean = mean_s;
  for (uint32_t j = a + threadIdx.x; j < e; j += NT) b[idx[j]] -= mean;
}
void launch_seal_project(uint32_t ncomp, const uint32_t *off, const uint32_t *idx, double *b, hipStream_t s) {
  if (ncomp) k_seal_project<<<ncomp, NT, 0, s>>>(off, idx, b);
}

// changed |= (f != prev); prev = f
__global__ __launch_bounds__(NT) void k_flags_changed(uint32_t n, const uint8_t *__restrict__ f, uint8_t *prev,
                                                      unsigned int *changed) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const uint8_t v = f[i];
  if (prev[i] != v) {
    prev[i] = v;
    *changed = 1u;
  }
}
void launch_flags_changed(uint32_t n, const uint8_t *f, uint8_t *prev, unsigned int *changed, hipStream_t s) {
  k_flags_changed<<<nblocks(n), NT, 0, s>>>(n, f, prev, changed);
}

}  // namespace pyro
