// k_net.hip -- kernels of the learned pressure preconditioner (the U-Net in
// net.hpp): MFMA implicit-GEMM 3x3x3 convolutions (forward and input
// gradient), MFMA weight gradients, and the VALU glue (first layer, head,
// pooling, relu masks, loss, Adam).
//
// MFMA v_mfma_f32_16x16x32_bf16 lane layout (64 lanes, j = lane / 16,
// c = lane % 16):  A[16 x 32]: lane holds A[c][8j .. 8j+7];
// B[32 x 16]: lane holds B[8j .. 8j+7][c];  D[16 x 16]: lane holds
// D[4j .. 4j+3][c].  (Checked on the device by the probe kernel below.)
#include "common.hpp"
#include "net.hpp"

namespace pyro {

namespace {
__device__ __forceinline__ f32x4 mfma16x16x32(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float bf(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 tobf(float f) { return (bf16)f; }
inline unsigned nblk(uint64_t n, int nt = 256) { return (unsigned)((n + nt - 1) / nt); }
}  // namespace

// ------------------------------------------------------------ probe
__global__ __launch_bounds__(64) void k_mfma_probe(const bf16 *a, const bf16 *b, float *d) {
  const int l = threadIdx.x;
  const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a + 8 * l);
  const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(b + 8 * l);
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = mfma16x16x32(av, bv, acc);
  for (int i = 0; i < 4; i++) d[4 * l + i] = acc[i];
}
void launch_mfma_probe(const bf16 *a, const bf16 *b, float *d, hipStream_t s) { k_mfma_probe<<<1, 64, 0, s>>>(a, b, d); }

// ------------------------------------------------------------ packing
// The upsampled input of a layer with one 16-row output block (cout = 16) is
// staged in 16-channel chunks (smaller LDS footprint, higher occupancy);
// otherwise in 32-channel chunks.
int conv_a_chunk(int cout) { return cout == 16 ? 16 : 32; }
int conv_groups(int ca, int cb, int cout) {
  const int cca = conv_a_chunk(cout);
  return (ca / cca) * (cca == 32 ? 27 : 14) + (cb == 16 ? 14 : (cb / 32) * 27);
}

__global__ __launch_bounds__(256) void k_pack_conv(const float *__restrict__ W, int cout, int cin, int ca_e, int cb_e,
                                                   int cout_e, int transposed, uint32_t total, bf16 *__restrict__ wp) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  const int kk = (int)(e % 32u);
  const int co_e = (int)((e / 32u) % (uint32_t)cout_e);
  int grp = (int)(e / (32u * (uint32_t)cout_e));
  const int cca = cout_e == 16 ? 16 : 32, ga = cca == 32 ? 27 : 14;
  const int nA = ca_e / cca;
  int tap, ci_e;
  if (grp < nA * ga) {
    if (cca == 32) {
      tap = grp % 27;
      ci_e = 32 * (grp / 27) + kk;
    } else {
      tap = 2 * (grp % 14) + kk / 16;
      ci_e = 16 * (grp / 14) + (kk % 16);
    }
  } else {
    grp -= nA * ga;
    if (cb_e == 16) {
      tap = 2 * grp + kk / 16;
      ci_e = ca_e + (kk % 16);
    } else {
      tap = grp % 27;
      ci_e = ca_e + 32 * (grp / 27) + kk;
    }
  }
  float v = 0.0f;
  if (tap < 27) {
    if (!transposed)
      v = W[((size_t)co_e * cin + ci_e) * 27 + tap];
    else  // adjoint: input channel = original output channel, taps mirrored
      v = W[((size_t)ci_e * cin + co_e) * 27 + (26 - tap)];
  }
  wp[e] = tobf(v);
}
void launch_pack_conv(const float *W, const ConvSpec &c, bool transposed, bf16 *wp, hipStream_t s) {
  const int ca_e = transposed ? 0 : c.ca, cb_e = transposed ? c.cout : c.cb;
  const int cout_e = transposed ? c.cin() : c.cout;
  const uint32_t total = (uint32_t)conv_groups(ca_e, cb_e, cout_e) * (uint32_t)cout_e * 32u;
  k_pack_conv<<<nblk(total), 256, 0, s>>>(W, c.cout, c.cin(), ca_e, cb_e, cout_e, transposed ? 1 : 0, total, wp);
}
__global__ __launch_bounds__(256) void k_pack_bf16(const float *__restrict__ w, uint32_t n, bf16 *__restrict__ o) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) o[i] = tobf(w[i]);
}
void launch_pack_bf16(const float *w, uint32_t n, bf16 *out, hipStream_t s) {
  k_pack_bf16<<<nblk(n), 256, 0, s>>>(w, n, out);
}

// ------------------------------------------------------------ conv (MFMA)
// Output tile 16 (x) x 4 (y) x 4 (z) voxels per 256-thread block: wave w
// takes row y0 + w, its four 16-voxel column blocks are z0 .. z0+3, lane c
// is x0 + c.  The input footprint (18 x 6 x 6 voxels) of one channel chunk
// (32 channels, or 16 for a 16-channel source) is staged in LDS, then every
// K32 group issues (cout/16) x 4 MFMAs with A = packed weights (global, L1
// resident) and B = 8 channels of one footprint voxel per lane.
constexpr int CV_TX = 16, CV_TY = 4, CV_TZ = 4;
constexpr int CV_FX = CV_TX + 2, CV_FY = CV_TY + 2, CV_FZ = CV_TZ + 2;
constexpr int CV_FN = CV_FX * CV_FY * CV_FZ;

struct ConvArgs {
  int nx, ny, nz;
  int anx, any;  // half-resolution grid of `a`
  const bf16 *a, *b, *wp;
  const float *bias;
  bf16 *out;
};

// footprint offset (in voxels) of tap t relative to footprint voxel (col, w, nb)
__device__ __forceinline__ constexpr int tap_off(int t) {
  return (t / 9) * (CV_FX * CV_FY) + ((t / 3) % 3) * CV_FX + (t % 3);
}

// Stage one channel chunk (CC channels from `a` through nearest upsampling
// when FROMA, else from `b`) of the 18 x 6 x 6 footprint into LDS.
template <int CC, int CSRC, bool FROMA>
__device__ __forceinline__ void conv_stage(const ConvArgs &p, int c0, int x0, int y0, int z0, bf16 *lds) {
  constexpr int PIECES = CC / 8, STRIDE = CC + 8;
  for (int s = threadIdx.x; s < CV_FN * PIECES; s += 256) {
    const int vox = s / PIECES, pc = s - vox * PIECES;
    const int fx = vox % CV_FX, r2 = vox / CV_FX, fy = r2 % CV_FY, fz = r2 / CV_FY;
    const int gx = x0 + fx - 1, gy = y0 + fy - 1, gz = z0 + fz - 1;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (gx >= 0 && gx < p.nx && gy >= 0 && gy < p.ny && gz >= 0 && gz < p.nz) {
      size_t v;
      if (FROMA)
        v = (size_t)(gx >> 1) + (size_t)p.anx * ((size_t)(gy >> 1) + (size_t)p.any * (size_t)(gz >> 1));
      else
        v = (size_t)gx + (size_t)p.nx * ((size_t)gy + (size_t)p.ny * (size_t)gz);
      val = *reinterpret_cast<const uint4 *>((FROMA ? p.a : p.b) + v * CSRC + c0 + 8 * pc);
    }
    *reinterpret_cast<uint4 *>(lds + vox * STRIDE + 8 * pc) = val;
  }
}

// MFMAs of one staged chunk: NG K32 groups (27 taps of 32 channels, or 14
// tap pairs of 16 channels).  With one 16-row output block (MB = 1) every
// weight fragment of the chunk is loaded into registers up front and the
// group loop is fully unrolled, so the loop body is LDS reads and MFMAs only.
template <int CC, int MB, int COUT>
__device__ __forceinline__ void conv_chunk(const bf16 *__restrict__ wp, const bf16 *lds, f32x4 (&acc)[MB][4], int w,
                                           int j, int col) {
  constexpr int NG = CC == 32 ? 27 : 14, STRIDE = CC + 8;
  // output voxel (col, w, nb) sits at footprint (col+1, w+1, nb+1); tap t reads
  // footprint (col+kx, w+ky, nb+kz) = base + nb*FX*FY + tap_off(t)
  const int base = w * CV_FX + col;
  auto group = [&](int g, const bf16x8 (&A)[MB]) {
    int off, ch;
    if (CC == 32) {
      off = tap_off(g);
      ch = 8 * j;
    } else {
      int t = 2 * g + (j >> 1);
      t = t > 26 ? 26 : t;  // the 28th tap has zero weights
      off = tap_off(t);
      ch = 8 * (j & 1);
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const bf16x8 B = *reinterpret_cast<const bf16x8 *>(lds + (base + nb * (CV_FX * CV_FY) + off) * STRIDE + ch);
#pragma unroll
      for (int m = 0; m < MB; m++) acc[m][nb] = mfma16x16x32(A[m], B, acc[m][nb]);
    }
  };
  if constexpr (MB == 1) {
    // batches of BS groups, the next batch's weight fragments loaded while
    // the current batch computes
    constexpr int BS = CC == 32 ? 9 : 7, NBATCH = NG / BS;
    static_assert(NBATCH * BS == NG, "group batches");
    bf16x8 Acur[BS][1], Anext[BS][1];
#pragma unroll
    for (int g = 0; g < BS; g++)
      Acur[g][0] = *reinterpret_cast<const bf16x8 *>(wp + ((size_t)g * COUT + col) * 32 + 8 * j);
#pragma unroll 1
    for (int bt = 0; bt < NBATCH; bt++) {
      if (bt + 1 < NBATCH) {
#pragma unroll
        for (int g = 0; g < BS; g++)
          Anext[g][0] =
              *reinterpret_cast<const bf16x8 *>(wp + ((size_t)((bt + 1) * BS + g) * COUT + col) * 32 + 8 * j);
      }
#pragma unroll
      for (int g = 0; g < BS; g++) group(bt * BS + g, Acur[g]);
      if (bt + 1 < NBATCH) {
#pragma unroll
        for (int g = 0; g < BS; g++) Acur[g][0] = Anext[g][0];
      }
    }
  } else {
#pragma unroll 3
    for (int g = 0; g < NG; g++) {
      bf16x8 A[MB];
#pragma unroll
      for (int m = 0; m < MB; m++)
        A[m] = *reinterpret_cast<const bf16x8 *>(wp + ((size_t)g * COUT + m * 16 + col) * 32 + 8 * j);
      group(g, A);
    }
  }
}

template <int CA, int CB, int COUT, bool RELU>
__global__ __launch_bounds__(256, 3) void k_conv3(ConvArgs p) {
  constexpr int MB = COUT / 16;
  constexpr int CCA = MB == 1 ? 16 : 32;  // conv_a_chunk
  constexpr int NCA = CA / CCA;
  constexpr int NCB = CB == 16 ? 1 : CB / 32;
  constexpr int CCB = CB == 16 ? 16 : 32;
  constexpr int CCMAX = ((NCA > 0 && CCA == 32) || CCB == 32) ? 32 : 16;
  __shared__ __align__(16) bf16 lds[CV_FN * (CCMAX + 8)];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = lane >> 4, col = lane & 15;
  const int ntx = (p.nx + CV_TX - 1) / CV_TX, nty = (p.ny + CV_TY - 1) / CV_TY;
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = bid % ntx, rr = bid / ntx, ty = rr % nty, tz = rr / nty;
  const int x0 = tx * CV_TX, y0 = ty * CV_TY, z0 = tz * CV_TZ;
  f32x4 acc[MB][4];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[m][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const bf16 *wp = p.wp;
#pragma unroll 1
  for (int q = 0; q < NCA; q++) {
    __syncthreads();
    conv_stage<CCA, CA, true>(p, CCA * q, x0, y0, z0, lds);
    __syncthreads();
    conv_chunk<CCA, MB, COUT>(wp, lds, acc, w, j, col);
    wp += (size_t)(CCA == 32 ? 27 : 14) * COUT * 32;
  }
#pragma unroll 1
  for (int q = 0; q < NCB; q++) {
    __syncthreads();
    conv_stage<CCB, CB, false>(p, 32 * q, x0, y0, z0, lds);
    __syncthreads();
    conv_chunk<CCB, MB, COUT>(wp, lds, acc, w, j, col);
    wp += (size_t)(CCB == 32 ? 27 : 14) * COUT * 32;
  }
  const int gx = x0 + col, gy = y0 + w;
  if (gx >= p.nx || gy >= p.ny) return;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const int gz = z0 + nb;
    if (gz >= p.nz) continue;
    const size_t v = (size_t)gx + (size_t)p.nx * ((size_t)gy + (size_t)p.ny * (size_t)gz);
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int co = m * 16 + 4 * j;
      bf16x4 o;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        float t = acc[m][nb][i] + (p.bias ? p.bias[co + i] : 0.0f);
        if (RELU) t = t > 0.0f ? t : 0.0f;
        o[i] = tobf(t);
      }
      *reinterpret_cast<bf16x4 *>(p.out + v * COUT + co) = o;
    }
  }
}

bool launch_conv3(const G3 &g, const bf16 *a, const G3 &ga, int ca, const bf16 *b, int cb, const bf16 *wp,
                  const float *bias, int cout, bool relu, bf16 *out, hipStream_t s) {
  ConvArgs p;
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
  p.anx = ga.nx; p.any = ga.ny;
  p.a = a; p.b = b; p.wp = wp; p.bias = bias; p.out = out;
  const unsigned nb = (unsigned)(((g.nx + CV_TX - 1) / CV_TX) * ((g.ny + CV_TY - 1) / CV_TY) *
                                 ((g.nz + CV_TZ - 1) / CV_TZ));
#define PYRO_CONV(CA_, CB_, CO_, R_)                                          \
  if (ca == CA_ && cb == CB_ && cout == CO_ && relu == R_) {                  \
    k_conv3<CA_, CB_, CO_, R_><<<nb, 256, 0, s>>>(p);                         \
    return true;                                                              \
  }
  PYRO_CONV(0, 16, 16, true)
  PYRO_CONV(0, 16, 32, true)
  PYRO_CONV(0, 32, 32, true)
  PYRO_CONV(0, 32, 64, true)
  PYRO_CONV(0, 64, 64, true)
  PYRO_CONV(64, 32, 32, true)
  PYRO_CONV(32, 16, 16, true)
  PYRO_CONV(0, 16, 16, false)
  PYRO_CONV(0, 32, 16, false)
  PYRO_CONV(0, 32, 32, false)
  PYRO_CONV(0, 64, 32, false)
  PYRO_CONV(0, 64, 64, false)
  PYRO_CONV(0, 32, 96, false)
  PYRO_CONV(0, 16, 48, false)
#undef PYRO_CONV
  return false;
}

// ------------------------------------------------------------ first layer, head, pool
// First layer on MFMA: 27 taps x 2 channels = 54 of the K = 64 of two K32
// groups (group g holds taps 16g .. 16g+15, channel-minor; taps past 26 have
// zero weights).  Lane group j of a B fragment holds 4 taps x 2 channels, one
// 4-byte LDS word per tap.  Packed weights: [2][16][32].
__global__ __launch_bounds__(256) void k_pack_conv_in(const float *__restrict__ W, bf16 *__restrict__ wp) {
  const int e = threadIdx.x + 256 * blockIdx.x;
  if (e >= 2 * 16 * 32) return;
  const int kk = e % 32, co = (e / 32) % 16, g = e / 512;
  const int tap = 16 * g + kk / 2, ci = kk % 2;
  wp[e] = tobf(tap < 27 ? W[(co * 2 + ci) * 27 + tap] : 0.0f);
}
void launch_pack_conv_in(const float *W, bf16 *wp, hipStream_t s) { k_pack_conv_in<<<4, 256, 0, s>>>(W, wp); }

__global__ __launch_bounds__(256) void k_conv_in(int nx, int ny, int nz, const bf16 *__restrict__ x0,
                                                 const bf16 *__restrict__ wp, const float *__restrict__ bias,
                                                 bf16 *__restrict__ out) {
  __shared__ uint32_t lds[CV_FN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = lane >> 4, col = lane & 15;
  const int ntx = (nx + CV_TX - 1) / CV_TX, nty = (ny + CV_TY - 1) / CV_TY;
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = bid % ntx, rr = bid / ntx, ty = rr % nty, tz = rr / nty;
  const int x0v = tx * CV_TX, y0 = ty * CV_TY, z0 = tz * CV_TZ;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(x0);
  for (int s = threadIdx.x; s < CV_FN; s += 256) {
    const int fx = s % CV_FX, r2 = s / CV_FX, fy = r2 % CV_FY, fz = r2 / CV_FY;
    const int gx = x0v + fx - 1, gy = y0 + fy - 1, gz = z0 + fz - 1;
    uint32_t v = 0u;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
      v = src[(size_t)gx + (size_t)nx * ((size_t)gy + (size_t)ny * (size_t)gz)];
    lds[s] = v;
  }
  bf16x8 A[2];
#pragma unroll
  for (int g = 0; g < 2; g++) A[g] = *reinterpret_cast<const bf16x8 *>(wp + (g * 16 + col) * 32 + 8 * j);
  __syncthreads();
  f32x4 acc[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int base = w * CV_FX + col;
#pragma unroll
  for (int g = 0; g < 2; g++) {
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      uint4 u;
      uint32_t t4[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int t = 16 * g + 4 * j + i;
        t = t > 26 ? 26 : t;  // zero weights beyond the 27 taps
        t4[i] = lds[base + nb * (CV_FX * CV_FY) + tap_off(t)];
      }
      u.x = t4[0]; u.y = t4[1]; u.z = t4[2]; u.w = t4[3];
      acc[nb] = mfma16x16x32(A[g], __builtin_bit_cast(bf16x8, u), acc[nb]);
    }
  }
  const int gx = x0v + col, gy = y0 + w;
  if (gx >= nx || gy >= ny) return;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const int gz = z0 + nb;
    if (gz >= nz) continue;
    const size_t v = (size_t)gx + (size_t)nx * ((size_t)gy + (size_t)ny * (size_t)gz);
    const int co = 4 * j;
    bf16x4 o;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float t = acc[nb][i] + bias[co + i];
      o[i] = tobf(t > 0.0f ? t : 0.0f);
    }
    *reinterpret_cast<bf16x4 *>(out + v * 16 + co) = o;
  }
}
void launch_conv_in(const G3 &g, const bf16 *x0, const bf16 *wp, const float *bias, bf16 *out, hipStream_t s) {
  const unsigned nb = (unsigned)(((g.nx + CV_TX - 1) / CV_TX) * ((g.ny + CV_TY - 1) / CV_TY) *
                                 ((g.nz + CV_TZ - 1) / CV_TZ));
  k_conv_in<<<nb, 256, 0, s>>>(g.nx, g.ny, g.nz, x0, wp, bias, out);
}

__global__ __launch_bounds__(256) void k_head(uint32_t n, const bf16 *__restrict__ d0, const float *__restrict__ wh,
                                              const float *__restrict__ bh, float *__restrict__ y) {
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(d0 + 16 * (size_t)v);
  const bf16x8 b = *reinterpret_cast<const bf16x8 *>(d0 + 16 * (size_t)v + 8);
  float t = bh[0];
#pragma unroll
  for (int c = 0; c < 8; c++) t += wh[c] * bf(a[c]);
#pragma unroll
  for (int c = 0; c < 8; c++) t += wh[8 + c] * bf(b[c]);
  y[v] = t;
}
void launch_head(const G3 &g, const bf16 *d0, const float *wh, const float *bh, float *y, hipStream_t s) {
  k_head<<<nblk(g.n()), 256, 0, s>>>(g.n(), d0, wh, bh, y);
}

// children of coarse voxel (ux, uy, uz) in the fine grid, in z, y, x order
__global__ __launch_bounds__(256) void k_pool(int nx, int ny, int nz, int c, const bf16 *__restrict__ in,
                                              bf16 *__restrict__ out) {
  const int cnx = (nx + 1) / 2, cny = (ny + 1) / 2, cnz = (nz + 1) / 2;
  const int ng = c / 8;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t total = (uint32_t)cnx * cny * cnz * ng;
  if (t >= total) return;
  const int cg = (int)(t % (uint32_t)ng);
  const uint32_t u = t / (uint32_t)ng;
  const int ux = (int)(u % (uint32_t)cnx), r = (int)(u / (uint32_t)cnx), uy = r % cny, uz = r / cny;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int cnt = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const int x = 2 * ux + dx, y = 2 * uy + dy, z = 2 * uz + dz;
        if (x >= nx || y >= ny || z >= nz) continue;
        const size_t v = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(in + v * c + 8 * cg);
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += bf(a[i]);
        cnt++;
      }
  const float inv = 1.0f / (float)cnt;
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; i++) o[i] = tobf(s[i] * inv);
  *reinterpret_cast<bf16x8 *>(out + (size_t)u * c + 8 * cg) = o;
}
void launch_pool(const G3 &g, const bf16 *in, int c, bf16 *out, hipStream_t s) {
  const G3 h = half_of(g);
  k_pool<<<nblk((uint64_t)h.n() * (c / 8)), 256, 0, s>>>(g.nx, g.ny, g.nz, c, in, out);
}

// ------------------------------------------------------------ backward glue
__global__ __launch_bounds__(256) void k_relu_bwd(uint32_t n8, const bf16 *__restrict__ act, bf16 *__restrict__ g) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n8) return;
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + 8 * (size_t)i);
  bf16x8 v = *reinterpret_cast<const bf16x8 *>(g + 8 * (size_t)i);
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (!(bf(a[k]) > 0.0f)) v[k] = tobf(0.0f);
  *reinterpret_cast<bf16x8 *>(g + 8 * (size_t)i) = v;
}
void launch_relu_bwd(uint32_t n, const bf16 *act, bf16 *g, hipStream_t s) {
  k_relu_bwd<<<nblk(n / 8), 256, 0, s>>>(n / 8, act, g);
}

__global__ __launch_bounds__(256) void k_unup_relu(int nx, int ny, int nz, const bf16 *__restrict__ gcat, int ctot,
                                                   int ca, const bf16 *__restrict__ act, bf16 *__restrict__ gc) {
  const int cnx = (nx + 1) / 2, cny = (ny + 1) / 2, cnz = (nz + 1) / 2;
  const int ng = ca / 8;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)cnx * cny * cnz * ng) return;
  const int cg = (int)(t % (uint32_t)ng);
  const uint32_t u = t / (uint32_t)ng;
  const int ux = (int)(u % (uint32_t)cnx), r = (int)(u / (uint32_t)cnx), uy = r % cny, uz = r / cny;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const int x = 2 * ux + dx, y = 2 * uy + dy, z = 2 * uz + dz;
        if (x >= nx || y >= ny || z >= nz) continue;
        const size_t v = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(gcat + v * ctot + 8 * cg);
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += bf(a[i]);
      }
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + (size_t)u * ca + 8 * cg);
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; i++) o[i] = tobf(bf(a[i]) > 0.0f ? s[i] : 0.0f);
  *reinterpret_cast<bf16x8 *>(gc + (size_t)u * ca + 8 * cg) = o;
}
void launch_unup_relu(const G3 &g, const bf16 *gcat, int ctot, int ca, const G3 &gc, const bf16 *act, bf16 *gcoarse,
                      hipStream_t s) {
  k_unup_relu<<<nblk((uint64_t)gc.n() * (ca / 8)), 256, 0, s>>>(g.nx, g.ny, g.nz, gcat, ctot, ca, act, gcoarse);
}

__global__ __launch_bounds__(256) void k_skip_grad(int nx, int ny, int nz, const bf16 *__restrict__ gcat, int ctot,
                                                   int off, const bf16 *__restrict__ gp, const bf16 *__restrict__ sa,
                                                   int c, bf16 *__restrict__ gs) {
  const int ng = c / 8;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)nx * ny * nz * ng) return;
  const int cg = (int)(t % (uint32_t)ng);
  const uint32_t v = t / (uint32_t)ng;
  const int x = (int)(v % (uint32_t)nx), r = (int)(v / (uint32_t)nx), y = r % ny, z = r / ny;
  const int cnx = (nx + 1) / 2, cny = (ny + 1) / 2;
  const int ux = x >> 1, uy = y >> 1, uz = z >> 1;
  // children of the parent inside the grid
  const int cnt = (2 * ux + 1 < nx ? 2 : 1) * (2 * uy + 1 < ny ? 2 : 1) * (2 * uz + 1 < nz ? 2 : 1);
  const float inv = 1.0f / (float)cnt;
  const size_t u = (size_t)ux + (size_t)cnx * ((size_t)uy + (size_t)cny * (size_t)uz);
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(gcat + (size_t)v * ctot + off + 8 * cg);
  const bf16x8 b = *reinterpret_cast<const bf16x8 *>(gp + u * c + 8 * cg);
  const bf16x8 m = *reinterpret_cast<const bf16x8 *>(sa + (size_t)v * c + 8 * cg);
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; i++) o[i] = tobf(bf(m[i]) > 0.0f ? bf(a[i]) + bf(b[i]) * inv : 0.0f);
  *reinterpret_cast<bf16x8 *>(gs + (size_t)v * c + 8 * cg) = o;
}
void launch_skip_grad(const G3 &g, const bf16 *gcat, int ctot, int off, const G3 &gc, const bf16 *gp, const bf16 *s,
                      int c, bf16 *gs, hipStream_t st) {
  (void)gc;
  k_skip_grad<<<nblk((uint64_t)g.n() * (c / 8)), 256, 0, st>>>(g.nx, g.ny, g.nz, gcat, ctot, off, gp, s, c, gs);
}

// head backward: gd0 and per-block partial sums of dwh (16) and dbh (1)
constexpr int HB_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_head_bwd(uint32_t n, const bf16 *__restrict__ d0, const float *__restrict__ gy,
                                                  const float *__restrict__ wh, bf16 *__restrict__ gd0,
                                                  float *__restrict__ part) {
  float s[17];
#pragma unroll
  for (int c = 0; c < 17; c++) s[c] = 0.0f;
  for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v < n; v += HB_BLOCKS * 256u) {
    const float g = gy[v];
    bf16x8 a = *reinterpret_cast<const bf16x8 *>(d0 + 16 * (size_t)v);
    bf16x8 b = *reinterpret_cast<const bf16x8 *>(d0 + 16 * (size_t)v + 8);
    bf16x8 oa, ob;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      s[c] += g * bf(a[c]);
      s[8 + c] += g * bf(b[c]);
      oa[c] = tobf(bf(a[c]) > 0.0f ? g * wh[c] : 0.0f);
      ob[c] = tobf(bf(b[c]) > 0.0f ? g * wh[8 + c] : 0.0f);
    }
    s[16] += g;
    *reinterpret_cast<bf16x8 *>(gd0 + 16 * (size_t)v) = oa;
    *reinterpret_cast<bf16x8 *>(gd0 + 16 * (size_t)v + 8) = ob;
  }
  __shared__ float sh[256];
  for (int c = 0; c < 17; c++) {
    sh[threadIdx.x] = s[c];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 17 + c] = sh[0];
    __syncthreads();
  }
}
__global__ __launch_bounds__(32) void k_head_bwd_reduce(const float *__restrict__ part, float *__restrict__ dwh,
                                                        float *__restrict__ dbh, int accumulate) {
  const int c = threadIdx.x;
  if (c >= 17) return;
  float s = 0.0f;
  for (int b = 0; b < HB_BLOCKS; b++) s += part[b * 17 + c];
  float *o = c < 16 ? dwh + c : dbh;
  *o = accumulate ? *o + s : s;
}
void launch_head_bwd(const G3 &g, const bf16 *d0, const float *gy, const float *wh, bf16 *gd0, float *part,
                     float *dwh, float *dbh, bool accumulate, hipStream_t s) {
  k_head_bwd<<<HB_BLOCKS, 256, 0, s>>>(g.n(), d0, gy, wh, gd0, part);
  k_head_bwd_reduce<<<1, 32, 0, s>>>(part, dwh, dbh, accumulate ? 1 : 0);
}

// ------------------------------------------------------------ weight gradients
// dW[co][ci][t] = sum_v g[v][co] * X[v + off(t)][ci].  The MFMA contracts over
// voxels, so both operands are first written channel-major: GT[co][z][y][x]
// and, for each x shift dx in {-1, 0, 1}, XT_dx[ci][z+1][y+1][x] =
// X[x+dx][y][z] (zero outside the grid), rows padded to a pitch P multiple of
// 32.  Every fragment is then one aligned 16-byte load.
DwPlan dw_plan(const G3 &g, const ConvSpec &c) {
  DwPlan d;
  d.P = (g.nx + 31) / 32 * 32;
  d.rows_x = c.cin() < 16 ? 16 : c.cin();
  const int tasks = 27 * (d.rows_x / 16);
  int S = (2048 + tasks - 1) / tasks;
  const int rows = g.ny * g.nz;
  const int smax = rows / 4 > 0 ? rows / 4 : 1;
  d.S = S < 1 ? 1 : (S > smax ? smax : S);
  return d;
}
size_t dw_xt_elems(const G3 &g, const ConvSpec &c) {
  const DwPlan d = dw_plan(g, c);
  return (size_t)3 * d.rows_x * (size_t)(g.nz + 2) * (size_t)(g.ny + 2) * (size_t)d.P;
}
size_t dw_gt_elems(const G3 &g, const ConvSpec &c) {
  const DwPlan d = dw_plan(g, c);
  return (size_t)c.cout * (size_t)g.nz * (size_t)g.ny * (size_t)d.P;
}
size_t dw_part_elems(const G3 &g, const ConvSpec &c) {
  const DwPlan d = dw_plan(g, c);
  const size_t dwp = (size_t)d.S * 27 * c.cout * d.rows_x;
  const size_t bp = (size_t)256 * c.cout;
  return dwp > bp ? dwp : bp;
}

// XT from the concatenated conv input [up(a) (ca ch), b (cb ch)]; 8 channels per thread
__global__ __launch_bounds__(256) void k_to_xt(int nx, int ny, int nz, int P, int anx, int any, const bf16 *__restrict__ a,
                                               int ca, const bf16 *__restrict__ b, int cb, int rows_x,
                                               bf16 *__restrict__ xt) {
  const int ng = (ca + cb) / 8;
  const size_t plane = (size_t)(nz + 2) * (ny + 2) * P;
  const size_t total = 3 * (size_t)ng * plane;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  const int xp = (int)(t % P);
  size_t r = t / P;
  const int yy = (int)(r % (ny + 2));
  r /= (ny + 2);
  const int zz = (int)(r % (nz + 2));
  r /= (nz + 2);
  const int cg = (int)(r % ng);
  const int d = (int)(r / ng);
  const int x = xp + d - 1, y = yy - 1, z = zz - 1;
  bf16x8 v;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = tobf(0.0f);
  if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
    const int c = 8 * cg;
    if (c < ca) {
      const size_t u = (size_t)(x >> 1) + (size_t)anx * ((size_t)(y >> 1) + (size_t)any * (size_t)(z >> 1));
      v = *reinterpret_cast<const bf16x8 *>(a + u * ca + c);
    } else {
      const size_t w = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
      v = *reinterpret_cast<const bf16x8 *>(b + w * cb + (c - ca));
    }
  }
  const size_t base = ((size_t)d * rows_x + 8 * cg) * plane + ((size_t)zz * (ny + 2) + yy) * P + xp;
#pragma unroll
  for (int i = 0; i < 8; i++) xt[base + (size_t)i * plane] = v[i];
}
// XT of the 2-channel network input (rows 2..15 zero)
__global__ __launch_bounds__(256) void k_to_xt_in(int nx, int ny, int nz, int P, const bf16 *__restrict__ x0,
                                                  bf16 *__restrict__ xt) {
  const size_t plane = (size_t)(nz + 2) * (ny + 2) * P;
  const size_t total = 3 * plane;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  const int xp = (int)(t % P);
  size_t r = t / P;
  const int yy = (int)(r % (ny + 2));
  r /= (ny + 2);
  const int zz = (int)(r % (nz + 2));
  const int d = (int)(r / (nz + 2));
  const int x = xp + d - 1, y = yy - 1, z = zz - 1;
  bf16 v0 = tobf(0.0f), v1 = tobf(0.0f);
  if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
    const size_t w = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
    v0 = x0[2 * w];
    v1 = x0[2 * w + 1];
  }
  const size_t base = (size_t)d * 16 * plane + ((size_t)zz * (ny + 2) + yy) * P + xp;
  xt[base] = v0;
  xt[base + plane] = v1;
  for (int i = 2; i < 16; i++) xt[base + (size_t)i * plane] = tobf(0.0f);
}
// GT[co][z][y][x] (zero for x >= nx)
__global__ __launch_bounds__(256) void k_to_gt(int nx, int ny, int nz, int P, int cout, const bf16 *__restrict__ g,
                                               bf16 *__restrict__ gt) {
  const int ng = cout / 8;
  const size_t plane = (size_t)nz * ny * P;
  const size_t total = (size_t)ng * plane;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  const int xp = (int)(t % P);
  size_t r = t / P;
  const int y = (int)(r % ny);
  r /= ny;
  const int z = (int)(r % nz);
  const int cg = (int)(r / nz);
  bf16x8 v;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = tobf(0.0f);
  if (xp < nx) {
    const size_t w = (size_t)xp + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
    v = *reinterpret_cast<const bf16x8 *>(g + w * cout + 8 * cg);
  }
  const size_t base = (size_t)(8 * cg) * plane + ((size_t)z * ny + y) * P + xp;
#pragma unroll
  for (int i = 0; i < 8; i++) gt[base + (size_t)i * plane] = v[i];
}

// One block per (tap, 16-channel input block, row split); its 4 waves take
// interleaved quarters of the split's rows and are summed through LDS.
template <int COUT>
__global__ __launch_bounds__(256) void k_dw(int nx, int ny, int nz, int P, int rows_x, int S,
                                            const bf16 *__restrict__ gt, const bf16 *__restrict__ xt,
                                            float *__restrict__ part) {
  constexpr int MB = COUT / 16;
  const int ncib = rows_x / 16;
  const int task = blockIdx.x % (27 * ncib), split = blockIdx.x / (27 * ncib);
  const int tap = task / ncib, cib = task % ncib;
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, j = lane >> 4, col = lane & 15;
  const int rows = nz * ny;
  const int pid = split * 4 + w, np = S * 4;
  const int r0 = (int)((int64_t)rows * pid / np), r1 = (int)((int64_t)rows * (pid + 1) / np);
  const size_t xplane = (size_t)(nz + 2) * (ny + 2) * P;
  const size_t gplane = (size_t)nz * ny * P;
  const bf16 *xbase = xt + ((size_t)(dx + 1) * rows_x + cib * 16 + col) * xplane;
  f32x4 acc[MB];
#pragma unroll
  for (int m = 0; m < MB; m++) acc[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int r = r0; r < r1; r++) {
    const int z = r / ny, y = r - z * ny;
    const bf16 *xrow = xbase + ((size_t)(z + dz + 1) * (ny + 2) + (y + dy + 1)) * P;
    const bf16 *grow = gt + ((size_t)z * ny + y) * P;
    for (int xc = 0; xc < P; xc += 32) {
      const bf16x8 B = *reinterpret_cast<const bf16x8 *>(xrow + xc + 8 * j);
#pragma unroll
      for (int m = 0; m < MB; m++) {
        const bf16x8 A = *reinterpret_cast<const bf16x8 *>(grow + (size_t)(m * 16 + col) * gplane + xc + 8 * j);
        acc[m] = mfma16x16x32(A, B, acc[m]);
      }
    }
  }
  __shared__ float red[4][COUT * 16];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int i = 0; i < 4; i++) red[w][(m * 16 + 4 * j + i) * 16 + col] = acc[m][i];
  __syncthreads();
  for (int e = threadIdx.x; e < COUT * 16; e += 256) {
    const float s = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    const int co = e / 16, c = e % 16;
    part[(((size_t)split * 27 + tap) * COUT + co) * rows_x + cib * 16 + c] = s;
  }
}
__global__ __launch_bounds__(256) void k_dw_reduce(int cout, int cin, int rows_x, int S, const float *__restrict__ part,
                                                   float *__restrict__ dW, int accumulate) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= cout * cin * 27) return;
  const int t = e % 27, ci = (e / 27) % cin, co = e / (27 * cin);
  float s = 0.0f;
  for (int k = 0; k < S; k++) s += part[(((size_t)k * 27 + t) * cout + co) * rows_x + ci];
  dW[e] = accumulate ? dW[e] + s : s;
}
// bias gradient: db[c] = sum_v g[v][c]
constexpr int BG_BLOCKS = 256;
template <int C>
__global__ __launch_bounds__(256) void k_bias_grad(uint32_t n, const bf16 *__restrict__ g, float *__restrict__ part) {
  constexpr int VL = 256 / C;
  const int c = threadIdx.x % C, vl = threadIdx.x / C;
  float s = 0.0f;
  for (uint32_t v = blockIdx.x * VL + vl; v < n; v += BG_BLOCKS * VL) s += bf(g[(size_t)v * C + c]);
  __shared__ float sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (vl == 0) {
    float t = 0.0f;
    for (int k = 0; k < VL; k++) t += sh[k * C + c];
    part[blockIdx.x * C + c] = t;
  }
}
__global__ __launch_bounds__(64) void k_bias_reduce(int C, const float *__restrict__ part, float *__restrict__ db,
                                                    int accumulate) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float s = 0.0f;
  for (int b = 0; b < BG_BLOCKS; b++) s += part[b * C + c];
  db[c] = accumulate ? db[c] + s : s;
}
static void launch_bias_grad(uint32_t n, int c, const bf16 *g, float *part, float *db, bool acc, hipStream_t s) {
  if (c == 16) k_bias_grad<16><<<BG_BLOCKS, 256, 0, s>>>(n, g, part);
  else if (c == 32) k_bias_grad<32><<<BG_BLOCKS, 256, 0, s>>>(n, g, part);
  else k_bias_grad<64><<<BG_BLOCKS, 256, 0, s>>>(n, g, part);
  k_bias_reduce<<<1, 64, 0, s>>>(c, part, db, acc ? 1 : 0);
}

static void launch_dw_core(const G3 &g, int cout, int cin_real, const DwPlan &d, const bf16 *gt, const bf16 *xt,
                           float *part, float *dW, bool acc, hipStream_t s) {
  const unsigned nb = (unsigned)(27 * (d.rows_x / 16) * d.S);
  if (cout == 16) k_dw<16><<<nb, 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, d.rows_x, d.S, gt, xt, part);
  else if (cout == 32) k_dw<32><<<nb, 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, d.rows_x, d.S, gt, xt, part);
  else k_dw<64><<<nb, 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, d.rows_x, d.S, gt, xt, part);
  k_dw_reduce<<<nblk((uint64_t)cout * cin_real * 27), 256, 0, s>>>(cout, cin_real, d.rows_x, d.S, part, dW, acc ? 1 : 0);
}

void launch_conv_dw(const G3 &g, const ConvSpec &c, const bf16 *a, const G3 &ga, const bf16 *b, const bf16 *gpre,
                    bf16 *xt, bf16 *gt, float *part, float *dW, float *db, bool accumulate, hipStream_t s) {
  const DwPlan d = dw_plan(g, c);
  const size_t xplane = (size_t)(g.nz + 2) * (g.ny + 2) * d.P;
  k_to_xt<<<nblk(3 * (size_t)(c.cin() / 8) * xplane), 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, ga.nx, ga.ny, a, c.ca, b,
                                                                    c.cb, d.rows_x, xt);
  const size_t gplane = (size_t)g.nz * g.ny * d.P;
  k_to_gt<<<nblk((size_t)(c.cout / 8) * gplane), 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, c.cout, gpre, gt);
  launch_dw_core(g, c.cout, c.cin(), d, gt, xt, part, dW, accumulate, s);
  launch_bias_grad(g.n(), c.cout, gpre, part, db, accumulate, s);
}

void launch_conv_in_dw(const G3 &g, const bf16 *x0, const bf16 *gpre, bf16 *xt, bf16 *gt, float *part, float *dW,
                       float *db, bool accumulate, hipStream_t s) {
  const ConvSpec c{0, 2, 16};
  const DwPlan d = dw_plan(g, c);
  const size_t xplane = (size_t)(g.nz + 2) * (g.ny + 2) * d.P;
  k_to_xt_in<<<nblk(3 * xplane), 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, x0, xt);
  const size_t gplane = (size_t)g.nz * g.ny * d.P;
  k_to_gt<<<nblk(2 * gplane), 256, 0, s>>>(g.nx, g.ny, g.nz, d.P, 16, gpre, gt);
  launch_dw_core(g, 16, 2, d, gt, xt, part, dW, accumulate, s);
  launch_bias_grad(g.n(), 16, gpre, part, db, accumulate, s);
}

// ------------------------------------------------------------ solver transforms
__global__ __launch_bounds__(256) void k_net_input64(uint32_t n, const double *__restrict__ r,
                                                     const uint8_t *__restrict__ f, double inv, bf16 *__restrict__ x0) {
  const uint32_t v = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (v >= n) return;
  const bool fluid = f[v] != 1;
  x0[2 * (size_t)v] = tobf(fluid ? (float)(r[v] * inv) : 0.0f);
  x0[2 * (size_t)v + 1] = tobf(fluid ? 1.0f : 0.0f);
}
void launch_net_input64(const G3 &g, const double *r, const uint8_t *flags, double inv, bf16 *x0, hipStream_t s) {
  k_net_input64<<<nblk(g.n()), 256, 0, s>>>(g.n(), r, flags, inv, x0);
}
__global__ __launch_bounds__(256) void k_net_output64(uint32_t n, const float *__restrict__ y,
                                                      const uint8_t *__restrict__ f, double scale,
                                                      double *__restrict__ z) {
  const uint32_t v = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  if (v >= n) return;
  z[v] = f[v] != 1 ? (double)y[v] * scale : 0.0;
}
void launch_net_output64(const G3 &g, const float *y, const uint8_t *flags, double scale, double *z, hipStream_t s) {
  k_net_output64<<<nblk(g.n()), 256, 0, s>>>(g.n(), y, flags, scale, z);
}

// training target of a solved pair: t[v] = fluid ? phi[v] * factor : 0
__global__ __launch_bounds__(256) void k_net_target64(uint32_t n, const double *__restrict__ phi,
                                                      const uint8_t *__restrict__ f, double factor,
                                                      float *__restrict__ t) {
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  t[v] = f[v] != 1 ? (float)(phi[v] * factor) : 0.0f;
}
void launch_net_target64(const G3 &g, const double *phi, const uint8_t *flags, double factor, float *t,
                         hipStream_t s) {
  k_net_target64<<<nblk(g.n()), 256, 0, s>>>(g.n(), phi, flags, factor, t);
}
// host data: x0 (2 floats per voxel) -> bf16
__global__ __launch_bounds__(256) void k_f32_to_bf16(uint32_t n, const float *__restrict__ a, bf16 *__restrict__ b) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) b[i] = tobf(a[i]);
}
__global__ __launch_bounds__(256) void k_bf16_to_f32(uint32_t n, const bf16 *__restrict__ a, float *__restrict__ b) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) b[i] = bf(a[i]);
}
void launch_f32_to_bf16(uint32_t n, const float *a, bf16 *b, hipStream_t s) { k_f32_to_bf16<<<nblk(n), 256, 0, s>>>(n, a, b); }
void launch_bf16_to_f32(uint32_t n, const bf16 *a, float *b, hipStream_t s) { k_bf16_to_f32<<<nblk(n), 256, 0, s>>>(n, a, b); }

// ------------------------------------------------------------ loss, Adam
__global__ __launch_bounds__(256) void k_loss(uint32_t n, const float *__restrict__ y, const float *__restrict__ t,
                                              const bf16 *__restrict__ x0, double *__restrict__ part) {
  double e2 = 0.0, t2 = 0.0;
  for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v < n; v += LOSS_BLOCKS * 256u) {
    if (bf(x0[2 * (size_t)v + 1]) > 0.5f) {
      const double d = (double)y[v] - (double)t[v];
      e2 += d * d;
      t2 += (double)t[v] * (double)t[v];
    }
  }
  __shared__ double sh[256 / 64];
  const double a = block_sum<256>(e2, sh);
  const double b = block_sum<256>(t2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[LOSS_BLOCKS + blockIdx.x] = b;
  }
}
__global__ __launch_bounds__(256) void k_loss_final(double *__restrict__ part) {
  // part[2*NB], part[2*NB+1] = totals (fixed order)
  __shared__ double sh[256 / 64];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < LOSS_BLOCKS; i += 256) {
    a += part[i];
    b += part[LOSS_BLOCKS + i];
  }
  a = block_sum<256>(a, sh);
  b = block_sum<256>(b, sh);
  if (threadIdx.x == 0) {
    part[2 * LOSS_BLOCKS] = a;
    part[2 * LOSS_BLOCKS + 1] = b;
  }
}
__global__ __launch_bounds__(64) void k_loss_record(const double *__restrict__ part, float *__restrict__ rec) {
  if (threadIdx.x == 0) {
    const double d = part[2 * LOSS_BLOCKS + 1];
    rec[0] = (float)(d > 0.0 ? part[2 * LOSS_BLOCKS] / d : part[2 * LOSS_BLOCKS]);
  }
}
void launch_loss(const G3 &g, const float *y, const float *t, const bf16 *x0, double *part, float *rec,
                 hipStream_t s) {
  k_loss<<<LOSS_BLOCKS, 256, 0, s>>>(g.n(), y, t, x0, part);
  k_loss_final<<<1, 256, 0, s>>>(part);
  if (rec) k_loss_record<<<1, 64, 0, s>>>(part, rec);
}
__global__ __launch_bounds__(256) void k_loss_grad(uint32_t n, const float *__restrict__ y, const float *__restrict__ t,
                                                   const bf16 *__restrict__ x0, const double *__restrict__ den,
                                                   float *__restrict__ gy) {
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  const double d = den[0] > 0.0 ? den[0] : 1.0;
  gy[v] = bf(x0[2 * (size_t)v + 1]) > 0.5f ? (float)(2.0 * ((double)y[v] - (double)t[v]) / d) : 0.0f;
}
void launch_loss_grad(const G3 &g, const float *y, const float *t, const bf16 *x0, const double *den, float *gy,
                      hipStream_t s) {
  k_loss_grad<<<nblk(g.n()), 256, 0, s>>>(g.n(), y, t, x0, den, gy);
}

__global__ __launch_bounds__(256) void k_adam(uint32_t n, float *__restrict__ p, const float *__restrict__ gr,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2, float gscale) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i] * gscale;
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * (g * g);
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
void launch_adam(uint32_t n, float *p, const float *gr, float *m, float *v, float lr, float b1, float b2, float eps,
                 float c1, float c2, float gscale, hipStream_t s) {
  k_adam<<<nblk(n), 256, 0, s>>>(n, p, gr, m, v, lr, b1, b2, eps, c1, c2, gscale);
}

}  // namespace pyro
