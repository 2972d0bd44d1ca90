// net.hip -- the learned pressure preconditioner (C ABI pyro_net_*): FP32
// master weights with Adam moments, BF16 packed copies for the MFMA kernels,
// per-grid workspaces, a device-resident training set of solved pressure
// pairs and the training loop.  No reference counterpart: the reference only
// runs fixed V-cycles (sim.rs:405); the network follows the U-Net in net.hpp.
//
// Flat parameter vector (FP32), in this order:
//   W0a[16][2][27]  b0a[16]   W0b[16][16][27] b0b[16]   W1a[32][16][27] b1a[32]
//   W1b[32][32][27] b1b[32]   W2a[64][32][27] b2a[64]   W2b[64][64][27] b2b[64]
//   W1c[32][96][27] b1c[32]   W0c[16][48][27] b0c[16]   wh[16] bh[1]
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "net.hpp"
#include "pyro_hip.h"

using namespace pyro;

namespace {
thread_local std::string g_nerr;
int nfail(int code, const std::string &m) {
  g_nerr = m;
  return code;
}
#define NCHK(x)                                                                                      \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return nfail(PYRO_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

// conv layers after the first, in forward order
enum { L0B = 0, L1A, L1B, L2A, L2B, L1C, L0C, NCONV };
const ConvSpec kSpec[NCONV] = {{0, 16, 16}, {0, 16, 32}, {0, 32, 32}, {0, 32, 64},
                               {0, 64, 64}, {64, 32, 32}, {32, 16, 16}};

struct Work {
  G3 g0, g1, g2;
  bool train = false;
  std::vector<void *> allocs;
  bf16 *x0 = nullptr, *a0, *s0, *p1, *a1, *s1, *p2, *a2, *c2, *d1, *d0;
  float *y = nullptr;
  // training
  float *gy = nullptr;
  bf16 *gd0, *gcat0, *gd1, *gcat1, *gc2, *ga2, *gp2, *gs1, *ga1, *gp1, *gs0, *ga0;
  bf16 *xt, *gt;
  float *part;
  double *lpart;
};

struct Sample {
  bf16 *x0;
  float *t;
};
}  // namespace

struct PyroNet {
  hipStream_t stream = nullptr;
  size_t np = 0;
  float *params = nullptr, *grads = nullptr, *am = nullptr, *av = nullptr;
  size_t w0a = 0, b0a = 0, wh = 0, bh = 0;
  size_t w[NCONV], b[NCONV];
  bf16 *w0a_bf = nullptr;
  bf16 *wp[NCONV], *wpt[NCONV];
  float out_scale = 0.0f;
  int adam_t = 0;
  Work inf, tr;
  G3 data_g;
  std::vector<Sample> data;
  float *losses = nullptr;
  int losses_cap = 0;
  std::vector<void *> allocs;
  double *dpart = nullptr, *dout = nullptr;
};

namespace {
template <class T>
int dalloc(std::vector<void *> &list, T **p, size_t count) {
  void *v = nullptr;
  const size_t bytes = count * sizeof(T) < 256 ? 256 : count * sizeof(T);
  if (hipMalloc(&v, bytes) != hipSuccess) return nfail(PYRO_E_OOM, "hipMalloc failed");
  list.push_back(v);
  if (hipMemset(v, 0, bytes) != hipSuccess) return nfail(PYRO_E_HIP, "hipMemset failed");
  *p = static_cast<T *>(v);
  return 0;
}
void free_list(std::vector<void *> &list) {
  for (void *v : list) (void)hipFree(v);
  list.clear();
}

int work_alloc(Work &W, const G3 &g, bool train) {
  if (!W.allocs.empty() && W.g0 == g && (W.train || !train)) return 0;
  free_list(W.allocs);
  W.g0 = g;
  W.g1 = half_of(g);
  W.g2 = half_of(W.g1);
  W.train = train;
  const size_t n0 = g.n(), n1 = W.g1.n(), n2 = W.g2.n();
  auto &L = W.allocs;
  int rc = 0;
  rc |= dalloc(L, &W.x0, n0 * 2);
  if (!rc) rc |= dalloc(L, &W.a0, n0 * 16);
  if (!rc) rc |= dalloc(L, &W.s0, n0 * 16);
  if (!rc) rc |= dalloc(L, &W.p1, n1 * 16);
  if (!rc) rc |= dalloc(L, &W.a1, n1 * 32);
  if (!rc) rc |= dalloc(L, &W.s1, n1 * 32);
  if (!rc) rc |= dalloc(L, &W.p2, n2 * 32);
  if (!rc) rc |= dalloc(L, &W.a2, n2 * 64);
  if (!rc) rc |= dalloc(L, &W.c2, n2 * 64);
  if (!rc) rc |= dalloc(L, &W.d1, n1 * 32);
  if (!rc) rc |= dalloc(L, &W.d0, n0 * 16);
  if (!rc) rc |= dalloc(L, &W.y, n0);
  if (!rc && train) {
    rc |= dalloc(L, &W.gy, n0);
    if (!rc) rc |= dalloc(L, &W.gd0, n0 * 16);
    if (!rc) rc |= dalloc(L, &W.gcat0, n0 * 48);
    if (!rc) rc |= dalloc(L, &W.gd1, n1 * 32);
    if (!rc) rc |= dalloc(L, &W.gcat1, n1 * 96);
    if (!rc) rc |= dalloc(L, &W.gc2, n2 * 64);
    if (!rc) rc |= dalloc(L, &W.ga2, n2 * 64);
    if (!rc) rc |= dalloc(L, &W.gp2, n2 * 32);
    if (!rc) rc |= dalloc(L, &W.gs1, n1 * 32);
    if (!rc) rc |= dalloc(L, &W.ga1, n1 * 32);
    if (!rc) rc |= dalloc(L, &W.gp1, n1 * 16);
    if (!rc) rc |= dalloc(L, &W.gs0, n0 * 16);
    if (!rc) rc |= dalloc(L, &W.ga0, n0 * 16);
    // weight-gradient scratch: the largest layer at each level
    const G3 lg[NCONV] = {W.g0, W.g1, W.g1, W.g2, W.g2, W.g1, W.g0};
    size_t xt = dw_xt_elems(W.g0, ConvSpec{0, 2, 16}), gt = dw_gt_elems(W.g0, ConvSpec{0, 2, 16});
    size_t part = dw_part_elems(W.g0, ConvSpec{0, 2, 16});
    for (int l = 0; l < NCONV; l++) {
      xt = std::max(xt, dw_xt_elems(lg[l], kSpec[l]));
      gt = std::max(gt, dw_gt_elems(lg[l], kSpec[l]));
      part = std::max(part, dw_part_elems(lg[l], kSpec[l]));
    }
    part = std::max(part, (size_t)512 * 17);
    if (!rc) rc |= dalloc(L, &W.xt, xt);
    if (!rc) rc |= dalloc(L, &W.gt, gt);
    if (!rc) rc |= dalloc(L, &W.part, part);
    if (!rc) rc |= dalloc(L, &W.lpart, 2 * LOSS_BLOCKS + 2);
  }
  if (rc) {
    free_list(W.allocs);
    return rc;
  }
  return 0;
}

void repack(PyroNet *N) {
  launch_pack_conv_in(N->params + N->w0a, N->w0a_bf, N->stream);
  for (int l = 0; l < NCONV; l++) {
    launch_pack_conv(N->params + N->w[l], kSpec[l], false, N->wp[l], N->stream);
    launch_pack_conv(N->params + N->w[l], kSpec[l], true, N->wpt[l], N->stream);
  }
}

void forward(PyroNet *N, Work &W, const bf16 *x0) {
  hipStream_t s = N->stream;
  const float *P = N->params;
  launch_conv_in(W.g0, x0, N->w0a_bf, P + N->b0a, W.a0, s);
  launch_conv3(W.g0, nullptr, W.g1, 0, W.a0, 16, N->wp[L0B], P + N->b[L0B], 16, true, W.s0, s);
  launch_pool(W.g0, W.s0, 16, W.p1, s);
  launch_conv3(W.g1, nullptr, W.g2, 0, W.p1, 16, N->wp[L1A], P + N->b[L1A], 32, true, W.a1, s);
  launch_conv3(W.g1, nullptr, W.g2, 0, W.a1, 32, N->wp[L1B], P + N->b[L1B], 32, true, W.s1, s);
  launch_pool(W.g1, W.s1, 32, W.p2, s);
  launch_conv3(W.g2, nullptr, W.g2, 0, W.p2, 32, N->wp[L2A], P + N->b[L2A], 64, true, W.a2, s);
  launch_conv3(W.g2, nullptr, W.g2, 0, W.a2, 64, N->wp[L2B], P + N->b[L2B], 64, true, W.c2, s);
  launch_conv3(W.g1, W.c2, W.g2, 64, W.s1, 32, N->wp[L1C], P + N->b[L1C], 32, true, W.d1, s);
  launch_conv3(W.g0, W.d1, W.g1, 32, W.s0, 16, N->wp[L0C], P + N->b[L0C], 16, true, W.d0, s);
  launch_head(W.g0, W.d0, P + N->wh, P + N->bh, W.y, s);
}

// gradients of all parameters for the upstream gradient W.gy (overwritten)
void backward(PyroNet *N, Work &W, const bf16 *x0) {
  hipStream_t s = N->stream;
  const float *P = N->params;
  float *G = N->grads;
  const G3 &g0 = W.g0, &g1 = W.g1, &g2 = W.g2;
  launch_head_bwd(g0, W.d0, W.gy, P + N->wh, W.gd0, W.part, G + N->wh, G + N->bh, false, s);
  // d0 = conv([up(d1), s0])
  launch_conv_dw(g0, kSpec[L0C], W.d1, g1, W.s0, W.gd0, W.xt, W.gt, W.part, G + N->w[L0C], G + N->b[L0C], false, s);
  launch_conv3(g0, nullptr, g1, 0, W.gd0, 16, N->wpt[L0C], nullptr, 48, false, W.gcat0, s);
  launch_unup_relu(g0, W.gcat0, 48, 32, g1, W.d1, W.gd1, s);
  // d1 = conv([up(c2), s1])
  launch_conv_dw(g1, kSpec[L1C], W.c2, g2, W.s1, W.gd1, W.xt, W.gt, W.part, G + N->w[L1C], G + N->b[L1C], false, s);
  launch_conv3(g1, nullptr, g2, 0, W.gd1, 32, N->wpt[L1C], nullptr, 96, false, W.gcat1, s);
  launch_unup_relu(g1, W.gcat1, 96, 64, g2, W.c2, W.gc2, s);
  // c2 = conv(a2)
  launch_conv_dw(g2, kSpec[L2B], nullptr, g2, W.a2, W.gc2, W.xt, W.gt, W.part, G + N->w[L2B], G + N->b[L2B], false, s);
  launch_conv3(g2, nullptr, g2, 0, W.gc2, 64, N->wpt[L2B], nullptr, 64, false, W.ga2, s);
  launch_relu_bwd(g2.n() * 64, W.a2, W.ga2, s);
  // a2 = conv(p2)
  launch_conv_dw(g2, kSpec[L2A], nullptr, g2, W.p2, W.ga2, W.xt, W.gt, W.part, G + N->w[L2A], G + N->b[L2A], false, s);
  launch_conv3(g2, nullptr, g2, 0, W.ga2, 64, N->wpt[L2A], nullptr, 32, false, W.gp2, s);
  // s1: decoder concat + pooling
  launch_skip_grad(g1, W.gcat1, 96, 64, g2, W.gp2, W.s1, 32, W.gs1, s);
  launch_conv_dw(g1, kSpec[L1B], nullptr, g1, W.a1, W.gs1, W.xt, W.gt, W.part, G + N->w[L1B], G + N->b[L1B], false, s);
  launch_conv3(g1, nullptr, g1, 0, W.gs1, 32, N->wpt[L1B], nullptr, 32, false, W.ga1, s);
  launch_relu_bwd(g1.n() * 32, W.a1, W.ga1, s);
  launch_conv_dw(g1, kSpec[L1A], nullptr, g1, W.p1, W.ga1, W.xt, W.gt, W.part, G + N->w[L1A], G + N->b[L1A], false, s);
  launch_conv3(g1, nullptr, g1, 0, W.ga1, 32, N->wpt[L1A], nullptr, 16, false, W.gp1, s);
  // s0
  launch_skip_grad(g0, W.gcat0, 48, 32, g1, W.gp1, W.s0, 16, W.gs0, s);
  launch_conv_dw(g0, kSpec[L0B], nullptr, g0, W.a0, W.gs0, W.xt, W.gt, W.part, G + N->w[L0B], G + N->b[L0B], false, s);
  launch_conv3(g0, nullptr, g0, 0, W.gs0, 16, N->wpt[L0B], nullptr, 16, false, W.ga0, s);
  launch_relu_bwd(g0.n() * 16, W.a0, W.ga0, s);
  launch_conv_in_dw(g0, x0, W.ga0, W.xt, W.gt, W.part, G + N->w0a, G + N->b0a, false, s);
}

int nsync(PyroNet *N) {
  NCHK(hipStreamSynchronize(N->stream));
  NCHK(hipGetLastError());
  return 0;
}

int check_dims(uint32_t nx, uint32_t ny, uint32_t nz) {
  if (nx < 1 || ny < 1 || nz < 1) return nfail(PYRO_E_INVALID, "empty grid");
  if ((uint64_t)nx * ny * nz * 96ull >= (1ull << 40)) return nfail(PYRO_E_INVALID, "grid too large");
  return 0;
}

// deterministic normal deviates (splitmix64 + Box-Muller)
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};
}  // namespace

namespace pyro {
const char *net_error() { return g_nerr.c_str(); }

int net_apply64(PyroNet *N, const G3 &g, const double *r, const uint8_t *flags, double rr, double nfluid, double h,
                double *z, hipStream_t s) {
  int rc = work_alloc(N->inf, g, false);
  if (rc) return rc;
  const double rms = nfluid > 0.0 ? std::sqrt(rr / nfluid) : 0.0;
  if (!(rms > 0.0) || !std::isfinite(rms)) return nfail(PYRO_E_INVALID, "zero or non-finite residual");
  hipStream_t keep = N->stream;
  N->stream = s;
  launch_net_input64(g, r, flags, 1.0 / rms, N->inf.x0, s);
  forward(N, N->inf, N->inf.x0);
  launch_net_output64(g, N->inf.y, flags, (double)N->out_scale * (double)h * (double)h * rms, z, s);
  N->stream = keep;
  return 0;
}

int net_add_pair64(PyroNet *N, const G3 &g, const double *b, const double *phi, const uint8_t *flags, double bb,
                   double nfluid, double h, hipStream_t s) {
  if (!N->data.empty() && !(N->data_g == g)) return nfail(PYRO_E_INVALID, "training pairs must share one grid");
  const double rms = nfluid > 0.0 ? std::sqrt(bb / nfluid) : 0.0;
  if (!(rms > 0.0) || !std::isfinite(rms)) return nfail(PYRO_E_INVALID, "zero right-hand side");
  const double hh = (double)h * (double)h;
  if (N->out_scale == 0.0f) {
    // first pair fixes the output scale: rms of phi / (h^2 rms(b)) over fluid cells
    const double *a[1] = {phi}, *bq[1] = {phi};
    launch_dots(a, bq, 1, g.n(), N->dpart, N->dout, s);
    double pp = 0.0;
    NCHK(hipMemcpyAsync(&pp, N->dout, sizeof(double), hipMemcpyDeviceToHost, s));
    NCHK(hipStreamSynchronize(s));
    const double sc = std::sqrt(pp / nfluid) / (hh * rms);
    N->out_scale = sc > 0.0 && std::isfinite(sc) ? (float)std::exp2(std::round(std::log2(sc))) : 1.0f;
  }
  Sample S;
  std::vector<void *> &L = N->allocs;
  int rc = dalloc(L, &S.x0, (size_t)g.n() * 2);
  if (!rc) rc = dalloc(L, &S.t, (size_t)g.n());
  if (rc) return rc;
  launch_net_input64(g, b, flags, 1.0 / rms, S.x0, s);
  launch_net_target64(g, phi, flags, 1.0 / (hh * rms * (double)N->out_scale), S.t, s);
  N->data.push_back(S);
  N->data_g = g;
  return 0;
}
}  // namespace pyro

extern "C" {

const char *pyro_net_last_error(void) { return g_nerr.c_str(); }

size_t pyro_net_param_count(void) {
  size_t n = 16 * 2 * 27 + 16 + 16 + 1;
  for (int l = 0; l < NCONV; l++) n += (size_t)kSpec[l].cout * kSpec[l].cin() * 27 + kSpec[l].cout;
  return n;
}

int pyro_net_create(uint64_t seed, PyroNet **out) {
  if (!out) return nfail(PYRO_E_INVALID, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return nfail(PYRO_E_NODEVICE, "no HIP device available");
  PyroNet *N = new PyroNet();
  auto bail = [&](int rc) {
    pyro_net_destroy(N);
    return rc;
  };
  if (hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(nfail(PYRO_E_HIP, "hipStreamCreate failed"));
  // layout of the flat parameter vector (see the file header)
  size_t o = 0;
  N->w0a = o; o += 16 * 2 * 27;
  N->b0a = o; o += 16;
  for (int l = 0; l < NCONV; l++) {
    N->w[l] = o; o += (size_t)kSpec[l].cout * kSpec[l].cin() * 27;
    N->b[l] = o; o += kSpec[l].cout;
  }
  N->wh = o; o += 16;
  N->bh = o; o += 1;
  N->np = o;
  int rc = 0;
  rc |= dalloc(N->allocs, &N->params, o);
  if (!rc) rc |= dalloc(N->allocs, &N->grads, o);
  if (!rc) rc |= dalloc(N->allocs, &N->am, o);
  if (!rc) rc |= dalloc(N->allocs, &N->av, o);
  if (!rc) rc |= dalloc(N->allocs, &N->w0a_bf, 2 * 16 * 32);
  for (int l = 0; l < NCONV && !rc; l++) {
    const ConvSpec &c = kSpec[l];
    rc |= dalloc(N->allocs, &N->wp[l], (size_t)conv_groups(c.ca, c.cb, c.cout) * c.cout * 32);
    if (!rc) rc |= dalloc(N->allocs, &N->wpt[l], (size_t)conv_groups(0, c.cout, c.cin()) * c.cin() * 32);
  }
  if (!rc) rc |= dalloc(N->allocs, &N->dpart, DOT_PART);
  if (!rc) rc |= dalloc(N->allocs, &N->dout, 4);
  if (rc) return bail(rc);
  // He-normal weights (std sqrt(2 / fan_in)), zero biases, small head
  std::vector<float> h(o, 0.0f);
  Rng R{seed};
  for (size_t i = 0; i < 16 * 2 * 27; i++) h[N->w0a + i] = (float)(R.normal() * std::sqrt(2.0 / 54.0));
  for (int l = 0; l < NCONV; l++) {
    const size_t nw = (size_t)kSpec[l].cout * kSpec[l].cin() * 27;
    const double sd = std::sqrt(2.0 / (27.0 * kSpec[l].cin()));
    for (size_t i = 0; i < nw; i++) h[N->w[l] + i] = (float)(R.normal() * sd);
  }
  for (int i = 0; i < 16; i++) h[N->wh + i] = (float)(R.normal() * 0.25);
  if (hipMemcpy(N->params, h.data(), o * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return bail(nfail(PYRO_E_HIP, "upload failed"));
  repack(N);
  if (nsync(N)) return bail(PYRO_E_HIP);
  *out = N;
  return 0;
}

int pyro_net_destroy(PyroNet *N) {
  if (!N) return 0;
  if (N->stream) (void)hipStreamSynchronize(N->stream);
  free_list(N->inf.allocs);
  free_list(N->tr.allocs);
  free_list(N->allocs);
  if (N->stream) (void)hipStreamDestroy(N->stream);
  delete N;
  return 0;
}

int pyro_net_get_params(PyroNet *N, float *dst, size_t n) {
  if (!N || !dst || n != N->np) return nfail(PYRO_E_INVALID, "bad argument (n must be pyro_net_param_count())");
  int rc = nsync(N);
  if (rc) return rc;
  NCHK(hipMemcpy(dst, N->params, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pyro_net_set_params(PyroNet *N, const float *src, size_t n) {
  if (!N || !src || n != N->np) return nfail(PYRO_E_INVALID, "bad argument (n must be pyro_net_param_count())");
  int rc = nsync(N);
  if (rc) return rc;
  NCHK(hipMemcpy(N->params, src, n * sizeof(float), hipMemcpyHostToDevice));
  repack(N);
  return nsync(N);
}

int pyro_net_get_grads(PyroNet *N, float *dst, size_t n) {
  if (!N || !dst || n != N->np) return nfail(PYRO_E_INVALID, "bad argument (n must be pyro_net_param_count())");
  int rc = nsync(N);
  if (rc) return rc;
  NCHK(hipMemcpy(dst, N->grads, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pyro_net_get_output_scale(PyroNet *N, float *s) {
  if (!N || !s) return nfail(PYRO_E_INVALID, "null argument");
  *s = N->out_scale;
  return 0;
}

int pyro_net_set_output_scale(PyroNet *N, float s) {
  if (!N || !(s > 0.0f) || !std::isfinite(s)) return nfail(PYRO_E_INVALID, "bad argument");
  if (!N->data.empty()) return nfail(PYRO_E_STATE, "the training set already fixes the output scale");
  N->out_scale = s;
  return 0;
}

int pyro_net_forward(PyroNet *N, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, float *y) {
  if (!N || !x0 || !y) return nfail(PYRO_E_INVALID, "null argument");
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const G3 g{(int)nx, (int)ny, (int)nz};
  rc = work_alloc(N->inf, g, false);
  if (rc) return rc;
  float *tmp = N->inf.y;  // host x0 staged through y (n floats) twice
  const size_t n = g.n();
  for (int half = 0; half < 2; half++) {
    NCHK(hipMemcpyAsync(tmp, x0 + half * n, n * sizeof(float), hipMemcpyHostToDevice, N->stream));
    launch_f32_to_bf16((uint32_t)n, tmp, N->inf.x0 + half * n, N->stream);
  }
  forward(N, N->inf, N->inf.x0);
  NCHK(hipMemcpyAsync(y, N->inf.y, n * sizeof(float), hipMemcpyDeviceToHost, N->stream));
  return nsync(N);
}

int pyro_net_backward(PyroNet *N, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, const float *gy) {
  if (!N || !x0 || !gy) return nfail(PYRO_E_INVALID, "null argument");
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const G3 g{(int)nx, (int)ny, (int)nz};
  rc = work_alloc(N->tr, g, true);
  if (rc) return rc;
  const size_t n = g.n();
  for (int half = 0; half < 2; half++) {
    NCHK(hipMemcpyAsync(N->tr.gy, x0 + half * n, n * sizeof(float), hipMemcpyHostToDevice, N->stream));
    launch_f32_to_bf16((uint32_t)n, N->tr.gy, N->tr.x0 + half * n, N->stream);
  }
  forward(N, N->tr, N->tr.x0);
  NCHK(hipMemcpyAsync(N->tr.gy, gy, n * sizeof(float), hipMemcpyHostToDevice, N->stream));
  backward(N, N->tr, N->tr.x0);
  return nsync(N);
}

int pyro_net_dataset_clear(PyroNet *N) {
  if (!N) return nfail(PYRO_E_INVALID, "null handle");
  int rc = nsync(N);
  if (rc) return rc;
  for (Sample &S : N->data) {
    for (void *p : {(void *)S.x0, (void *)S.t}) {
      for (size_t i = 0; i < N->allocs.size(); i++)
        if (N->allocs[i] == p) {
          (void)hipFree(p);
          N->allocs.erase(N->allocs.begin() + i);
          break;
        }
    }
  }
  N->data.clear();
  return 0;
}

int pyro_net_dataset_size(PyroNet *N, int *count) {
  if (!N || !count) return nfail(PYRO_E_INVALID, "null argument");
  *count = (int)N->data.size();
  return 0;
}

int pyro_net_dataset_add(PyroNet *N, uint32_t nx, uint32_t ny, uint32_t nz, const float *x0, const float *target) {
  if (!N || !x0 || !target) return nfail(PYRO_E_INVALID, "null argument");
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const G3 g{(int)nx, (int)ny, (int)nz};
  if (!N->data.empty() && !(N->data_g == g)) return nfail(PYRO_E_INVALID, "training pairs must share one grid");
  const size_t n = g.n();
  Sample S;
  rc = dalloc(N->allocs, &S.x0, n * 2);
  if (!rc) rc = dalloc(N->allocs, &S.t, n);
  if (rc) return rc;
  float *tmp = nullptr;
  if (hipMalloc(&tmp, n * sizeof(float)) != hipSuccess) return nfail(PYRO_E_OOM, "hipMalloc failed");
  for (int half = 0; half < 2; half++) {
    if (hipMemcpyAsync(tmp, x0 + half * n, n * sizeof(float), hipMemcpyHostToDevice, N->stream) != hipSuccess) break;
    launch_f32_to_bf16((uint32_t)n, tmp, S.x0 + half * n, N->stream);
  }
  (void)hipStreamSynchronize(N->stream);
  (void)hipFree(tmp);
  NCHK(hipMemcpy(S.t, target, n * sizeof(float), hipMemcpyHostToDevice));
  N->data.push_back(S);
  N->data_g = g;
  return nsync(N);
}

int pyro_net_dataset_get(PyroNet *N, int index, float *x0, float *target) {
  if (!N || index < 0 || index >= (int)N->data.size()) return nfail(PYRO_E_INVALID, "bad index");
  const size_t n = N->data_g.n();
  int rc = nsync(N);
  if (rc) return rc;
  if (x0) {
    float *tmp = nullptr;
    if (hipMalloc(&tmp, 2 * n * sizeof(float)) != hipSuccess) return nfail(PYRO_E_OOM, "hipMalloc failed");
    launch_bf16_to_f32((uint32_t)(2 * n), N->data[index].x0, tmp, N->stream);
    (void)hipStreamSynchronize(N->stream);
    hipError_t e = hipMemcpy(x0, tmp, 2 * n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return nfail(PYRO_E_HIP, "download failed");
  }
  if (target) NCHK(hipMemcpy(target, N->data[index].t, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// Adam on the relative squared error of random training pairs from
// [first, first + count), one pair per step; losses[k] = the loss of step k
// before its update.
int pyro_net_train(PyroNet *N, int first, int count, int steps, float lr, uint64_t seed, float *losses) {
  if (!N || steps < 0 || !(lr > 0.0f)) return nfail(PYRO_E_INVALID, "bad argument");
  if (N->data.empty()) return nfail(PYRO_E_STATE, "empty training set");
  if (first < 0 || count < 1 || first + count > (int)N->data.size()) return nfail(PYRO_E_INVALID, "bad pair range");
  int rc = work_alloc(N->tr, N->data_g, true);
  if (rc) return rc;
  if (N->losses_cap < steps) {
    if (N->losses) {
      for (size_t i = 0; i < N->allocs.size(); i++)
        if (N->allocs[i] == N->losses) {
          (void)hipFree(N->losses);
          N->allocs.erase(N->allocs.begin() + i);
          break;
        }
    }
    rc = dalloc(N->allocs, &N->losses, (size_t)steps);
    if (rc) return rc;
    N->losses_cap = steps;
  }
  Rng R{seed};
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  Work &W = N->tr;
  for (int k = 0; k < steps; k++) {
    const Sample &S = N->data[first + (int)(R.next() % (uint64_t)count)];
    forward(N, W, S.x0);
    launch_loss(W.g0, W.y, S.t, S.x0, W.lpart, N->losses + k, N->stream);
    launch_loss_grad(W.g0, W.y, S.t, S.x0, W.lpart + 2 * LOSS_BLOCKS + 1, W.gy, N->stream);
    backward(N, W, S.x0);
    N->adam_t++;
    const float c1 = 1.0f - std::pow(b1, (float)N->adam_t), c2 = 1.0f - std::pow(b2, (float)N->adam_t);
    launch_adam((uint32_t)N->np, N->params, N->grads, N->am, N->av, lr, b1, b2, eps, c1, c2, 1.0f, N->stream);
    repack(N);
  }
  rc = nsync(N);
  if (rc) return rc;
  if (losses && steps > 0) NCHK(hipMemcpy(losses, N->losses, steps * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// mean loss over the training set (or over pairs [first, first + count))
int pyro_net_eval(PyroNet *N, int first, int count, double *mean_loss) {
  if (!N || !mean_loss || first < 0 || count < 1 || first + count > (int)N->data.size())
    return nfail(PYRO_E_INVALID, "bad argument");
  int rc = work_alloc(N->tr, N->data_g, true);
  if (rc) return rc;
  double sum = 0.0;
  for (int k = first; k < first + count; k++) {
    forward(N, N->tr, N->data[k].x0);
    launch_loss(N->tr.g0, N->tr.y, N->data[k].t, N->data[k].x0, N->tr.lpart, nullptr, N->stream);
    double tot[2];
    NCHK(hipMemcpyAsync(tot, N->tr.lpart + 2 * LOSS_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, N->stream));
    rc = nsync(N);
    if (rc) return rc;
    sum += tot[1] > 0.0 ? tot[0] / tot[1] : tot[0];
  }
  *mean_loss = sum / count;
  return 0;
}

// Average device time of `reps` forward passes on an nx x ny x nz grid (the
// inference workspace's current input; zero after allocation).
int pyro_net_time_forward(PyroNet *N, uint32_t nx, uint32_t ny, uint32_t nz, int reps, float *avg_ms) {
  if (!N || !avg_ms || reps < 1) return nfail(PYRO_E_INVALID, "bad argument");
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const G3 g{(int)nx, (int)ny, (int)nz};
  rc = work_alloc(N->inf, g, false);
  if (rc) return rc;
  forward(N, N->inf, N->inf.x0);
  hipEvent_t a, b;
  NCHK(hipEventCreate(&a));
  NCHK(hipEventCreate(&b));
  NCHK(hipEventRecord(a, N->stream));
  for (int k = 0; k < reps; k++) forward(N, N->inf, N->inf.x0);
  NCHK(hipEventRecord(b, N->stream));
  NCHK(hipEventSynchronize(b));
  float ms = 0.0f;
  NCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *avg_ms = ms / reps;
  return nsync(N);
}

int pyro_debug_mfma(const float *a, const float *b, float *d) {
  if (!a || !b || !d) return nfail(PYRO_E_INVALID, "null argument");
  std::vector<void *> L;
  bf16 *da, *db;
  float *dd, *tmp;
  int rc = dalloc(L, &da, 512);
  if (!rc) rc = dalloc(L, &db, 512);
  if (!rc) rc = dalloc(L, &dd, 256);
  if (!rc) rc = dalloc(L, &tmp, 512);
  if (!rc) {
    if (hipMemcpy(tmp, a, 512 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = PYRO_E_HIP;
    launch_f32_to_bf16(512, tmp, da, nullptr);
    (void)hipDeviceSynchronize();
    if (hipMemcpy(tmp, b, 512 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = PYRO_E_HIP;
    launch_f32_to_bf16(512, tmp, db, nullptr);
    launch_mfma_probe(da, db, dd, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) rc = PYRO_E_HIP;
    if (!rc && hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = PYRO_E_HIP;
  }
  free_list(L);
  return rc ? nfail(rc, "mfma probe failed") : 0;
}

// One conv of the network's kinds on host data (FP32 in, rounded to BF16):
// forward (out n x cout) or, with transposed = 1, its adjoint applied to
// b = an n x cout gradient (out n x cin; a unused).  a: half-grid x ca.
int pyro_debug_conv(int ca, int cb, int cout, int transposed, int relu, uint32_t nx, uint32_t ny, uint32_t nz,
                    const float *a, const float *b, const float *W, const float *bias, float *out) {
  const ConvSpec c{ca, cb, cout};
  const G3 g{(int)nx, (int)ny, (int)nz}, gh = half_of(g);
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  const bool first = ca == 0 && cb == 2 && cout == 16 && !transposed && relu && bias;
  bool known = first;
  for (int l = 0; l < NCONV; l++)
    if (kSpec[l].ca == ca && kSpec[l].cb == cb && kSpec[l].cout == cout) known = true;
  if (!known || !b || !W || !out || (ca && !a && !transposed)) return nfail(PYRO_E_INVALID, "unsupported conv");
  const int in_b = transposed ? cout : cb, out_c = transposed ? c.cin() : cout;
  std::vector<void *> L;
  bf16 *da = nullptr, *dbb, *wp, *dout;
  float *tmp, *dW, *dbias = nullptr;
  const size_t n = g.n(), nh = gh.n();
  const size_t nwp = first ? 1024 : (size_t)(transposed ? conv_groups(0, cout, c.cin()) * c.cin() : conv_groups(ca, cb, cout) * cout) * 32;
  const size_t big = std::max(std::max(n * in_b, nh * (size_t)(ca ? ca : 1)), n * out_c);
  rc = dalloc(L, &dbb, n * in_b);
  if (!rc && ca) rc = dalloc(L, &da, nh * ca);
  if (!rc) rc = dalloc(L, &wp, nwp);
  if (!rc) rc = dalloc(L, &dout, n * out_c);
  if (!rc) rc = dalloc(L, &tmp, big);
  if (!rc) rc = dalloc(L, &dW, (size_t)cout * c.cin() * 27);
  if (!rc && bias && !transposed) rc = dalloc(L, &dbias, cout);
  auto up = [&](const float *h, size_t cnt, bf16 *dst) {
    if (hipMemcpy(tmp, h, cnt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return PYRO_E_HIP;
    launch_f32_to_bf16((uint32_t)cnt, tmp, dst, nullptr);
    return hipDeviceSynchronize() == hipSuccess ? 0 : PYRO_E_HIP;
  };
  if (!rc) rc = up(b, n * in_b, dbb);
  if (!rc && ca && !transposed) rc = up(a, nh * ca, da);
  if (!rc && hipMemcpy(dW, W, (size_t)cout * c.cin() * 27 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    rc = PYRO_E_HIP;
  if (!rc && dbias && hipMemcpy(dbias, bias, cout * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = PYRO_E_HIP;
  if (!rc) {
    bool ok = true;
    if (first) {
      launch_pack_conv_in(dW, wp, nullptr);
      launch_conv_in(g, dbb, wp, dbias, dout, nullptr);
    } else {
      launch_pack_conv(dW, c, transposed != 0, wp, nullptr);
      ok = transposed ? launch_conv3(g, nullptr, gh, 0, dbb, cout, wp, nullptr, c.cin(), false, dout, nullptr)
                      : launch_conv3(g, da, gh, ca, dbb, cb, wp, dbias, cout, relu != 0, dout, nullptr);
    }
    if (!ok) rc = PYRO_E_INVALID;
    launch_bf16_to_f32((uint32_t)(n * out_c), dout, tmp, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) rc = PYRO_E_HIP;
    if (!rc && hipMemcpy(out, tmp, n * out_c * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = PYRO_E_HIP;
  }
  free_list(L);
  return rc ? nfail(rc, "debug conv failed") : 0;
}

// Weight and bias gradients of one conv: dW[cout][cin][27], db[cout] for input
// [up(a), b] and output gradient g (n x cout).  ca = 0, cb = 2 is the first layer.
int pyro_debug_conv_dw(int ca, int cb, int cout, uint32_t nx, uint32_t ny, uint32_t nz, const float *a, const float *b,
                       const float *g, float *dW, float *db) {
  const ConvSpec c{ca, cb, cout};
  const G3 gg{(int)nx, (int)ny, (int)nz}, gh = half_of(gg);
  int rc = check_dims(nx, ny, nz);
  if (rc) return rc;
  bool known = ca == 0 && cb == 2 && cout == 16;
  for (int l = 0; l < NCONV; l++)
    if (kSpec[l].ca == ca && kSpec[l].cb == cb && kSpec[l].cout == cout) known = true;
  if (!known || !b || !g || !dW || !db || (ca && !a)) return nfail(PYRO_E_INVALID, "unsupported conv");
  std::vector<void *> L;
  const size_t n = gg.n(), nh = gh.n();
  bf16 *da = nullptr, *dbb, *dg, *xt, *gt;
  float *tmp, *ddw, *ddb, *part;
  const size_t big = std::max(std::max(n * cb, n * cout), nh * (size_t)(ca ? ca : 1));
  rc = dalloc(L, &dbb, n * cb);
  if (!rc && ca) rc = dalloc(L, &da, nh * ca);
  if (!rc) rc = dalloc(L, &dg, n * cout);
  if (!rc) rc = dalloc(L, &xt, dw_xt_elems(gg, c));
  if (!rc) rc = dalloc(L, &gt, dw_gt_elems(gg, c));
  if (!rc) rc = dalloc(L, &part, dw_part_elems(gg, c));
  if (!rc) rc = dalloc(L, &tmp, big);
  if (!rc) rc = dalloc(L, &ddw, (size_t)cout * c.cin() * 27);
  if (!rc) rc = dalloc(L, &ddb, cout);
  auto up = [&](const float *h, size_t cnt, bf16 *dst) {
    if (hipMemcpy(tmp, h, cnt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return PYRO_E_HIP;
    launch_f32_to_bf16((uint32_t)cnt, tmp, dst, nullptr);
    return hipDeviceSynchronize() == hipSuccess ? 0 : PYRO_E_HIP;
  };
  if (!rc) rc = up(b, n * cb, dbb);
  if (!rc && ca) rc = up(a, nh * ca, da);
  if (!rc) rc = up(g, n * cout, dg);
  if (!rc) {
    if (cb == 2)
      launch_conv_in_dw(gg, dbb, dg, xt, gt, part, ddw, ddb, false, nullptr);
    else
      launch_conv_dw(gg, c, da, gh, dbb, dg, xt, gt, part, ddw, ddb, false, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) rc = PYRO_E_HIP;
    if (!rc && hipMemcpy(dW, ddw, (size_t)cout * c.cin() * 27 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      rc = PYRO_E_HIP;
    if (!rc && hipMemcpy(db, ddb, cout * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = PYRO_E_HIP;
  }
  free_list(L);
  return rc ? nfail(rc, "debug conv dw failed") : 0;
}

}  // extern "C"
