// net.hpp -- the learned pressure preconditioner: a 3-level 3D U-Net in BF16
// on MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), with forward,
// backward and Adam kernels.
//
// Topology (D0 = grid, D1 = ceil(D0/2), D2 = ceil(D1/2); conv = 3x3x3,
// zero padding, stride 1; relu after every conv but the head):
//   x0  D0 x 2   (scaled residual, fluid mask)
//   a0 = conv(x0)            2 -> 16     s0 = conv(a0)  16 -> 16   (skip 0)
//   p1 = avgpool2(s0)                    a1 = conv(p1)  16 -> 32
//   s1 = conv(a1)           32 -> 32     (skip 1)       p2 = avgpool2(s1)
//   a2 = conv(p2)           32 -> 64     c2 = conv(a2)  64 -> 64
//   d1 = conv([up(c2), s1]) 96 -> 32     d0 = conv([up(d1), s0]) 48 -> 16
//   y  = head(d0)  1x1x1    16 -> 1  (linear),  z = y * mask
// Activations are channels-last ([voxel][channel], voxel = x + nx*(y + ny*z)).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pyro {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct G3 {
  int nx = 0, ny = 0, nz = 0;
  uint32_t n() const { return (uint32_t)nx * (uint32_t)ny * (uint32_t)nz; }
  bool operator==(const G3 &o) const { return nx == o.nx && ny == o.ny && nz == o.nz; }
};
inline G3 half_of(const G3 &g) { return G3{(g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2}; }

// One 3x3x3 conv layer of the U-Net.  Input channels [0, ca) come from a
// half-resolution tensor through nearest upsampling, [ca, ca + cb) from a
// same-resolution tensor.  Master weights are FP32 in [cout][cin][27] order
// (tap = kz*9 + ky*3 + kx, offset (kx-1, ky-1, kz-1)), bias [cout].
struct ConvSpec {
  int ca, cb, cout;
  int cin() const { return ca + cb; }
};

// Packed BF16 weights: K32 groups of [cout_eff][32].  A 32-channel chunk
// contributes 27 groups (one per tap), a 16-channel chunk 14 groups (two taps
// each, the 28th tap zero); the upsampled input is chunked by
// conv_a_chunk(cout_eff) channels.  `transposed` packs the adjoint conv
// (cout_eff = cin, taps mirrored) used for input gradients.
int conv_a_chunk(int cout);
int conv_groups(int ca, int cb, int cout);
void launch_pack_conv(const float *W, const ConvSpec &c, bool transposed, bf16 *wp, hipStream_t s);
void launch_pack_bf16(const float *w, uint32_t n, bf16 *out, hipStream_t s);

// out[v][0..cout) = act(bias + sum_taps W * in), in = [up(a), b]; false (and
// nothing launched) for a channel combination the network does not use
bool launch_conv3(const G3 &g, const bf16 *a, const G3 &ga, int ca, const bf16 *b, int cb, const bf16 *wp,
                  const float *bias, int cout, bool relu, bf16 *out, hipStream_t s);
// first layer (2 input channels): a0 = relu(b + W * x0) on MFMA, weights
// packed by launch_pack_conv_in into [2][16][32] BF16
void launch_pack_conv_in(const float *W, bf16 *wp, hipStream_t s);
void launch_conv_in(const G3 &g, const bf16 *x0, const bf16 *wp, const float *bias, bf16 *out, hipStream_t s);
// head: y[v] = bh + sum_c wh[c] * d0[v][c]
void launch_head(const G3 &g, const bf16 *d0, const float *wh, const float *bh, float *y, hipStream_t s);
void launch_pool(const G3 &g, const bf16 *in, int c, bf16 *out, hipStream_t s);

// weight gradients: partials of dW over row splits, then a fixed-order sum
struct DwPlan {
  int P;       // row pitch (round-up of nx to 32)
  int rows_x;  // cin rows of the transposed input
  int S;       // row splits
};
DwPlan dw_plan(const G3 &g, const ConvSpec &c);
size_t dw_xt_elems(const G3 &g, const ConvSpec &c);
size_t dw_gt_elems(const G3 &g, const ConvSpec &c);
size_t dw_part_elems(const G3 &g, const ConvSpec &c);
void launch_conv_dw(const G3 &g, const ConvSpec &c, const bf16 *a, const G3 &ga, const bf16 *b, const bf16 *gpre,
                    bf16 *xt, bf16 *gt, float *part, float *dW, float *db, bool accumulate, hipStream_t s);
// first-layer weight gradient (x0 has 2 channels)
void launch_conv_in_dw(const G3 &g, const bf16 *x0, const bf16 *gpre, bf16 *xt, bf16 *gt, float *part, float *dW,
                       float *db, bool accumulate, hipStream_t s);

// backward glue
// gd0[v][c] = bf16(gy[v] * wh[c]) where d0[v][c] > 0 else 0; dwh, dbh partial sums
void launch_head_bwd(const G3 &g, const bf16 *d0, const float *gy, const float *wh, bf16 *gd0, float *part,
                     float *dwh, float *dbh, bool accumulate, hipStream_t s);
// gcoarse[u][c] = bf16(sum over children v of u of gcat[v][c] (c < ca)) where act[u][c] > 0
void launch_unup_relu(const G3 &g, const bf16 *gcat, int ctot, int ca, const G3 &gc, const bf16 *act, bf16 *gcoarse,
                      hipStream_t s);
// gs[v][c] = bf16((gcat[v][off + c] + gp[parent][c] / cnt(parent)) ) where s[v][c] > 0 else 0
void launch_skip_grad(const G3 &g, const bf16 *gcat, int ctot, int off, const G3 &gc, const bf16 *gp, const bf16 *s,
                      int c, bf16 *gs, hipStream_t st);
// g[v][c] = act[v][c] > 0 ? g : 0   (in place)
void launch_relu_bwd(uint32_t n, const bf16 *act, bf16 *g, hipStream_t s);

// input / output transforms of the solver
// x0[v] = (bf16(r[v] * inv), fluid ? 1 : 0)
void launch_net_input64(const G3 &g, const double *r, const uint8_t *flags, double inv, bf16 *x0, hipStream_t s);
// z[v] = fluid ? y[v] * scale : 0
void launch_net_output64(const G3 &g, const float *y, const uint8_t *flags, double scale, double *z, hipStream_t s);
// t[v] = fluid ? phi[v] * factor : 0
void launch_net_target64(const G3 &g, const double *phi, const uint8_t *flags, double factor, float *t,
                         hipStream_t s);
void launch_f32_to_bf16(uint32_t n, const float *a, bf16 *b, hipStream_t s);
void launch_bf16_to_f32(uint32_t n, const bf16 *a, float *b, hipStream_t s);

// training: relative squared error on fluid cells
//   part[0..NB) = sum m (y - t)^2, part[NB..2NB) = sum m t^2 (FP64), totals in
//   part[2NB], part[2NB + 1]; rec (if not null) = their ratio as float
constexpr int LOSS_BLOCKS = 512;
void launch_loss(const G3 &g, const float *y, const float *t, const bf16 *x0, double *part, float *rec,
                 hipStream_t s);
// gy = 2 m (y - t) / den   (den = part[2NB + 1] of launch_loss, read on the device)
void launch_loss_grad(const G3 &g, const float *y, const float *t, const bf16 *x0, const double *den, float *gy,
                      hipStream_t s);
void launch_adam(uint32_t n, float *p, const float *gr, float *m, float *v, float lr, float b1, float b2, float eps,
                 float c1, float c2, float gscale, hipStream_t s);

// a single MFMA 16x16x32 on given fragments (lane layout check)
void launch_mfma_probe(const bf16 *a, const bf16 *b, float *d, hipStream_t s);

// internal entry points shared with the pressure solver (poisson.hip)
}  // namespace pyro
struct PyroNet;
namespace pyro {
// z = fluid ? N(r) : 0, scaled back to the units of the solution: the input is
// r / rms(r) (rms over nfluid fluid cells, rr = sum r^2), the output is
// multiplied by out_scale * h^2 * rms(r).  Returns 0 or a PYRO_E_* code.
int net_apply64(PyroNet *net, const G3 &g, const double *r, const uint8_t *flags, double rr, double nfluid, double h,
                double *z, hipStream_t s);
// store the solved pair (b, phi) as a training sample: input b / rms(b),
// target phi / (h^2 rms(b) out_scale); bb = sum b^2
int net_add_pair64(PyroNet *net, const G3 &g, const double *b, const double *phi, const uint8_t *flags, double bb,
                   double nfluid, double h, hipStream_t s);
const char *net_error();
}  // namespace pyro
