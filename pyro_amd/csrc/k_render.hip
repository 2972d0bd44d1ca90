// k_render.hip -- the renderer's pack pass (render.wgsl:16-37) on gfx950.
//
// The reference renderer starts every frame by packing the simulation
// buffers it binds (render.rs:255-277: s0, s1, sol[sol_cur], flags,
// vel[cur]) into three rgba16float volume textures.  Running that pass here
// turns the hand-off into 24 bytes per cell (three half4 volumes, x fastest:
// the row layout queue.write_texture takes) instead of the ~64 bytes per
// cell of raw fields, and leaves the renderer only the upload.
//
// One thread per cell: reads s0, s1, sol (16 B each), the u8 solid flag and
// the six faces around the cell (each face is shared with a neighbour, so
// 12 B per cell of compulsory velocity traffic), writes 24 B.  The f32 -> f16
// conversion is a plain fptrunc, which rounds to nearest even (the mode
// register's default), as a storage-texture store of rgba16float does on
// hardware that rounds to nearest.
#include "kernels.hpp"

namespace pyro {

typedef _Float16 half4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ half4v to_h4(float a, float b, float c, float d) {
  half4v h;
  h.x = (_Float16)a;
  h.y = (_Float16)b;
  h.z = (_Float16)c;
  h.w = (_Float16)d;
  return h;
}

__global__ __launch_bounds__(NT) void k_render_pack(const Params *__restrict__ P, Geo g,
                                                     const float4 *__restrict__ s0, const float4 *__restrict__ s1,
                                                     const float4 *__restrict__ sol,
                                                     const uint8_t *__restrict__ flags,
                                                     const float *__restrict__ vu, const float *__restrict__ vv,
                                                     const float *__restrict__ vw, half4v *__restrict__ out) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= g.n) return;
  int x, y, z;
  decode_cell(g, i, x, y, z);
  const float4 a = s0[i], b = s1[i], so = sol[i];
  const bool solid = flags[i] == 1;
  const float rho = rhoOf(P, a.x);
  // render.wgsl:24-26
  const half4v vol = solid ? to_h4(0.0f, a.x, 0.0f, 0.0f)
                           : to_h4(rho * a.w / RHO_SOOT * 1e6f, a.x, rho * b.x * 1e3f, b.y * 1e-3f);
  // render.wgsl:28-31: everything but the occupancy premultiplied by it
  const float tot = so.y + so.z;
  const float oc = solid ? 1.0f : 0.0f;
  const half4v sl = to_h4(oc, so.w * oc, so.z / fmax_(tot, 1e-6f) * oc, so.x * oc);
  // render.wgsl:32-35: cell-centre velocity
  const uint32_t f = vidx(g, x, y, z);
  const half4v ve = to_h4(0.5f * (vu[f] + vu[f + 1]), 0.5f * (vv[f] + vv[f + (uint32_t)g.vx]),
                          0.5f * (vw[f] + vw[f + (uint32_t)g.vx * (uint32_t)g.vy]), 0.0f);
  out[i] = vol;
  out[g.n + i] = sl;
  out[2u * g.n + i] = ve;
}

void launch_render_pack(const Fields &f, void *out, hipStream_t s) {
  k_render_pack<<<nblocks(f.g.n), NT, 0, s>>>(f.P, f.g, f.s0, f.s1, f.sol_in, f.flags, f.va[0], f.va[1], f.va[2],
                                               static_cast<half4v *>(out));
}

}  // namespace pyro
