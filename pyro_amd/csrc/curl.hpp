// curl.hpp -- the vorticity of one cell (fluid.wgsl curl, :3-29), shared by
// k_curl (k_fluid.hip) and the step's fused solid + curl launch
// (k_combust.hip).
#pragma once
#include "common.hpp"

namespace pyro {

// Face-averaged velocity at the clamped cell c (fluid.wgsl:134-138).
__device__ __forceinline__ void uc(const Geo &g, const float *__restrict__ U, const float *__restrict__ V,
                                   const float *__restrict__ W, int x, int y, int z, float &ox, float &oy,
                                   float &oz) {
  x = clampi(x, 0, g.nx - 1);
  y = clampi(y, 0, g.ny - 1);
  z = clampi(z, 0, g.nz - 1);
  uint32_t i = vidx(g, x, y, z);
  ox = 0.5f * (U[i] + U[i + 1]);
  oy = 0.5f * (V[i] + V[i + g.vx]);
  oz = 0.5f * (W[i] + W[i + (uint32_t)g.vx * (uint32_t)g.vy]);
}

// curl[i] = (w, |w|) of cell i < n; with STATS, the cell's speed is returned
// (else 0).
template <bool STATS>
__device__ __forceinline__ float curl_cell(const Params *__restrict__ P, const Geo &g, const float *__restrict__ U,
                                           const float *__restrict__ V, const float *__restrict__ W,
                                           float4 *__restrict__ curl, uint32_t i) {
  int x, y, z;
  decode_cell(g, i, x, y, z);
  float tdx = 2.0f * P->dx;
  float pyx, pyy, pyz, myx, myy, myz, pzx, pzy, pzz, mzx, mzy, mzz, pxx, pxy, pxz, mxx, mxy, mxz;
  uc(g, U, V, W, x, y + 1, z, pyx, pyy, pyz);
  uc(g, U, V, W, x, y - 1, z, myx, myy, myz);
  uc(g, U, V, W, x, y, z + 1, pzx, pzy, pzz);
  uc(g, U, V, W, x, y, z - 1, mzx, mzy, mzz);
  uc(g, U, V, W, x + 1, y, z, pxx, pxy, pxz);
  uc(g, U, V, W, x - 1, y, z, mxx, mxy, mxz);
  float wx = ((pyz - myz) - (pzy - mzy)) / tdx;
  float wy = ((pzx - mzx) - (pxz - mxz)) / tdx;
  float wz = ((pxy - mxy) - (pyx - myx)) / tdx;
  curl[i] = make_float4(wx, wy, wz, len3(wx, wy, wz));
  float spd = 0.0f;
  if (STATS) {
    float cx, cy, cz;
    uc(g, U, V, W, x, y, z, cx, cy, cz);
    spd = len3(cx, cy, cz);
  }
  return spd;
}

}  // namespace pyro
