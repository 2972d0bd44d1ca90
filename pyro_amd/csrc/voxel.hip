// voxel.hip -- Scene::voxelize (scene.rs:621-660) on the device.
//
// One thread per cell tests the cell centre against every primitive of the
// scene's expanded shape list (pyro_scene_shapes), in order; the last one
// that contains it sets SOL = (slot, mass, 0, T0), mass = density x jitter
// (hash of the cell index, scene.rs:662-667) or 1 for inert materials.  The
// inside tests of scene.rs:670-727 are restated in the same f32 operations
// (no contraction; device sqrt and division are correctly rounded), with the
// box yaw's sine and cosine taken from the host, so the result equals the
// host voxelizer bit for bit.  It removes the O(cells x shapes) host loop of
// large grids (1.4 s for the bonfire at 10 cm, 10 s at 5 cm on the host).
#include "kernels.hpp"
#include "pyro_hip.h"
#include "pyro_scene.h"

namespace pyro {

__device__ __forceinline__ bool shape_inside(const PyroShape &sh, float px, float py, float pz, float dx) {
  switch (sh.kind) {
    case 0: {  // box, yawed about y
      const float d0 = px - sh.c[0], d1 = py - sh.c[1], d2 = pz - sh.c[2];
      const float lx = sh.cos_yaw * d0 + sh.sin_yaw * d2;
      const float lz = -sh.sin_yaw * d0 + sh.cos_yaw * d2;
      const float h0 = fmaxf(sh.h[0], 0.5f * dx), h1 = fmaxf(sh.h[1], 0.5f * dx), h2 = fmaxf(sh.h[2], 0.5f * dx);
      return fabsf(lx) <= h0 && fabsf(d1) <= h1 && fabsf(lz) <= h2;
    }
    case 1: {  // cylinder around the segment a-b
      const float ba0 = sh.b[0] - sh.a[0], ba1 = sh.b[1] - sh.a[1], ba2 = sh.b[2] - sh.a[2];
      const float pa0 = px - sh.a[0], pa1 = py - sh.a[1], pa2 = pz - sh.a[2];
      const float l2 = ba0 * ba0 + ba1 * ba1 + ba2 * ba2;
      const float t = (pa0 * ba0 + pa1 * ba1 + pa2 * ba2) / l2;
      if (!(t >= 0.0f && t <= 1.0f)) return false;
      const float q0 = pa0 - ba0 * t, q1 = pa1 - ba1 * t, q2 = pa2 - ba2 * t;
      return sqrtf(q0 * q0 + q1 * q1 + q2 * q2) <= fmaxf(sh.r, 0.5f * dx);
    }
    case 2: {  // sphere
      const float d0 = px - sh.c[0], d1 = py - sh.c[1], d2 = pz - sh.c[2];
      return sqrtf(d0 * d0 + d1 * d1 + d2 * d2) <= sh.r;
    }
    case 3: {  // pool: vertical cylinder from the ground
      const float ex = px - sh.c[0], ez = pz - sh.c[2];
      const float dxz = sqrtf(ex * ex + ez * ez);
      return dxz <= sh.r && py <= fmaxf(sh.depth, dx * 0.99f);
    }
  }
  return false;
}

__device__ __forceinline__ float scene_hash(uint32_t v) {
  uint32_t x = v * 747796405u + 2891336453u;
  x = ((x >> ((x >> 28) + 4)) ^ x) * 277803737u;
  x = (x >> 22) ^ x;
  return (float)x / (float)4294967295u;
}

__global__ __launch_bounds__(NT) void k_voxelize(const PyroShape *__restrict__ sh, int count, uint32_t nx,
                                                  uint32_t ny, uint32_t nz, FastDiv dnx, FastDiv dny, float dx,
                                                  float t0, uint32_t seed, float4 *__restrict__ out) {
  const uint32_t i = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= nx * ny * nz) return;
  const uint32_t r = fdiv(i, dnx), x = i - r * nx, z = fdiv(r, dny), y = r - z * ny;
  const float px = ((float)x + 0.5f) * dx, py = ((float)y + 0.5f) * dx, pz = ((float)z + 0.5f) * dx;
  float4 v = make_float4(0.0f, 0.0f, 0.0f, t0);
  for (int k = 0; k < count; k++) {
    if (!shape_inside(sh[k], px, py, pz, dx)) continue;
    const float h = scene_hash(i ^ (seed * 0x9e3779b9u));
    const float jit = 0.85f + 0.3f * h;
    v = make_float4((float)sh[k].slot, sh[k].inert ? 1.0f : sh[k].density * jit, 0.0f, t0);
  }
  out[i] = v;
}

int launch_voxelize(const PyroShape *d_shapes, int count, const uint32_t dims[3], float dx, float t0, uint32_t seed,
                    float4 *out, hipStream_t s) {
  const uint32_t n = dims[0] * dims[1] * dims[2];
  k_voxelize<<<nblocks(n), NT, 0, s>>>(d_shapes, count, dims[0], dims[1], dims[2], make_fastdiv(dims[0]),
                                       make_fastdiv(dims[1]), dx, t0, seed, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pyro
