// sim.hip -- PyroSim: device state, step schedule, hipGraph replay and the
// C ABI of include/pyro_hip.h.
//
// Mirrors the reference's Sim (src/sim.rs): buffers and MG hierarchy of
// Sim::new (:108-232), the dispatch order of Sim::step (:246-423), the
// V-cycle of Sim::vcycles (:451-491) and the stats readback of
// Sim::read_stats (:510-530).  Each step is one hipGraph replay on the
// handle's stream, preceded by a one-thread kernel that writes dt, time,
// ign_on and frame into the device copy of Params (the reference's 576-byte
// write_buffer, sim.rs:250-251).  Graphs are keyed by the ping-pong roles
// (vel[cur], sol[sol_cur]) and by want_stats, so at most 8 exist.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "pyro_hip.h"
#include "pyro_scene.h"

using namespace pyro;

namespace {
thread_local std::string g_err;
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}
#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(PYRO_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int MG_LEVELS = 4;     // sim.rs:13
constexpr int COARSE_ITERS = 80;  // sim.rs:14
}  // namespace

struct PyroSim {
  Params hp;
  Params *dP = nullptr;
  Geo g;
  float *vel[2][3] = {};
  float *hat[3] = {};
  float2 *mm[3] = {};
  float4 *s0 = nullptr, *s1 = nullptr, *s0b = nullptr, *s1b = nullptr, *h0 = nullptr, *h1 = nullptr;
  float4 *sol[2] = {}, *pyr = nullptr, *curl = nullptr;
  uint8_t *flags = nullptr;
  MgLevel lv[8];
  int nl = 0;
  StatParts st;
  int cur = 0, sol_cur = 0;
  uint32_t cycles = 4;
  float ign_until = 0.0f, time = 0.0f, dt = 0.0f;
  uint32_t frame = 0;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  hipGraphExec_t gexec[2][2][2] = {};
  bool in_step = false, step_stats = false;
  uint32_t stats_ran = 0;  // STATS_RAN_* bits of the stages run in the open step
  std::vector<void *> allocs;
  int *d_flag = nullptr;
  // renderer hand-off: packed volumes, the stream that copies them out and
  // the events that order the copy against the steps
  void *render = nullptr;
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_packed = nullptr, ev_copied = nullptr;
  bool copy_issued = false;
  bool pack_issued = false, last_copied = false;  // what pyro_sim_render_wait waits for
  // the volumes as an exportable allocation (pyro_sim_render_export): the
  // physical allocation, its mapping and size; render then points at it
  hipMemGenericAllocationHandle_t render_vmm = nullptr;
  void *render_vmm_ptr = nullptr;
  size_t render_vmm_size = 0;
};

namespace {

int dalloc(PyroSim *s, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) return fail(PYRO_E_OOM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  s->allocs.push_back(*p);
  e = hipMemset(*p, 0, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  return 0;
}
template <class T>
int dalloc_t(PyroSim *s, T **p, size_t count) {
  return dalloc(s, reinterpret_cast<void **>(p), count * sizeof(T));
}

Fields fields_of(PyroSim *s) {
  Fields f;
  f.P = s->dP;
  f.g = s->g;
  for (int k = 0; k < 3; k++) {
    f.va[k] = s->vel[s->cur][k];
    f.vb[k] = s->vel[1 - s->cur][k];
    f.hat[k] = s->hat[k];
    f.mm[k] = s->mm[k];
  }
  f.s0 = s->s0; f.s1 = s->s1; f.s0b = s->s0b; f.s1b = s->s1b; f.h0 = s->h0; f.h1 = s->h1;
  f.sol_in = s->sol[s->sol_cur];
  f.sol_out = s->sol[1 - s->sol_cur];
  f.pyr = s->pyr; f.curl = s->curl; f.flags = s->flags;
  f.phi0 = s->lv[0].phi; f.b0 = s->lv[0].b;
  return f;
}

// sim.rs:451-491
void enqueue_vcycles(PyroSim *s, hipStream_t st) {
  launch_mg_setup(s->lv, s->nl, st);
  for (uint32_t c = 0; c < s->cycles; c++) enqueue_vcycle(s->lv, s->nl, COARSE_ITERS, st);
}

void enqueue_stage(PyroSim *s, int stage, bool stats, hipStream_t st) {
  Fields f = fields_of(s);
  StatParts *sp = stats ? &s->st : nullptr;
  switch (stage) {
    case PYRO_STAGE_SOLID: launch_solid(f, sp, st); break;
    case PYRO_STAGE_CURL: launch_curl(f, sp, st); break;
    case PYRO_STAGE_ADV_VEL1: launch_adv_vel1(f, st); break;
    case PYRO_STAGE_ADV_VEL2: launch_adv_vel2(f, st); break;
    case PYRO_STAGE_ADV_S1: launch_adv_s1(f, st); break;
    case PYRO_STAGE_ADV_S2: launch_adv_s2(f, st); break;
    case PYRO_STAGE_REACT: launch_react(f, sp, st); break;
    case PYRO_STAGE_FORCES: launch_forces(f, st); break;
    case PYRO_STAGE_DIVERGENCE: launch_divergence(f, st); break;
    case PYRO_STAGE_VCYCLES: enqueue_vcycles(s, st); break;
    case PYRO_STAGE_RESID: if (stats) launch_mg_resid(s->lv[0], &s->st, st); break;
    case PYRO_STAGE_PROJECT: launch_project(f, st); break;
    default: break;
  }
}

// sim.rs:394-414, plus the FP64 stats combination on stats steps.  Solid and
// curl run as one launch, and the four advection stages as two (adv_vel1 with
// adv_s1, then adv_vel2 with adv_s2): in each group no stage reads what
// another writes, so the outputs are those of the reference's order.
void enqueue_step_body(PyroSim *s, bool stats, hipStream_t st) {
  for (int k = PYRO_STAGE_SOLID; k <= PYRO_STAGE_VCYCLES; k++) {
    if (k == PYRO_STAGE_SOLID) {
      launch_solid_curl(fields_of(s), stats ? &s->st : nullptr, st);
      k = PYRO_STAGE_CURL;
      continue;
    }
    if (k == PYRO_STAGE_ADV_VEL1) {
      const Fields f = fields_of(s);
      launch_adv_pass1(f, st);
      launch_adv_pass2(f, st);
      k = PYRO_STAGE_ADV_S2;
      continue;
    }
    enqueue_stage(s, k, stats, st);
  }
  if (stats) enqueue_stage(s, PYRO_STAGE_RESID, true, st);
  enqueue_stage(s, PYRO_STAGE_PROJECT, stats, st);
  if (stats) launch_stats_finalize(s->g, s->st, STATS_RAN_ALL, st);
}

void drop_graphs(PyroSim *s) {
  for (auto &a : s->gexec)
    for (auto &b : a)
      for (auto &c : b)
        if (c) {
          (void)hipGraphExecDestroy(c);
          c = nullptr;
        }
}

uint32_t stats_bit(int stage) {
  switch (stage) {
    case PYRO_STAGE_SOLID: return STATS_RAN_SOLID;
    case PYRO_STAGE_CURL: return STATS_RAN_CURL;
    case PYRO_STAGE_REACT: return STATS_RAN_REACT;
    case PYRO_STAGE_RESID: return STATS_RAN_RESID;
    default: return 0u;
  }
}

int get_graph(PyroSim *s, bool stats, hipGraphExec_t *out) {
  hipGraphExec_t &ge = s->gexec[s->cur][s->sol_cur][stats ? 1 : 0];
  if (!ge) {
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step_body(s, stats, s->stream);
    hipError_t le = hipGetLastError();
    HIPCHK(hipStreamEndCapture(s->stream, &graph));
    if (le != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(PYRO_E_HIP, std::string("launch during capture: ") + hipGetErrorString(le));
    }
    hipError_t ie = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      ge = nullptr;
      return fail(PYRO_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
  }
  *out = ge;
  return 0;
}

int sync_check(PyroSim *s) {
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  e = hipGetLastError();
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

struct FieldRef {
  void *ptr = nullptr;
  size_t dev_bytes = 0, host_bytes = 0;
  bool u8flags = false;
};
FieldRef field_ref(PyroSim *s, int field) {
  FieldRef r;
  const size_t n4 = (size_t)s->g.n * 16, nv = (size_t)s->g.nv * 4;
  auto v4 = [&](float4 *p) { r.ptr = p; r.dev_bytes = r.host_bytes = n4; };
  auto vf = [&](float *p) { r.ptr = p; r.dev_bytes = r.host_bytes = nv; };
  switch (field) {
    case PYRO_FIELD_S0: v4(s->s0); return r;
    case PYRO_FIELD_S1: v4(s->s1); return r;
    case PYRO_FIELD_S0B: v4(s->s0b); return r;
    case PYRO_FIELD_S1B: v4(s->s1b); return r;
    case PYRO_FIELD_H0: v4(s->h0); return r;
    case PYRO_FIELD_H1: v4(s->h1); return r;
    case PYRO_FIELD_SOL: v4(s->sol[s->sol_cur]); return r;
    case PYRO_FIELD_SOL_OUT: v4(s->sol[1 - s->sol_cur]); return r;
    case PYRO_FIELD_PYR: v4(s->pyr); return r;
    case PYRO_FIELD_CURL: v4(s->curl); return r;
    case PYRO_FIELD_FLAGS:
      r.ptr = s->flags; r.dev_bytes = s->g.n; r.host_bytes = (size_t)s->g.n * 4; r.u8flags = true;
      return r;
    case PYRO_FIELD_U: case PYRO_FIELD_V: case PYRO_FIELD_W: vf(s->vel[s->cur][field - PYRO_FIELD_U]); return r;
    case PYRO_FIELD_UB: case PYRO_FIELD_VB: case PYRO_FIELD_WB: vf(s->vel[1 - s->cur][field - PYRO_FIELD_UB]); return r;
    case PYRO_FIELD_HU: case PYRO_FIELD_HV: case PYRO_FIELD_HW: vf(s->hat[field - PYRO_FIELD_HU]); return r;
    case PYRO_FIELD_RENDER:
      if (s->render) { r.ptr = s->render; r.dev_bytes = r.host_bytes = (size_t)s->g.n * 24; }
      return r;
    default: break;
  }
  if (field >= PYRO_FIELD_LEVEL_BASE) {
    int l = (field - PYRO_FIELD_LEVEL_BASE) / 4, q = (field - PYRO_FIELD_LEVEL_BASE) % 4;
    if (l < s->nl && q < 3) {
      const MgLevel &L = s->lv[l];
      if (q == 0) { r.ptr = L.phi; r.dev_bytes = r.host_bytes = (size_t)L.n * 4; }
      else if (q == 1) { r.ptr = L.b; r.dev_bytes = r.host_bytes = (size_t)L.n * 4; }
      else { r.ptr = L.f; r.dev_bytes = L.n; r.host_bytes = (size_t)L.n * 4; r.u8flags = true; }
    }
  }
  return r;
}

__global__ void k_nonfinite(const float *p, size_t n, int *flag) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  int bad = 0;
  for (; i < n; i += stride) {
    float v = p[i];
    bad |= !(v - v == 0.0f);
  }
  if (bad) flag[0] = 1;
}

}  // namespace

// Shape staging of pyro_voxelize_noalloc: one array per GPU, held until the
// process ends, shared by every call on that GPU.  Its host-side name is not
// a device address: host code takes the address from hipGetSymbolAddress,
// once per device with that device current (scratch_ptr), and passes that
// pointer to copies and kernels.
constexpr size_t VOX_NOALLOC_MAX_SHAPES = PYRO_VOXELIZE_NOALLOC_MAX_SHAPES;  // x 80 B = 320 KiB
static __device__ __align__(256) PyroShape scratch[VOX_NOALLOC_MAX_SHAPES];
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES] = {};

// The current device's copy of `scratch`.
int scratch_ptr(void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PYRO_E_NODEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(PYRO_E_INVALID, "device index " + std::to_string(dev) + " too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PYRO_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

extern "C" {

const char *pyro_last_error(void) { return g_err.c_str(); }

int pyro_sim_create(const void *params, size_t params_bytes, const float *sol_init, const float *s0_init,
                    uint32_t pressure_cycles, float ign_until, uint32_t cflags, PyroSim **out) {
  if (!params || !out) return fail(PYRO_E_INVALID, "pyro_sim_create: null argument");
  if (params_bytes != PYRO_PARAMS_BYTES) return fail(PYRO_E_INVALID, "pyro_sim_create: Params blob must be 576 bytes");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PYRO_E_NODEVICE, "no HIP device available");
  Params hp;
  std::memcpy(&hp, params, sizeof(Params));
  if (hp.nx < 2 || hp.ny < 2 || hp.nz < 2) return fail(PYRO_E_INVALID, "grid dims must be >= 2");
  uint64_t n64 = (uint64_t)hp.nx * hp.ny * hp.nz, nv64 = (uint64_t)(hp.nx + 1) * (hp.ny + 1) * (hp.nz + 1);
  if (nv64 >= (1ull << 31)) return fail(PYRO_E_INVALID, "grid too large for 32-bit indexing");
  // vec4 fields are gathered through buffer resources with 32-bit byte offsets
  if (n64 * 16ull >= (1ull << 32)) return fail(PYRO_E_INVALID, "grid too large: a vec4 field must stay below 4 GiB");
  PyroSim *s = new PyroSim();
  s->hp = hp;
  s->hp.dt = 0.005f;
  Geo &g = s->g;
  g.nx = (int)hp.nx; g.ny = (int)hp.ny; g.nz = (int)hp.nz;
  g.vx = g.nx + 1; g.vy = g.ny + 1;
  g.n = (uint32_t)n64; g.nv = (uint32_t)nv64;
  g.dnx = make_fastdiv((uint32_t)g.nx); g.dny = make_fastdiv((uint32_t)g.ny);
  g.dvx = make_fastdiv((uint32_t)g.vx); g.dvy = make_fastdiv((uint32_t)g.vy);
  {
    auto cdiv = [](uint32_t a, uint32_t b) { return (a + b - 1) / b; };
    const uint32_t cbx = cdiv(g.nx, BR_X), cby = cdiv(g.ny, BR_Y), cbz = cdiv(g.nz, BR_Z);
    const uint32_t fbx = cdiv(g.vx, BR_X), fby = cdiv(g.vy, BR_Y), fbz = cdiv(g.nz + 1, BR_Z);
    g.cbricks = cbx * cby * cbz;
    g.fbricks = fbx * fby * fbz;
    g.dcbx = make_fastdiv(cbx); g.dcby = make_fastdiv(cby);
    g.dfbx = make_fastdiv(fbx); g.dfby = make_fastdiv(fby);
  }
  s->cycles = pressure_cycles;
  s->ign_until = ign_until;
  s->use_graph = !(cflags & PYRO_CREATE_NO_GRAPH);
  auto bail = [&](int rc) { pyro_sim_destroy(s); return rc; };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PYRO_E_HIP, "hipStreamCreate failed"));
  int rc = 0;
  rc |= dalloc_t(s, &s->dP, 1);
  for (int c = 0; c < 2 && !rc; c++)
    for (int k = 0; k < 3 && !rc; k++) rc |= dalloc_t(s, &s->vel[c][k], g.nv);
  for (int k = 0; k < 3 && !rc; k++) rc |= dalloc_t(s, &s->hat[k], g.nv);
  for (int k = 0; k < 3 && !rc; k++) rc |= dalloc_t(s, &s->mm[k], g.nv);
  float4 **f4[] = {&s->s0, &s->s1, &s->s0b, &s->s1b, &s->h0, &s->h1, &s->sol[0], &s->sol[1], &s->pyr, &s->curl};
  for (auto p : f4)
    if (!rc) rc |= dalloc_t(s, p, g.n);
  if (!rc) rc |= dalloc_t(s, &s->flags, g.n);
  if (!rc) rc |= dalloc_t(s, &s->d_flag, 1);
  // MG hierarchy (sim.rs:142-168)
  int d[3] = {g.nx, g.ny, g.nz};
  float h = hp.dx;
  for (int l = 0; l < MG_LEVELS && !rc; l++) {
    MgLevel &L = s->lv[s->nl++];
    L.d[0] = d[0]; L.d[1] = d[1]; L.d[2] = d[2];
    L.h = h;
    L.hh = h * h;
    L.n = (uint32_t)d[0] * (uint32_t)d[1] * (uint32_t)d[2];
    rc |= dalloc_t(s, &L.phi, L.n);
    if (!rc) rc |= dalloc_t(s, &L.b, L.n);
    if (!rc) rc |= dalloc_t(s, &L.tmp, L.n);
    if (!rc) rc |= dalloc_t(s, &L.nc, L.n);
    if (l == 0) L.f = s->flags;
    else if (!rc) rc |= dalloc_t(s, &L.f, L.n);
    int mn = d[0] < d[1] ? d[0] : d[1];
    mn = d[2] < mn ? d[2] : mn;
    if (mn < 16) break;
    for (int q = 0; q < 3; q++) d[q] = (d[q] + 1) / 2;
    h *= 2.0f;
  }
  // stats partials
  StatParts &st = s->st;
  st.n_solid = nblocks(g.n);
  st.n_curl = nblocks(g.n);
  st.react_gx = (g.nx * g.nz + NT - 1) / NT;
  st.n_resid = mg_resid_blocks(s->lv[0]);
  if (!rc) rc |= dalloc_t(s, &st.solid, (size_t)st.n_solid * 3);
  if (!rc) rc |= dalloc_t(s, &st.curl, (size_t)st.n_curl);
  if (!rc) rc |= dalloc_t(s, &st.react, (size_t)g.ny * st.react_gx * 6);
  if (!rc) rc |= dalloc_t(s, &st.react_t, (size_t)g.ny * st.react_gx);
  if (!rc) rc |= dalloc_t(s, &st.resid, (size_t)st.n_resid * 2);
  if (!rc) rc |= dalloc_t(s, &st.out, STATS_N);
  if (rc) return bail(rc < 0 ? rc : PYRO_E_OOM);
  // initial fields (sim.rs:133-140)
  std::vector<float> amb;
  const float *s0src = s0_init;
  if (!s0src) {
    amb.resize((size_t)g.n * 4);
    for (size_t i = 0; i < g.n; i++) {
      amb[4 * i] = hp.t0; amb[4 * i + 1] = 0.0f; amb[4 * i + 2] = hp.yo2; amb[4 * i + 3] = 0.0f;
    }
    s0src = amb.data();
  }
  if (hipMemcpy(s->dP, &s->hp, sizeof(Params), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->s0, s0src, (size_t)g.n * 16, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PYRO_E_HIP, "initial upload failed"));
  if (sol_init) {
    if (hipMemcpy(s->sol[0], sol_init, (size_t)g.n * 16, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->sol[1], sol_init, (size_t)g.n * 16, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PYRO_E_HIP, "initial upload failed"));
  } else {
    // no solids: SOL = (0, 0, 0, T0) everywhere (the voxelizer with no shapes)
    const uint32_t dims[3] = {hp.nx, hp.ny, hp.nz};
    if (launch_voxelize(nullptr, 0, dims, hp.dx, hp.t0, 0u, static_cast<float4 *>(s->sol[0]), s->stream) ||
        hipMemcpyAsync(s->sol[1], s->sol[0], (size_t)g.n * 16, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
      return bail(fail(PYRO_E_HIP, "initial SOL fill failed"));
  }
  launch_init_flags(s->dP, s->g, s->sol[0], s->flags, s->stream);
  if (sync_check(s)) return bail(PYRO_E_HIP);
  *out = s;
  return 0;
}

namespace {
// The voxelization into device memory dout, stream-ordered on st and
// finished when it returns: shapes uploaded, the kernel run, st
// synchronised (no device-wide synchronisation, so other streams of the
// process keep running), the shape buffer freed.
int voxelize_on(const struct PyroShape *shapes, size_t count, const uint32_t dims[3], float dx, float t0,
                uint32_t seed, float4 *dout, hipStream_t st) {
  PyroShape *dsh = nullptr;
  int rc = 0;
  if (count && (hipMalloc(&dsh, count * sizeof(PyroShape)) != hipSuccess ||
                hipMemcpyAsync(dsh, shapes, count * sizeof(PyroShape), hipMemcpyHostToDevice, st) != hipSuccess))
    rc = fail(PYRO_E_HIP, "pyro_voxelize: shape upload failed");
  if (!rc && launch_voxelize(dsh, (int)count, dims, dx, t0, seed, dout, st))
    rc = fail(PYRO_E_HIP, "pyro_voxelize: kernel launch failed");
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(PYRO_E_HIP, "pyro_voxelize: kernel failed");
  if (dsh) (void)hipFree(dsh);
  return rc;
}
}  // namespace

int pyro_voxelize(const struct PyroShape *shapes, size_t count, const uint32_t dims[3], float dx, float t0,
                  uint32_t seed, float *dst, int dst_on_device) {
  if (!dims || !dst || (count && !shapes)) return fail(PYRO_E_INVALID, "pyro_voxelize: null argument");
  const uint64_t n64 = (uint64_t)dims[0] * dims[1] * dims[2];
  if (!n64 || n64 * 16ull >= (1ull << 32) || count > (1u << 20)) return fail(PYRO_E_INVALID, "pyro_voxelize: bad size");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PYRO_E_NODEVICE, "no HIP device available");
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return fail(PYRO_E_HIP, "pyro_voxelize: hipStreamCreate failed");
  float4 *dout = nullptr;
  int rc = 0;
  if (dst_on_device) dout = reinterpret_cast<float4 *>(dst);
  else if (hipMalloc(&dout, (size_t)n64 * 16) != hipSuccess) rc = fail(PYRO_E_OOM, "pyro_voxelize: hipMalloc failed");
  if (!rc) rc = voxelize_on(shapes, count, dims, dx, t0, seed, dout, st);
  if (!rc && !dst_on_device &&
      (hipMemcpyAsync(dst, dout, (size_t)n64 * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(PYRO_E_HIP, "pyro_voxelize: download failed");
  if (dout && !dst_on_device) (void)hipFree(dout);
  (void)hipStreamDestroy(st);
  return rc;
}

int pyro_voxelize_noalloc(const struct PyroShape *shapes, size_t count, const uint32_t dims[3], float dx, float t0,
                          uint32_t seed, float *dst_device) {
  if (!dims || !dst_device || (count && !shapes)) return fail(PYRO_E_INVALID, "pyro_voxelize_noalloc: null argument");
  const uint64_t n64 = (uint64_t)dims[0] * dims[1] * dims[2];
  if (!n64 || n64 * 16ull >= (1ull << 32)) return fail(PYRO_E_INVALID, "pyro_voxelize_noalloc: bad grid size");
  if (count > VOX_NOALLOC_MAX_SHAPES)
    return fail(PYRO_E_INVALID, "pyro_voxelize_noalloc: " + std::to_string(count) + " shapes, at most " +
                                    std::to_string(VOX_NOALLOC_MAX_SHAPES));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PYRO_E_NODEVICE, "no HIP device available");
  // the kernel writes through dst_device: it must be device memory
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, dst_device) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return fail(PYRO_E_INVALID, "pyro_voxelize_noalloc: dst_device is not device memory");
  }
  void *sh = nullptr;
  int rc = scratch_ptr(&sh);
  if (rc) return rc;
  if (count && hipMemcpy(sh, shapes, count * sizeof(PyroShape), hipMemcpyHostToDevice) != hipSuccess)
    return fail(PYRO_E_HIP, "pyro_voxelize_noalloc: shape upload failed");
  if (launch_voxelize(static_cast<const PyroShape *>(sh), (int)count, dims, dx, t0, seed,
                      reinterpret_cast<float4 *>(dst_device), nullptr) ||
      hipStreamSynchronize(nullptr) != hipSuccess)
    return fail(PYRO_E_HIP, "pyro_voxelize_noalloc: kernel failed");
  return 0;
}

int pyro_sim_voxelize(PyroSim *s, const struct PyroShape *shapes, size_t count, float t0, uint32_t seed) {
  if (!s || (count && !shapes)) return fail(PYRO_E_INVALID, "pyro_sim_voxelize: null argument");
  int rc = sync_check(s);
  if (rc) return rc;
  if (count > (1u << 20)) return fail(PYRO_E_INVALID, "pyro_sim_voxelize: too many shapes");
  const uint32_t dims[3] = {s->hp.nx, s->hp.ny, s->hp.nz};
  rc = voxelize_on(shapes, count, dims, s->hp.dx, t0, seed, static_cast<float4 *>(s->sol[s->sol_cur]), s->stream);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(s->sol[1 - s->sol_cur], s->sol[s->sol_cur], (size_t)s->g.n * 16, hipMemcpyDeviceToDevice,
                        s->stream));
  launch_init_flags(s->dP, s->g, static_cast<const float4 *>(s->sol[s->sol_cur]), s->flags, s->stream);
  return sync_check(s);
}

int pyro_sim_render_pack(PyroSim *s, void *host_dst, size_t bytes) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (s->in_step) return fail(PYRO_E_STATE, "render_pack inside a step");
  const size_t need = (size_t)s->g.n * 24;
  if (host_dst && bytes != need)
    return fail(PYRO_E_INVALID, "pyro_sim_render_pack: destination must hold 3 x N x 8 bytes (" +
                                    std::to_string(need) + ")");
  if (!s->render) {
    int rc = dalloc(s, &s->render, need);
    if (rc) return rc;
  }
  if (!s->copy_stream) HIPCHK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  if (!s->ev_packed) HIPCHK(hipEventCreateWithFlags(&s->ev_packed, hipEventDisableTiming));
  if (!s->ev_copied) HIPCHK(hipEventCreateWithFlags(&s->ev_copied, hipEventDisableTiming));
  // the previous hand-off's copy must have read the buffer before it is
  // packed again
  if (s->copy_issued) HIPCHK(hipStreamWaitEvent(s->stream, s->ev_copied, 0));
  launch_render_pack(fields_of(s), s->render, s->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(s->ev_packed, s->stream));
  s->pack_issued = true;
  s->last_copied = host_dst != nullptr;
  if (!host_dst) return 0;
  HIPCHK(hipStreamWaitEvent(s->copy_stream, s->ev_packed, 0));
  HIPCHK(hipMemcpyAsync(host_dst, s->render, need, hipMemcpyDeviceToHost, s->copy_stream));
  HIPCHK(hipEventRecord(s->ev_copied, s->copy_stream));
  s->copy_issued = true;
  return 0;
}

int pyro_sim_render_wait(PyroSim *s) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (!s->pack_issued) return 0;
  hipError_t e = hipEventSynchronize(s->last_copied ? s->ev_copied : s->ev_packed);
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("render hand-off: ") + hipGetErrorString(e));
  return 0;
}

namespace {
hipMemAllocationProp export_prop(int dev) {
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = dev;
  return prop;
}
// reserve, map and open for read/write an allocation of `size` bytes
int map_allocation(hipMemGenericAllocationHandle_t h, size_t size, int dev, void **out) {
  void *p = nullptr;
  hipError_t e = hipMemAddressReserve(&p, size, 0, nullptr, 0);
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("hipMemAddressReserve: ") + hipGetErrorString(e));
  e = hipMemMap(p, size, 0, h, 0);
  if (e != hipSuccess) {
    (void)hipMemAddressFree(p, size);
    return fail(PYRO_E_HIP, std::string("hipMemMap: ") + hipGetErrorString(e));
  }
  hipMemAccessDesc acc = {};
  acc.location.type = hipMemLocationTypeDevice;
  acc.location.id = dev;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  e = hipMemSetAccess(p, size, &acc, 1);
  if (e != hipSuccess) {
    (void)hipMemUnmap(p, size);
    (void)hipMemAddressFree(p, size);
    return fail(PYRO_E_HIP, std::string("hipMemSetAccess: ") + hipGetErrorString(e));
  }
  *out = p;
  return 0;
}
}  // namespace

int pyro_sim_render_export(PyroSim *s, int *fd, uint64_t *bytes) {
  if (!s || !fd || !bytes) return fail(PYRO_E_INVALID, "pyro_sim_render_export: null argument");
  if (s->in_step) return fail(PYRO_E_STATE, "render_export inside a step");
  *fd = -1;
  *bytes = 0;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (!s->render_vmm) {
    int vmm = 0;
    HIPCHK(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, dev));
    if (!vmm) return fail(PYRO_E_HIP, "pyro_sim_render_export: the device has no virtual memory management");
    const hipMemAllocationProp prop = export_prop(dev);
    size_t gran = 0;
    HIPCHK(hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum));
    const size_t need = (size_t)s->g.n * 24;
    const size_t size = gran ? (need + gran - 1) / gran * gran : need;
    // packs and copies still in flight use the current buffer
    if (s->copy_stream) HIPCHK(hipStreamSynchronize(s->copy_stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    hipMemGenericAllocationHandle_t h = nullptr;
    hipError_t e = hipMemCreate(&h, size, &prop, 0);
    if (e != hipSuccess) return fail(PYRO_E_OOM, std::string("hipMemCreate: ") + hipGetErrorString(e));
    void *p = nullptr;
    int rc = map_allocation(h, size, dev, &p);
    if (rc) {
      (void)hipMemRelease(h);
      return rc;
    }
    s->render_vmm = h;
    s->render_vmm_ptr = p;
    s->render_vmm_size = size;
    s->render = p;
  }
  int f = -1;
  hipError_t e = hipMemExportToShareableHandle(&f, s->render_vmm, hipMemHandleTypePosixFileDescriptor, 0);
  if (e != hipSuccess || f < 0)
    return fail(PYRO_E_HIP, std::string("hipMemExportToShareableHandle: ") + hipGetErrorString(e));
  *fd = f;
  *bytes = (uint64_t)s->render_vmm_size;
  return 0;
}

int pyro_shared_read(int fd, uint64_t alloc_bytes, void *host_dst, size_t bytes) {
  // Importing an exported descriptor back into this process (import, map,
  // copy) ended in a host segmentation fault inside the runtime when run
  // after a long sequence of other work; no cause has been shown, so the
  // import is switched off and this entry point only reports that.
  (void)fd;
  (void)alloc_bytes;
  (void)host_dst;
  (void)bytes;
  return fail(PYRO_E_STATE, "pyro_shared_read is disabled: importing an exported descriptor in-process is not verified");
}

int pyro_host_alloc(size_t bytes, void **out) {
  if (!out || !bytes) return fail(PYRO_E_INVALID, "pyro_host_alloc: bad argument");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return fail(PYRO_E_OOM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  return 0;
}

int pyro_host_free(void *p) {
  if (!p) return 0;
  hipError_t e = hipHostFree(p);
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("hipHostFree: ") + hipGetErrorString(e));
  return 0;
}

int pyro_sim_destroy(PyroSim *s) {
  if (!s) return 0;
  if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->ev_packed) (void)hipEventDestroy(s->ev_packed);
  if (s->ev_copied) (void)hipEventDestroy(s->ev_copied);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  drop_graphs(s);
  for (void *p : s->allocs) (void)hipFree(p);
  if (s->render_vmm) {
    (void)hipMemUnmap(s->render_vmm_ptr, s->render_vmm_size);
    (void)hipMemAddressFree(s->render_vmm_ptr, s->render_vmm_size);
    (void)hipMemRelease(s->render_vmm);
  }
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int pyro_sim_begin_step(PyroSim *s, float dt, int want_stats) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (s->in_step) return fail(PYRO_E_STATE, "begin_step inside a step");
  // sim.rs:247-251
  float ign_on = s->time < s->ign_until ? 1.0f : 0.0f;
  s->hp.dt = dt;
  s->hp.time = s->time;
  s->hp.ign_on = ign_on;
  s->hp.frame = s->frame;
  launch_set_step(s->dP, dt, s->time, ign_on, s->frame, s->stream);
  s->dt = dt;
  s->in_step = true;
  s->step_stats = want_stats != 0;
  s->stats_ran = 0;
  return 0;
}

int pyro_sim_run_stage(PyroSim *s, int stage) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (!s->in_step) return fail(PYRO_E_STATE, "run_stage outside begin_step/end_step");
  if (stage < 0 || stage >= PYRO_STAGE_COUNT) return fail(PYRO_E_INVALID, "bad stage id");
  enqueue_stage(s, stage, s->step_stats, s->stream);
  s->stats_ran |= stats_bit(stage);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return 0;
}

int pyro_sim_end_step(PyroSim *s) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (!s->in_step) return fail(PYRO_E_STATE, "end_step without begin_step");
  if (s->step_stats) launch_stats_finalize(s->g, s->st, s->stats_ran, s->stream);
  // sim.rs:419-422
  s->cur = 1 - s->cur;
  s->sol_cur = 1 - s->sol_cur;
  s->time = s->time + s->dt;
  s->frame += 1;
  s->in_step = false;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PYRO_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return 0;
}

int pyro_sim_abort_step(PyroSim *s) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (!s->in_step) return 0;
  // roles, time and frame stay those before begin_step
  s->in_step = false;
  s->step_stats = false;
  s->stats_ran = 0;
  return sync_check(s);
}

int pyro_sim_step(PyroSim *s, float dt, int want_stats) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  int rc = pyro_sim_begin_step(s, dt, want_stats);
  if (rc) return rc;
  bool stats = want_stats != 0;
  if (s->use_graph) {
    hipGraphExec_t ge;
    rc = get_graph(s, stats, &ge);
    if (rc) { s->in_step = false; return rc; }
    hipError_t e = hipGraphLaunch(ge, s->stream);
    if (e != hipSuccess) { s->in_step = false; return fail(PYRO_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e)); }
  } else {
    enqueue_step_body(s, stats, s->stream);
  }
  // the finalize is part of the body; end the step without a second one
  s->step_stats = false;
  return pyro_sim_end_step(s);
}

int pyro_sim_read_stats(PyroSim *s, double *out, size_t n) {
  if (!s || !out) return fail(PYRO_E_INVALID, "null argument");
  if (n != PYRO_STATS_N) return fail(PYRO_E_INVALID, "stats buffer must hold 1040 doubles");
  int rc = sync_check(s);
  if (rc) return rc;
  HIPCHK(hipMemcpy(out, s->st.out, PYRO_STATS_N * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pyro_sim_read_field(PyroSim *s, int field, void *dst, size_t bytes) {
  if (!s || !dst) return fail(PYRO_E_INVALID, "null argument");
  FieldRef r = field_ref(s, field);
  if (!r.ptr) return fail(PYRO_E_INVALID, "unknown field id " + std::to_string(field));
  if (bytes != r.host_bytes) return fail(PYRO_E_INVALID, "field size mismatch: expected " + std::to_string(r.host_bytes));
  int rc = sync_check(s);
  if (rc) return rc;
  if (r.u8flags) {
    std::vector<uint8_t> tmp(r.dev_bytes);
    HIPCHK(hipMemcpy(tmp.data(), r.ptr, r.dev_bytes, hipMemcpyDeviceToHost));
    uint32_t *o = static_cast<uint32_t *>(dst);
    for (size_t i = 0; i < r.dev_bytes; i++) o[i] = tmp[i] == 1 ? 1u : 0u;
  } else {
    HIPCHK(hipMemcpy(dst, r.ptr, bytes, hipMemcpyDeviceToHost));
  }
  return 0;
}

int pyro_sim_write_field(PyroSim *s, int field, const void *src, size_t bytes) {
  if (!s || !src) return fail(PYRO_E_INVALID, "null argument");
  FieldRef r = field_ref(s, field);
  if (!r.ptr) return fail(PYRO_E_INVALID, "unknown field id " + std::to_string(field));
  if (bytes != r.host_bytes) return fail(PYRO_E_INVALID, "field size mismatch: expected " + std::to_string(r.host_bytes));
  int rc = sync_check(s);
  if (rc) return rc;
  if (r.u8flags) {
    std::vector<uint8_t> tmp(r.dev_bytes);
    const uint32_t *in = static_cast<const uint32_t *>(src);
    for (size_t i = 0; i < r.dev_bytes; i++) tmp[i] = in[i] == 1u ? 1 : 0;
    HIPCHK(hipMemcpy(r.ptr, tmp.data(), r.dev_bytes, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(r.ptr, src, bytes, hipMemcpyHostToDevice));
  }
  if (field == PYRO_FIELD_SOL || field == PYRO_FIELD_SOL_OUT) {
    // a new SOL state: both ping-pong buffers and the solid flags follow it
    void *other = field == PYRO_FIELD_SOL ? (void *)s->sol[1 - s->sol_cur] : (void *)s->sol[s->sol_cur];
    HIPCHK(hipMemcpy(other, src, bytes, hipMemcpyHostToDevice));
    launch_init_flags(s->dP, s->g, static_cast<const float4 *>(r.ptr), s->flags, s->stream);
    return sync_check(s);
  }
  return 0;
}

// ---------------------------------------------------------------- checkpoint
// A step reads only vel[cur], s0, s1, sol[sol_cur] (flags follow from it),
// level-0 phi (the warm start), time and frame; everything else is written
// before it is read.  The file: a fixed header (magic, version, dims, the
// 576-byte Params, time, frame, cycles, ign_until) followed by the seven
// fields as raw little-endian FP32 in the device layout.
}  // extern "C"
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version, nx, ny, nz, params_bytes;
  float time, ign_until;
  uint32_t frame, cycles, nfields, pad;
};
constexpr char kCkptMagic[8] = {'P', 'Y', 'R', 'O', 'C', 'K', 'P', 'T'};

struct CkptField {
  void *ptr;
  size_t bytes;
};
std::vector<CkptField> ckpt_fields(PyroSim *s) {
  const size_t n4 = (size_t)s->g.n * 16, nv = (size_t)s->g.nv * 4;
  return {{s->vel[s->cur][0], nv}, {s->vel[s->cur][1], nv}, {s->vel[s->cur][2], nv}, {s->s0, n4},
          {s->s1, n4},           {s->sol[s->sol_cur], n4}, {s->lv[0].phi, (size_t)s->g.n * 4}};
}
}  // namespace
extern "C" {

int pyro_sim_save(PyroSim *s, const char *path) {
  if (!s || !path) return fail(PYRO_E_INVALID, "null argument");
  if (s->in_step) return fail(PYRO_E_STATE, "save inside a step");
  int rc = sync_check(s);
  if (rc) return rc;
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.nx = s->hp.nx; h.ny = s->hp.ny; h.nz = s->hp.nz;
  h.params_bytes = (uint32_t)sizeof(Params);
  h.time = s->time; h.ign_until = s->ign_until;
  h.frame = s->frame; h.cycles = s->cycles;
  const std::vector<CkptField> fl = ckpt_fields(s);
  h.nfields = (uint32_t)fl.size();
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return fail(PYRO_E_INVALID, std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(&s->hp, sizeof(Params), 1, f) == 1;
  std::vector<char> buf;
  for (const CkptField &c : fl) {
    if (!ok) break;
    buf.resize(c.bytes);
    ok = hipMemcpy(buf.data(), c.ptr, c.bytes, hipMemcpyDeviceToHost) == hipSuccess &&
         std::fwrite(buf.data(), 1, c.bytes, f) == c.bytes;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(PYRO_E_HIP, std::string("writing checkpoint ") + path + " failed");
  return 0;
}

int pyro_sim_load(PyroSim *s, const char *path) {
  if (!s || !path) return fail(PYRO_E_INVALID, "null argument");
  if (s->in_step) return fail(PYRO_E_STATE, "load inside a step");
  int rc = sync_check(s);
  if (rc) return rc;
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return fail(PYRO_E_INVALID, std::string("cannot open ") + path);
  CkptHeader h{};
  Params hp;
  const std::vector<CkptField> fl = ckpt_fields(s);
  std::string err;
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, 8) != 0)
    err = "not a checkpoint file";
  else if (h.version != 1)
    err = "unsupported checkpoint version " + std::to_string(h.version);
  else if (h.params_bytes != sizeof(Params) || h.nfields != fl.size())
    err = "checkpoint layout mismatch";
  else if (h.nx != s->hp.nx || h.ny != s->hp.ny || h.nz != s->hp.nz)
    err = "checkpoint grid " + std::to_string(h.nx) + "x" + std::to_string(h.ny) + "x" + std::to_string(h.nz) +
          " does not match the simulation's";
  else if (std::fread(&hp, sizeof(Params), 1, f) != 1)
    err = "truncated checkpoint";
  else if (hp.nx != h.nx || hp.ny != h.ny || hp.nz != h.nz)
    err = "checkpoint Params grid does not match its header";
  else if (std::memcmp(&hp.dx, &s->hp.dx, sizeof(float)) != 0)
    // the multigrid levels' h and h^2 are derived from dx at creation
    err = "checkpoint cell size dx = " + std::to_string(hp.dx) + " differs from the simulation's " +
          std::to_string(s->hp.dx);
  std::vector<std::vector<char>> data(fl.size());
  for (size_t k = 0; k < fl.size() && err.empty(); k++) {
    data[k].resize(fl[k].bytes);
    if (std::fread(data[k].data(), 1, fl[k].bytes, f) != fl[k].bytes) err = "truncated checkpoint";
  }
  std::fclose(f);
  if (!err.empty()) return fail(PYRO_E_INVALID, err + " (" + path + ")");
  // nothing on the device changes before the whole file has been read.
  // Captured graphs hold the V-cycle count and the ping-pong roles of the
  // running simulation: they are captured again at the next step.
  drop_graphs(s);
  s->hp = hp;
  HIPCHK(hipMemcpy(s->dP, &s->hp, sizeof(Params), hipMemcpyHostToDevice));
  for (size_t k = 0; k < fl.size(); k++) HIPCHK(hipMemcpy(fl[k].ptr, data[k].data(), fl[k].bytes, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->sol[1 - s->sol_cur], data[5].data(), fl[5].bytes, hipMemcpyHostToDevice));
  launch_init_flags(s->dP, s->g, s->sol[s->sol_cur], s->flags, s->stream);
  s->time = h.time;
  s->frame = h.frame;
  s->cycles = h.cycles;
  s->ign_until = h.ign_until;
  return sync_check(s);
}

int pyro_sim_info(const PyroSim *s, uint32_t dims[3], uint64_t *cells, double *time, uint32_t *frame,
                  uint32_t *levels) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  if (dims) { dims[0] = (uint32_t)s->g.nx; dims[1] = (uint32_t)s->g.ny; dims[2] = (uint32_t)s->g.nz; }
  if (cells) *cells = s->g.n;
  if (time) *time = s->time;
  if (frame) *frame = s->frame;
  if (levels) *levels = (uint32_t)s->nl;
  return 0;
}

int pyro_sim_level_dims(const PyroSim *s, int level, uint32_t dims[3]) {
  if (!s || !dims) return fail(PYRO_E_INVALID, "null argument");
  if (level < 0 || level >= s->nl) return fail(PYRO_E_INVALID, "bad level");
  for (int q = 0; q < 3; q++) dims[q] = (uint32_t)s->lv[level].d[q];
  return 0;
}

int pyro_sim_sync(PyroSim *s) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  return sync_check(s);
}

int pyro_sim_device_ptr(PyroSim *s, int field, void **dptr, size_t *bytes) {
  if (!s || !dptr) return fail(PYRO_E_INVALID, "null argument");
  FieldRef r = field_ref(s, field);
  if (!r.ptr) return fail(PYRO_E_INVALID, "unknown field id");
  *dptr = r.ptr;
  if (bytes) *bytes = r.dev_bytes;
  return 0;
}

int pyro_sim_time_stage(PyroSim *s, int stage, int reps, float *avg_ms) {
  if (!s || !avg_ms || reps < 1) return fail(PYRO_E_INVALID, "bad argument");
  if (stage < 0 || stage >= PYRO_STAGE_COUNT) return fail(PYRO_E_INVALID, "bad stage id");
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, s->stream));
  for (int r = 0; r < reps; r++) enqueue_stage(s, stage, false, s->stream);
  HIPCHK(hipEventRecord(b, s->stream));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *avg_ms = ms / (float)reps;
  return sync_check(s);
}

int pyro_sim_time_kernel(PyroSim *s, int kernel, int level, int reps, float *avg_ms) {
  if (!s || !avg_ms || reps < 1) return fail(PYRO_E_INVALID, "bad argument");
  if (level < 0 || level >= s->nl) return fail(PYRO_E_INVALID, "bad level");
  if ((kernel == PYRO_KERNEL_MG_RESTRICT || kernel == PYRO_KERNEL_MG_PROLONG) && level >= s->nl - 1)
    return fail(PYRO_E_INVALID, "restrict/prolong need a coarser level");
  if (kernel < 0 || kernel > PYRO_KERNEL_SOLID_CURL) return fail(PYRO_E_INVALID, "bad kernel id");
  if (kernel == PYRO_KERNEL_RENDER_PACK && !s->render)
    return fail(PYRO_E_STATE, "render_pack timing needs a pyro_sim_render_pack call first");
  if (kernel == PYRO_KERNEL_MG_SMOOTH && (reps & 1)) reps++;  // even: phi ends in its own buffer
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, s->stream));
  for (int r = 0; r < reps; r++) {
    switch (kernel) {
      case PYRO_KERNEL_MG_SMOOTH: {
        const MgLevel &L = s->lv[level];
        if (r & 1) launch_mg_sweep(L, L.tmp, L.phi, 1.0f, s->stream, level == 0);
        else launch_mg_sweep(L, L.phi, L.tmp, 1.0f, s->stream, level == 0);
        break;
      }
      case PYRO_KERNEL_MG_RESTRICT: launch_mg_restrict(s->lv[level], s->lv[level + 1], s->stream); break;
      case PYRO_KERNEL_MG_PROLONG: launch_mg_prolong(s->lv[level + 1], s->lv[level], s->stream); break;
      case PYRO_KERNEL_RENDER_PACK: launch_render_pack(fields_of(s), s->render, s->stream); break;
      case PYRO_KERNEL_ADV_PASS1: launch_adv_pass1(fields_of(s), s->stream); break;
      case PYRO_KERNEL_ADV_PASS2: launch_adv_pass2(fields_of(s), s->stream); break;
      case PYRO_KERNEL_SOLID_CURL: launch_solid_curl(fields_of(s), nullptr, s->stream); break;
      default: launch_mg_coarse(s->lv[s->nl - 1], COARSE_ITERS, 1.6f, s->stream); break;
    }
  }
  HIPCHK(hipEventRecord(b, s->stream));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *avg_ms = ms / (float)reps;
  return sync_check(s);
}

int pyro_sim_add_velocity_noise(PyroSim *s, float amp, uint32_t seed) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  Fields f = fields_of(s);
  launch_velocity_noise(f, amp, seed, s->stream);
  return sync_check(s);
}

int pyro_sim_check_finite(PyroSim *s) {
  if (!s) return fail(PYRO_E_INVALID, "null sim");
  HIPCHK(hipMemsetAsync(s->d_flag, 0, sizeof(int), s->stream));
  auto scan = [&](const void *p, size_t nfloat) {
    k_nonfinite<<<1024, 256, 0, s->stream>>>(static_cast<const float *>(p), nfloat, s->d_flag);
  };
  const size_t n4 = (size_t)s->g.n * 4;
  scan(s->s0, n4);
  scan(s->s1, n4);
  scan(s->sol[s->sol_cur], n4);
  for (int k = 0; k < 3; k++) scan(s->vel[s->cur][k], s->g.nv);
  scan(s->lv[0].phi, s->g.n);
  int flag = 0;
  int rc = sync_check(s);
  if (rc) return rc;
  HIPCHK(hipMemcpy(&flag, s->d_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) return fail(PYRO_E_NONFINITE, "non-finite value in the simulation state");
  return 0;
}

int pyro_debug_copy(size_t n_floats, int reps, float *avg_ms) {
  if (n_floats == 0 || n_floats % 4 || n_floats / 4 >= (1ull << 31) || reps < 1 || !avg_ms)
    return fail(PYRO_E_INVALID, "bad argument");
  float *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n_floats * 4));
  HIPCHK(hipMalloc(&b, n_floats * 4));
  HIPCHK(hipMemset(a, 0, n_floats * 4));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  launch_copy(reinterpret_cast<const float4 *>(a), reinterpret_cast<float4 *>(b), (uint32_t)(n_floats / 4), nullptr);
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++)
    launch_copy(reinterpret_cast<const float4 *>(a), reinterpret_cast<float4 *>(b), (uint32_t)(n_floats / 4), nullptr);
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / (float)reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int pyro_debug_math(int fn, const float *in, const float *in2, float *out, size_t n) {
  if (!in || !out || (fn == 2 && !in2) || fn < 0 || fn > 2) return fail(PYRO_E_INVALID, "bad argument");
  if (n == 0) return 0;
  float *di = nullptr, *di2 = nullptr, *dout = nullptr;
  HIPCHK(hipMalloc(&di, n * 4));
  HIPCHK(hipMalloc(&dout, n * 4));
  HIPCHK(hipMemcpy(di, in, n * 4, hipMemcpyHostToDevice));
  if (fn == 2) {
    HIPCHK(hipMalloc(&di2, n * 4));
    HIPCHK(hipMemcpy(di2, in2, n * 4, hipMemcpyHostToDevice));
  }
  launch_debug_math(fn, di, di2, dout, (uint32_t)n, nullptr);
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
  (void)hipFree(di);
  (void)hipFree(dout);
  if (di2) (void)hipFree(di2);
  return 0;
}

}  // extern "C"
