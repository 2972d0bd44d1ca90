// scene.cpp -- host-side scene plumbing: TOML subset -> Scene -> Params + SOL.
//
// Restates src/scene.rs of the reference (schema and defaults :8-215, material
// table :240-406, Params :426-477, grid :485-493, params :523-618, voxelize
// :621-660, hash :662-667, shapes :670-879).  All arithmetic is f32 in the
// order of the Rust source (compiled with -ffp-contract=off, as rustc never
// contracts), so the Params blob and SOL array are the ones scene.rs produces.
#include "pyro_scene.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct SceneError : std::runtime_error {
  int code;
  SceneError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// ----------------------------------------------------------------------------
// Minimal TOML: tables, arrays of tables, key = string | number | bool | array.
// ----------------------------------------------------------------------------
struct TVal {
  enum Kind { Int, Float, Str, Bool, Arr } kind = Int;
  double num = 0.0;
  long long ival = 0;
  std::string str;
  std::vector<TVal> arr;
};
using TTable = std::map<std::string, TVal>;
struct TDoc {
  std::map<std::string, TTable> tables;
  std::map<std::string, std::vector<TTable>> arrays;
};

struct Parser {
  const std::string &src;
  size_t pos = 0;
  int line = 1;
  explicit Parser(const std::string &s) : src(s) {}

  [[noreturn]] void fail(const std::string &m) {
    throw SceneError(PYRO_SCENE_E_PARSE, "scene file format error: line " + std::to_string(line) + ": " + m);
  }
  bool eof() const { return pos >= src.size(); }
  char peek() const { return eof() ? '\0' : src[pos]; }
  void skip_ws() {
    while (!eof() && (src[pos] == ' ' || src[pos] == '\t' || src[pos] == '\r')) pos++;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!eof() && src[pos] != '\n') pos++;
  }
  // whitespace, newlines and comments (inside arrays)
  void skip_all() {
    for (;;) {
      skip_ws();
      if (peek() == '#') { skip_comment(); continue; }
      if (peek() == '\n') { pos++; line++; continue; }
      break;
    }
  }
  void end_of_line() {
    skip_ws();
    skip_comment();
    if (eof()) return;
    if (peek() != '\n') fail("unexpected trailing characters");
    pos++;
    line++;
  }
  std::string key() {
    skip_ws();
    std::string k;
    if (peek() == '"') return basic_string();
    while (!eof()) {
      char c = src[pos];
      if (std::isalnum((unsigned char)c) || c == '_' || c == '-') { k += c; pos++; }
      else break;
    }
    if (k.empty()) fail("expected a key");
    return k;
  }
  std::string basic_string() {
    pos++;  // opening quote
    std::string s;
    while (!eof() && src[pos] != '"') {
      char c = src[pos++];
      if (c == '\n') fail("newline in string");
      if (c == '\\') {
        if (eof()) fail("bad escape");
        char e = src[pos++];
        switch (e) {
          case 'n': s += '\n'; break;
          case 't': s += '\t'; break;
          case '"': s += '"'; break;
          case '\\': s += '\\'; break;
          default: fail("unsupported escape");
        }
      } else {
        s += c;
      }
    }
    if (eof()) fail("unterminated string");
    pos++;
    return s;
  }
  TVal value() {
    skip_ws();
    TVal v;
    char c = peek();
    if (c == '"') { v.kind = TVal::Str; v.str = basic_string(); return v; }
    if (c == '[') {
      pos++;
      v.kind = TVal::Arr;
      skip_all();
      if (peek() == ']') { pos++; return v; }
      for (;;) {
        skip_all();
        v.arr.push_back(value());
        skip_all();
        if (peek() == ',') { pos++; skip_all(); if (peek() == ']') { pos++; return v; } continue; }
        if (peek() == ']') { pos++; return v; }
        fail("expected , or ] in array");
      }
    }
    if (src.compare(pos, 4, "true") == 0) { pos += 4; v.kind = TVal::Bool; v.ival = 1; return v; }
    if (src.compare(pos, 5, "false") == 0) { pos += 5; v.kind = TVal::Bool; v.ival = 0; return v; }
    std::string t;
    while (!eof()) {
      char d = src[pos];
      if (std::isdigit((unsigned char)d) || d == '+' || d == '-' || d == '.' || d == 'e' || d == 'E' || d == '_') {
        if (d != '_') t += d;
        pos++;
      } else {
        break;
      }
    }
    if (t.empty()) fail("expected a value");
    bool is_float = t.find_first_of(".eE") != std::string::npos;
    char *endp = nullptr;
    if (is_float) {
      v.kind = TVal::Float;
      v.num = std::strtod(t.c_str(), &endp);
    } else {
      v.kind = TVal::Int;
      v.ival = std::strtoll(t.c_str(), &endp, 10);
      v.num = (double)v.ival;
    }
    if (!endp || *endp != '\0') fail("invalid number '" + t + "'");
    return v;
  }
  TDoc parse() {
    TDoc doc;
    TTable *cur = &doc.tables[""];
    while (!eof()) {
      skip_ws();
      char c = peek();
      if (c == '\n') { pos++; line++; continue; }
      if (c == '#') { skip_comment(); continue; }
      if (c == '\0') break;
      if (c == '[') {
        bool arr = src.compare(pos, 2, "[[") == 0;
        pos += arr ? 2 : 1;
        std::string name = key();
        skip_ws();
        if (arr) {
          if (src.compare(pos, 2, "]]") != 0) fail("expected ]]");
          pos += 2;
          if (doc.tables.count(name)) fail("table " + name + " redefined as array");
          doc.arrays[name].emplace_back();
          cur = &doc.arrays[name].back();
        } else {
          if (peek() != ']') fail("expected ]");
          pos++;
          if (doc.tables.count(name) || doc.arrays.count(name)) fail("duplicate table [" + name + "]");
          cur = &doc.tables[name];
        }
        end_of_line();
        continue;
      }
      std::string k = key();
      skip_ws();
      if (peek() != '=') fail("expected = after key " + k);
      pos++;
      TVal v = value();
      if (cur->count(k)) fail("duplicate key " + k);
      (*cur)[k] = v;
      end_of_line();
    }
    return doc;
  }
};

// ----------------------------------------------------------------------------
// Schema (scene.rs:8-215)
// ----------------------------------------------------------------------------
struct V3 { float v[3] = {0.f, 0.f, 0.f}; };

struct Object {
  std::string kind, material;
  V3 center, size, a, b;
  float radius = 0.f, depth = 0.f, yaw = 0.f, stick = 0.f;
  uint32_t count = 0, sticks = 0;
};
struct Ignition {
  V3 pos;
  float radius = 0.25f, flux = 40e3f, temp = 1300.f, until = 4.0f, fuel = 0.f;
};
struct Scene {
  V3 domain_size;
  float cell = 0.f;
  // SimCfg defaults scene.rs:47-60
  float fps = 24.f, duration = 10.f, cfl = 1.5f, vorticity = 3.f, smagorinsky = 0.17f;
  uint32_t pressure_cycles = 4, seed = 1;
  float ambient_temp = 293.f;
  float wind_speed = 0.f, wind_dir = 0.f;
  bool has_ign = false;
  Ignition ign;
  bool fuel_material_set = false;
  std::string fuel_material;
  bool has_dhc = false, has_soot = false, has_ratio = false, has_ox = false, has_rf = false;
  float fuel_dhc = 0, fuel_soot = 0, fuel_ratio = 0, fuel_ox = 0, fuel_rf = 0;
  std::vector<Object> objects;
  float preroll = 0.f;  // OutputCfg (scene.rs:193-209)
};

[[noreturn]] void type_err(const std::string &key, const char *want) {
  throw SceneError(PYRO_SCENE_E_PARSE, "scene file format error: invalid type for `" + key + "`, expected " + want);
}
float as_f32(const TVal &v, const std::string &key) {
  if (v.kind != TVal::Float && v.kind != TVal::Int) type_err(key, "f32");
  return (float)v.num;  // serde: f64 -> f32 / i64 -> f32 via `as`
}
uint32_t as_u32(const TVal &v, const std::string &key) {
  if (v.kind != TVal::Int) type_err(key, "u32");
  if (v.ival < 0 || v.ival > 0xffffffffLL) type_err(key, "u32 in range");
  return (uint32_t)v.ival;
}
std::string as_str(const TVal &v, const std::string &key) {
  if (v.kind != TVal::Str) type_err(key, "a string");
  return v.str;
}
V3 as_v3(const TVal &v, const std::string &key) {
  if (v.kind != TVal::Arr || v.arr.size() != 3) type_err(key, "an array of 3 numbers");
  V3 r;
  for (int i = 0; i < 3; i++) r.v[i] = as_f32(v.arr[i], key);
  return r;
}
template <class F>
void get(const TTable &t, const char *k, F f) {
  auto it = t.find(k);
  if (it != t.end()) f(it->second, std::string(k));
}
[[noreturn]] void missing(const std::string &what) {
  throw SceneError(PYRO_SCENE_E_SCHEMA, "scene file format error: missing field `" + what + "`");
}

Scene from_doc(const TDoc &doc) {
  Scene s;
  auto dt = doc.tables.find("domain");
  if (dt == doc.tables.end()) missing("domain");
  if (!dt->second.count("size")) missing("size");
  if (!dt->second.count("cell")) missing("cell");
  s.domain_size = as_v3(dt->second.at("size"), "size");
  s.cell = as_f32(dt->second.at("cell"), "cell");
  auto st = doc.tables.find("sim");
  if (st != doc.tables.end()) {
    const TTable &t = st->second;
    get(t, "fps", [&](const TVal &v, const std::string &k) { s.fps = as_f32(v, k); });
    get(t, "duration", [&](const TVal &v, const std::string &k) { s.duration = as_f32(v, k); });
    get(t, "cfl", [&](const TVal &v, const std::string &k) { s.cfl = as_f32(v, k); });
    get(t, "vorticity", [&](const TVal &v, const std::string &k) { s.vorticity = as_f32(v, k); });
    get(t, "smagorinsky", [&](const TVal &v, const std::string &k) { s.smagorinsky = as_f32(v, k); });
    get(t, "pressure_cycles", [&](const TVal &v, const std::string &k) { s.pressure_cycles = as_u32(v, k); });
    get(t, "seed", [&](const TVal &v, const std::string &k) { s.seed = as_u32(v, k); });
    get(t, "ambient_temp", [&](const TVal &v, const std::string &k) { s.ambient_temp = as_f32(v, k); });
  }
  auto wt = doc.tables.find("wind");
  if (wt != doc.tables.end()) {
    get(wt->second, "speed", [&](const TVal &v, const std::string &k) { s.wind_speed = as_f32(v, k); });
    get(wt->second, "dir", [&](const TVal &v, const std::string &k) { s.wind_dir = as_f32(v, k); });
  }
  auto it = doc.tables.find("ignition");
  if (it != doc.tables.end()) {
    const TTable &t = it->second;
    s.has_ign = true;
    if (!t.count("pos")) missing("pos");
    s.ign.pos = as_v3(t.at("pos"), "pos");
    get(t, "radius", [&](const TVal &v, const std::string &k) { s.ign.radius = as_f32(v, k); });
    get(t, "flux", [&](const TVal &v, const std::string &k) { s.ign.flux = as_f32(v, k); });
    get(t, "temp", [&](const TVal &v, const std::string &k) { s.ign.temp = as_f32(v, k); });
    get(t, "until", [&](const TVal &v, const std::string &k) { s.ign.until = as_f32(v, k); });
    get(t, "fuel", [&](const TVal &v, const std::string &k) { s.ign.fuel = as_f32(v, k); });
  }
  auto ft = doc.tables.find("fuel");
  if (ft != doc.tables.end()) {
    const TTable &t = ft->second;
    get(t, "material", [&](const TVal &v, const std::string &k) { s.fuel_material = as_str(v, k); s.fuel_material_set = true; });
    get(t, "heat_of_combustion", [&](const TVal &v, const std::string &k) { s.fuel_dhc = as_f32(v, k); s.has_dhc = true; });
    get(t, "soot_yield", [&](const TVal &v, const std::string &k) { s.fuel_soot = as_f32(v, k); s.has_soot = true; });
    get(t, "soot_formation_ratio", [&](const TVal &v, const std::string &k) { s.fuel_ratio = as_f32(v, k); s.has_ratio = true; });
    get(t, "soot_oxidation", [&](const TVal &v, const std::string &k) { s.fuel_ox = as_f32(v, k); s.has_ox = true; });
    get(t, "radiative_fraction", [&](const TVal &v, const std::string &k) { s.fuel_rf = as_f32(v, k); s.has_rf = true; });
  }
  auto ot = doc.tables.find("output");
  if (ot != doc.tables.end())
    get(ot->second, "preroll", [&](const TVal &v, const std::string &k) { s.preroll = as_f32(v, k); });
  auto oa = doc.arrays.find("object");
  if (oa != doc.arrays.end())
    for (const TTable &t : oa->second) {
      Object o;
      if (!t.count("kind")) missing("kind");
      if (!t.count("material")) missing("material");
      o.kind = as_str(t.at("kind"), "kind");
      o.material = as_str(t.at("material"), "material");
      get(t, "center", [&](const TVal &v, const std::string &k) { o.center = as_v3(v, k); });
      get(t, "size", [&](const TVal &v, const std::string &k) { o.size = as_v3(v, k); });
      get(t, "a", [&](const TVal &v, const std::string &k) { o.a = as_v3(v, k); });
      get(t, "b", [&](const TVal &v, const std::string &k) { o.b = as_v3(v, k); });
      get(t, "radius", [&](const TVal &v, const std::string &k) { o.radius = as_f32(v, k); });
      get(t, "depth", [&](const TVal &v, const std::string &k) { o.depth = as_f32(v, k); });
      get(t, "yaw", [&](const TVal &v, const std::string &k) { o.yaw = as_f32(v, k); });
      get(t, "count", [&](const TVal &v, const std::string &k) { o.count = as_u32(v, k); });
      get(t, "sticks", [&](const TVal &v, const std::string &k) { o.sticks = as_u32(v, k); });
      get(t, "stick", [&](const TVal &v, const std::string &k) { o.stick = as_f32(v, k); });
      s.objects.push_back(o);
    }
  return s;
}

// ----------------------------------------------------------------------------
// Material table scene.rs:240-406 (name, kind, density, lg, tp, char_yield,
// emissivity, cp, k, smoke, dhc, soot)
// ----------------------------------------------------------------------------
struct Material {
  const char *name, *label;
  uint32_t kind;
  float density, lg, tp, char_yield, emissivity, cp, k, smoke, dhc, soot;
};
const Material MATERIALS[] = {
    {"air", "air", 2, 0.0f, 1.0f, 1e4f, 0.0f, 0.9f, 1000.0f, 0.03f, 0.0f, 0.0f, 0.0f},
    {"wood", "wood", 0, 480.0f, 1.5e6f, 600.0f, 0.25f, 0.9f, 1500.0f, 0.15f, 0.02f, 15.0e6f, 0.015f},
    {"cardboard", "cardboard", 0, 150.0f, 1.8e6f, 580.0f, 0.2f, 0.9f, 1400.0f, 0.08f, 0.02f, 14.0e6f, 0.012f},
    {"pu_foam", "polyurethane foam (sofa/mattress)", 3, 35.0f, 1.4e6f, 570.0f, 0.0f, 0.9f, 1400.0f, 0.04f, 0.0f, 24.0e6f, 0.10f},
    {"pmma", "PMMA acrylic", 3, 1190.0f, 1.6e6f, 630.0f, 0.0f, 0.9f, 1500.0f, 0.19f, 0.0f, 24.9e6f, 0.022f},
    {"rubber", "rubber (tyres)", 3, 1100.0f, 1.7e6f, 620.0f, 0.0f, 0.95f, 1700.0f, 0.2f, 0.0f, 30.0e6f, 0.15f},
    {"heptane", "n-heptane pool", 1, 684.0f, 0.48e6f, 371.0f, 0.0f, 0.95f, 2240.0f, 0.14f, 0.0f, 44.6e6f, 0.037f},
    {"gasoline", "gasoline pool", 1, 740.0f, 0.42e6f, 345.0f, 0.0f, 0.95f, 2200.0f, 0.13f, 0.0f, 43.7e6f, 0.059f},
    {"diesel", "diesel pool", 1, 830.0f, 0.6e6f, 460.0f, 0.0f, 0.95f, 2100.0f, 0.13f, 0.0f, 44.4e6f, 0.07f},
    {"steel", "steel (non-combustible)", 2, 7850.0f, 1.0f, 1e4f, 0.0f, 0.7f, 500.0f, 45.0f, 0.0f, 0.0f, 0.0f},
    {"concrete", "concrete (non-combustible)", 2, 2300.0f, 1.0f, 1e4f, 0.0f, 0.9f, 900.0f, 1.4f, 0.0f, 0.0f, 0.0f},
};
constexpr int N_MATERIALS = sizeof(MATERIALS) / sizeof(MATERIALS[0]);

// scene.rs:408-421
int material(const std::string &name) {
  for (int i = 0; i < N_MATERIALS; i++)
    if (name == MATERIALS[i].name) return i;
  std::string all;
  for (int i = 1; i < N_MATERIALS; i++) {
    if (i > 1) all += ", ";
    all += std::string(MATERIALS[i].name) + " (" + MATERIALS[i].label + ")";
  }
  throw SceneError(PYRO_SCENE_E_MATERIAL, "unknown material " + name + ", choices: " + all);
}

struct Params {
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
  float mats[8][12];
};
static_assert(sizeof(Params) == 576, "Params must match scene.rs:426-477");

// Rust f32::to_radians: self * (PI_f32 / 180.0_f32)
inline float to_radians(float d) {
  const float RADS_PER_DEG = 3.14159265358979323846f / 180.0f;
  return d * RADS_PER_DEG;
}

// scene.rs:485-493
void grid(const Scene &s, uint32_t dims[3], float *dx) {
  float c = s.cell;
  for (int i = 0; i < 3; i++) {
    float q = std::round((s.domain_size.v[i] / c) / 16.0f);
    if (!(q >= 1.0f)) q = 1.0f;  // f32::max(1.0)
    dims[i] = (uint32_t)q * 16u;
  }
  *dx = c;
}

// scene.rs:496-508
std::vector<int> used_materials(const Scene &s) {
  std::vector<int> v{0};
  for (const Object &o : s.objects) {
    int i = material(o.material);
    bool seen = false;
    for (int u : v) seen |= (u == i);
    if (!seen) v.push_back(i);
  }
  if (v.size() > 8) throw SceneError(PYRO_SCENE_E_MATERIAL, "a scene may use at most 7 materials");
  return v;
}

// scene.rs:510-521
const Material &fuel_material(const Scene &s) {
  if (s.fuel_material_set) return MATERIALS[material(s.fuel_material)];
  for (const Object &o : s.objects) {
    const Material &m = MATERIALS[material(o.material)];
    if (m.kind != 2) return m;
  }
  throw SceneError(PYRO_SCENE_E_MATERIAL, "the scene has no combustible object");
}

// scene.rs:523-618
Params params(const Scene &s) {
  uint32_t dims[3];
  float dx;
  grid(s, dims, &dx);
  const Material &fm = fuel_material(s);
  float dhc = s.has_dhc ? s.fuel_dhc : fm.dhc;
  float soot = s.has_soot ? s.fuel_soot : fm.soot;
  float ratio = s.has_ratio ? s.fuel_ratio : 2.5f;
  Params p;
  std::memset(&p, 0, sizeof(p));
  std::vector<int> used = used_materials(s);
  for (size_t slot = 0; slot < used.size(); slot++) {
    const Material &m = MATERIALS[used[slot]];
    float row[12] = {m.density, m.lg, m.tp, (float)m.kind, m.char_yield, m.emissivity,
                     m.cp, m.k, m.smoke, 1.0f, 0.0f, 0.0f};
    std::memcpy(p.mats[slot], row, sizeof(row));
  }
  float wr = to_radians(s.wind_dir);
  float span = 0.5f;
  for (const Object &o : s.objects) {
    float sp;
    if (o.kind == "pool" || o.kind == "sphere") sp = 2.0f * o.radius;
    else sp = std::fmax(std::fmax(o.size.v[0], o.size.v[2]), 0.3f);
    span = std::fmax(span, sp);
  }
  p.nx = dims[0]; p.ny = dims[1]; p.nz = dims[2]; p.frame = 0;
  p.dx = dx; p.dt = 0.005f; p.time = 0.0f; p.t0 = s.ambient_temp;
  p.rho0 = 1.204f * 293.0f / s.ambient_temp;
  p.g = 9.81f; p.cp = 1100.0f; p.yo2 = 0.233f;
  p.vort = s.vorticity; p.cs = s.smagorinsky; p.buoy = 1.0f; p.smoke_w = 0.0f;
  p.wind_x = s.wind_speed * std::cos(wr);
  p.wind_z = s.wind_speed * std::sin(wr);
  p.wind_relax = s.wind_speed > 0.0f ? 0.3f : 0.0f;
  p.sponge = 6.0f; p.t_ign = 750.0f; p.s_o2 = dhc / 13.1e6f; p.dhc = dhc; p.tau_min = 2e-3f;
  p.soot_form = soot * ratio;
  p.soot_ox = s.has_ox ? s.fuel_ox : 4.0f;
  p.soot_ox_t = 1150.0f; p.smoke_conv = 0.5f; p.kap_gas = 0.1f; p.h_conv = 15.0f;
  p.ign_on = s.has_ign ? 1.0f : 0.0f;
  p.ign_temp = s.has_ign ? s.ign.temp : 0.0f;
  p.ign_x = s.has_ign ? s.ign.pos.v[0] : 0.0f;
  p.ign_y = s.has_ign ? s.ign.pos.v[1] : 0.0f;
  p.ign_z = s.has_ign ? s.ign.pos.v[2] : 0.0f;
  p.ign_r = s.has_ign ? s.ign.radius : 0.0f;
  p.ign_flux = s.has_ign ? s.ign.flux : 0.0f;
  p.ign_fuel = s.has_ign ? s.ign.fuel : 0.0f;
  p.char_a = 465.0f; p.char_e = 68e3f; p.dh_char = 32.8e6f; p.burnout = 2.0f; p.layer = 0.005f;
  p.rad_len = 0.5f * span; p.h_mass = 0.02f; p.vort_t = 500.0f;
  float rf;
  if (s.has_rf) rf = s.fuel_rf;
  else {
    std::string n = fm.name;
    if (n == "heptane") rf = 0.33f;
    else if (n == "gasoline" || n == "diesel") rf = 0.35f;
    else if (n == "pu_foam" || n == "rubber") rf = 0.42f;
    else if (n == "pmma") rf = 0.31f;
    else rf = 0.30f;
  }
  p.rad_frac = rf;
  p.seed = (float)s.seed;
  return p;
}

// scene.rs:662-667
float hash(uint32_t v) {
  uint32_t x = v * 747796405u + 2891336453u;
  x = ((x >> ((x >> 28) + 4)) ^ x) * 277803737u;
  x = (x >> 22) ^ x;
  return (float)x / (float)4294967295u;
}

// ----------------------------------------------------------------------------
// Shapes scene.rs:670-879
// ----------------------------------------------------------------------------
struct Shape {
  enum Kind { Box, Cyl, Sphere, Pool } kind;
  float c[3], h[3], yaw;   // Box / Sphere / Pool centre
  float a[3], b[3], r;     // Cyl
  float depth;
};
inline void sincos_(float x, float *s, float *c) { *s = std::sin(x); *c = std::cos(x); }

bool inside(const Shape &sh, const float p[3], float dx) {
  switch (sh.kind) {
    case Shape::Box: {
      float s, co;
      sincos_(sh.yaw, &s, &co);
      float d[3] = {p[0] - sh.c[0], p[1] - sh.c[1], p[2] - sh.c[2]};
      float lx = co * d[0] + s * d[2];
      float lz = -s * d[0] + co * d[2];
      float hh[3];
      for (int i = 0; i < 3; i++) hh[i] = std::fmax(sh.h[i], 0.5f * dx);
      return std::fabs(lx) <= hh[0] && std::fabs(d[1]) <= hh[1] && std::fabs(lz) <= hh[2];
    }
    case Shape::Cyl: {
      float ba[3] = {sh.b[0] - sh.a[0], sh.b[1] - sh.a[1], sh.b[2] - sh.a[2]};
      float pa[3] = {p[0] - sh.a[0], p[1] - sh.a[1], p[2] - sh.a[2]};
      float l2 = ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2];
      float t = (pa[0] * ba[0] + pa[1] * ba[1] + pa[2] * ba[2]) / l2;
      if (!(t >= 0.0f && t <= 1.0f)) return false;
      float q[3] = {pa[0] - ba[0] * t, pa[1] - ba[1] * t, pa[2] - ba[2] * t};
      return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) <= std::fmax(sh.r, 0.5f * dx);
    }
    case Shape::Sphere: {
      float d[3] = {p[0] - sh.c[0], p[1] - sh.c[1], p[2] - sh.c[2]};
      return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= sh.r;
    }
    case Shape::Pool: {
      float ex = p[0] - sh.c[0], ez = p[2] - sh.c[2];
      float dxz = std::sqrt(ex * ex + ez * ez);
      return dxz <= sh.r && p[1] <= std::fmax(sh.depth, dx * 0.99f);
    }
  }
  return false;
}

Shape make_box(const float c[3], const float h[3], float yaw) {
  Shape s{};
  s.kind = Shape::Box;
  for (int i = 0; i < 3; i++) { s.c[i] = c[i]; s.h[i] = h[i]; }
  s.yaw = yaw;
  return s;
}
Shape make_cyl(const float a[3], const float b[3], float r) {
  Shape s{};
  s.kind = Shape::Cyl;
  for (int i = 0; i < 3; i++) { s.a[i] = a[i]; s.b[i] = b[i]; }
  s.r = r;
  return s;
}

// scene.rs:871-879
void rot(const float p[3], float yaw, const float c[3], float out[3]) {
  float s, co;
  sincos_(yaw, &s, &co);
  out[0] = c[0] + co * p[0] - s * p[2];
  out[1] = c[1] + p[1];
  out[2] = c[2] + s * p[0] + co * p[2];
}

// scene.rs:730-869
std::vector<Shape> expand(const Object &o, float dx) {
  float yaw = to_radians(o.yaw);
  const float *c = o.center.v;
  std::vector<Shape> v;
  if (o.kind == "box") {
    float h[3] = {o.size.v[0] * 0.5f, o.size.v[1] * 0.5f, o.size.v[2] * 0.5f};
    v.push_back(make_box(c, h, yaw));
  } else if (o.kind == "cylinder") {
    v.push_back(make_cyl(o.a.v, o.b.v, o.radius));
  } else if (o.kind == "sphere") {
    Shape s{};
    s.kind = Shape::Sphere;
    for (int i = 0; i < 3; i++) s.c[i] = c[i];
    s.r = o.radius;
    v.push_back(s);
  } else if (o.kind == "pool") {
    Shape s{};
    s.kind = Shape::Pool;
    for (int i = 0; i < 3; i++) s.c[i] = c[i];
    s.r = o.radius;
    s.depth = o.depth;
    v.push_back(s);
  } else if (o.kind == "crib") {
    uint32_t n = o.sticks > 2 ? o.sticks : 2;
    float w = std::fmax(o.stick, dx);
    float len = std::fmax(o.size.v[0], 0.2f);
    uint32_t cnt = o.count > 1 ? o.count : 1;
    for (uint32_t l = 0; l < cnt; l++) {
      float y = c[1] + w * ((float)l + 0.5f);
      for (uint32_t s = 0; s < n; s++) {
        float off = -0.5f * len + 0.5f * w + (len - w) * (float)s / (float)(n - 1);
        float hx, hz, ox, oz;
        if (l % 2 == 0) { hx = 0.5f * len; hz = 0.5f * w; ox = 0.0f; oz = off; }
        else { hx = 0.5f * w; hz = 0.5f * len; ox = off; oz = 0.0f; }
        float p[3] = {ox, 0.0f, oz}, cc[3] = {c[0], y, c[2]}, out[3];
        rot(p, yaw, cc, out);
        float h[3] = {hx, 0.5f * w, hz};
        v.push_back(make_box(out, h, yaw));
      }
    }
  } else if (o.kind == "pallets") {
    float ph = 0.144f;
    uint32_t cnt = o.count > 1 ? o.count : 1;
    for (uint32_t l = 0; l < cnt; l++) {
      float y0 = c[1] + ph * (float)l;
      const float zs[3] = {-0.45f, 0.0f, 0.45f};
      for (float z : zs) {
        float p[3] = {0.0f, 0.0f, z}, out[3];
        float c1[3] = {c[0], y0 + 0.011f, c[2]};
        rot(p, yaw, c1, out);
        float h1[3] = {0.6f, 0.011f, 0.05f};
        v.push_back(make_box(out, h1, yaw));
        float c2[3] = {c[0], y0 + 0.061f, c[2]};
        rot(p, yaw, c2, out);
        float h2[3] = {0.6f, 0.039f, 0.05f};
        v.push_back(make_box(out, h2, yaw));
      }
      for (int s = 0; s < 7; s++) {
        float x = -0.55f + 1.1f * (float)s / 6.0f;
        float p[3] = {x, 0.0f, 0.0f}, out[3];
        float c3[3] = {c[0], y0 + 0.122f, c[2]};
        rot(p, yaw, c3, out);
        float h3[3] = {0.05f, 0.011f, 0.5f};
        v.push_back(make_box(out, h3, yaw));
      }
    }
  } else if (o.kind == "bonfire") {
    float r = std::fmax(o.radius, 0.5f);
    float h = o.size.v[1] > 0.0f ? o.size.v[1] : 1.3f * r;
    uint32_t n = o.count > 6 ? o.count : 6;
    float d = o.stick > 0.0f ? o.stick : 0.08f * r;
    uint32_t k = 0;
    auto rnd = [&]() {
      k += 1;
      return hash((k * 2654435761u) ^ 0x5bd1e995u);
    };
    const float tau = 6.28318530717958647692f;
    const float PI = 3.14159265358979323846f;
    struct Ring { int ring; uint32_t cnt; float rr, hh; } rings[2] = {{0, n, r, h}, {1, n / 2, 0.6f * r, 0.75f * h}};
    for (const Ring &R : rings) {
      for (uint32_t i = 0; i < R.cnt; i++) {
        float th = tau * ((float)i + 0.5f * (float)R.ring) / (float)R.cnt + 0.25f * (rnd() - 0.5f);
        float rb = R.rr * (0.88f + 0.17f * rnd());
        float a[3] = {c[0] + rb * std::cos(th), c[1] + 0.5f * d, c[2] + rb * std::sin(th)};
        float tt = th + PI + 0.6f * (rnd() - 0.5f);
        float off = 0.07f * r * (0.5f + rnd());
        float top[3];
        top[0] = c[0] + off * std::cos(tt);
        top[1] = c[1] + R.hh * (0.9f + 0.15f * rnd());
        top[2] = c[2] + off * std::sin(tt);
        float ext = 1.08f + 0.1f * rnd();
        float b[3] = {a[0] + (top[0] - a[0]) * ext, a[1] + (top[1] - a[1]) * ext, a[2] + (top[2] - a[2]) * ext};
        float rad = 0.5f * d * (0.75f + 0.5f * rnd());
        v.push_back(make_cyl(a, b, rad));
      }
    }
    for (int l = 0; l < 3; l++) {
      float y = c[1] + d * ((float)l + 0.5f);
      for (int s = 0; s < 5; s++) {
        float off = -0.5f * r + r * (float)s / 4.0f + 0.1f * r * (rnd() - 0.5f);
        float len = 0.7f * r;
        float ang = (l % 2 == 0 ? 0.0f : 1.57079632679489661923f) + 0.3f * (rnd() - 0.5f);
        float sa, ca;
        sincos_(ang, &sa, &ca);
        float p0[3] = {c[0] - sa * off, y, c[2] + ca * off};
        float a[3] = {p0[0] - ca * len, y, p0[2] - sa * len};
        float b[3] = {p0[0] + ca * len, y, p0[2] + sa * len};
        float rad = 0.5f * d * (0.7f + 0.4f * rnd());
        v.push_back(make_cyl(a, b, rad));
      }
    }
  } else {
    throw SceneError(PYRO_SCENE_E_SCHEMA, "unknown object kind " + o.kind +
                                              " (choices: box/cylinder/sphere/pool/crib/pallets/bonfire)");
  }
  return v;
}

// scene.rs:621-660
void voxelize(const Scene &s, float *out) {
  uint32_t dims[3];
  float dx;
  grid(s, dims, &dx);
  size_t nx = dims[0], ny = dims[1], nz = dims[2];
  float t0 = s.ambient_temp;
  std::vector<int> used = used_materials(s);
  for (size_t i = 0; i < nx * ny * nz; i++) {
    out[4 * i] = 0.0f; out[4 * i + 1] = 0.0f; out[4 * i + 2] = 0.0f; out[4 * i + 3] = t0;
  }
  std::vector<std::pair<Shape, int>> shapes;
  for (const Object &o : s.objects) {
    int mi = material(o.material);
    int slot = 0;
    for (size_t u = 0; u < used.size(); u++)
      if (used[u] == mi) { slot = (int)u; break; }
    for (const Shape &sh : expand(o, dx)) shapes.emplace_back(sh, slot);
  }
  uint32_t seed = s.seed;
  for (size_t k = 0; k < nz; k++)
    for (size_t j = 0; j < ny; j++)
      for (size_t i = 0; i < nx; i++) {
        float p[3] = {((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx};
        for (const auto &sp : shapes) {
          if (!inside(sp.first, p, dx)) continue;
          const Material &m = MATERIALS[used[sp.second]];
          size_t idx = i + nx * (j + ny * k);
          float h = hash((uint32_t)idx ^ (seed * 0x9e3779b9u));
          float jit = 0.85f + 0.3f * h;
          float mass = m.kind == 2 ? 1.0f : m.density * jit;
          out[4 * idx] = (float)sp.second;
          out[4 * idx + 1] = mass;
          out[4 * idx + 2] = 0.0f;
          out[4 * idx + 3] = t0;
        }
      }
}

template <class F>
int guard(F f) {
  try {
    f();
    return 0;
  } catch (const SceneError &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    return PYRO_SCENE_E_ARG;
  }
}

}  // namespace

struct PyroScene {
  Scene s;
};

extern "C" {

const char *pyro_scene_last_error(void) { return g_err.c_str(); }

int pyro_scene_parse(const char *text, PyroScene **out) {
  if (!text || !out) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  return guard([&] {
    std::string src(text);
    Parser p(src);
    TDoc doc = p.parse();
    std::unique_ptr<PyroScene> sc(new PyroScene);
    sc->s = from_doc(doc);
    *out = sc.release();
  });
}

int pyro_scene_load(const char *path, PyroScene **out) {
  if (!path || !out) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  std::ifstream f(path, std::ios::binary);
  if (!f) { g_err = std::string("failed to read scene ") + path; return PYRO_SCENE_E_ARG; }
  std::stringstream ss;
  ss << f.rdbuf();
  std::string txt = ss.str();
  return pyro_scene_parse(txt.c_str(), out);
}

void pyro_scene_free(PyroScene *s) { delete s; }

int pyro_scene_get(const PyroScene *sc, const char *key, double *out) {
  if (!sc || !key || !out) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  const Scene &s = sc->s;
  std::string k(key);
  if (k == "domain.cell") *out = s.cell;
  else if (k == "sim.fps") *out = s.fps;
  else if (k == "sim.duration") *out = s.duration;
  else if (k == "sim.cfl") *out = s.cfl;
  else if (k == "sim.vorticity") *out = s.vorticity;
  else if (k == "sim.smagorinsky") *out = s.smagorinsky;
  else if (k == "sim.pressure_cycles") *out = s.pressure_cycles;
  else if (k == "sim.seed") *out = s.seed;
  else if (k == "sim.ambient_temp") *out = s.ambient_temp;
  else if (k == "ignition.until") *out = s.has_ign ? s.ign.until : 0.0;  // sim.rs:204
  else if (k == "output.preroll") *out = s.preroll;
  else if (k == "objects") *out = (double)s.objects.size();
  else if (k.rfind("object.", 0) == 0) {
    // object.<i>.radius, object.<i>.pool (1 if kind == "pool"; main.rs:264-272 uses both)
    const size_t dot = k.find('.', 7);
    size_t i = 0;
    bool ok = dot != std::string::npos && dot > 7;
    for (size_t c = 7; ok && c < dot; c++) ok = std::isdigit((unsigned char)k[c]) != 0;
    if (ok) i = (size_t)std::stoul(k.substr(7, dot - 7));
    if (!ok || i >= s.objects.size()) { g_err = "unknown scene key " + k; return PYRO_SCENE_E_ARG; }
    const std::string f = k.substr(dot + 1);
    if (f == "radius") *out = s.objects[i].radius;
    else if (f == "pool") *out = s.objects[i].kind == "pool" ? 1.0 : 0.0;
    else { g_err = "unknown scene key " + k; return PYRO_SCENE_E_ARG; }
  }
  else { g_err = "unknown scene key " + k; return PYRO_SCENE_E_ARG; }
  return 0;
}

int pyro_scene_set(PyroScene *sc, const char *key, double v) {
  if (!sc || !key) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  Scene &s = sc->s;
  std::string k(key);
  if (k == "domain.cell") s.cell = (float)v;
  else if (k == "sim.fps") s.fps = (float)v;
  else if (k == "sim.duration") s.duration = (float)v;
  else if (k == "sim.cfl") s.cfl = (float)v;
  else if (k == "sim.pressure_cycles") s.pressure_cycles = (uint32_t)v;
  else if (k == "sim.seed") s.seed = (uint32_t)v;
  else if (k == "output.preroll") s.preroll = (float)v;
  else { g_err = "scene key " + k + " cannot be overridden"; return PYRO_SCENE_E_ARG; }
  return 0;
}

int pyro_scene_has_ignition(const PyroScene *sc) { return sc && sc->s.has_ign ? 1 : 0; }

int pyro_scene_grid(const PyroScene *sc, uint32_t dims[3], float *dx) {
  if (!sc || !dims || !dx) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  grid(sc->s, dims, dx);
  return 0;
}

int pyro_scene_params(const PyroScene *sc, void *out, size_t bytes) {
  if (!sc || !out) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  if (bytes != sizeof(Params)) { g_err = "Params blob must be 576 bytes"; return PYRO_SCENE_E_ARG; }
  return guard([&] {
    Params p = params(sc->s);
    std::memcpy(out, &p, sizeof(p));
  });
}

int pyro_scene_voxelize(const PyroScene *sc, float *out, size_t n_floats) {
  if (!sc || !out) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  uint32_t d[3];
  float dx;
  grid(sc->s, d, &dx);
  if (n_floats != (size_t)d[0] * d[1] * d[2] * 4) { g_err = "voxelize: output must hold N*4 floats"; return PYRO_SCENE_E_ARG; }
  return guard([&] { voxelize(sc->s, out); });
}

int pyro_material_index(const char *name) {
  if (!name) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  int r = 0;
  int rc = guard([&] { r = material(name); });
  return rc ? rc : r;
}

float pyro_scene_hash(uint32_t v) { return hash(v); }

int pyro_scene_shapes(const PyroScene *sc, PyroShape *out, size_t cap, size_t *count) {
  if (!sc || !count) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  return guard([&] {
    const Scene &s = sc->s;
    uint32_t dims[3];
    float dx;
    grid(s, dims, &dx);
    std::vector<int> used = used_materials(s);
    std::vector<PyroShape> v;
    for (const Object &o : s.objects) {
      const int mi = material(o.material);
      int slot = 0;
      for (size_t u = 0; u < used.size(); u++)
        if (used[u] == mi) { slot = (int)u; break; }
      const Material &m = MATERIALS[used[slot]];
      for (const Shape &sh : expand(o, dx)) {
        PyroShape p{};
        p.kind = (uint32_t)sh.kind;
        p.slot = (uint32_t)slot;
        p.inert = m.kind == 2 ? 1u : 0u;
        p.density = m.density;
        for (int q = 0; q < 3; q++) { p.c[q] = sh.c[q]; p.h[q] = sh.h[q]; p.a[q] = sh.a[q]; p.b[q] = sh.b[q]; }
        sincos_(sh.yaw, &p.sin_yaw, &p.cos_yaw);
        p.r = sh.r;
        p.depth = sh.depth;
        v.push_back(p);
      }
    }
    *count = v.size();
    if (out && cap >= v.size() && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(PyroShape));
  });
}

int pyro_scene_voxel_spec(const PyroScene *sc, uint32_t dims[3], float *dx, float *t0, uint32_t *seed) {
  if (!sc || !dims || !dx || !t0 || !seed) { g_err = "null argument"; return PYRO_SCENE_E_ARG; }
  return guard([&] {
    grid(sc->s, dims, dx);
    *t0 = sc->s.ambient_temp;
    *seed = sc->s.seed;
  });
}

}  // extern "C"
