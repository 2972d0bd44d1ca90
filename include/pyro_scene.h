/*
 * pyro_scene.h -- C ABI of the host-side scene plumbing (CPU only).
 *
 * Restates the reference's scene/TOML layer (src/scene.rs) so that the HIP
 * solver can be driven without the Rust host.  In the reference build the
 * Rust scene.rs stays and produces the same Params blob and SOL array; this
 * library exists for the C++/Python hosts and the tests.
 *
 * Return convention: 0 = ok, negative = error (message from
 * pyro_scene_last_error(), thread-local).
 */
#ifndef PYRO_SCENE_H
#define PYRO_SCENE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PYRO_SCENE_E_PARSE (-10)  /* malformed TOML / wrong value type */
#define PYRO_SCENE_E_SCHEMA (-11) /* missing required field, unknown kind */
#define PYRO_SCENE_E_MATERIAL (-12) /* unknown material, >7 materials, no fuel */
#define PYRO_SCENE_E_ARG (-13)

typedef struct PyroScene PyroScene;

/* scene.rs:211-215 load(path): read + parse. */
int pyro_scene_load(const char *path, PyroScene **out);
/* toml::from_str::<Scene>(text) (scene.rs:213, tests scene.rs:885-890). */
int pyro_scene_parse(const char *toml_text, PyroScene **out);
void pyro_scene_free(PyroScene *s);
const char *pyro_scene_last_error(void);

/* Read / override a scalar scene field by dotted name, e.g. "domain.cell",
 * "sim.duration", "sim.fps", "sim.cfl", "sim.seed", "sim.pressure_cycles",
 * "ignition.until", "output.preroll" (the CLI overrides of main.rs:58-82). */
int pyro_scene_get(const PyroScene *s, const char *key, double *out);
int pyro_scene_set(PyroScene *s, const char *key, double value);
int pyro_scene_has_ignition(const PyroScene *s);

/* scene.rs:485-493 Scene::grid. */
int pyro_scene_grid(const PyroScene *s, uint32_t dims[3], float *dx);
/* scene.rs:523-618 Scene::params -> the 576-byte Params blob (scene.rs:426-477). */
int pyro_scene_params(const PyroScene *s, void *out, size_t bytes);
/* scene.rs:621-660 Scene::voxelize -> N x vec4 SOL = (slot, mass, char, Ts). */
int pyro_scene_voxelize(const PyroScene *s, float *out, size_t n_floats);

/* The primitives Scene::voxelize tests each cell against (scene.rs:670-879,
 * every object expanded), in order -- the last one containing a cell wins --
 * with the material's SOL slot and mass rule, and the yaw's sine and cosine
 * computed here, so that a device voxelizer (pyro_voxelize in pyro_hip.h)
 * reproduces pyro_scene_voxelize bit for bit with plain f32 arithmetic. */
typedef struct PyroShape {
  uint32_t kind;   /* 0 box (yawed about y), 1 cylinder (segment a-b), 2 sphere, 3 pool */
  uint32_t slot;   /* SOL slot of the material */
  uint32_t inert;  /* 1: mass 1 (non-combustible material); else density x jitter */
  float density;
  float c[3], h[3];        /* box / sphere / pool centre, box half sizes */
  float sin_yaw, cos_yaw;  /* box yaw */
  float a[3], b[3], r;     /* cylinder ends and radius; sphere / pool radius in r */
  float depth;             /* pool depth */
} PyroShape;
/* count = number of primitives; they are written to out when cap >= count. */
int pyro_scene_shapes(const PyroScene *s, PyroShape *out, size_t cap, size_t *count);
/* What voxelize needs besides the shapes: grid dims, cell size, ambient
 * temperature, seed of the mass jitter. */
int pyro_scene_voxel_spec(const PyroScene *s, uint32_t dims[3], float *dx, float *t0, uint32_t *seed);

/* scene.rs:408-421 material(name): index into the 11-entry table, or error
 * listing the choices. */
int pyro_material_index(const char *name);
/* scene.rs:662-667 hash(v) in [0,1]. */
float pyro_scene_hash(uint32_t v);

#ifdef __cplusplus
}
#endif
#endif
