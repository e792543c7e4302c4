// px_napi.cc - N-API addon: a thin binding of the C ABI (include/pixels.h)
// for the reference's TypeScript host (src/gpu/GpuWorld.ts, Simulation.ts).
// Every function maps 1:1 onto a px_* entry point; typed arrays are passed as
// host buffers.  Errors become thrown JS Errors carrying px_last_error().
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pixels.h"

namespace {

#define NAPI_CALL(env, call)                                   \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                          \
    }                                                          \
  } while (0)

napi_value throw_px(napi_env env, px_ctx* ctx, int rc) {
  std::string m = std::string("pixels error ") + std::to_string(rc) + ": " + px_last_error(ctx);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t n, napi_value* argv) {
  size_t argc = n;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t i = argc; i < n; i++) napi_get_undefined(env, &argv[i]);
  return true;
}

// A context handle: the external holds this box, so destroy() can clear it
// and any later call fails instead of touching freed memory; the finalizer
// frees the box (and the context, if the script never destroyed it).
struct CtxBox {
  px_ctx* ctx;
};

void finalize_box(napi_env, void* data, void*) {
  CtxBox* b = static_cast<CtxBox*>(data);
  if (b->ctx) px_destroy(b->ctx);
  delete b;
}

CtxBox* get_box(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a pixels context");
    return nullptr;
  }
  return static_cast<CtxBox*>(p);
}

px_ctx* get_ctx(napi_env env, napi_value v) {
  CtxBox* b = get_box(env, v);
  if (!b) return nullptr;
  if (!b->ctx) {
    napi_throw_error(env, nullptr, "the pixels context was destroyed");
    return nullptr;
  }
  return b->ctx;
}

// pointer + byte length of a TypedArray / ArrayBuffer / Buffer
bool get_bytes(napi_env env, napi_value v, void** data, size_t* bytes) {
  bool is_ta = false, is_ab = false;
  napi_is_typedarray(env, v, &is_ta);
  if (is_ta) {
    napi_typedarray_type t;
    size_t len = 0, off = 0;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &t, &len, data, &ab, &off) != napi_ok) return false;
    size_t el = 1;
    switch (t) {
      case napi_int16_array: case napi_uint16_array: el = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
      case napi_float64_array: el = 8; break;
      default: el = 1;
    }
    *bytes = len * el;
    return true;
  }
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) return napi_get_arraybuffer_info(env, v, data, bytes) == napi_ok;
  return false;
}

uint32_t get_u32(napi_env env, napi_value v, uint32_t dflt) {
  uint32_t x = dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_uint32(env, v, &x);
  return x;
}

double get_f64(napi_env env, napi_value v, double dflt) {
  double x = dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_double(env, v, &x);
  return x;
}

napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) napi_get_undefined(env, &v);
  return v;
}

napi_value new_typed(napi_env env, napi_typedarray_type t, size_t count, size_t el, void** data) {
  napi_value ab, out;
  if (napi_create_arraybuffer(env, count * el, data, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, count, ab, 0, &out) != napi_ok) return nullptr;
  return out;
}

napi_value wrap_ctx(napi_env env, px_ctx* ctx) {
  CtxBox* b = new CtxBox{ctx};
  napi_value out;
  if (napi_create_external(env, b, finalize_box, nullptr, &out) != napi_ok) {
    px_destroy(ctx);
    delete b;
    napi_throw_error(env, nullptr, "napi_create_external failed");
    return nullptr;
  }
  return out;
}

// create({dims, grid, pool, poolUsed, holes, mats, poolCapacity, voxelSize, device}) -> context
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  napi_value d = argv[0];
  px_world_desc desc;
  std::memset(&desc, 0, sizeof desc);
  napi_value dims = prop(env, d, "dims");
  for (uint32_t a = 0; a < 3; a++) {
    napi_value e;
    napi_get_element(env, dims, a, &e);
    desc.dims[a] = get_u32(env, e, 0);
  }
  void* p = nullptr;
  size_t nb = 0;
  const size_t ncell = (size_t)(desc.dims[0] / 8u) * (desc.dims[1] / 8u) * (desc.dims[2] / 8u);
  if (!get_bytes(env, prop(env, d, "grid"), &p, &nb) || nb < ncell * 4) {
    napi_throw_type_error(env, nullptr, "grid: Uint32Array of (dims/8) cells");
    return nullptr;
  }
  desc.grid = static_cast<const uint32_t*>(p);
  desc.pool_used = get_u32(env, prop(env, d, "poolUsed"), 0);
  size_t pool_bytes = 0;
  if (get_bytes(env, prop(env, d, "pool"), &p, &pool_bytes)) desc.pool = static_cast<const uint16_t*>(p);
  if (desc.pool_used && (!desc.pool || pool_bytes < (size_t)desc.pool_used * 1024)) {
    napi_throw_type_error(env, nullptr, "pool: Uint16Array of poolUsed * 512 voxels");
    return nullptr;
  }
  if (get_bytes(env, prop(env, d, "holes"), &p, &nb)) {
    desc.holes = static_cast<const uint32_t*>(p);
    desc.n_holes = (uint32_t)(nb / 4);
  }
  if (!get_bytes(env, prop(env, d, "mats"), &p, &nb) || nb < PX_MAT_COUNT * PX_MAT_FLOATS * 4) {
    napi_throw_type_error(env, nullptr, "mats: Float32Array(256*20)");
    return nullptr;
  }
  desc.mats = static_cast<const float*>(p);
  desc.pool_capacity = get_u32(env, prop(env, d, "poolCapacity"), 0);
  desc.voxel_size = (float)get_f64(env, prop(env, d, "voxelSize"), 1.0 / 80.0);
  desc.device = (int32_t)get_u32(env, prop(env, d, "device"), 0);
  desc.limits = nullptr;
  px_ctx* ctx = nullptr;
  int rc = px_create(&desc, &ctx);
  if (rc != PX_OK) return throw_px(env, nullptr, rc);
  return wrap_ctx(env, ctx);
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  CtxBox* b = get_box(env, argv[0]);
  if (!b) return nullptr;
  if (b->ctx) px_destroy(b->ctx);
  b->ctx = nullptr;  // later calls throw instead of using freed memory
  return nullptr;
}

napi_value SetUniforms(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb) || nb < PX_UNIFORM_FLOATS * 4) {
    napi_throw_type_error(env, nullptr, "uniforms: Float32Array(52)");
    return nullptr;
  }
  int rc = px_set_uniforms(c, static_cast<const float*>(p));
  return rc ? throw_px(env, c, rc) : nullptr;
}

napi_value SetLights(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb)) { napi_throw_type_error(env, nullptr, "lights: Float32Array"); return nullptr; }
  int rc = px_set_lights(c, static_cast<const float*>(p), (uint32_t)(nb / 16));
  return rc ? throw_px(env, c, rc) : nullptr;
}

// trace(ctx, kind, rays: ArrayBuffer of 32-byte records) -> Uint8Array of 16-byte hits
napi_value Trace(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  uint32_t kind = get_u32(env, argv[1], 0);
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[2], &p, &nb) || nb % sizeof(px_ray)) {
    napi_throw_type_error(env, nullptr, "rays: 32-byte records");
    return nullptr;
  }
  uint32_t n = (uint32_t)(nb / sizeof(px_ray));
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_uint8_array, (size_t)n * sizeof(px_hit), 1, &outp);
  if (!out) return nullptr;
  int rc = px_trace(c, kind, static_cast<const px_ray*>(p), n, static_cast<px_hit*>(outp));
  return rc ? throw_px(env, c, rc) : out;
}

// probe(ctx, [ox,oy,oz], [sx,sy,sz]) -> Uint32Array (Simulation.probe)
napi_value Probe(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int32_t o[3], s[3];
  for (uint32_t a = 0; a < 3; a++) {
    napi_value e;
    napi_get_element(env, argv[1], a, &e);
    napi_get_value_int32(env, e, &o[a]);
    napi_get_element(env, argv[2], a, &e);
    napi_get_value_int32(env, e, &s[a]);
  }
  if (s[0] < 0 || s[1] < 0 || s[2] < 0) { napi_throw_range_error(env, nullptr, "negative size"); return nullptr; }
  size_t n = (size_t)s[0] * s[1] * s[2];
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_uint32_array, n, 4, &outp);
  if (!out) return nullptr;
  int rc = px_probe(c, o, s, static_cast<uint32_t*>(outp));
  return rc ? throw_px(env, c, rc) : out;
}

// readCtrl(ctx) -> {ctrl: Uint32Array(36), args: Uint32Array(44), cam: Float32Array(64)}
napi_value ReadCtrl(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void *pc, *pa, *pm;
  napi_value ctrl = new_typed(env, napi_uint32_array, PX_CTRL_WORDS, 4, &pc);
  napi_value args = new_typed(env, napi_uint32_array, PX_ARGS_WORDS, 4, &pa);
  napi_value cam = new_typed(env, napi_float32_array, PX_CAMERA_FLOATS, 4, &pm);
  if (!ctrl || !args || !cam) return nullptr;
  int rc = px_read_ctrl(c, static_cast<uint32_t*>(pc), static_cast<uint32_t*>(pa), static_cast<float*>(pm));
  if (rc) return throw_px(env, c, rc);
  napi_value obj;
  NAPI_CALL(env, napi_create_object(env, &obj));
  napi_set_named_property(env, obj, "ctrl", ctrl);
  napi_set_named_property(env, obj, "args", args);
  napi_set_named_property(env, obj, "cam", cam);
  return obj;
}

napi_value PushImpacts(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb) || nb % 64) {
    napi_throw_type_error(env, nullptr, "impacts: Float32Array of 16-float records");
    return nullptr;
  }
  int rc = px_push_impacts(c, static_cast<const float*>(p), (uint32_t)(nb / 64));
  return rc ? throw_px(env, c, rc) : nullptr;
}

napi_value IslandPass(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int rc = px_island_pass(c);
  if (rc == PX_OK) rc = px_sync(c);
  return rc ? throw_px(env, c, rc) : nullptr;
}

// islandLabels(ctx) -> {keys: Uint32Array, counts: Uint32Array}
napi_value IslandLabels(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  uint32_t n = 0;
  int rc = px_island_labels(c, nullptr, nullptr, 0, &n);
  if (rc) return throw_px(env, c, rc);
  void *pk, *pc;
  napi_value keys = new_typed(env, napi_uint32_array, n, 4, &pk);
  napi_value counts = new_typed(env, napi_uint32_array, n, 4, &pc);
  if (!keys || !counts) return nullptr;
  uint32_t m = 0;
  rc = px_island_labels(c, static_cast<uint32_t*>(pk), static_cast<uint32_t*>(pc), n, &m);
  if (rc) return throw_px(env, c, rc);
  napi_value obj;
  NAPI_CALL(env, napi_create_object(env, &obj));
  napi_set_named_property(env, obj, "keys", keys);
  napi_set_named_property(env, obj, "counts", counts);
  return obj;
}

// createFromPrims({dims, mats, poolCapacity, voxelSize, device}, prims: 168-byte
// px_prim records) -> context: the scene rasterised on the GPU (shapes.ts:334-414)
napi_value CreateFromPrims(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  napi_value d = argv[0];
  px_world_desc desc;
  std::memset(&desc, 0, sizeof desc);
  napi_value dims = prop(env, d, "dims");
  for (uint32_t a = 0; a < 3; a++) {
    napi_value e;
    napi_get_element(env, dims, a, &e);
    desc.dims[a] = get_u32(env, e, 0);
  }
  void* p = nullptr;
  size_t nb = 0;
  if (!get_bytes(env, prop(env, d, "mats"), &p, &nb) || nb < PX_MAT_COUNT * PX_MAT_FLOATS * 4) {
    napi_throw_type_error(env, nullptr, "mats: Float32Array(256*20)");
    return nullptr;
  }
  desc.mats = static_cast<const float*>(p);
  desc.pool_capacity = get_u32(env, prop(env, d, "poolCapacity"), 0);
  desc.voxel_size = (float)get_f64(env, prop(env, d, "voxelSize"), 1.0 / 80.0);
  desc.device = (int32_t)get_u32(env, prop(env, d, "device"), 0);
  void* pp = nullptr;
  size_t pb = 0;
  if (!get_bytes(env, argv[1], &pp, &pb) || pb % sizeof(px_prim)) {
    napi_throw_type_error(env, nullptr, "prims: 168-byte px_prim records");
    return nullptr;
  }
  px_ctx* ctx = nullptr;
  int rc = px_create_from_prims(&desc, static_cast<const px_prim*>(pp), (uint32_t)(pb / sizeof(px_prim)), &ctx);
  if (rc != PX_OK) return throw_px(env, nullptr, rc);
  return wrap_ctx(env, ctx);
}

// simStep(ctx, stagesMask): Simulation.encode's stages in the reference's order
napi_value SimStep(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int rc = px_sim_step(c, get_u32(env, argv[1], 0));
  if (rc == PX_OK) rc = px_sync(c);
  return rc ? throw_px(env, c, rc) : nullptr;
}

napi_value ParticlesInit(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int rc = px_particles_init(c, get_u32(env, argv[1], 0));
  return rc ? throw_px(env, c, rc) : nullptr;
}

// spawn(ctx, records: 32-byte particle records)
napi_value Spawn(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb) || nb % sizeof(px_particle)) {
    napi_throw_type_error(env, nullptr, "particles: 32-byte records");
    return nullptr;
  }
  int rc = px_spawn(c, static_cast<const px_particle*>(p), (uint32_t)(nb / sizeof(px_particle)));
  if (rc == PX_OK) rc = px_sync(c);
  return rc ? throw_px(env, c, rc) : nullptr;
}

// readParticles(ctx, first, count) -> Uint8Array of 32-byte records
napi_value ReadParticles(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  const uint32_t first = get_u32(env, argv[1], 0), count = get_u32(env, argv[2], 0);
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_uint8_array, (size_t)count * sizeof(px_particle), 1, &outp);
  if (!out) return nullptr;
  int rc = px_read_particles(c, first, count, static_cast<px_particle*>(outp));
  return rc ? throw_px(env, c, rc) : out;
}

// readAlive(ctx, cap) -> Uint32Array of ring slots (the particle draw list)
napi_value ReadAlive(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  const uint32_t cap = get_u32(env, argv[1], 0);
  std::vector<uint32_t> buf(cap ? cap : 1);
  uint32_t n = 0;
  int rc = px_read_alive(c, buf.data(), cap, &n);
  if (rc) return throw_px(env, c, rc);
  const uint32_t m = n < cap ? n : cap;
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_uint32_array, m, 4, &outp);
  if (!out) return nullptr;
  if (m) std::memcpy(outp, buf.data(), (size_t)m * 4);
  return out;
}

// pushWrites(ctx, pairs: Uint32Array of (key, value))
napi_value PushWrites(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb) || nb % 8) {
    napi_throw_type_error(env, nullptr, "writes: Uint32Array of (key, value) pairs");
    return nullptr;
  }
  int rc = px_push_writes(c, static_cast<const uint32_t*>(p), (uint32_t)(nb / 8));
  return rc ? throw_px(env, c, rc) : nullptr;
}

// readChunks(ctx, arenaCap) -> {table: Uint32Array(maxChunks * 16), arena: Uint32Array}
napi_value ReadChunks(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  px_limits lim;
  px_default_limits(&lim);
  const uint32_t cap = get_u32(env, argv[1], lim.max_arena);
  void *pt, *pa;
  napi_value table = new_typed(env, napi_uint32_array, (size_t)lim.max_chunks * 16u, 4, &pt);
  std::vector<uint32_t> arena((size_t)cap * 2u + 2u);
  uint32_t n = 0;
  if (!table) return nullptr;
  int rc = px_read_chunks(c, static_cast<uint32_t*>(pt), arena.data(), cap, &n);
  if (rc) return throw_px(env, c, rc);
  const uint32_t m = n < cap ? n : cap;
  napi_value ar = new_typed(env, napi_uint32_array, (size_t)m * 2u, 4, &pa);
  if (!ar) return nullptr;
  if (m) std::memcpy(pa, arena.data(), (size_t)m * 8u);
  napi_value obj;
  NAPI_CALL(env, napi_create_object(env, &obj));
  napi_set_named_property(env, obj, "table", table);
  napi_set_named_property(env, obj, "arena", ar);
  return obj;
}

// render(ctx, rays: 32-byte records) -> Float32Array of RGBA per ray (fsMain)
napi_value Render(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* p;
  size_t nb;
  if (!get_bytes(env, argv[1], &p, &nb) || nb % sizeof(px_ray)) {
    napi_throw_type_error(env, nullptr, "rays: 32-byte records");
    return nullptr;
  }
  const uint32_t n = (uint32_t)(nb / sizeof(px_ray));
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_float32_array, (size_t)n * 4u, 4, &outp);
  if (!out) return nullptr;
  int rc = px_render(c, static_cast<const px_ray*>(p), n, static_cast<float*>(outp));
  return rc ? throw_px(env, c, rc) : out;
}

// readDensity(ctx, cells) -> Uint8Array (the AO density texture)
napi_value ReadDensity(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  px_ctx* c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  void* outp = nullptr;
  napi_value out = new_typed(env, napi_uint8_array, get_u32(env, argv[1], 0), 1, &outp);
  if (!out) return nullptr;
  int rc = px_read_density(c, static_cast<uint8_t*>(outp));
  return rc ? throw_px(env, c, rc) : out;
}

napi_value LastError(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, argv[0], &p);
  napi_value s;
  NAPI_CALL(env, napi_create_string_utf8(env, px_last_error(static_cast<px_ctx*>(p)), NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback fn;
  } fns[] = {
      {"create", Create}, {"destroy", Destroy}, {"setUniforms", SetUniforms}, {"setLights", SetLights},
      {"trace", Trace}, {"probe", Probe}, {"readCtrl", ReadCtrl}, {"pushImpacts", PushImpacts},
      {"islandPass", IslandPass}, {"islandLabels", IslandLabels}, {"lastError", LastError},
      {"createFromPrims", CreateFromPrims}, {"simStep", SimStep}, {"particlesInit", ParticlesInit},
      {"spawn", Spawn}, {"readParticles", ReadParticles}, {"readAlive", ReadAlive}, {"pushWrites", PushWrites},
      {"readChunks", ReadChunks}, {"render", Render}, {"readDensity", ReadDensity},
  };
  for (auto& f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
