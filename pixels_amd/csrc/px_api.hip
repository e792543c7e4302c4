// px_api.hip - the C ABI of libpixels_hip.so (declared in include/pixels.h).
//
// Context creation mirrors `new GpuWorld(device, world, scene)`
// (src/gpu/GpuWorld.ts:46-137): pool capacity rule (:60), super-grid counts
// (:90-100), free list = holes then the top of the pool in descending order
// (:103-113), C_FREE_TOP (:111-113), indirect draw args (:114-117).
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>

#include "px_internal.h"

static thread_local std::string g_create_error;

int px_fail(px_ctx* ctx, int code, const char* what, const char* detail) {
  std::string msg = std::string(what) + ": " + (detail ? detail : "");
  if (ctx) ctx->err = msg;
  else g_create_error = msg;
  return code;
}

static hipEvent_t take_event(px_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

PxProfScope::PxProfScope(px_ctx* ctx, int s) : c(ctx), slot(s) {
  if (!c->prof) return;
  a = take_event(c);
  if (a) hipEventRecord(a, c->stream);
}

PxProfScope::~PxProfScope() {
  if (!a) return;
  hipEvent_t b = take_event(c);
  if (!b) {
    c->ev_pool.push_back(a);
    return;
  }
  hipEventRecord(b, c->stream);
  c->prof_rec.push_back({slot, a, b});
}

extern "C" {

void px_default_limits(px_limits* l) {
  if (!l) return;
  // src/gpu/layout.ts:11-33
  l->max_dirty = 65535;
  l->max_alloc = 65535;
  l->max_writes = 131072;
  l->max_impacts = 64;
  l->max_jobs = 65535;
  l->max_flood = 131072;
  l->max_seeds = 262144;
  l->hash_size = 1u << 20;
  l->max_chunks = 256;
  l->max_arena = 1u << 20;
  l->max_lights = 32;
  l->max_anchors = 256;
  l->max_flashes = 8;
  l->spawn_budget = 60000;
  l->component_budget = 400000;
  l->search_iters_per_frame = 16;
  l->flood_iters_per_frame = 12;
  l->island_mode = PX_ISLANDS_ALL;
  l->uf_walk_cap = 0;
}

const char* px_last_error(const px_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_create_error.c_str();
}

static void release(px_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);  // no queued kernel still uses what is freed
  for (auto& r : c->prof_rec) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->batch_exec) hipGraphExecDestroy(c->batch_exec);
  void* bufs[] = {c->d_grid, c->d_pool, c->d_super, c->d_mats, c->d_ctrl, c->d_args, c->d_camera, c->d_free,
                  c->d_super_mask, c->d_brick_mask, c->d_super_bm, c->d_sub_mask, c->d_shadow_queue,
                  c->d_wave_count, c->d_counters,
                  c->d_batch_cursors, c->d_density, c->d_rprim, c->d_rtg, c->d_rsh, c->d_rrq, c->d_rrc,
                  c->b_grid, c->b_pool, c->b_super, c->b_ctrl, c->b_free};
  for (void* p : bufs)
    if (p) hipFree(p);
  px_edit_free(c);
  px_particles_free(c);
  px_writes_free(c);
  px_chunks_free(c);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

// Validation and the context shell (device, limits, dims, stream).
static int ctx_begin(const px_world_desc* d, const char* what, px_ctx** out) {
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (d->dims[a] == 0 || (d->dims[a] & 7u)) return px_fail(nullptr, PX_ERR_ARG, what, "dims must be positive multiples of 8");
  if (!d->mats) return px_fail(nullptr, PX_ERR_ARG, what, "missing mats");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return px_fail(nullptr, PX_ERR_NODEVICE, what, "no HIP device");
  if (d->device < 0 || d->device >= ndev) return px_fail(nullptr, PX_ERR_ARG, what, "bad device ordinal");
  uint64_t cells64 = (uint64_t)(d->dims[0] / 8) * (d->dims[1] / 8) * (d->dims[2] / 8);
  uint64_t vox64 = (uint64_t)d->dims[0] * d->dims[1] * d->dims[2];
  if (vox64 > 0xFFFFFFFFull) return px_fail(nullptr, PX_ERR_CAPACITY, what, "voxel keys must fit in 32 bits");
  px_ctx* c = new (std::nothrow) px_ctx();
  if (!c) return px_fail(nullptr, PX_ERR_ARG, what, "out of host memory");
  c->device = d->device;
  if (hipSetDevice(c->device) != hipSuccess) { release(c); return px_fail(nullptr, PX_ERR_HIP, what, "hipSetDevice"); }
  if (d->limits) c->lim = *d->limits;
  else px_default_limits(&c->lim);
  if (c->lim.island_mode > PX_ISLANDS_SEEDED) {
    release(c);
    return px_fail(nullptr, PX_ERR_ARG, what, "unknown island_mode");
  }
  c->vs = d->voxel_size > 0.0f ? d->voxel_size : 1.0f;
  for (int a = 0; a < 3; a++) {
    c->dims[a] = d->dims[a];
    c->bd[a] = d->dims[a] >> 3;
    c->sd[a] = (c->bd[a] + 3u) >> 2;
  }
  c->ncell = (uint32_t)cells64;
  c->nsuper = c->sd[0] * c->sd[1] * c->sd[2];
  c->vox_total = (size_t)vox64;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    release(c);
    return px_fail(nullptr, PX_ERR_HIP, what, "hipStreamCreateWithFlags");
  }
  // the fixed scratch of the no-allocation frame entry, once per device,
  // with the device current
  if (px_fixed_scratch_init(c) != PX_OK) {
    const std::string m = c->err;
    release(c);
    return px_fail(nullptr, PX_ERR_HIP, what, m.c_str());
  }
  *out = c;
  return PX_OK;
}

// World buffers for `used` pool bricks (capacity GpuWorld.ts:57-60), the
// material table, the counters and the free list (holes first, then the top
// of the pool descending, GpuWorld.ts:103-113).  Grid, pool contents and the
// super grid are the caller's.
static int ctx_world(px_ctx* c, const px_world_desc* d, uint32_t used, const uint32_t* holes, uint32_t n_holes,
                     const char* what) {
  // the storage-binding limit becomes 4 GiB of pool (4M bricks of 1 KiB), so
  // every pool byte offset fits the 32-bit offsets the traversal kernels use
  uint64_t max_bricks = (1ull << 32) / 1024ull;
  uint64_t cap = d->pool_capacity ? d->pool_capacity : std::max<uint64_t>((uint64_t)used + 98304ull, 2ull * used);
  if (cap > max_bricks) cap = max_bricks;
  if (used > cap) return px_fail(nullptr, PX_ERR_CAPACITY, what, "scene needs more bricks than the pool holds");
  c->pool_cap = (uint32_t)cap;
  c->pool_used = used;
#define TRY(expr)                                                    \
  do {                                                               \
    hipError_t e_ = (expr);                                          \
    if (e_ != hipSuccess) {                                          \
      std::string m_ = std::string(#expr) + ": " + hipGetErrorString(e_); \
      return px_fail(nullptr, PX_ERR_HIP, what, m_.c_str());         \
    }                                                                \
  } while (0)
  TRY(hipMalloc(&c->d_grid, (size_t)c->ncell * 4));
  TRY(hipMalloc(&c->d_pool, (size_t)c->pool_cap * 1024));
  TRY(hipMalloc(&c->d_super, (size_t)c->nsuper * 4));
  TRY(hipMalloc(&c->d_mats, PX_MAT_COUNT * PX_MAT_FLOATS * 4));
  TRY(hipMalloc(&c->d_ctrl, PX_CTRL_WORDS * 4));
  TRY(hipMalloc(&c->d_args, PX_ARGS_WORDS * 4));
  TRY(hipMalloc(&c->d_camera, PX_CAMERA_FLOATS * 4));
  TRY(hipMalloc(&c->d_free, (size_t)c->pool_cap * 4));
  TRY(hipMalloc(&c->d_brick_mask, (size_t)((c->ncell + 31u) / 32u) * 4 + 256));
  TRY(hipMalloc(&c->d_super_mask, (size_t)((c->nsuper + 31u) / 32u) * 4 + 256));
  TRY(hipMalloc(&c->d_super_bm, (size_t)c->nsuper * 8 + 256));
  TRY(hipMalloc(&c->d_sub_mask, (size_t)c->ncell * 8 + 256));
  TRY(hipMalloc(&c->d_counters, 256));
  TRY(hipMalloc(&c->d_batch_cursors, 8 * 256));
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0)
      c->n_cu = (uint32_t)ncu;
  }
  // ---- one-time upload (GpuWorld.ts:85-117) ----
  TRY(hipMemcpy(c->d_mats, d->mats, PX_MAT_COUNT * PX_MAT_FLOATS * 4, hipMemcpyHostToDevice));
  TRY(hipMemset(c->d_pool, 0, (size_t)c->pool_cap * 1024));
  std::vector<uint32_t> fl((size_t)n_holes + (c->pool_cap - used));
  for (uint32_t i = 0; i < n_holes; i++) fl[i] = holes[i];
  for (uint32_t i = 0; i < c->pool_cap - used; i++) fl[n_holes + i] = c->pool_cap - 1u - i;
  if (!fl.empty()) TRY(hipMemcpy(c->d_free, fl.data(), fl.size() * 4, hipMemcpyHostToDevice));
  uint32_t ctrl[PX_CTRL_WORDS] = {0};
  ctrl[PX_C_FREE_TOP] = (uint32_t)fl.size();
  TRY(hipMemcpy(c->d_ctrl, ctrl, sizeof ctrl, hipMemcpyHostToDevice));
  uint32_t args[PX_ARGS_WORDS] = {0};
  args[36] = 36;  // A_DRAWP
  args[40] = 36;  // A_DRAWC
  TRY(hipMemcpy(c->d_args, args, sizeof args, hipMemcpyHostToDevice));
  float cam[PX_CAMERA_FLOATS] = {0};
  TRY(hipMemcpy(c->d_camera, cam, sizeof cam, hipMemcpyHostToDevice));
#undef TRY
  return PX_OK;
}

// Uniform defaults, traversal masks, edit-chain buffers; waits for the uploads.
static int ctx_finish(px_ctx* c, const char* what) {
  // the uploads ran on the null stream; the context's stream does not wait for it
  if (hipDeviceSynchronize() != hipSuccess) return px_fail(nullptr, PX_ERR_HIP, what, "upload failed");
  c->h_uniforms[0] = (float)c->dims[0];
  c->h_uniforms[1] = (float)c->dims[1];
  c->h_uniforms[2] = (float)c->dims[2];
  c->h_uniforms[3] = c->vs;
  c->h_lights.assign(4, 0.0f);
  c->shadow.slots = 1;
  c->shadow.vs = c->vs;
  c->masks_dirty = c->density_dirty = true;
  int rc = px_launch_masks(c);
  if (rc == PX_OK) rc = px_edit_alloc(c);
  if (rc != PX_OK) return px_fail(nullptr, rc, what, c->err.c_str());
  if (hipStreamSynchronize(c->stream) != hipSuccess) return px_fail(nullptr, PX_ERR_HIP, what, "initial upload failed");
  return PX_OK;
}

int px_create(const px_world_desc* d, px_ctx** out) {
  if (!d || !out) return px_fail(nullptr, PX_ERR_ARG, "px_create", "null argument");
  *out = nullptr;
  if (!d->grid || (d->pool_used && !d->pool) || (d->n_holes && !d->holes))
    return px_fail(nullptr, PX_ERR_ARG, "px_create", "missing grid/pool/mats/holes");
  px_ctx* c = nullptr;
  if (int rc = ctx_begin(d, "px_create", &c)) return rc;
  int rc = ctx_world(c, d, d->pool_used, d->holes, d->n_holes, "px_create");
  if (rc == PX_OK) {
    bool ok = hipMemcpy(c->d_grid, d->grid, (size_t)c->ncell * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && d->pool_used)
      ok = hipMemcpy(c->d_pool, d->pool, (size_t)d->pool_used * 1024, hipMemcpyHostToDevice) == hipSuccess;
    std::vector<uint32_t> sc(c->nsuper, 0u);
    bool grid_ok = true;
    for (uint32_t z = 0; z < c->bd[2]; z++)
      for (uint32_t y = 0; y < c->bd[1]; y++)
        for (uint32_t x = 0; x < c->bd[0]; x++) {
          const uint32_t g = d->grid[x + c->bd[0] * (y + c->bd[1] * z)];
          if (g == 0u) continue;
          // a pool brick index must name an uploaded brick
          if ((g & PX_UNIFORM_BIT) == 0u && g - 1u >= d->pool_used) grid_ok = false;
          sc[(x >> 2) + c->sd[0] * ((y >> 2) + c->sd[1] * (z >> 2))]++;
        }
    for (uint32_t i = 0; i < d->n_holes; i++)
      if (d->holes[i] >= d->pool_used) grid_ok = false;
    if (ok) ok = hipMemcpy(c->d_super, sc.data(), (size_t)c->nsuper * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!grid_ok)
      rc = px_fail(nullptr, PX_ERR_ARG, "px_create", "the grid or a hole names a pool brick at or above pool_used");
    else if (!ok)
      rc = px_fail(nullptr, PX_ERR_HIP, "px_create", "world upload failed");
  }
  if (rc == PX_OK) rc = ctx_finish(c, "px_create");
  if (rc != PX_OK) {
    release(c);
    return rc;
  }
  *out = c;
  return PX_OK;
}

int px_create_from_prims(const px_world_desc* d, const px_prim* prims, uint32_t n, px_ctx** out) {
  if (!d || !out || (n && !prims)) return px_fail(nullptr, PX_ERR_ARG, "px_create_from_prims", "null argument");
  *out = nullptr;
  for (uint32_t k = 0; k < n; k++)
    if (prims[k].type > PX_PRIM_AABB)
      return px_fail(nullptr, PX_ERR_ARG, "px_create_from_prims", "unknown primitive type");
  px_ctx* c = nullptr;
  if (int rc = ctx_begin(d, "px_create_from_prims", &c)) return rc;
  int rc = PX_OK;
  {
    PxRaster R;
    uint32_t used = 0;
    rc = px_raster_begin(c, prims, n, &R, &used);
    if (rc != PX_OK) rc = px_fail(nullptr, rc, "px_create_from_prims", c->err.c_str());
    if (rc == PX_OK) rc = ctx_world(c, d, used, nullptr, 0, "px_create_from_prims");
    if (rc == PX_OK) {
      rc = px_raster_finish(c, &R);
      if (rc != PX_OK) rc = px_fail(nullptr, rc, "px_create_from_prims", c->err.c_str());
    }
  }
  if (rc == PX_OK) rc = ctx_finish(c, "px_create_from_prims");
  if (rc != PX_OK) {
    release(c);
    return rc;
  }
  *out = c;
  return PX_OK;
}

int px_destroy(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  release(c);
  return PX_OK;
}

static void decode_shadow(px_ctx* c) {
  PxShadowDesc& s = c->shadow;
  const float* u = c->h_uniforms;
  s.sun[0] = u[4]; s.sun[1] = u[5]; s.sun[2] = u[6]; s.sun_w = u[7];
  s.vs = u[3];
  s.lamp_shadows = u[21];
  float cf = c->h_lights.empty() ? 0.0f : c->h_lights[0];
  int count = (int)cf;
  if (count < 0) count = 0;
  int avail = c->h_lights.size() >= 4 ? (int)((c->h_lights.size() / 4 - 1) / 2) : 0;
  if (count > avail) count = avail;
  if (count > 32) count = 32;
  s.n_lights = (uint32_t)count;
  s.slots = 1u + (uint32_t)count;
  for (int i = 0; i < count; i++) {
    for (int a = 0; a < 4; a++) s.lpv[i][a] = c->h_lights[(1 + 2 * i) * 4 + a] / s.vs;
    s.lc_w[i] = c->h_lights[(2 + 2 * i) * 4 + 3];
  }
}

int px_set_uniforms(px_ctx* c, const float u[PX_UNIFORM_FLOATS]) {
  if (!c || !u) return PX_ERR_ARG;
  std::memcpy(c->h_uniforms, u, sizeof c->h_uniforms);
  decode_shadow(c);
  return PX_OK;
}

int px_set_lights(px_ctx* c, const float* lights, uint32_t n_vec4) {
  if (!c || (!lights && n_vec4)) return PX_ERR_ARG;
  if (n_vec4 == 0) c->h_lights.assign(4, 0.0f);
  else c->h_lights.assign(lights, lights + (size_t)n_vec4 * 4);
  if (c->h_lights[0] > 32.0f) return px_fail(c, PX_ERR_CAPACITY, "px_set_lights", "at most 32 lights");
  decode_shadow(c);
  return PX_OK;
}

uint32_t px_frame_slots(const px_ctx* c) { return c ? c->shadow.slots : 0u; }

int px_set_camera(px_ctx* c, const float cam[PX_CAMERA_FLOATS]) {
  if (!c || !cam) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_camera, cam, PX_CAMERA_FLOATS * 4, hipMemcpyHostToDevice, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_read_flashes(px_ctx* c, float out[PX_MAX_FLASHES * 8]) {
  if (!c || !out) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipMemcpyAsync(out, c->d_flashes, PX_MAX_FLASHES * 8 * 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_sync(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

void* px_stream(px_ctx* c) { return c ? (void*)c->stream : nullptr; }

int px_trace_device(px_ctx* c, uint32_t kind, const px_ray* d_rays, uint32_t n, px_hit* d_out) {
  if (!c || (n && (!d_rays || !d_out))) return PX_ERR_ARG;
  if (kind > 4u) return px_fail(c, PX_ERR_ARG, "px_trace_device", "unknown ray kind");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_trace(c, kind, d_rays, n, d_out);
}

int px_trace_perm_device(px_ctx* c, uint32_t kind, const px_ray* d_rays, const uint32_t* d_perm, uint32_t n,
                         px_hit* d_out) {
  if (!c || (n && (!d_rays || !d_perm || !d_out))) return PX_ERR_ARG;
  if (kind > 3u) return px_fail(c, PX_ERR_ARG, "px_trace_perm_device", "unknown ray kind");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_trace_perm(c, kind, d_rays, d_perm, n, d_out);
}

int px_trace(px_ctx* c, uint32_t kind, const px_ray* rays, uint32_t n, px_hit* out) {
  if (!c || (n && (!rays || !out))) return PX_ERR_ARG;
  if (kind > 4u) return px_fail(c, PX_ERR_ARG, "px_trace", "unknown ray kind");
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  px_ray* dr = nullptr;
  px_hit* dh = nullptr;
  PX_HIP_CHECK(c, hipMalloc(&dr, (size_t)n * sizeof(px_ray)));
  if (hipMalloc(&dh, (size_t)n * sizeof(px_hit)) != hipSuccess) {
    hipFree(dr);
    return px_fail(c, PX_ERR_HIP, "px_trace", "hipMalloc");
  }
  int rc = PX_OK;
  if (hipMemcpyAsync(dr, rays, (size_t)n * sizeof(px_ray), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = px_fail(c, PX_ERR_HIP, "px_trace", "upload");
  if (rc == PX_OK) rc = px_launch_trace(c, kind, dr, n, dh);
  if (rc == PX_OK && hipMemcpyAsync(out, dh, (size_t)n * sizeof(px_hit), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = px_fail(c, PX_ERR_HIP, "px_trace", "download");
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PX_OK) rc = px_fail(c, PX_ERR_HIP, "px_trace", "sync");
  hipFree(dr);
  hipFree(dh);
  return rc;
}

int px_trace_frame_device(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                          uint32_t* n_cast) {
  if (!c || (n && (!d_primary || !d_prim || !d_shadow))) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_frame(c, d_primary, n, d_prim, d_shadow, n_cast);
}

int px_trace_frame_device_noalloc(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim,
                                  px_hit* d_shadow, uint32_t* n_cast) {
  if (!c || (n && (!d_primary || !d_prim || !d_shadow))) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_frame_noalloc(c, d_primary, n, d_prim, d_shadow, n_cast);
}

int px_set_frame_layout(px_ctx* c, uint32_t layout) {
  if (!c) return PX_ERR_ARG;
  if (layout != PX_FRAME_PIXEL_MAJOR && layout != PX_FRAME_SLOT_PLANAR)
    return px_fail(c, PX_ERR_ARG, "px_set_frame_layout", "unknown layout");
  c->frame_layout = layout;
  return PX_OK;
}

int px_render_frame_device(px_ctx* c, const px_ray* d_rays, uint32_t n, float* d_rgba) {
  if (!c || (n && (!d_rays || !d_rgba))) return PX_ERR_ARG;
  if (reinterpret_cast<uintptr_t>(d_rgba) & 15u)
    return px_fail(c, PX_ERR_ARG, "px_render_frame_device", "d_rgba must be 16-byte aligned");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_render(c, d_rays, n, d_rgba);
}

int px_render(px_ctx* c, const px_ray* rays, uint32_t n, float* rgba) {
  if (!c || (n && (!rays || !rgba))) return PX_ERR_ARG;
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  px_ray* dr = nullptr;
  float* dc = nullptr;
  PX_HIP_CHECK(c, hipMalloc(&dr, (size_t)n * sizeof(px_ray)));
  if (hipMalloc(&dc, (size_t)n * 16u) != hipSuccess) {
    hipFree(dr);
    return px_fail(c, PX_ERR_HIP, "px_render", "hipMalloc");
  }
  int rc = PX_OK;
  if (hipMemcpyAsync(dr, rays, (size_t)n * sizeof(px_ray), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = px_fail(c, PX_ERR_HIP, "px_render", "upload");
  if (rc == PX_OK) rc = px_launch_render(c, dr, n, dc);
  if (rc == PX_OK && hipMemcpyAsync(rgba, dc, (size_t)n * 16u, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = px_fail(c, PX_ERR_HIP, "px_render", "download");
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PX_OK) rc = px_fail(c, PX_ERR_HIP, "px_render", "sync");
  hipFree(dr);
  hipFree(dc);
  return rc;
}

int px_read_density(px_ctx* c, uint8_t* out) {
  if (!c || !out) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  int rc = px_refresh_density(c);
  if (rc) return rc;
  PX_HIP_CHECK(c, hipMemcpyAsync(out, c->d_density, c->ncell, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

static int check_query_set(px_ctx* c, const px_query_set* q, const char* what) {
  if (!q) return px_fail(c, PX_ERR_ARG, what, "query set is NULL");
  uint64_t total = 0;
  for (int k = 0; k < 4; k++) {
    if (q->n[k] && (!q->rays[k] || !q->out[k])) return px_fail(c, PX_ERR_ARG, what, "NULL buffer for a non-empty kind");
    total += q->n[k];
  }
  if (total > 0xFFFFFFFFull) return px_fail(c, PX_ERR_CAPACITY, what, "more than 2^32-1 queries");
  return PX_OK;
}

int px_trace_batch(px_ctx* c, const px_query_set* q) {
  if (!c) return PX_ERR_ARG;
  int rc = check_query_set(c, q, "px_trace_batch");
  if (rc) return rc;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_batch(c, q, nullptr);
}

int px_trace_batch_perm(px_ctx* c, const px_query_set* q, const uint32_t* const d_perm[4]) {
  if (!c) return PX_ERR_ARG;
  int rc = check_query_set(c, q, "px_trace_batch_perm");
  if (rc) return rc;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_batch(c, q, d_perm);
}

int px_set_batch_shape(px_ctx* c, uint32_t block, uint32_t blocks, uint32_t chunk) {
  if (!c) return PX_ERR_ARG;
  if (block != 64u && block != 128u && block != 256u)
    return px_fail(c, PX_ERR_ARG, "px_set_batch_shape", "block must be 64, 128 or 256");
  if (chunk == 0u) return px_fail(c, PX_ERR_ARG, "px_set_batch_shape", "chunk must be at least 1");
  if (blocks > (1u << 20)) return px_fail(c, PX_ERR_ARG, "px_set_batch_shape", "blocks: at most 2^20");
  c->batch_block = block;
  c->batch_blocks = blocks;
  c->batch_chunk = chunk;
  return PX_OK;
}

int px_get_batch_shape(const px_ctx* c, uint32_t out[3]) {
  if (!c || !out) return PX_ERR_ARG;
  out[0] = c->batch_block;
  out[1] = c->batch_blocks ? c->batch_blocks : c->n_cu * 6u * 4u * 64u / c->batch_block;
  out[2] = c->batch_chunk;
  return PX_OK;
}

int px_autotune_batch(px_ctx* c, const px_query_set* q, uint32_t repeats, uint32_t out[3], float* best_ms) {
  if (!c || !out) return PX_ERR_ARG;
  int rc = check_query_set(c, q, "px_autotune_batch");
  if (rc) return rc;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_autotune_batch_impl(c, q, nullptr, repeats ? repeats : 3u, out, best_ms);
}

int px_autotune_batch_perm(px_ctx* c, const px_query_set* q, const uint32_t* const d_perm[4], uint32_t repeats,
                           uint32_t out[3], float* best_ms) {
  if (!c || !out) return PX_ERR_ARG;
  int rc = check_query_set(c, q, "px_autotune_batch_perm");
  if (rc) return rc;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_autotune_batch_impl(c, q, d_perm, repeats ? repeats : 3u, out, best_ms);
}

static bool box_ok(const px_ctx* c, const int32_t s[3]) {
  (void)c;
  return s[0] >= 0 && s[1] >= 0 && s[2] >= 0 && (uint64_t)s[0] * s[1] * s[2] <= (1ull << 34);
}

int px_probe_device(px_ctx* c, const int32_t o[3], const int32_t s[3], uint32_t* d_out) {
  if (!c || !o || !s) return PX_ERR_ARG;
  if (!box_ok(c, s)) return px_fail(c, PX_ERR_ARG, "px_probe_device", "bad box size");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_probe(c, o, s, d_out);
}

int px_probe(px_ctx* c, const int32_t o[3], const int32_t s[3], uint32_t* out) {
  if (!c || !o || !s || !out) return PX_ERR_ARG;
  if (!box_ok(c, s)) return px_fail(c, PX_ERR_ARG, "px_probe", "bad box size");
  size_t n = (size_t)s[0] * s[1] * s[2];
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  uint32_t* d = nullptr;
  PX_HIP_CHECK(c, hipMalloc(&d, n * 4));
  int rc = px_launch_probe(c, o, s, d);
  if (rc == PX_OK && hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = px_fail(c, PX_ERR_HIP, "px_probe", "download");
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PX_OK) rc = px_fail(c, PX_ERR_HIP, "px_probe", "sync");
  hipFree(d);
  return rc;
}

int px_read_ctrl(px_ctx* c, uint32_t* ctrl, uint32_t* args, float* cam) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (ctrl) PX_HIP_CHECK(c, hipMemcpyAsync(ctrl, c->d_ctrl, PX_CTRL_WORDS * 4, hipMemcpyDeviceToHost, c->stream));
  if (args) PX_HIP_CHECK(c, hipMemcpyAsync(args, c->d_args, PX_ARGS_WORDS * 4, hipMemcpyDeviceToHost, c->stream));
  if (cam) PX_HIP_CHECK(c, hipMemcpyAsync(cam, c->d_camera, PX_CAMERA_FLOATS * 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_read_grid(px_ctx* c, uint32_t* grid, uint32_t* sgrid) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (grid) PX_HIP_CHECK(c, hipMemcpyAsync(grid, c->d_grid, (size_t)c->ncell * 4, hipMemcpyDeviceToHost, c->stream));
  if (sgrid) PX_HIP_CHECK(c, hipMemcpyAsync(sgrid, c->d_super, (size_t)c->nsuper * 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_read_pool(px_ctx* c, uint32_t first, uint32_t count, uint32_t* words) {
  if (!c || (count && !words)) return PX_ERR_ARG;
  if ((uint64_t)first + count > c->pool_cap) return px_fail(c, PX_ERR_ARG, "px_read_pool", "range outside the pool");
  if (!count) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipMemcpyAsync(words, c->d_pool + (size_t)first * 256, (size_t)count * 1024, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_world_info(const px_ctx* c, uint32_t dims[3], uint32_t bricks[3], uint32_t supers[3], uint32_t* pool_capacity) {
  if (!c) return PX_ERR_ARG;
  for (int a = 0; a < 3; a++) {
    if (dims) dims[a] = c->dims[a];
    if (bricks) bricks[a] = c->bd[a];
    if (supers) supers[a] = c->sd[a];
  }
  if (pool_capacity) *pool_capacity = c->pool_cap;
  return PX_OK;
}

int px_push_impacts(px_ctx* c, const float* impacts16, uint32_t n) {
  if (!c || (n && !impacts16)) return PX_ERR_ARG;
  uint32_t maxi = c->lim.max_impacts < 64u ? c->lim.max_impacts : 64u;
  if (n > maxi)
    return px_fail(c, PX_ERR_CAPACITY, "px_push_impacts", "more impacts than max_impacts (at most 64 per pass)");
  if (!n) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_impact_stage, impacts16, (size_t)n * 64, hipMemcpyHostToDevice, c->stream));
  if (int rc = px_launch_push_impacts(c, n)) return rc;
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_snapshot(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (!c->b_grid) {
    PX_HIP_CHECK(c, hipMalloc(&c->b_grid, (size_t)c->ncell * 4));
    PX_HIP_CHECK(c, hipMalloc(&c->b_pool, (size_t)c->pool_cap * 1024));
    PX_HIP_CHECK(c, hipMalloc(&c->b_super, (size_t)c->nsuper * 4));
    PX_HIP_CHECK(c, hipMalloc(&c->b_ctrl, PX_CTRL_WORDS * 4));
    PX_HIP_CHECK(c, hipMalloc(&c->b_free, (size_t)c->pool_cap * 4));
  }
  hipStream_t s = c->stream;
  PX_HIP_CHECK(c, hipMemcpyAsync(c->b_grid, c->d_grid, (size_t)c->ncell * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->b_pool, c->d_pool, (size_t)c->pool_cap * 1024, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->b_super, c->d_super, (size_t)c->nsuper * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->b_ctrl, c->d_ctrl, PX_CTRL_WORDS * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->b_free, c->d_free, (size_t)c->pool_cap * 4, hipMemcpyDeviceToDevice, s));
  return PX_OK;
}

int px_restore(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  if (!c->b_grid) return px_fail(c, PX_ERR_STATE, "px_restore", "no snapshot taken");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_grid, c->b_grid, (size_t)c->ncell * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_pool, c->b_pool, (size_t)c->pool_cap * 1024, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_super, c->b_super, (size_t)c->nsuper * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_ctrl, c->b_ctrl, PX_CTRL_WORDS * 4, hipMemcpyDeviceToDevice, s));
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_free, c->b_free, (size_t)c->pool_cap * 4, hipMemcpyDeviceToDevice, s));
  c->labels_valid = false;
  c->masks_dirty = c->density_dirty = true;
  if (int rc = px_clear_seeds(c)) return rc;  // pending seeds belong to the world being dropped
  if (int rc = px_clear_shots(c)) return rc;  // so do shoot's flood entries and dirty cells
  return px_refresh_masks(c);
}

int px_profile_begin(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  c->prof = true;
  return PX_OK;
}

int px_profile_end(px_ctx* c, double out[2 * PX_PROF_SLOTS]) {
  if (!c || !out) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 2 * PX_PROF_SLOTS; k++) out[k] = 0.0;
  for (auto& r : c->prof_rec) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess && r.slot >= 0 && r.slot < PX_PROF_SLOTS) {
      out[2 * r.slot] += ms;
      out[2 * r.slot + 1] += 1.0;
    }
    c->ev_pool.push_back(r.a);
    c->ev_pool.push_back(r.b);
  }
  c->prof_rec.clear();
  c->prof = false;
  return PX_OK;
}

int px_island_pass(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_island_pass(c, true);
}

static int sim_step(px_ctx* c, uint32_t stages_mask) {
  if (!c) return PX_ERR_ARG;
  const uint32_t known = PX_STAGE_EDIT | PX_STAGE_PARTICLES | PX_STAGE_FRAME_BEGIN | PX_STAGE_WRITES |
                         PX_STAGE_RESOLVE | PX_STAGE_CHUNKS | PX_STAGE_SHOOT;
  if (stages_mask & ~known) return px_fail(c, PX_ERR_ARG, "px_sim_step", "unknown stage bits");
  if ((stages_mask & PX_STAGE_PARTICLES) && !c->d_particles)
    return px_fail(c, PX_ERR_STATE, "px_sim_step", "particle stage without px_particles_init");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (stages_mask & PX_STAGE_FRAME_BEGIN) {
    if (int rc = px_launch_frame_begin(c)) return rc;
  }
  if (stages_mask & PX_STAGE_SHOOT) {
    if (int rc = px_launch_shoot(c)) return rc;
  }
  if (stages_mask & PX_STAGE_EDIT) {
    // runs on the device only with queued impacts (or pending seeds, seeded mode)
    if (int rc = px_launch_island_pass(c, false)) return rc;
  }
  if ((stages_mask & PX_STAGE_RESOLVE) && c->labels_valid) {
    if (int rc = px_launch_resolve(c)) return rc;
  }
  if (stages_mask & PX_STAGE_CHUNKS) {
    if (int rc = px_launch_chunk_frame(c)) return rc;
  }
  if (stages_mask & PX_STAGE_PARTICLES) {
    if (int rc = px_launch_particle_step(c)) return rc;
  }
  if (stages_mask & PX_STAGE_WRITES) return px_launch_writes(c);
  return PX_OK;
}

int px_sim_step(px_ctx* c, uint32_t stages_mask) {
  const int rc = sim_step(c, stages_mask);
  if (rc == PX_OK && !c->capturing) c->stages_warm |= stages_mask;
  return rc;
}

struct px_graph {
  px_ctx* ctx;
  hipGraph_t graph;
  hipGraphExec_t exec;
  bool labels_valid;  // the host's label state after the recorded stages
};

int px_capture_sim_step(px_ctx* c, uint32_t stages_mask, px_graph** out) {
  if (!c || !out) return PX_ERR_ARG;
  *out = nullptr;
  if (stages_mask & ~c->stages_warm)
    return px_fail(c, PX_ERR_STATE, "px_capture_sim_step",
                   "run these stages once with px_sim_step first (their buffers are allocated on first use)");
  if (c->prof) return px_fail(c, PX_ERR_STATE, "px_capture_sim_step", "profiling is on");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  // host state the recording changes although nothing runs yet
  const bool lv = c->labels_valid, md = c->masks_dirty, dd = c->density_dirty;
  const uint32_t mg = c->mask_gen, smg = c->sub_mask_gen;
  PX_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int rc = sim_step(c, stages_mask);
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  const bool lv_after = c->labels_valid;
  c->labels_valid = lv;
  c->masks_dirty = md;
  c->density_dirty = dd;
  c->mask_gen = mg;
  c->sub_mask_gen = smg;
  if (rc != PX_OK || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != PX_OK) return rc;
    return px_fail(c, PX_ERR_HIP, "px_capture_sim_step", hipGetErrorString(ec));
  }
  // memset nodes are refused: every clear of the recorded stages is a kernel
  size_t nn = 0;
  (void)hipGraphGetNodes(g, nullptr, &nn);
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn) (void)hipGraphGetNodes(g, nodes.data(), &nn);
  for (size_t k = 0; k < nn; k++) {
    hipGraphNodeType t;
    if (hipGraphNodeGetType(nodes[k], &t) != hipSuccess || t == hipGraphNodeTypeMemset) {
      (void)hipGraphDestroy(g);
      return px_fail(c, PX_ERR_STATE, "px_capture_sim_step", "the recorded stages hold a memset node");
    }
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    return px_fail(c, PX_ERR_HIP, "px_capture_sim_step", hipGetErrorString(ei));
  }
  px_graph* G = new (std::nothrow) px_graph;
  if (!G) {
    (void)hipGraphExecDestroy(x);
    (void)hipGraphDestroy(g);
    return px_fail(c, PX_ERR_HIP, "px_capture_sim_step", "out of host memory");
  }
  G->ctx = c;
  G->graph = g;
  G->exec = x;
  G->labels_valid = lv_after;
  *out = G;
  return PX_OK;
}

int px_graph_launch(px_ctx* c, px_graph* G) {
  if (!c || !G || G->ctx != c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  {
    PxProfScope prof(c, PX_PROF_GRAPH);
    PX_HIP_CHECK(c, hipGraphLaunch(G->exec, c->stream));
  }
  // the frame changed the world: masks and density are rebuilt before the
  // next eager use
  c->labels_valid = G->labels_valid;
  c->masks_dirty = c->density_dirty = true;
  return PX_OK;
}

int px_graph_destroy(px_graph* G) {
  if (!G) return PX_OK;
  (void)hipSetDevice(G->ctx->device);
  (void)hipGraphExecDestroy(G->exec);
  (void)hipGraphDestroy(G->graph);
  delete G;
  return PX_OK;
}

int px_resolve_islands(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_resolve(c);
}

int px_chunk_step(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_chunk_frame(c);
}

int px_read_chunks(px_ctx* c, uint32_t* table, uint32_t* arena, uint32_t arena_cap, uint32_t* n_arena) {
  if (!c || (arena && arena_cap && !n_arena)) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (int rc = px_chunks_alloc(c)) return rc;
  uint32_t na = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&na, c->d_ctrl + PX_C_ARENA, 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  na = std::min(na, c->lim.max_arena);
  if (n_arena) *n_arena = na;
  if (table)
    PX_HIP_CHECK(c, hipMemcpyAsync(table, c->d_chunks, (size_t)std::min<uint32_t>(c->lim.max_chunks, 511u) * 64,
                                   hipMemcpyDeviceToHost, c->stream));
  const uint32_t m = std::min(na, arena_cap);
  if (arena && m) PX_HIP_CHECK(c, hipMemcpyAsync(arena, c->d_arena, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_read_device_impacts(px_ctx* c, float* out16, uint32_t cap, uint32_t* n) {
  if (!c || !n || (cap && !out16)) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (int rc = px_chunks_alloc(c)) return rc;
  uint32_t k = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&k, c->d_ctrl + PX_C_IMPACTS, 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  *n = k;
  const uint32_t m = std::min(std::min(k, cap), std::min<uint32_t>(c->lim.max_impacts, 64u));
  if (m) {
    PX_HIP_CHECK(c, hipMemcpyAsync(out16, c->d_impacts, (size_t)m * 64, hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return PX_OK;
}

int px_apply_writes(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_writes(c);
}

int px_run_to_quiescence(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (int rc = px_launch_island_pass(c, false)) return rc;
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return px_check_island_pass(c, "px_run_to_quiescence");
}

int px_island_labels(px_ctx* c, uint32_t* min_key, uint32_t* count, uint32_t cap, uint32_t* n) {
  if (!c || !n || (cap && (!min_key || !count))) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_read_island_labels(c, min_key, count, cap, n);
}

}  // extern "C"
