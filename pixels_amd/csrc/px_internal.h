// px_internal.h - context layout and device helpers shared by the HIP sources
// of libpixels_hip.so.  Everything here is gfx950 (CDNA4) code: wave64,
// 160 KB LDS per CU, FP32 rounded once per operation (built with
// -ffp-contract=off so the arithmetic order of the reference's WGSL is kept).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pixels.h"

#pragma clang fp contract(off)

#define PX_HIP_CHECK(ctx, expr)                                                   \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return px_fail((ctx), PX_ERR_HIP, #expr, hipGetErrorString(e_)); \
  } while (0)

// Read-only view of the world as the traversal kernels see it.
struct PxWorldView {
  const uint32_t* grid;
  const uint32_t* pool;
  const uint32_t* super_grid;
  const float* mats;
  const uint32_t* super_mask;  // 1 bit per super cell: count != 0
  const uint32_t* brick_mask;  // 1 bit per brick cell: grid != 0
  const uint64_t* super_bm;    // per super cell: 64-bit brick occupancy
  const uint64_t* sub_mask;    // per brick cell with a pool brick: bit (x/2) + 4 (y/2) + 16 (z/2) = that 2x2x2 block has a voxel
  int32_t dims[3];
  int32_t bd[3];
  int32_t sd[3];
  float fdims[3];
};

// Per-frame shadow-ray description (decoded from the uniforms + light list).
struct PxShadowDesc {
  float sun[3];
  float sun_w;
  float vs;
  float lamp_shadows;  // U.flags.y
  uint32_t n_lights;
  uint32_t slots;
  float lpv[32][4];  // light position and radius in voxels: lp / vs, divided on the host (IEEE FP32, as the WGSL)
  float lc_w[32];   // light shadow flag
};

struct px_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t dims[3] = {0, 0, 0}, bd[3] = {0, 0, 0}, sd[3] = {0, 0, 0};
  uint32_t ncell = 0, nsuper = 0, pool_cap = 0, pool_used = 0;
  px_limits lim{};
  float vs = 1.0f;
  // world
  uint32_t* d_grid = nullptr;
  uint32_t* d_pool = nullptr;
  uint32_t* d_super = nullptr;
  float* d_mats = nullptr;
  uint32_t* d_ctrl = nullptr;
  uint32_t* d_args = nullptr;
  float* d_camera = nullptr;
  uint32_t* d_free = nullptr;
  // traversal acceleration (derived from grid / super grid)
  uint32_t* d_super_mask = nullptr;
  uint32_t* d_brick_mask = nullptr;
  uint64_t* d_super_bm = nullptr;
  // 2x2x2-block occupancy of each pool brick, by grid cell; rebuilt before a
  // march when the masks have been rebuilt since (mask_gen != sub_mask_gen)
  uint64_t* d_sub_mask = nullptr;
  uint32_t mask_gen = 0, sub_mask_gen = 0xFFFFFFFFu;
  bool masks_dirty = true;
  // frame
  uint32_t* d_shadow_queue = nullptr;  // output slot of each queued shadow ray
  size_t shadow_cap = 0;
  uint32_t* d_wave_count = nullptr;  // shadow rays queued per wavefront region
  uint32_t wave_count_cap = 0;
  uint32_t frame_layout = PX_FRAME_PIXEL_MAJOR;  // layout of the frame's shadow outputs
  // device address of the library's fixed scratch array on this context's
  // device (queue + counts of px_trace_frame_device_noalloc)
  uint32_t* d_fixed_scratch = nullptr;
  // batched queries: launch shape, per-XCD claim cursors, the captured graph
  uint32_t n_cu = 256;
  uint32_t batch_block = 256, batch_blocks = 0, batch_chunk = 256;
  uint32_t* d_batch_cursors = nullptr;
  hipGraphExec_t batch_exec = nullptr;
  unsigned char batch_key[192] = {0};
  uint32_t* d_counters = nullptr;  // small scratch counters
  // edit chain
  float* d_impacts = nullptr;       // the impact queue: 64 records of 16 floats, C_IMPACTS of them used
  float* d_impact_stage = nullptr;  // host impacts on their way into the queue
  uint32_t* d_seeds = nullptr;      // seeded search: one bit per voxel key
  uint32_t* d_shots = nullptr;      // shoot's pane flood entries and dirtied cells (kShotWords)
  uint32_t* d_edit_work = nullptr;  // cells the carve / flood count passes found work in (2 x ncell)
  uint2* d_carve_rows = nullptr;     // carve's removals per row of its first job entries (count -> apply pass)
  float* d_flashes = nullptr;       // muzzle flashes: PX_MAX_FLASHES records of 8 floats
  uint64_t* d_cell_imask = nullptr;  // per cell: bit k = impact k has a job here
  uint32_t* d_dirty_flags = nullptr; // per cell bitset: dirty
  uint32_t* d_job_list = nullptr;    // cells with at least one job
  uint32_t* d_alloc_list = nullptr;
  uint32_t* d_part = nullptr;        // per-block partial sums of the edit pass
  uint32_t* d_bcount = nullptr;      // per-block list lengths of the edit pass
  uint32_t* d_pane_cells = nullptr;  // pool cells holding pane voxels
  bool labels_valid = false;
  uint32_t stages_warm = 0;  // px_sim_step stages run eagerly at least once (buffers exist)
  bool capturing = false;    // px_capture_sim_step is recording the stream
  uint32_t* d_edit_ctr = nullptr;    // edit-pass counters
  // component labelling
  uint32_t* d_par = nullptr;     // union-find parent per voxel key (node ids are keys)
  uint32_t* d_ncnt = nullptr;    // per node: voxel count, then per-root totals
  uint32_t* d_nflag = nullptr;   // per node flags
  uint16_t* d_lab = nullptr;     // per pool brick: local root of each voxel (0xFFFF = none)
  uint32_t* d_bnode = nullptr;   // per pool brick: its one island node, none, or several (labelling)
  uint64_t* d_bface = nullptr;   // per pool brick: node-class and structural voxels of its 6 faces
  uint32_t* d_nodes = nullptr;   // node list
  uint32_t node_cap = 0;
  uint32_t* d_out_labels = nullptr;  // (key, count) pairs
  uint32_t label_cap = 0;
  size_t vox_total = 0;
  // voxel writes (write list shared by the debris and chunk stages)
  uint32_t* d_write_list = nullptr; // (key, value) pairs, lim.max_writes of them
  uint32_t* d_cfirst = nullptr;     // per cell: first write index needing a brick (0xFFFFFFFF = none)
  uint64_t* d_wkeys = nullptr;      // per write: column << 32 | index, then sorted
  uint64_t* d_wsorted = nullptr;
  uint32_t* d_wdirty = nullptr;     // cells the puts dirtied
  void* d_sort_tmp = nullptr;       // radix-sort temporary storage
  size_t sort_tmp_bytes = 0;
  // island extraction and falling chunks
  uint32_t* d_chunks = nullptr;       // max_chunks x 16 words
  uint32_t* d_arena = nullptr;        // (key, value | chunk << 16) pairs, max_arena of them
  uint2* d_cpair = nullptr;           // per cell / extract cell / arena entry (x, y) counts
  uint2* d_coff = nullptr;            // their exclusive offsets
  uint2* d_tiles = nullptr;           // per 1024-element tile sums / offsets
  uint32_t* d_rlist = nullptr;        // extract cells
  uint32_t* d_rdirty = nullptr;       // cells the extraction dirtied
  uint32_t* d_rtot = nullptr;         // scan totals
  uint32_t* d_best = nullptr;         // per chunk: smallest candidate index of a selection round
  size_t scan_cap = 0;                // elements d_cpair / d_coff hold
  // debris particles
  uint32_t max_particles = 0;
  float4* d_particles = nullptr;    // ring: 2 float4 per particle (pos, value) (vel, life)
  uint32_t* d_alive = nullptr;      // alive list (slot indices)
  uint8_t* d_pstatus = nullptr;     // per slot: bit 0 alive, bit 1 write request
  uint32_t* d_pblock = nullptr;     // per tile: alive / write counts, then offsets
  float4* d_spawn_stage = nullptr;  // host spawn records staged on the device
  uint32_t spawn_cap = 0;
  // rendering: AO density per cell, per-pixel hits, transmittance / glow and
  // shadow results of the last frame
  uint8_t* d_density = nullptr;
  bool density_dirty = true;
  px_hit* d_rprim = nullptr;
  size_t rprim_cap = 0;
  float4* d_rtg = nullptr;
  size_t rtg_cap = 0;
  float4* d_rsh = nullptr;
  size_t rsh_cap = 0;
  uint32_t* d_rrq = nullptr;  // reflection queue: 64 entries per region, then one count per region
  size_t rrq_cap = 0;
  float4* d_rrc = nullptr;    // reflection colour per pixel
  size_t rrc_cap = 0;
  // device-side checkpoint (px_snapshot / px_restore)
  uint32_t* b_grid = nullptr;
  uint32_t* b_pool = nullptr;
  uint32_t* b_super = nullptr;
  uint32_t* b_ctrl = nullptr;
  uint32_t* b_free = nullptr;
  // profiling: event pairs recorded around kernels while `prof` is set
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct ProfRec {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> prof_rec;
  // host mirrors
  float h_uniforms[PX_UNIFORM_FLOATS] = {0};
  std::vector<float> h_lights;
  PxShadowDesc shadow{};
  std::string err;
};

int px_fail(px_ctx* ctx, int code, const char* what, const char* detail);

// Records an event pair around the kernels launched in its scope when
// profiling is on (px_profile_begin).
struct PxProfScope {
  px_ctx* c;
  int slot;
  hipEvent_t a = nullptr;
  PxProfScope(px_ctx* ctx, int s);
  ~PxProfScope();
};
PxWorldView px_view(const px_ctx* c);
int px_refresh_masks(px_ctx* c);
// px_refresh_masks, then the sub-block masks march() reads; every launcher that
// runs march() calls this instead.
int px_refresh_march(px_ctx* c);
// Dynamic LDS of march()'s tables: 8 glass words, then the super-cell bits
// when they fit in 64 KB (*sup_lds; else the kernels read the global copy).
inline size_t px_table_lds(const px_ctx* c, bool* sup_lds) {
  const size_t full = (8u + (c->nsuper + 31u) / 32u) * 4u;
  *sup_lds = full <= 64u * 1024u;
  return *sup_lds ? full : 32u;
}

// launchers (px_trace.hip)
int px_launch_trace(px_ctx* c, uint32_t kind, const px_ray* d_rays, uint32_t n, px_hit* d_out);
int px_launch_frame(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                    uint32_t* n_cast);
int px_launch_probe(px_ctx* c, const int32_t o[3], const int32_t s[3], uint32_t* d_out);
// Address of the fixed scratch array on the current device, looked up once
// per device (null + PX_ERR_HIP if the lookup fails).
int px_fixed_scratch_init(px_ctx* c);
int px_launch_frame_noalloc(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                            uint32_t* n_cast);
int px_launch_batch(px_ctx* c, const px_query_set* q, const uint32_t* const* perm);
// px_render.hip
int px_refresh_density(px_ctx* c);
int px_launch_render(px_ctx* c, const px_ray* d_rays, uint32_t n, float* d_rgba);
int px_launch_trace_perm(px_ctx* c, uint32_t kind, const px_ray* d_rays, const uint32_t* d_perm, uint32_t n,
                         px_hit* d_out);
int px_autotune_batch_impl(px_ctx* c, const px_query_set* q, const uint32_t* const* perm, uint32_t repeats,
                           uint32_t out[3], float* best_ms);
int px_launch_masks(px_ctx* c);
// edit chain (px_edit.hip)
int px_edit_alloc(px_ctx* c);
void px_edit_free(px_ctx* c);
// force: label even without impacts or pending seeds (px_island_pass)
int px_launch_island_pass(px_ctx* c, bool force);
int px_launch_push_impacts(px_ctx* c, uint32_t n);
int px_launch_shoot(px_ctx* c);
int px_clear_shots(px_ctx* c);

// Shot buffer (d_shots): word 0 = pane flood entries shoot queued, word 1 =
// cells its takes dirtied, then kShotFloodCap (voxel key, material id) pairs,
// then kShotDirtyCap cell indices.  The next island pass that runs consumes
// both lists and empties them.
constexpr uint32_t kShotFloodCap = 64u;
constexpr uint32_t kShotDirtyCap = 256u;
constexpr uint32_t kShotFloodBase = 4u;
constexpr uint32_t kShotDirtyBase = kShotFloodBase + 2u * kShotFloodCap;
constexpr uint32_t kShotWords = kShotDirtyBase + kShotDirtyCap;
int px_seeds_alloc(px_ctx* c);
int px_clear_seeds(px_ctx* c);
int px_read_island_labels(px_ctx* c, uint32_t* min_key, uint32_t* count, uint32_t cap, uint32_t* n);
// Error bits of the edit counters as a status; the blocking form reads them
// from the device after the stream drains.
int px_edit_errors(px_ctx* c, const char* what, uint32_t bits);
int px_check_island_pass(px_ctx* c, const char* what);
// scene rasteriser (px_scene.hip): px_raster_begin rasterises prims into
// these temporaries and reports the pool bricks the world needs;
// px_raster_finish writes the grid, pool and super grid of the context,
// whose buffers exist by then.
struct PxRaster {
  hipStream_t stream = nullptr;  // the stream its kernels run on: drained before the frees
  px_prim* d_prims = nullptr;
  uint64_t* d_pairs = nullptr;
  uint64_t* d_sorted = nullptr;
  uint32_t* d_counts = nullptr;
  uint64_t* d_offs = nullptr;
  uint8_t* d_state = nullptr;
  uint16_t* d_val = nullptr;
  uint32_t* d_slot = nullptr;
  uint32_t* d_temp = nullptr;
  uint2* d_flags = nullptr;
  uint2* d_pidx = nullptr;
  uint2* d_tiles = nullptr;
  uint32_t* d_tot = nullptr;
  ~PxRaster() {
    if (stream) (void)hipStreamSynchronize(stream);
    void* bufs[] = {d_prims, d_pairs, d_sorted, d_counts, d_offs, d_state, d_val,
                    d_slot, d_temp, d_flags, d_pidx, d_tiles, d_tot};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
  }
};

int px_raster_begin(px_ctx* c, const px_prim* prims, uint32_t n, PxRaster* R, uint32_t* used);
int px_raster_finish(px_ctx* c, PxRaster* R);
// voxel writes (px_edit.hip) and the radix sort they use (px_sort.hip)
int px_writes_alloc(px_ctx* c);
void px_writes_free(px_ctx* c);
int px_launch_writes(px_ctx* c);
// stable radix sort of n 64-bit keys by bits [begin_bit, end_bit) (px_sort.hip)
int px_sort_u64(px_ctx* c, const uint64_t* d_in, uint64_t* d_out, uint32_t n, uint32_t begin_bit,
                uint32_t end_bit);
// island extraction and falling chunks (px_edit.hip)
int px_chunks_alloc(px_ctx* c);
void px_chunks_free(px_ctx* c);
int px_launch_resolve(px_ctx* c);
int px_launch_chunk_frame(px_ctx* c);
// debris particles (px_particles.hip)
void px_particles_free(px_ctx* c);
int px_launch_frame_begin(px_ctx* c);
int px_launch_particle_step(px_ctx* c);

// ---------------------------------------------------------------------------
// device helpers (shared.wgsl:44-81, render_common.wgsl:25-31)

#define PX_DEV __device__ __forceinline__

PX_DEV uint32_t px_hashu(uint32_t x) {
  uint32_t h = x;
  h ^= h >> 16u;
  h *= 0x7feb352du;
  h ^= h >> 15u;
  h *= 0x846ca68bu;
  h ^= h >> 16u;
  return h;
}
PX_DEV uint32_t px_hash3i(int32_t x, int32_t y, int32_t z) {
  return px_hashu(((uint32_t)x * 0x9E3779B1u) ^
                  px_hashu(((uint32_t)y * 0x85EBCA77u) ^ px_hashu((uint32_t)z * 0xC2B2AE3Du)));
}
PX_DEV float px_rnd(uint32_t seed) { return (float)(px_hashu(seed) & 0xffffffu) / 16777216.0f; }

// WGSL i32(f32): truncation toward zero, saturating (v_cvt_i32_f32 saturates;
// NaN gives 0).
PX_DEV int32_t px_f2i(float x) { return (int32_t)__builtin_amdgcn_fmed3f(x, -2147483648.0f, 2147483520.0f); }
PX_DEV int32_t px_clampi(int32_t v, int32_t lo, int32_t hi) { return min(max(v, lo), hi); }
// Makes the compiler forget what it knows of a VGPR value (emits nothing):
// used after a loop so that tests of its live-out values are evaluated there
// once, not tracked as masks on every turn of the loop.
#define PX_OPAQUE(x) asm volatile("" : "+v"(x))

// a * b + c for a, b < 2^24 as one v_mad_u32_u24 (written out: the compiler
// otherwise widens an index that addresses memory into a 64-bit multiply).
// b must be wave-uniform (an SGPR operand).
PX_DEV uint32_t px_mad24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}
PX_DEV float px_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
PX_DEV float px_canon0(float t) { return t == 0.0f ? 0.0f : t; }
PX_DEV float px_pow5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x;
}
PX_DEV bool px_is_glass(const float* mats, uint32_t id) { return mats[id * 20u + 13u] > 0.0f; }

PX_DEV uint32_t px_brick_voxel(const uint32_t* pool, uint32_t g, int32_t lx, int32_t ly, int32_t lz) {
  if ((g & PX_UNIFORM_BIT) != 0u) return g & 0xffffu;
  uint32_t li = (uint32_t)(lx | (ly << 3) | (lz << 6));
  uint32_t idx = (g - 1u) * 512u + li;
  uint32_t w = pool[idx >> 1u];
  return (w >> ((idx & 1u) * 16u)) & 0xffffu;
}
