/*
 * pixels.h - C ABI of the MI355X (gfx950) brick-map voxel library.
 *
 * The library owns one HIP stream and all device memory of a world: the brick
 * grid, the voxel pool, the super grid, the material table, the control
 * counters and the edit-chain heap.  Every entry point returns an int status
 * (PX_OK = 0, negative = error); px_last_error() gives the message.  No C++
 * exception crosses this boundary.  A context is used from one host thread at
 * a time (the reference drives its GPUDevice from one JS thread as well).
 *
 * Citations are into the reference repository (Crunchy000/pixels).
 */
#ifndef PIXELS_H
#define PIXELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PX_OK 0
#define PX_ERR_ARG (-1)      /* bad argument (null pointer, size, kind) */
#define PX_ERR_HIP (-2)      /* a HIP runtime call failed */
#define PX_ERR_CAPACITY (-3) /* a size exceeds a configured limit */
#define PX_ERR_STATE (-4)    /* call not valid in the current state */
#define PX_ERR_NODEVICE (-5) /* no HIP device */

/* ---- encoding constants (src/world/VoxelWorld.ts:4-13, shared.wgsl:44) -- */
#define PX_NONE 0xFFFFFFFFu
#define PX_UNIFORM_BIT 0x80000000u
#define PX_BRICK 8
#define PX_BRICK_VOXELS 512
#define PX_UNIFORM_FLOATS 52 /* src/gpu/GpuWorld.ts:8, shared.wgsl:5-19 */
#define PX_MAT_FLOATS 20     /* src/world/materials.ts:193 */
#define PX_MAT_COUNT 256
#define PX_CTRL_WORDS 36     /* layout.ts:48-58 rounded up to 4, layout.ts:67 */
#define PX_ARGS_WORDS 44     /* layout.ts:41-45 */
#define PX_CAMERA_FLOATS 64  /* GpuWorld.ts:10 (256 bytes) */

/* Control-counter word indices, in the order of CTRL_NAMES (layout.ts:48-58). */
enum px_ctrl_word {
  PX_C_FREE_TOP = 0, PX_C_IMPACTS, PX_C_JOBS, PX_C_ALLOCS, PX_C_ALLOC_N,
  PX_C_FLOOD_CUR, PX_C_FLOOD_NEXT, PX_C_FLOOD_PAR,
  PX_C_S_STATE, PX_C_S_CUR, PX_C_S_NEXT, PX_C_S_PAR, PX_C_NODES, PX_C_PENDING,
  PX_C_S_ITERS, PX_C_SEED_N, PX_C_CHUNKS, PX_C_ARENA, PX_C_WRITES, PX_C_WRITE_N,
  PX_C_DIRTY, PX_C_PCURSOR, PX_C_FLASH, PX_C_SPAWNED,
  PX_S_BILL, PX_S_DESTROYED, PX_S_ALLOC_FAIL, PX_S_FALLING, PX_S_SEARCHES, PX_S_ISLANDS,
  PX_S_SEED_DROP, PX_S_BUDGET, PX_S_HASH_FULL,
  PX_CTRL_USED
};

/* Face ids follow dirOf() order (sim.wgsl:27-32): the face is k with
 * normal == dirOf(k).  6 = zero normal (the ray started inside the volume and
 * stopped in its first cell), 0xFF = miss, 0xFE = shadow ray not cast. */
#define PX_FACE_PX 0
#define PX_FACE_NX 1
#define PX_FACE_PY 2
#define PX_FACE_NY 3
#define PX_FACE_PZ 4
#define PX_FACE_NZ 5
#define PX_FACE_ZERO 6
#define PX_FACE_NOT_CAST 0xFE
#define PX_FACE_MISS 0xFF

/* ---- ray queries ---------------------------------------------------------
 * PRIMARY: march(ro, rd, maxT, shadow=false)      raymarch.wgsl:44-162
 * SHADOW : march(ro, rd, maxT, shadow=true)       raymarch.wgsl:326,354
 * SHOOT  : traceSolid(ro, rd, maxT)               sim.wgsl:296-357
 * SEGMENT: particle segment DDA from pos=ro along the displacement rd
 *          (max_t unused)                         sim.wgsl:1110-1145       */
#define PX_RAY_PRIMARY 0u
#define PX_RAY_SHADOW 1u
#define PX_RAY_SHOOT 2u
#define PX_RAY_SEGMENT 3u
#define PX_RAY_KIND_MASK 3u

/* 32-byte ray record: origin and direction in voxel units, FP32. `flags`
 * carries the query kind in its low 2 bits for the batched entry point. */
typedef struct px_ray {
  float ro[3];
  float rd[3];
  float max_t;
  uint32_t flags;
} px_ray;

/* 16-byte result: voxel key x + DX*(y + DY*z) (sim.wgsl:50) or PX_NONE, t of
 * the hit (0 on a miss, as the zero-initialised Hit of raymarch.wgsl:45),
 * step count (gOuterSteps + gInnerSteps, raymarch.wgsl:18-19,84,93,107),
 * face id, and the low byte of the hit voxel's material id. */
typedef struct px_hit {
  uint32_t key;
  float t;
  uint32_t steps;
  uint8_t face;
  uint8_t mat;
  uint16_t value_hi; /* high byte of the 16-bit voxel value ("shade") */
} px_hit;

/* ---- limits (layout.ts:11-33) ------------------------------------------- */
typedef struct px_limits {
  uint32_t max_dirty, max_alloc, max_writes, max_impacts, max_jobs, max_flood, max_seeds;
  uint32_t hash_size, max_chunks, max_arena, max_lights, max_anchors, max_flashes;
  uint32_t spawn_budget, component_budget, search_iters_per_frame, flood_iters_per_frame;
  /* Beyond the reference's LIMITS: which components the island pass
   * reports (PX_ISLANDS_ALL or PX_ISLANDS_SEEDED, see px_island_pass), and
   * the step bound of its union-find walks (0 = the node capacity of the
   * pass, which no walk over well-formed nodes reaches; a walk that passes
   * it, or meets a parent above its child, stops and the labels are
   * refused with PX_ERR_STATE). */
  uint32_t island_mode, uf_walk_cap;
} px_limits;
#define PX_ISLANDS_ALL 0u
#define PX_ISLANDS_SEEDED 1u

/* ---- world description: what `new GpuWorld(device, world, scene)` takes
 * (GpuWorld.ts:46-137).  All host pointers; the library copies them. */
typedef struct px_world_desc {
  uint32_t dims[3];        /* voxel dims, multiples of 8 (config.ts:26-30) */
  const uint32_t* grid;    /* (dx/8)*(dy/8)*(dz/8) entries, VoxelWorld.ts:6-8 */
  const uint16_t* pool;    /* pool_used * 512 voxels, local index x|y<<3|z<<6 */
  uint32_t pool_used;      /* bricks present in `pool` */
  const uint32_t* holes;   /* free bricks inside [0, pool_used) (VoxelWorld.freeBricks) */
  uint32_t n_holes;
  const float* mats;       /* 256 * 20 floats, packMaterials() layout */
  uint32_t pool_capacity;  /* 0 = min(limit, max(used + 98304, 2 * used)), GpuWorld.ts:60;
                              limit = 4M bricks (4 GiB of pool) */
  const px_limits* limits; /* NULL = the reference LIMITS */
  float voxel_size;        /* metres per voxel, U.dims.w */
  int32_t device;          /* HIP device ordinal */
} px_world_desc;

typedef struct px_ctx px_ctx;

/* ---- SDF primitives of the scene rasteriser (src/world/shapes.ts:24-143),
 * in voxel units.  p[] by type:
 *   PX_PRIM_BOX       centre[3], half extents[3]
 *   PX_PRIM_ROUNDBOX  centre[3], half extents[3], corner radius
 *   PX_PRIM_SPHERE    centre[3], radius
 *   PX_PRIM_CYLINDER  centre[3], radius, half height, axis (0 x, 1 y, 2 z)
 *   PX_PRIM_CAPSULE   a[3], b[3], radius
 *   PX_PRIM_TORUS     centre[3], major radius, minor radius, axis
 *   PX_PRIM_AABB      lo[3], hi[3] (exact containment test, shapes.ts:47-48)
 * cy, sy: cosine and sine of a yaw about the vertical axis through the
 * centre (1, 0 for none; the unyawed box also has the exact containment
 * test).  lo / hi: world bounds of the shape (Shape.min / max); the
 * rasteriser visits the bricks of [lo - 0.5, hi + 0.5].  mat: the voxel
 * value painted. */
#define PX_PRIM_BOX 0u
#define PX_PRIM_ROUNDBOX 1u
#define PX_PRIM_SPHERE 2u
#define PX_PRIM_CYLINDER 3u
#define PX_PRIM_CAPSULE 4u
#define PX_PRIM_TORUS 5u
#define PX_PRIM_AABB 6u
typedef struct px_prim {
  uint32_t type;
  uint32_t mat;
  double p[12];
  double cy, sy;
  double lo[3], hi[3];
} px_prim;

/* Fills *l with the reference LIMITS (layout.ts:11-33). */
void px_default_limits(px_limits* l);

/* Replaces `new GpuWorld(device, world, scene)` (GpuWorld.ts:46-137): sizes
 * the pool, builds the super-grid counts (:90-100) and the free list
 * (:103-113), writes C_FREE_TOP (:111-113) and uploads everything once. */
int px_create(const px_world_desc* desc, px_ctx** out);
/* Builds the world on the GPU: apply() of shapes.ts:334-414 (Op.Add,
 * constant paint) for prims[0..n) in order over an empty world of desc->dims,
 * straight into the context's brick map (no host grid or pool: desc->grid,
 * pool, holes and pool_used are ignored; pool_capacity 0 = GpuWorld.ts:60 on
 * the bricks the scene needs).  Pool bricks are numbered in cell order.
 * Blocking.  New entry: the reference rasterises on the CPU at load time
 * (README.md:192 names the GPU version as its next step). */
int px_create_from_prims(const px_world_desc* desc, const px_prim* prims, uint32_t n, px_ctx** out);
/* Frees every device buffer of the context. */
int px_destroy(px_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed px_create when
 * ctx is NULL).  Never NULL. */
const char* px_last_error(const px_ctx* ctx);

/* Replaces the per-frame `writeBuffer(uniformBuf, 208 B)` (main.ts:218-237);
 * same field order as Uniforms (shared.wgsl:5-19). */
int px_set_uniforms(px_ctx* ctx, const float u[PX_UNIFORM_FLOATS]);
/* Replaces the renderer's `lights` buffer (render_common.wgsl:8-9,16-18):
 * lights[0].x = count; light i at [1+2i] = (pos m, radius m), [2+2i] =
 * (colour, shadow flag).  n_vec4 = 1 + 2*count. */
int px_set_lights(px_ctx* ctx, const float* lights_vec4, uint32_t n_vec4);

/* Blocks until all work queued on the context's stream has finished. */
int px_sync(px_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work with it. */
void* px_stream(px_ctx* ctx);

/* Ray queries over host buffers (blocking).  kind = PX_RAY_*.  New entry the
 * hot path needs: the reference has rays only inside fsMain/shoot/particleSim
 * (raymarch.wgsl:388-398, sim.wgsl:380, sim.wgsl:1119-1145). */
int px_trace(px_ctx* ctx, uint32_t kind, const px_ray* rays, uint32_t n, px_hit* out);
/* Same over device buffers, asynchronous on the context's stream. */
int px_trace_device(px_ctx* ctx, uint32_t kind, const px_ray* d_rays, uint32_t n, px_hit* d_out);

/* One frame of the fragment shader's ray work (fsMain raymarch.wgsl:398 and
 * the shadow rays of shade() :323-356, quality 1): the primary march of every
 * ray, then for every primary hit on a non-emissive voxel the sun shadow ray
 * (if dot(N, sunDir) > 0 and sun intensity > 0) and one shadow ray per light
 * (if within its radius, dot(N, L) > 0, the light casts shadows and
 * U.flags.y > 0.5).  Device buffers; d_shadow_out holds (1 + n_lights) slots
 * per primary ray, slot 0 = sun, slot 1+i = light i; slots not cast read
 * key PX_NONE, face PX_FACE_NOT_CAST.  *n_shadow_cast (host pointer, may be
 * NULL) receives the number of shadow rays actually traced (blocking then). */
int px_trace_frame_device(px_ctx* ctx, const px_ray* d_primary, uint32_t n, px_hit* d_primary_out,
                          px_hit* d_shadow_out, uint32_t* n_shadow_cast);

/* px_trace_frame_device without any allocation: the same kernels and
 * results (and the layout of px_set_frame_layout), with the shadow-ray
 * queue, its per-region counts and the cast total in a fixed scratch array
 * of 20,971,520 words (80 MiB) that the library holds on each device it
 * runs on until the process ends (the 1920x1080 frame with the sun and 8
 * lights needs 18.7M words).  It calls no hipMalloc / hipFree: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * Fails with PX_ERR_CAPACITY when 64 * ceil(n / 64) * slots + ceil(n / 64)
 * + 64 exceeds the array.  Every call on one device shares the array: calls
 * that overlap on a device (other streams, other contexts, other host
 * threads) race and must be serialised by the caller.  No counterpart in
 * the reference, which allocates every buffer once at load
 * (GpuWorld.ts:46-137). */
int px_trace_frame_device_noalloc(px_ctx* ctx, const px_ray* d_primary, uint32_t n, px_hit* d_primary_out,
                                  px_hit* d_shadow_out, uint32_t* n_shadow_cast);

/* fsMain (raymarch.wgsl:387-450) to an image: the colour of every primary
 * ray as RGBA floats (alpha 1; no depth), with shade() (:304-366): material
 * colour (render_common.wgsl:58-140), coarse AO from the density texture and
 * fine AO, sun and point lights with the shadow rays of
 * px_trace_frame_device, hemisphere ambient; glass transmittance and glow
 * along the primary ray; the reflection ray of reflective materials when
 * U.flags.x > 0.5 (shaded at quality 0); the sky (:212-264) for misses.
 * U.flags.w > 0 selects the debug views 1-9 (:399-423).  Uniforms and lights
 * come from px_set_uniforms / px_set_lights.  sin, cos, exp, pow and atan2
 * are evaluated in double precision and rounded once to FP32.  Device
 * buffers: d_rgba holds 4 floats per ray, 16-byte aligned; asynchronous on
 * the context's stream.  New entry: the reference shades inside its
 * fragment shader. */
int px_render_frame_device(px_ctx* ctx, const px_ray* d_rays, uint32_t n, float* d_rgba);
/* Same over host buffers, blocking. */
int px_render(px_ctx* ctx, const px_ray* rays, uint32_t n, float* rgba);
/* The AO density texture (GpuWorld.ts:77-101, maint sim.wgsl:1255-1271):
 * one byte per brick cell, round(opaque voxels / 512 * 255), rebuilt from the
 * brick map after edits; out holds bx*by*bz bytes. */
int px_read_density(px_ctx* ctx, uint8_t* out);
/* Slots per primary ray that px_trace_frame_device writes (1 + light count). */
uint32_t px_frame_slots(const px_ctx* ctx);
/* Layout of d_shadow_out for the frame entry points of this context:
 * PX_FRAME_PIXEL_MAJOR (the default) puts slot s of primary ray i at
 * i * slots + s; PX_FRAME_SLOT_PLANAR puts it at s * n + i (one plane of n
 * records per slot), which lets the shadow kernel store the records of
 * neighbouring rays as one contiguous run.  The records are the same.  New
 * entry: the reference consumes its shadow rays inside shade()
 * (raymarch.wgsl:323-356) and stores no per-ray result. */
#define PX_FRAME_PIXEL_MAJOR 0u
#define PX_FRAME_SLOT_PLANAR 1u
int px_set_frame_layout(px_ctx* ctx, uint32_t layout);

/* ---- batched queries ------------------------------------------------------
 * New entry the hot path needs (the reference runs each query kind inside its
 * own shader: fsMain raymarch.wgsl:388-398, shoot sim.wgsl:371-403, the
 * segment DDA of particleSim sim.wgsl:1110-1145).  Queries of all four kinds,
 * of any (uneven) counts, run in ONE persistent launch: wavefronts claim
 * chunks of `chunk` rays of one kind from per-XCD cursors and trace them with
 * the same code and results as px_trace_device (outputs are bit-identical for
 * every shape).  The launch is captured once into a HIP graph and replayed
 * while the query set and shape stay the same.  Device buffers; asynchronous. */
typedef struct px_query_set {
  const px_ray* rays[4]; /* indexed by kind PX_RAY_PRIMARY..PX_RAY_SEGMENT */
  px_hit* out[4];
  uint32_t n[4];
} px_query_set;
int px_trace_batch(px_ctx* ctx, const px_query_set* q);
/* The same launch with kind k traced in the order d_perm[k] (n[k] source
 * indices, a permutation; NULL = index order): position j of the kind's
 * chunk space traces ray d_perm[k][j] and writes its hit to out[k] at that
 * index, so the hits equal px_trace_batch's.  With the orders
 * px_mlp_sort_rays gives, this is the learned-binned batched traversal. */
int px_trace_batch_perm(px_ctx* ctx, const px_query_set* q, const uint32_t* const d_perm[4]);
/* Launch shape of px_trace_batch: block = threads per workgroup (64, 128 or
 * 256), blocks = workgroups (0 = six wavefronts per SIMD of the device),
 * chunk = rays per claim (>= 1). */
int px_set_batch_shape(px_ctx* ctx, uint32_t block, uint32_t blocks, uint32_t chunk);
/* out = {block, blocks, chunk} */
int px_get_batch_shape(const px_ctx* ctx, uint32_t out[3]);
/* Autotuner: times px_trace_batch on `q` over the shapes block {64, 128, 256}
 * x wavefronts per CU {8, 16, 24} x chunk {64, 256, 1024} (`repeats` timed
 * launches each, after one warm-up), leaves the fastest set and writes it to
 * out[3] and its mean time to *best_ms.  Blocking. */
int px_autotune_batch(px_ctx* ctx, const px_query_set* q, uint32_t repeats, uint32_t out[3], float* best_ms);
/* px_autotune_batch timed on px_trace_batch_perm(ctx, q, d_perm): the shape
 * for the launch in that order (no counterpart in the reference; for callers
 * of px_trace_batch_perm, whose best shape differs from the index order's). */
int px_autotune_batch_perm(px_ctx* ctx, const px_query_set* q, const uint32_t* const d_perm[4], uint32_t repeats,
                           uint32_t out[3], float* best_ms);

/* Replaces Simulation.probe (Simulation.ts:161-196, probe.wgsl:25-32):
 * decoded voxel values of the box [o, o+s), x fastest.  Host buffer. */
int px_probe(px_ctx* ctx, const int32_t o[3], const int32_t s[3], uint32_t* out);
/* Same into a device buffer, asynchronous. */
int px_probe_device(px_ctx* ctx, const int32_t o[3], const int32_t s[3], uint32_t* d_out);

/* Replaces afterSubmit's readback (Simulation.ts:275-322): ctrl counters in
 * CTRL_NAMES order, indirect args in ARGS_NAMES order and the 64-float camera
 * block.  Any pointer may be NULL.  Blocking. */
int px_read_ctrl(px_ctx* ctx, uint32_t* ctrl, uint32_t* args, float* cam);
/* Copies the brick grid (cells) and the super-grid counts to host buffers. */
int px_read_grid(px_ctx* ctx, uint32_t* grid, uint32_t* super_grid);
/* Copies pool bricks [first, first+count) as 256 words each. */
int px_read_pool(px_ctx* ctx, uint32_t first, uint32_t count, uint32_t* words);
/* World geometry: dims[3], bricks[3], supers[3], pool capacity. */
int px_world_info(const px_ctx* ctx, uint32_t dims[3], uint32_t bricks[3], uint32_t supers[3],
                  uint32_t* pool_capacity);

/* Queue impacts for the next edit pass: n records of 16 floats in the heap
 * layout of pushImpact (sim.wgsl:203-211): c.xyz, r, fr, impulse, n.xyz,
 * dir.xyz, 4 unused.  Fails with PX_ERR_CAPACITY beyond max_impacts. */
int px_push_impacts(px_ctx* ctx, const float* impacts16, uint32_t n);

/* The island pass of the edit chain, run to quiescence on the context's
 * stream: impactJobs -> allocBricks -> carve (sim.wgsl:408-584), the glass
 * pane flood to completion (:586-618), maint (:1205-1274), then a 6-connected
 * labelling of every non-air, non-rubble, non-structural voxel; a component
 * is supported if it holds a voxel with y == 0 or touches a structural voxel
 * (the anchor rules of visit/searchStep, :767-827).  Asynchronous; labels are
 * read with px_island_labels. */
int px_island_pass(px_ctx* ctx);
/* Replaces the per-frame `Simulation.encode(enc)` + submit of the hot-path
 * stages (Simulation.ts:200-273): runs the selected stages in the
 * reference's order on the context's stream, non-blocking.  PX_STAGE_EDIT =
 * the island pass above (a no-op without queued impacts).  Unknown bits are
 * an error. */
#define PX_STAGE_EDIT 1u
/* Stage bits, run in this order (Simulation.ts:223-268): PX_STAGE_FRAME_BEGIN
 * (px_frame_begin), PX_STAGE_EDIT (the island pass, when impacts are queued,
 * including those falling chunks queued), PX_STAGE_RESOLVE
 * (px_resolve_islands, when an island pass has run since the last one),
 * PX_STAGE_CHUNKS (px_chunk_step), PX_STAGE_PARTICLES (px_particle_step),
 * PX_STAGE_WRITES (px_apply_writes). */
#define PX_STAGE_PARTICLES 2u
#define PX_STAGE_FRAME_BEGIN 4u
int px_sim_step(px_ctx* ctx, uint32_t stages_mask);

/* ---- a frame of stages as a HIP graph (no counterpart in the reference,
 * whose WebGPU command buffers are encoded anew every frame; for callers that
 * replay the same frame without host work per launch).
 * px_capture_sim_step records px_sim_step(stages_mask) on the context's
 * stream into a graph and instantiates it.  Each stage must have run once
 * through px_sim_step (its buffers are allocated on first use), else
 * PX_ERR_STATE; a recording that holds a memset node is refused.  The
 * uniforms, the limits and the island mode are those at capture (kernel
 * arguments); everything else is read from device memory, so each launch
 * continues from the world, queues and counters the device holds.  Blocking.
 * px_graph_launch enqueues one replay on the context's stream
 * (asynchronous); px_graph_destroy frees the graph. */
typedef struct px_graph px_graph;
int px_capture_sim_step(px_ctx* ctx, uint32_t stages_mask, px_graph** out);
int px_graph_launch(px_ctx* ctx, px_graph* graph);
int px_graph_destroy(px_graph* graph);
/* The edit stages run until no work is pending (the flood and the island
 * search of the reference iterate over frames until their queues empty and
 * C_S_STATE is 0; here one island pass reaches that state), then waits for
 * the stream.  Blocking. */
int px_run_to_quiescence(px_ctx* ctx);
/* Unsupported components of the last island pass, sorted by label: the
 * canonical label (minimum voxel key of the component) and its voxel count.
 * *n receives the total count even when it exceeds cap.  Blocking. */
int px_island_labels(px_ctx* ctx, uint32_t* min_key, uint32_t* count, uint32_t cap, uint32_t* n);

/* ---- debris particles ------------------------------------------------------
 * Particle record (shared.wgsl:33-42): position and velocity in voxel units,
 * the 16-bit voxel value in the low bits of `value` with the flags above it,
 * seconds of life left (<= 0 = free slot). */
typedef struct px_particle {
  float pos[3];
  uint32_t value;
  float vel[3];
  float life;
} px_particle;
#define PX_P_SETTLED 0x10000u
#define PX_P_DEPOSIT 0x20000u
#define PX_P_DUST 0x40000u

/* Allocates the particle ring of max_particles slots (a power of two in
 * [2^12, 2^21], MAX_PARTICLES of config.ts:38; 0 = 2^18), the alive list
 * (H_ALIVE, layout.ts:105) and the write list (H_WRITE_LIST, max_writes
 * pairs, layout.ts:91), all zeroed, and zeroes C_PCURSOR.  Again to resize
 * (the ring is cleared).  Blocking. */
int px_particles_init(px_ctx* ctx, uint32_t max_particles);
/* frameBegin (sim.wgsl:363-369): C_SPAWNED, C_DIRTY, C_WRITES and C_ALLOCS
 * to 0, the particle draw args (A_DRAWP) to (36, 0, 0, 0).  Asynchronous. */
int px_frame_begin(px_ctx* ctx);
/* spawn() (sim.wgsl:166-170) for every record of src in order: a record past
 * the frame's spawn budget (C_SPAWNED >= spawn_budget) is dropped, the others
 * go to ring slot C_PCURSOR++ & (max_particles - 1); C_SPAWNED counts every
 * call.  px_spawn takes a host buffer, px_spawn_device a device buffer (both
 * asynchronous; the host buffer may be reused on return). */
int px_spawn(px_ctx* ctx, const px_particle* src, uint32_t n);
int px_spawn_device(px_ctx* ctx, const px_particle* d_src, uint32_t n);
/* particleSim (sim.wgsl:1088-1173) over every slot with dt = U.sim.x,
 * gravity = U.sim.y and the voxel size U.dims.w of the last px_set_uniforms:
 * life, gravity and drag, the pop out of solids, the 48-cell segment DDA with
 * its bounce, the rest check that settles a particle and, for a DEPOSIT
 * particle inside the world, requests a rubble voxel write (pushWrite
 * :186-192).  The reference appends alive slots (pushAlive :1082-1085, count
 * in A_DRAWP + 1) and write requests (C_WRITES) with atomics in schedule
 * order; here both lists are in slot order, and when requests exceed
 * max_writes the lowest slots get the writes (the rest settle with the
 * failed-write life, as the reference's losers do).  Asynchronous. */
int px_particle_step(px_ctx* ctx);
/* Copies ring slots [first, first + count) to the host.  Blocking. */
int px_read_particles(px_ctx* ctx, uint32_t first, uint32_t count, px_particle* out);
/* The alive list of the frame (slot indices, args[A_DRAWP + 1] of them) and
 * the write list (key, value pairs, min(C_WRITES, max_writes) of them, the
 * count writeArgs :1179-1184 takes).  *n receives the full count; at most cap
 * entries are copied.  Blocking. */
int px_read_alive(px_ctx* ctx, uint32_t* out, uint32_t cap, uint32_t* n);
int px_read_writes(px_ctx* ctx, uint32_t* pairs, uint32_t cap, uint32_t* n);
/* Device addresses of the ring (px_particle[max_particles]), the alive list
 * and the write list, for callers that draw or consume them on the GPU. */
int px_particle_buffers(px_ctx* ctx, void** particles, void** alive, void** writes);

/* ---- voxel writes ------------------------------------------------------------
 * writeArgs + writeRequest + allocBricks + writeApply (sim.wgsl:1178-1203,
 * requestBrick :94-102, putVoxel :127-146, allocBricks :459-483), then maint
 * (:1219-1274) of the cells touched: takes n = min(C_WRITES, max_writes)
 * entries of the write list (C_WRITE_N = n, C_WRITES = 0), gives every cell
 * that a write targets and that holds no brick a pool brick, and puts each
 * write's value into the first empty voxel of p, p + y, p + 2y, stopping at
 * a cell without a pool brick.  The reference serves requests and puts in
 * atomic order; here they follow the list order: the first max_alloc
 * distinct cells are served, and a later write to a taken voxel stacks above
 * the earlier one.  Asynchronous. */
#define PX_STAGE_WRITES 8u
int px_apply_writes(px_ctx* ctx);
/* Appends (key, value) pairs to the write list as pushWrite (sim.wgsl:186-192)
 * does: C_WRITES grows by n, entries past max_writes are dropped.  For callers
 * that deposit voxels from the host.  Blocking on the copy. */
int px_push_writes(px_ctx* ctx, const uint32_t* pairs, uint32_t n);

/* ---- island extraction and falling chunks -------------------------------------
 * px_resolve_islands: resolveChunks + resolveExpand + allocBricks +
 * resolveExtract (sim.wgsl:862-923) on the labels of the last island pass:
 * a component over component_budget voxels stays (S_BUDGET); of the others,
 * those with >= 48 voxels take the free chunk ids in increasing label order
 * (C_CHUNKS counts them all, S_ISLANDS the ones that got an id), the rest
 * crumble.  Uniform bricks holding such voxels get pool bricks first; then
 * every such voxel, in cell order and local index order, leaves the world:
 * the rubble stacked on it tumbles as debris (dropRubbleAbove :274-284), and
 * the voxel goes to its chunk's arena (C_ARENA) or, crumbling or past
 * max_arena, becomes debris (S_BILL, S_DESTROYED).  Debris is spawned into
 * the particle ring (allocated with the default size if px_particles_init was
 * not called), under the frame's spawn budget; maint runs on the touched
 * cells.  The reference orders these by its atomics.  Asynchronous; fails
 * with PX_ERR_STATE when no island pass has run since the last call. */
int px_resolve_islands(px_ctx* ctx);
/* One frame of the falling chunks: chunkArgs, chunkBegin, chunkCollide,
 * chunkStep, chunkLand (sim.wgsl:951-1075) with dt, gravity and the voxel size
 * of the uniforms.  A chunk crossing a fragile voxel at > 1.5 m/s breaks
 * through and queues impacts (at most 6 per chunk, in arena order; they join
 * the next island pass, which then reads them back); a landed chunk's voxels
 * are pushed to the write list in arena order, or shatter into debris when
 * the chunk is fragile and lands faster than 2 m/s.  Asynchronous. */
int px_chunk_step(px_ctx* ctx);
/* Chunk table (max_chunks x 16 words, the reference's layout: state, vy,
 * fallen, frac, fragile, voxel count, blockedAt, broke count, steps, landing
 * speed) and the arena (key, value | chunk << 16 pairs, min(C_ARENA,
 * max_arena) of them; *n_arena gets the count, at most arena_cap copied).
 * Either pointer may be NULL.  Blocking. */
int px_read_chunks(px_ctx* ctx, uint32_t* table, uint32_t* arena, uint32_t arena_cap, uint32_t* n_arena);
/* Impacts queued by falling chunks and not yet taken by an island pass
 * (16 floats each, pushImpact's layout); *n gets C_IMPACTS.  Blocking. */
int px_read_device_impacts(px_ctx* ctx, float* out16, uint32_t cap, uint32_t* n);
#define PX_STAGE_RESOLVE 16u
#define PX_STAGE_CHUNKS 32u

/* ---- shooting (sim.wgsl:371-403) ------------------------------------------
 * Replaces the player pass's writes of the Camera block (shared.wgsl:22-30,
 * player.wgsl:150-190), which this library does not run: 64 floats,
 * invViewProj, viewProj, eye (m) at [32..35], look at [36..39], feet, vel,
 * state.  The shoot stage reads eye and look.  Blocking copy. */
int px_set_camera(px_ctx* ctx, const float cam[PX_CAMERA_FLOATS]);
/* PX_STAGE_SHOOT runs shoot() after PX_STAGE_FRAME_BEGIN and before
 * PX_STAGE_EDIT (Simulation.ts:223-224): u32(U.fire.x) pellets (at most 16,
 * the reference's workgroup) trace traceSolid from C.eye / VS along C.look,
 * spread by U.fire2.x with seed U.fire2.y, up to 80 m.  Every hit queues an
 * impact (c = voxel + 0.5 - n * min(0.3 r, 2), r = U.fire.y, fr = U.fire.z,
 * impulse = U.fire.w); a pane voxel (fragile 2) in a pool brick is taken with
 * shardEffects and floods its pane in the next island pass; pellet 0's hit
 * writes a flash.  The pellets are traced on the world before any of them
 * takes a voxel and applied in pellet order (the reference's threads race;
 * this order fixes the impact slots and the debris order). */
#define PX_STAGE_SHOOT 64u
#define PX_MAX_FLASHES 8 /* layout.ts:24 */
/* The flash ring (H_FLASHES, sim.wgsl:395-401): PX_MAX_FLASHES records of 8
 * floats, world position (m) at [0..2], radius (m) at [3], U.time.x at [4];
 * record C_FLASH % PX_MAX_FLASHES is the next one written.  Blocking. */
int px_read_flashes(px_ctx* ctx, float out[PX_MAX_FLASHES * 8]);

/* ---- checkpoint of the device world (no counterpart in the reference, which
 * rebuilds its world from the seed on every load and never saves GPU state).
 * px_snapshot copies grid, pool, super grid, ctrl words and free list to a
 * device-side backup; px_restore copies them back and drops queued impacts
 * and labels.  Both are asynchronous on the context's stream. */
int px_snapshot(px_ctx* ctx);
int px_restore(px_ctx* ctx);

/* ---- profiling (no counterpart in the reference, which has no GPU timers;
 * for callers that report per-kernel time and the roofline) -------------
 * px_profile_begin starts recording HIP events around every kernel the
 * library launches on its stream; px_profile_end waits for them and writes,
 * per kernel slot (PX_PROF_*), out[2*k] = summed milliseconds and
 * out[2*k+1] = number of launches, then stops recording. */
#define PX_PROF_PRIMARY 0  /* frame: primary march + shadow-ray generation */
#define PX_PROF_SHADOW 1   /* frame: shadow-ray march */
#define PX_PROF_TRACE 2    /* px_trace / px_trace_device */
#define PX_PROF_EDIT 3     /* island pass, all kernels */
#define PX_PROF_BATCH 4    /* px_trace_batch (the graph launch) */
#define PX_PROF_MLP 5      /* ray-cost predictor: features, predict, train, bin */
#define PX_PROF_PARTICLES 6 /* spawn + particle step, all kernels */
#define PX_PROF_PARTICLE_SIM 7 /* the integration kernel of the step alone */
#define PX_PROF_WRITES 8   /* voxel-write stage, all kernels */
#define PX_PROF_CHUNKS 9   /* island extraction and chunk frames, all kernels */
#define PX_PROF_DENSITY 10 /* AO density rebuild */
#define PX_PROF_RENDER_PRIMARY 11 /* render: primary march with transmittance / glow + shadow queue */
#define PX_PROF_RENDER_SHADOW 12  /* render: queued shadow rays with transmittance */
#define PX_PROF_SHADE 13   /* render: shading (or the debug-view kernel) */
#define PX_PROF_REFLECT 14 /* render: queued reflection rays with their quality-0 shading */
#define PX_PROF_GRAPH 15   /* px_graph_launch: a recorded frame of stages, whole */
#define PX_PROF_SLOTS 16
int px_profile_begin(px_ctx* ctx);
int px_profile_end(px_ctx* ctx, double out[2 * PX_PROF_SLOTS]);

/* ---- learned ray-cost binning (no counterpart in the reference, whose
 * traversal is the plain DDA; for callers that reorder rays into
 * cost-homogeneous wavefronts).  A 16-64-64-1 MLP (ReLU) predicts log2 of a
 * ray's step count from 16 features: origin / dims, unit direction,
 * super-cell occupancy at 8 points along the clipped ray, log2 extent, kind.
 * BF16 weights and activations, FP32 accumulation on MFMA, FP32 Adam state.
 * Binning only reorders rays: px_trace_perm_device traces every ray with the
 * same kernel as px_trace_device, so outputs are bit-identical. */
typedef struct px_mlp px_mlp;
/* Number of FP32 parameters: W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3. */
uint32_t px_mlp_param_count(void);
/* Weights He-uniform from a splitmix64 stream of `seed`, biases 0. */
int px_mlp_create(px_ctx* ctx, uint64_t seed, px_mlp** out);
int px_mlp_destroy(px_mlp* m);
int px_mlp_get_params(px_mlp* m, float* h_params);
/* Sets the FP32 parameters (and their BF16 copies) and resets the Adam state. */
int px_mlp_set_params(px_mlp* m, const float* h_params);
/* d_feat: n * 16 BF16 features of rays of one kind (PX_RAY_*). */
int px_mlp_features(px_ctx* ctx, uint32_t kind, const px_ray* d_rays, uint32_t n, uint16_t* d_feat);
/* d_pred: n FP32 predictions of log2(steps). */
int px_mlp_predict(px_mlp* m, const uint16_t* d_feat, uint32_t n, float* d_pred);
/* Gradient of (1/n) sum (pred - target)^2 over a batch (blocking): h_grad
 * gets px_mlp_param_count() floats in the parameter order, *h_loss the loss. */
int px_mlp_grad(px_mlp* m, const uint16_t* d_feat, const float* d_target, uint32_t n, float* h_grad, float* h_loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; *h_loss (may be
 * NULL: then asynchronous) gets the batch loss before the step. */
int px_mlp_train_step(px_mlp* m, const uint16_t* d_feat, const float* d_target, uint32_t n, float lr, float* h_loss);
/* Stable counting sort of each window of `window` rays into `buckets` cost
 * buckets over [lo, hi) of the prediction; d_perm[k] = source index. */
int px_bin_rays(px_ctx* ctx, const float* d_pred, uint32_t n, uint32_t window, uint32_t buckets, float lo, float hi,
                uint32_t* d_perm);
/* Sort keys of px_mlp_sort_rays: the cost bucket alone (0), or the major
 * axis and sign of the ray direction (6 faces), then the cost bucket. */
#define PX_SORT_FACES 1u
/* Sort each window of 2^log2 consecutive rays (log2 8..30) on its own, the
 * windows in input order, instead of all n rays at once (0): rays that come
 * in a spatially coherent order keep part of it. */
#define PX_SORT_WINDOW(log2) ((uint32_t)(log2) << 8)
/* Features, prediction and a stable counting sort of ALL n rays by key, in
 * three launches.  The key is the cost bucket (`buckets` of the predicted
 * log2 steps over [lo, hi), 1..256), or with PX_SORT_FACES (face, bucket)
 * with face = 2 * major axis + (component < 0) of the direction (the first of
 * equal magnitudes wins; buckets 1..42); with PX_SORT_WINDOW(k), by (window
 * of 2^k rays, key).  Without flags d_perm is the order
 * px_bin_rays(window = n) gives for px_mlp_predict's output on
 * px_mlp_features(kind, d_rays).  Device scratch (keys, per-block counts)
 * grows with n and lives with the model.  Calls on one model are ordered on
 * its context's stream. */
int px_mlp_sort_rays(px_mlp* m, uint32_t kind, const px_ray* d_rays, uint32_t n, uint32_t buckets, float lo, float hi,
                     uint32_t flags, uint32_t* d_perm);
/* The learned-binned traversal in one call: px_mlp_sort_rays, then
 * px_trace_perm_device in that order (d_perm is the scratch order).  The
 * hits equal px_trace_device's for the same rays, bit for bit. */
int px_mlp_trace_sorted(px_mlp* m, uint32_t kind, const px_ray* d_rays, uint32_t n, uint32_t buckets, float lo,
                        float hi, uint32_t flags, uint32_t* d_perm, px_hit* d_out);
/* Traces d_rays[d_perm[k]] for k = 0..n-1 in that order and writes the hit to
 * d_out[d_perm[k]] (kinds as px_trace_device). */
int px_trace_perm_device(px_ctx* ctx, uint32_t kind, const px_ray* d_rays, const uint32_t* d_perm, uint32_t n,
                         px_hit* d_out);

#ifdef __cplusplus
}
#endif

#endif /* PIXELS_H */
