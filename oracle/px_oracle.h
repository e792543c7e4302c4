/*
 * px_oracle.h - CPU oracle for the brick-map voxel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * WGSL (Crunchy000/pixels) used as the checker of the HIP library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark call it.
 * The product path never links or calls it.
 *
 * Built with -O2 -ffp-contract=off and without fast-math so every FP32
 * operation rounds once, in the order the WGSL source writes it.
 */
#ifndef PX_ORACLE_H
#define PX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct o_world {
  int32_t dx, dy, dz;          /* voxel dims */
  const uint32_t* grid;        /* brick grid */
  const uint32_t* pool;        /* pool words: 256 per brick, 2 voxels per word */
  const uint32_t* super_grid;  /* non-empty cell count per 4^3-brick region */
  const float* mats;           /* 256 * 20 */
} o_world;

/* Same 16-byte record as px_hit in include/pixels.h. */
typedef struct o_hit {
  uint32_t key;
  float t;
  uint32_t steps;
  uint8_t face;
  uint8_t mat;
  uint16_t value_hi;
} o_hit;

/* march() of raymarch.wgsl:44-162 on one ray. trans_out (3 floats) may be NULL. */
void o_march(const o_world* w, const float ro[3], const float rd[3], float max_t, int shadow,
             o_hit* out, float* trans_out);
/* march() with the shading outputs of raymarch.wgsl:6-14: transmittance,
 * glow (camera rays: glassSurface adds the Fresnel-weighted sky and the sun
 * glint, which needs the 52 uniform floats u52) and the outer / inner step
 * counters (:18-20).  Any output pointer may be NULL. */
void o_march_ex(const o_world* w, const float* u52, const float ro[3], const float rd[3], float max_t, int shadow,
                o_hit* out, float* trans_out, float* glow_out, uint32_t* outer_out, uint32_t* inner_out);
/* traceSolid() of sim.wgsl:296-357. */
void o_trace_solid(const o_world* w, const float ro[3], const float rd[3], float max_t, o_hit* out);
/* Segment DDA of particleSim, sim.wgsl:1110-1145 (pos = ro, displacement = mv). */
void o_segment(const o_world* w, const float pos[3], const float mv[3], o_hit* out);

/* Batch of 32-byte rays (ro[3], rd[3], max_t, flags) of one kind (0 primary,
 * 1 shadow, 2 shoot, 3 segment; kind < 0 = per-ray flags & 3), threads >= 1. */
void o_trace_batch(const o_world* w, int kind, const float* rays, uint32_t n, o_hit* out, int threads);

/* One frame of fsMain + shade() shadow rays (raymarch.wgsl:304-356,388-398). */
void o_frame(const o_world* w, const float* u52, const float* lights_vec4, uint32_t n_vec4,
             const float* rays, uint32_t n, o_hit* prim_out, o_hit* shadow_out, uint32_t* n_cast,
             int threads);

/* Debris particle (shared.wgsl:33-38), 32 bytes. */
typedef struct o_particle {
  float pos[3];
  uint32_t value;
  float vel[3];
  float life;
} o_particle;

/* Edit chain: carve + pane flood + maint + island labels (sim.wgsl:194-944,1205-1274).
 * The world is given as a dense voxel array (x fastest) and a per-cell kind
 * (0 empty, 1 uniform, 2 pool brick) taken from the grid before the pass.
 * On return `vox` holds the decoded voxels after the pass and `kind` the
 * canonical kinds after maint.  Labels of unsupported components are written
 * (sorted by key) up to cap; *n_labels gets the total.  stats[0] = voxels
 * destroyed, stats[1] = bill in dimes, stats[2] = cells allocated,
 * stats[3] = carve jobs, stats[4] = voxels removed by the flood. */
int o_island_pass(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                  const float* impacts16, uint32_t n_impacts, uint32_t* lab_key, uint32_t* lab_count,
                  uint32_t cap, uint32_t* n_labels, uint64_t* stats);

/* The edit pass with its side effects and the seeded search.  mode 0 labels
 * every component (o_island_pass is mode 0 with the debris dropped), mode 1
 * only components that hold a seed (pushSeed, sim.wgsl:186, 524, 615, 1068),
 * with the floating ones kept in label order while their voxel total stays
 * within node_budget.  vs = U.dims.w, frame_seed = u32(U.time.y) *
 * 0x9E3779B9.  seed_keys: pending seeds of landed chunks.  flood_in:
 * n_flood_in (key, id) entries o_shoot pushed for struck pane voxels; the
 * flood takes each one's same-id neighbours (floodStep, sim.wgsl:599-618)
 * after carve, and in mode 1 its other neighbours are seeds.  dirty_in (one
 * byte per cell, may be NULL): cells o_shoot's takes touched, maintained
 * here with the pass's own.  Every spawn()
 * call (dropRubbleAbove, carve debris, shard debris) is written to spawns in
 * call order up to spawn_cap; *n_spawns counts them all.  stats: [0..4] as
 * o_island_pass, [5] seed voxels that are node voxels at the search,
 * [6] floating components over the node budget, [7] rubble voxels dropped. */
typedef struct o_edit_params {
  uint32_t mode;
  uint32_t node_budget;
  float vs;
  uint32_t frame_seed;
} o_edit_params;
int o_edit_pass(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                const o_edit_params* P, const float* impacts16, uint32_t n_impacts, const uint32_t* seed_keys,
                uint32_t n_seed_keys, const uint32_t* flood_in, uint32_t n_flood_in, const uint8_t* dirty_in,
                uint32_t* lab_key, uint32_t* lab_count, uint32_t cap, uint32_t* n_labels, o_particle* spawns,
                uint32_t spawn_cap, uint32_t* n_spawns, uint64_t* stats);

/* shoot (sim.wgsl:371-403) on the dense world of o_edit_pass, run before it:
 * min(u32(fire[0]), 16) pellets along look (spread by fire2[0] with seed
 * fire2[1]: rnd3 of u32(seed) * 7919 + i * 104729, then normalised) traced
 * with traceSolid from eye / vs to 80 / vs, every pellet on the world before
 * any pellet's effects, which then run in pellet order: pushImpact (slot
 * *n_impacts, stored below max_impacts), a struck pane voxel of a pool brick
 * taken with shardEffects (spawns in call order) and its flood entry (key,
 * id; 64 kept, *n_flood counts all), pellet 0's flash into record
 * (*flash_ctr)++ % 8 of flashes.  Cells the takes touch are set in dirty
 * (one byte per cell) for o_edit_pass's dirty_in.  stats: [0] destroyed,
 * [1] bill (dimes), [2] pellets that hit. */
typedef struct o_shoot_params {
  float fire[4];  /* pellets, radius, fragile radius, impulse */
  float fire2[2]; /* spread, seed */
  float time_x, vs;
  float eye[3], look[3];
  uint32_t frame_seed, max_impacts;
} o_shoot_params;
int o_shoot(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
            const o_shoot_params* P, float* impacts16, uint32_t* n_impacts, uint32_t* flood, uint32_t* n_flood,
            uint8_t* dirty, float* flashes, uint32_t* flash_ctr, o_particle* spawns, uint32_t spawn_cap,
            uint32_t* n_spawns, uint64_t* stats);

/* Unsupported-component labelling alone (the last stage of o_island_pass). */
int o_islands(int32_t dx, int32_t dy, int32_t dz, const uint16_t* vox, const float* mats,
              uint32_t* lab_key, uint32_t* lab_count, uint32_t cap, uint32_t* n_labels);

/* Shading (px_shade.c).  Elementary functions in double precision rounded
 * once to FP32 (the implementation's choice for WGSL's sin, cos, exp, exp2,
 * log2, pow, atan2). */
float o_sinf(float x);
float o_cosf(float x);
float o_expf(float x);
float o_exp2f(float x);
float o_log2f(float x);
float o_powf(float x, float y);
float o_atan2f(float y, float x);
/* skyColor (raymarch.wgsl:212-264) and materialColor (render_common.wgsl:58-140) */
void o_sky_color(const float* u52, const float rd[3], float out[3]);
void o_material_color(const float* mats, const float* u52, uint32_t v, const int32_t vp[3], float out[3]);
/* AO density texture, one byte per brick cell (VoxelWorld.ts:159-170). */
void o_density(const o_world* w, uint8_t* out);
/* fsMain colour (raymarch.wgsl:387-450) for each 32-byte ray: RGBA floats,
 * the debug views of U.flags.w included.  density from o_density. */
void o_render(const o_world* w, const float* u52, const float* lights_vec4, uint32_t n_vec4, const uint8_t* density,
              const float* rays, uint32_t n, float* rgba, int threads);

/* Integer hashes of shared.wgsl:65-81 (exported for tests). */
uint32_t o_hashu(uint32_t x);
uint32_t o_hash3i(int32_t x, int32_t y, int32_t z);
float o_rnd(uint32_t seed);
/* removes() of sim.wgsl:229-239 for one voxel (impact = 16 floats). */
int o_removes(const float* impact16, const float* mats, int32_t x, int32_t y, int32_t z, uint32_t v);


/* Particle counters: C_SPAWNED, C_PCURSOR, C_WRITES (sim.wgsl:167-168,187)
 * and the alive count args[A_DRAWP + 1] (:1083). */
#define O_P_SPAWNED 0
#define O_P_PCURSOR 1
#define O_P_WRITES 2
#define O_P_ALIVE 3

/* frameBegin (sim.wgsl:363-369): zeroes C_SPAWNED, C_WRITES and the alive count. */
void o_frame_begin(uint32_t* ctr);
/* spawn() (sim.wgsl:166-170) for src[0..n) in order. */
void o_spawn(o_particle* parts, uint32_t max_particles, uint32_t budget, uint32_t* ctr, const o_particle* src,
             uint32_t n);
/* particleSim (sim.wgsl:1088-1173) over every slot, serially in index order:
 * alive[] gets the alive slots, write_list (key, value) pairs the deposit
 * writes that fit max_writes. */
void o_particle_step(const o_world* w, o_particle* parts, uint32_t max_particles, float dt, float gravity, float vs,
                     uint32_t max_writes, uint32_t* ctr, uint32_t* write_list, uint32_t* alive);

/* Voxel writes (writeArgs / writeRequest / allocBricks / writeApply,
 * sim.wgsl:1178-1203, 459-483) and maint of the touched cells, on the dense
 * world of o_island_pass: writes = (key, value) pairs in list order.  Cells
 * of a write's first target without a brick get one (the first max_alloc
 * distinct cells, at most free_top of them); each write then puts its value
 * into the first empty voxel of p, p+y, p+2y, stopping at a cell without a
 * pool brick.  stats: [0] voxels written, [1] cells allocated, [2] requests
 * dropped, [3] cells dirtied. */
int o_apply_writes(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const uint32_t* writes,
                   uint32_t n, uint32_t max_alloc, uint32_t free_top, uint64_t* stats);

/* Island extraction and falling chunks (sim.wgsl:862-923, 951-1075): chunk
 * table words as the reference (0 state, 1 vy f32, 2 dy, 3 frac f32,
 * 4 fragile, 5 voxel count, 6 blockedAt, 7 broke count, 8 steps, 9 landing
 * speed f32); arena entries (key, value | chunk << 16). */
typedef struct o_resolve_params {
  uint32_t max_chunks, max_arena, component_budget, max_alloc, free_top;
  float vs;            /* U.dims.w */
  uint32_t frame_seed; /* u32(U.time.y) * 0x9E3779B9 */
} o_resolve_params;
/* counters of the chunk stages */
#define O_R_C_CHUNKS 0
#define O_R_C_ARENA 1
#define O_R_S_ISLANDS 2
#define O_R_S_BUDGET 3
#define O_R_S_BILL 4
#define O_R_S_DESTROYED 5
#define O_R_ALLOCS 6
#define O_R_ALLOC_FAIL 7
#define O_R_S_FALLING 8
#define O_R_C_WRITES 9
#define O_R_C_IMPACTS 10
#define O_R_WORDS 12
/* resolveChunks + resolveExpand + allocBricks + resolveExtract + maint on the
 * dense world after an island pass; labels sorted by key (o_islands).
 * Spawned records (dropRubbleAbove, crumbs) go to spawns in order. */
int o_resolve(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
              const uint32_t* lab_key, const uint32_t* lab_count, uint32_t n_labels, const o_resolve_params* prm,
              uint32_t* chunks, uint32_t* arena, uint32_t* ctr, o_particle* spawns, uint32_t spawn_cap,
              uint32_t* n_spawns);
/* One frame of chunkArgs / chunkBegin / chunkCollide / chunkStep / chunkLand:
 * writes (key, value) pushed at C_WRITES, impacts (16 floats) at C_IMPACTS,
 * shattered voxels spawned in arena order, and the keys of the voxels under
 * landed voxels that pushSeed re-checks (seeds; n_seeds counts them all).
 * The world is read, not written. */
int o_chunk_frame(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                  uint32_t* chunks, uint32_t* arena, uint32_t* ctr, const o_resolve_params* prm, float dt,
                  float gravity, uint32_t max_writes, uint32_t* writes, uint32_t max_impacts, float* impacts,
                  o_particle* spawns, uint32_t spawn_cap, uint32_t* n_spawns, uint32_t* seeds, uint32_t seed_cap,
                  uint32_t* n_seeds);

/* SDF primitive in voxel units, the px_prim layout of include/pixels.h:
 * type 0 box (c, h), 1 rounded box (c, h, r), 2 sphere (c, r), 3 capped
 * cylinder (c, r, half height, axis), 4 capsule (a, b, r), 5 torus (c, R, r,
 * axis), 6 axis-aligned box (lo, hi); cy, sy: cos / sin of the yaw; lo, hi:
 * world bounds. */
typedef struct o_prim {
  uint32_t type;
  uint32_t mat;
  double p[12];
  double cy, sy;
  double lo[3], hi[3];
} o_prim;
double o_prim_sdf(const o_prim* p, double x, double y, double z);
/* apply() of shapes.ts:334-414 for prims in order (Op.Add, constant paint)
 * on a dense world (vox, x fastest; kind per cell, compactCell semantics). */
void o_raster(int32_t dx, int32_t dy, int32_t dz, const o_prim* prims, uint32_t n, uint16_t* vox, uint8_t* kind);

#ifdef __cplusplus
}
#endif

#endif
