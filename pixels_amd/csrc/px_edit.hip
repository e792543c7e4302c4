// px_edit.hip - the destruction edit chain on the heap brick pool, gfx950.
//
// One island pass = the reference's chain run to quiescence:
//   impactJobs  sim.wgsl:414-443   cells in each impact's AABB -> jobs, uniform
//                                  cells within reach() -> alloc requests
//   allocBricks sim.wgsl:459-483   pop the free list, fill, publish (popFree :148-156)
//   carve       sim.wgsl:529-579   removes() (:229-239) per voxel of every job brick
//   pane flood  sim.wgsl:250-271,511-523,599-618
//   maint       sim.wgsl:1219-1274 collapse all-air / all-same pool bricks
//   islands     sim.wgsl:620-882   6-connected components of non-air,
//                                  non-rubble, non-structural voxels; anchored by
//                                  a voxel with y == 0 or a structural neighbour
//
// MI355X layout of the work:
//  * impactJobs ORs impact k into a 64-bit mask per cell (MAX_IMPACTS = 64), so
//    carve runs one wavefront per cell over all its impacts: each cell has a
//    single writer and needs no CAS loop (the reference CASes words shared by
//    several (impact, brick) workgroups).
//  * The pane flood is a connected-component pass instead of a ring-by-ring
//    BFS: over the voxels of pane materials (fragile 2) in pool bricks, same
//    material id connected; every component holding a voxel that carve removed
//    is removed whole.  That is the closure the reference's flood reaches
//    (each carved pane voxel takes its same-id neighbours, each taken voxel
//    takes its own), in a fixed number of launches.
//  * Components are labelled in two levels: a union-find in LDS over the 512
//    voxels of each brick (one row of 8 voxels per lane), then a lock-free
//    union-find over brick-local components whose node id is the voxel key of
//    the component's minimum voxel.  Links always point to the smaller key, so
//    every root is the minimum voxel key of its component: the canonical label.
//  * No counter is shared by the whole grid on the hot paths: the 8 XCDs have
//    separate L2 caches, and device-wide atomics on one address serialise
//    beyond them.  Node and pane-cell lists live in static per-block ranges,
//    destroyed/bill/dirty counts are per-block partials summed by the last
//    kernel, the free list is popped in one batch by request index, and maint
//    walks the job and pane lists with test-and-clear of the dirty bits.
//  * Every kernel is persistent over a device-side list or counter; the host
//    never waits inside the pass.
#include <algorithm>

#include "px_internal.h"
#include "px_march.h"
#include "px_scan.h"

#pragma clang fp contract(off)

namespace {

constexpr uint32_t kLabNone = 0xFFFFu;
constexpr uint32_t kFlagAnchor = 1u;
constexpr uint32_t kFlagHit = 2u;
// Grid of the edit kernels: 16384 single-wave blocks, 64 per CU (256 CUs),
// each taking every 16384th cell or list entry, with one node / pane-cell
// list per block.  More blocks than stay resident lets the dispatcher even
// out the blocks' uneven work: on the 64-impact island pass of the
// 512 x 256 x 512 floor scene 8192 blocks took 0.95 ms and 16384 0.91 ms
// (k_local_ccl<0> 180 -> 142 us, k_flood<0> 83 -> 56 us), 32768 0.95 ms.
constexpr uint32_t kEditBlocks = 16384;
// Carve's count pass keeps each row's removals (the impact loop, most of a
// carved cell's time) for the first kCarveRowCache job entries, 512 B each
// (16 MiB); the apply pass reads them instead of testing the impacts again.
constexpr uint32_t kCarveRowCache = 32768;
constexpr uint32_t kPartWords = 8;      // per-block partials
constexpr uint32_t kEditFinishBlocks = 64;
constexpr uint32_t kCountWords = 4;     // per-block list lengths

// edit-pass counters (ec[])
enum : uint32_t {
  EC_JOBCELLS = 0, EC_ALLOCS, EC_ACTIVE, EC_NIMP, EC_LABELS, EC_LABELS_OK, EC_JOBS, EC_CHUNKS,
  EC_SEEDS, EC_OVERFLOW, EC_NB_T, EC_NB_DROP, EC_REEMIT, EC_CARVE_N, EC_FLOOD_N, EC_WORDS = 16,
  // voxel-write stage (after the island pass's words)
  EC_W_N = 16, EC_W_DIRTY, EC_W_WRITTEN, EC_W_DROP,
  // extraction and chunk stages
  EC_R_CELLS = 24, EC_R_DIRTY, EC_R_BILL, EC_R_DESTROYED, EC_R_SPAWN_ACC, EC_C_LAND, EC_C_CAND, EC_R_ARENA0,
  // k_edit_finish's block sums (destroyed, bill, dirty, nodes) and its count of finished blocks
  EC_F_SUM = 32, EC_F_DONE = EC_F_SUM + 4
};
// component flags written by the extraction (beside kFlagAnchor / kFlagHit):
// kept in the world (over the component budget), crumbles into debris, and
// the chunk id + 1 in bits 8..16 (resolveChunks, sim.wgsl:862-880)
constexpr uint32_t kFlagKeep = 4u;
// component flag of the seeded search: the component holds a seed voxel
constexpr uint32_t kFlagSeed = 16u;
// k_impact_jobs: 64 blocks of 4 waves per impact row (MAX_IMPACTS = 64
// rows), each wave taking 64 cells of the impact's box at a time (32 and 128
// blocks were slower on the 64-impact island pass: 32 / 26 against 24 us)
constexpr uint32_t kJobBlocks = 64;
constexpr uint32_t kFlagCrumb = 8u;
// per pool brick (labelling): bnode = its island node if its node-class voxels
// form one component, kNoNode if it has none, kManyNodes otherwise; bface =
// 12 masks of 64 bits, node-class voxels then structural voxels, each of the
// faces x = 0, y = 0, z = 0 then x = 7, y = 7, z = 7 (face_masks())
constexpr uint32_t kNoNode = 0xFFFFFFFFu, kManyNodes = 0xFFFFFFFEu;
constexpr uint32_t kFaceMasks = 12;
constexpr uint32_t kNoBlock = 0xFFFFFFFFu;
constexpr uint32_t kMaxChunkIds = 511;  // chunk id + 1 fits bits 8..16
// per-block partials (part[b * kPartWords + k])
enum : uint32_t { PT_CARVE_D = 0, PT_CARVE_B, PT_FLOOD_D, PT_FLOOD_B, PT_DIRTY, PT_FREED };
// per-block list lengths (bcount[b * kCountWords + k])
enum : uint32_t { BC_NODES = 0, BC_PANENODES, BC_PANECELLS };

struct EditView {
  uint32_t* grid;
  uint32_t* pool;
  uint32_t* sgrid;
  const float* mats;
  uint32_t* ctrl;
  uint32_t* free_list;
  uint32_t pool_cap;
  int32_t dims[3], bd[3], sd[3];
  uint32_t ncell;
  uint64_t* imask;
  uint32_t* dirty_flags;
  uint32_t* job_list;
  uint32_t* alloc_list;
  uint32_t* pane_cells;  // kEditBlocks ranges of slots_cell
  uint32_t* par;
  uint32_t* cnt;
  uint32_t* flag;
  uint16_t* lab;
  uint32_t* bnode;  // per pool brick: its node if its island voxels are one component, else kNoNode / kManyNodes
  uint64_t* bface;  // per pool brick: kFaceMasks face masks (face_masks())
  uint32_t* nodes;  // kEditBlocks ranges of slots_node
  uint32_t slots_cell, slots_node;
  uint32_t walk_cap;  // step bound of the union-find walks (>= the node capacity)
  uint32_t* part;
  uint32_t* bcount;
  uint32_t* out;
  uint32_t out_cap;
  uint32_t* ec;
  float* impacts;     // the impact queue (pushImpact's records), C_IMPACTS of them
  uint32_t* shots;    // shoot's pane flood entries and dirtied cells (kShotWords)
  uint32_t* work;     // cells the count passes found work in: carve at [0, ncell), flood at [ncell, 2 ncell)
  uint2* carve_rows;  // per job entry i < kCarveRowCache, per lane: removed_row()'s rem | rk << 8
  float* flashes;     // flash ring, PX_MAX_FLASHES records of 8 floats
  uint32_t max_imp;   // min(max_impacts, 64)
  uint32_t* seeds;    // seeded mode: one bit per voxel key, pushSeed (sim.wgsl:180-184)
  uint32_t seed_words;
  uint32_t mode;      // PX_ISLANDS_ALL / PX_ISLANDS_SEEDED
  uint32_t node_budget;
  uint32_t* etot;     // totals of the effects scan: carve records, flood records
  // voxel writes
  uint32_t* wlist;
  uint32_t max_writes, max_alloc;
  uint32_t* cfirst;
  uint64_t* wkeys;
  const uint64_t* wsorted;
  uint32_t* wdirty;
  // extraction and chunks
  uint32_t* chunks;  // max_chunks x 16 words (sim.wgsl:941-946 layout)
  uint32_t* arena;   // (key, value | chunk << 16) pairs
  uint32_t max_chunks, max_arena, comp_budget, max_impacts;
  uint2* cpair;      // per cell, then per extract cell / arena entry: counts
  uint2* coff;       // exclusive offsets of cpair
  uint2* tiles;      // per 1024-element tile: sums, then offsets
  uint32_t* rlist;   // extract cells in cell order
  uint32_t* rdirty;  // cells dirtied by the extraction
  float4* parts;     // particle ring
  uint32_t pmask, spawn_budget;
  uint32_t fseed;    // frameSeed(), sim.wgsl:69
  float vs, dt, grav;

};

// Error bits of ec[EC_OVERFLOW]: a node list overflowed, a brick label was
// not a local index, a union-find walk met a parent above its child or ran
// past its step bound (the pass's labels are then not trusted).
constexpr uint32_t kErrNodes = 1u;
constexpr uint32_t kErrLabel = 2u;
constexpr uint32_t kErrWalk = 4u;
// Steps of an LDS walk inside one brick: a chain of strictly decreasing
// local indices has at most 512 entries.
constexpr uint32_t kLdsWalkCap = 512u;

// Union-find over voxel keys.  par[x] <= x always (links go to smaller keys),
// so chains strictly decrease.  Plain loads are enough: a stale parent is
// still an ancestor, and a stale root only makes the linking CAS fail and
// return the current parent.  Every walk checks that order (a parent above
// its child can only be memory no node wrote in this pass) and stops after a
// step bound that no chain of distinct nodes reaches (the bound derives from
// walk_cap >= the node capacity of the pass); either way it sets kErrWalk
// and stops on a key below the one it started from, so every access stays
// inside par[] and every walk ends.  The error goes to a register (`err`)
// that the kernel ORs into ec[EC_OVERFLOW] once, at its end.

// Unions of up to three pairs per lane (pair d when bit d of act is set),
// walked together: each round reads the parents of both nodes of every open
// pair at once, so the walks share their load latencies, and links two
// roots with one CAS (the larger root under the smaller).  Each round moves
// a pair's a or b (or both) to a smaller key, so a pair ends within the
// lengths of its two chains; the bound of 2 cap + 2 rounds (cap per walk)
// is never reached by chains of distinct nodes.  Once a pair is joined under
// root r, its two start nodes point at r (a store of an ancestor, like path
// halving).
PX_DEV void uf_union3(uint32_t* par, uint32_t a[3], uint32_t b[3], uint32_t act, uint32_t cap, uint32_t& err) {
  uint32_t a0[3], b0[3];
  for (int d = 0; d < 3; d++) {
    a0[d] = a[d];
    b0[d] = b[d];
  }
  const uint64_t rounds = 2ull * cap + 2ull;
  for (uint64_t it = 0; act != 0u && it < rounds; it++) {
    uint32_t pa[3], pb[3];
    for (int d = 0; d < 3; d++) {
      pa[d] = ((act >> d) & 1u) ? par[a[d]] : a[d];
      pb[d] = ((act >> d) & 1u) ? par[b[d]] : b[d];
    }
    for (int d = 0; d < 3; d++) {
      if (!((act >> d) & 1u)) continue;
      if (pa[d] > a[d] || pb[d] > b[d]) {
        err = kErrWalk;
        act &= ~(1u << d);
        continue;
      }
      if (pa[d] != a[d] || pb[d] != b[d]) {  // not both roots yet: one step up
        a[d] = pa[d];
        b[d] = pb[d];
        continue;
      }
      uint32_t r = a[d];
      if (a[d] != b[d]) {
        const uint32_t hi = max(a[d], b[d]), lo = min(a[d], b[d]);
        const uint32_t old = atomicCAS(par + hi, hi, lo);
        if (old != hi) {  // hi gained a parent meanwhile: go on from it
          if (old > hi) {
            err = kErrWalk;
            act &= ~(1u << d);
          }
          a[d] = old;
          b[d] = lo;
          continue;
        }
        r = lo;
      }
      if (a0[d] > r) par[a0[d]] = r;
      if (b0[d] > r) par[b0[d]] = r;
      act &= ~(1u << d);
    }
  }
  if (act != 0u) err = kErrWalk;
}

// LDS union-find over the 512 voxels of one brick (links to smaller indices,
// path halving; parents only decrease, so a racing halving store still
// writes an ancestor).  Same order check as the global walks, step bound
// 512.  The walk's reads and stores are relaxed workgroup-scope atomics: each read
// sees other lanes' links, and they stay LDS instructions (a volatile
// access through the generic pointer would be a flat access that also
// waits for every global load in flight).
PX_DEV uint32_t lds_get(uint32_t* lab, uint32_t i) {
  return __hip_atomic_load(lab + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

PX_DEV uint32_t lds_find(uint32_t* lab, uint32_t x, uint32_t& err) {
  for (uint32_t it = 0; it < kLdsWalkCap; it++) {
    const uint32_t p = lds_get(lab, x);
    if (p == x) return x;
    if (p > x) break;
    const uint32_t gp = lds_get(lab, p);
    if (gp > p) break;
    if (gp != p) __hip_atomic_store(lab + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = gp;
  }
  err = kErrWalk;
  return x;
}

PX_DEV void lds_union(uint32_t* lab, uint32_t a, uint32_t b, uint32_t& err) {
  for (uint32_t it = 0; it < kLdsWalkCap; it++) {
    a = lds_find(lab, a, err);
    b = lds_find(lab, b, err);
    if (a == b) return;
    if (a < b) {
      uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t old = atomicCAS(lab + a, a, b);
    if (old == a) return;
    if (old > a) break;
    a = old;
  }
  err = kErrWalk;
}

// inclusive prefix sum over the wavefront
PX_DEV uint32_t wave_scan(uint32_t v) {
  uint32_t lane = __lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

PX_DEV uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

PX_DEV void cell_coord(const EditView& e, uint32_t ci, int32_t& bx, int32_t& by, int32_t& bz) {
  bx = (int32_t)(ci % (uint32_t)e.bd[0]);
  by = (int32_t)((ci / (uint32_t)e.bd[0]) % (uint32_t)e.bd[1]);
  bz = (int32_t)(ci / ((uint32_t)e.bd[0] * (uint32_t)e.bd[1]));
}

PX_DEV uint32_t vkey(const EditView& e, int32_t x, int32_t y, int32_t z) {
  return (uint32_t)x + (uint32_t)e.dims[0] * ((uint32_t)y + (uint32_t)e.dims[1] * (uint32_t)z);
}

// key offset of local index li inside a brick
PX_DEV uint32_t loff(const EditView& e, uint32_t li) {
  return (li & 7u) + (uint32_t)e.dims[0] * (((li >> 3) & 7u) + (uint32_t)e.dims[1] * (li >> 6));
}

PX_DEV uint32_t super_of(const EditView& e, uint32_t ci) {
  int32_t bx, by, bz;
  cell_coord(e, ci, bx, by, bz);
  return (uint32_t)((bx >> 2) + e.sd[0] * ((by >> 2) + e.sd[1] * (bz >> 2)));
}

// Cell of step k of this block in a pass over every cell: step k covers
// cells [k G, (k + 1) G) (G = gridDim.x), so neighbouring cells are visited
// at about the same time, and block b takes cell k G + (b + k R) mod G of
// it.  The rotation R (odd, far from a multiple of a row or a slice) gives
// each block cells at many heights and columns: a world's solid cells sit in
// a few y layers, and with R = 0 the blocks whose column is in them take up
// to 3.4x the mean number of pool bricks.
constexpr uint32_t kStepRot = 1031u;
PX_DEV uint32_t step_cell(uint32_t k) { return k * gridDim.x + (blockIdx.x + k * kStepRot) % gridDim.x; }

// lowest set lane of a wave-uniform lane mask, and lane l's value of v
PX_DEV uint32_t first_lane(uint64_t m) { return (uint32_t)__ffsll((unsigned long long)m) - 1u; }
PX_DEV uint32_t lane_value(uint32_t v, uint32_t l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l); }
PX_DEV uint64_t lane_value64(uint64_t v, uint32_t l) {
  return ((uint64_t)lane_value((uint32_t)(v >> 32), l) << 32) | lane_value((uint32_t)v, l);
}

// lane's 16 bytes (8 voxels) of the brick grid word g names, zero unless g
// is a pool brick
PX_DEV uint4 brick_row(const EditView& e, uint32_t g, uint32_t lane) {
  if (g == 0u || (g & PX_UNIFORM_BIT) != 0u) return make_uint4(0u, 0u, 0u, 0u);
  return reinterpret_cast<const uint4*>(e.pool + (size_t)(g - 1u) * 256u)[lane];
}

PX_DEV void set_dirty(const EditView& e, uint32_t ci) { atomicOr(&e.dirty_flags[ci >> 5], 1u << (ci & 31u)); }

// flag[node] |= bits for every lane with node != PX_NONE, one atomic per
// distinct node of the wavefront, and none when a coherent read shows the
// bits already set: the lanes of a wavefront mostly name one component root,
// and same-address atomics from every lane would serialise.
PX_DEV void wave_flag_unique(uint32_t* flag, uint32_t node, uint32_t bits) {
  const uint32_t lane = __lane_id();
  for (;;) {
    const uint64_t act = __ballot(node != PX_NONE);
    if (act == 0ull) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
    const uint32_t n0 = __shfl(node, (int)leader);
    if (lane == leader) {
      const uint32_t f = __hip_atomic_load(flag + n0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((f & bits) != bits) atomicOr(flag + n0, bits);
    }
    if (node == n0) node = PX_NONE;
  }
}

// wave_flag_unique over eight nodes per lane: one atomic per distinct node
// of the wavefront.
PX_DEV void wave_flag_unique8(uint32_t* flag, uint32_t node[8], uint32_t bits) {
  const uint32_t lane = __lane_id();
  for (;;) {
    uint32_t mine = PX_NONE;
    for (int x = 7; x >= 0; x--)
      if (node[x] != PX_NONE) mine = node[x];
    const uint64_t act = __ballot(mine != PX_NONE);
    if (act == 0ull) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
    const uint32_t n0 = __shfl(mine, (int)leader);
    if (lane == leader) {
      const uint32_t f = __hip_atomic_load(flag + n0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((f & bits) != bits) atomicOr(flag + n0, bits);
    }
    for (int x = 0; x < 8; x++)
      if (node[x] == n0) node[x] = PX_NONE;
  }
}

PX_DEV uint32_t price_dimes(const float* mats, uint32_t v) {
  float p = mats[(v & 0xffu) * 20u + 15u];
  return (uint32_t)px_f2i(p * 10.0f + 0.5f);
}

// reach() of sim.wgsl:242-248
PX_DEV float reach(const float* im, const float* mats, uint32_t id) {
  const float* E = mats + id * 20u + 16u;
  float r = 0.0f;
  if (E[1] > 0.5f) r = im[4] * 1.2f;
  if (E[0] > 0.0f) r = fmaxf(r, fmaxf(im[3] * fminf(1.4f, E[0]), fminf(im[3], 2.5f)) * 1.2f);
  return r;
}

// class tables: MODE 0 islands (1 = non-air, non-rubble, non-structural),
// MODE 1 panes (material id of fragile-2 materials); str = structural, non-rubble
template <int MODE>
PX_DEV void load_tables(const float* mats, uint8_t* s_cls, uint8_t* s_str) {
  for (uint32_t id = threadIdx.x; id < 256u; id += blockDim.x) {
    const float* E = mats + id * 20u + 16u;
    bool rubble = (id & 0x80u) != 0u;
    bool structural = E[2] > 0.5f;
    if (MODE == 0) s_cls[id] = (id != 0u && !rubble && !structural) ? 1u : 0u;
    else s_cls[id] = (id != 0u && E[1] > 1.5f) ? (uint8_t)id : 0u;
    s_str[id] = (id != 0u && !rubble && structural) ? 1u : 0u;
  }
}

// The voxels of a brick's six faces as 64-bit masks, from its rows (lane l
// holds row y = l & 7, z = l >> 3, bit x of `bits`): out[0..2] the faces
// x = 0, y = 0, z = 0 and out[3..5] the faces x = 7, y = 7, z = 7, with bit
// y + 8 z on x faces, x + 8 z on y faces and x + 8 y on z faces, so the
// touching faces of two neighbouring bricks index their voxel pairs alike.
PX_DEV void face_masks(uint32_t bits, uint64_t out[6]) {
  // lane t of a y or z face stands for (x = t & 7, the other axis t >> 3):
  // it fetches that row's bits and the ballot packs the face
  const uint32_t t = __lane_id(), x = t & 7u, q = t >> 3;
  out[0] = __ballot((bits & 1u) != 0u);
  out[3] = __ballot(((bits >> 7) & 1u) != 0u);
  out[1] = __ballot(((__shfl(bits, (int)(8u * q)) >> x) & 1u) != 0u);        // y = 0, z = q
  out[4] = __ballot(((__shfl(bits, (int)(8u * q + 7u)) >> x) & 1u) != 0u);   // y = 7, z = q
  out[2] = __ballot(((__shfl(bits, (int)q) >> x) & 1u) != 0u);               // z = 0, y = q
  out[5] = __ballot(((__shfl(bits, (int)(56u + q)) >> x) & 1u) != 0u);       // z = 7, y = q
}

// defined with the stages they belong to, run inside other kernels of the
// island pass
PX_DEV void effects_finish(const EditView& e);
PX_DEV void shot_flood(const EditView& e);

// ---------------------------------------------------------------------------

// Start of a pass (impactArgs + searchBegin, sim.wgsl:408-412, 700-720): the
// impacts to take (min(C_IMPACTS, MAX_IMPACTS)) and whether the pass runs at
// all: when forced (px_island_pass), with impacts, or in seeded mode with
// pending seeds.  Every later kernel of the pass reads this word, so a frame
// can record the pass whatever the device holds.
// A pass that does not run keeps the labels, their status and the error
// bits of the last one that did; only its own work lists are emptied (k_alloc
// and k_alloc_commit, shared with other stages, run ungated).
__global__ void k_edit_begin(EditView e, uint32_t force) {
  if (threadIdx.x != 0) return;
  const uint32_t nimp = min(e.ctrl[PX_C_IMPACTS], e.max_imp);
  const uint32_t pend = e.ctrl[PX_C_PENDING];
  const bool active = force != 0u || nimp != 0u || (e.mode == PX_ISLANDS_SEEDED && pend != 0u);
  if (active) {
    for (uint32_t k = 0; k < EC_WORDS; k++) e.ec[k] = 0u;
    e.ctrl[PX_C_PENDING] = 0u;
  } else {
    e.ec[EC_JOBCELLS] = 0u;
    e.ec[EC_ALLOCS] = 0u;
    e.ec[EC_JOBS] = 0u;
  }
  e.ec[EC_ACTIVE] = active ? 1u : 0u;
  e.ec[EC_NIMP] = active ? nimp : 0u;
  e.ec[EC_R_DIRTY] = 0u;
  for (uint32_t k = EC_F_SUM; k <= EC_F_DONE; k++) e.ec[k] = 0u;
}

// pushImpact (sim.wgsl:203-211) for n staged records: record j goes to slot
// C_IMPACTS + j while that is below MAX_IMPACTS; C_IMPACTS grows by n in
// k_push_impacts_commit.
__global__ void k_push_impacts(EditView e, const float* __restrict__ stage, uint32_t n) {
  const uint32_t i = threadIdx.x;  // one block of 16 * 64 threads
  const uint32_t j = i >> 4;
  if (j >= n) return;
  const uint32_t k = e.ctrl[PX_C_IMPACTS] + j;
  if (k < e.max_imp) e.impacts[(size_t)k * 16u + (i & 15u)] = stage[i];
}

__global__ void k_push_impacts_commit(EditView e, uint32_t n) {
  if (threadIdx.x == 0) e.ctrl[PX_C_IMPACTS] += n;
}

// impactJobs, sim.wgsl:414-443, with one owner per cell and the impacts
// across the lanes.  Block row k walks impact k's box of bricks; a cell
// that an earlier impact's box also holds is left to that impact's row, so
// every cell is decided once.  Lane j of every wave holds impact j (its box
// of bricks, centre, radii and R = ceil(max(r, fr) 1.2) + 1), so one cell is
// tested against all impacts at once: box, distance from the impact centre
// to the brick against R, and for a uniform brick reach(); the ballot is the
// cell's impact mask.  A wave takes 64 cells at a time: lane l reads cell
// l's grid word and material, the 64 cells are decided in turn, and lane l
// then writes cell l's mask (a plain store: one writer per cell) and its
// list entries, one atomic add per list per wave.
__global__ __launch_bounds__(256) void k_impact_jobs(EditView e) {
  const uint32_t nimp = e.ec[EC_NIMP];
  const uint32_t k = blockIdx.y;
  if (k >= nimp) return;
  const uint32_t lane = __lane_id();
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  // impact `lane`
  const bool has = lane < nimp;
  const float* im = e.impacts + (size_t)(has ? lane : 0u) * 16u;
  const float ic[3] = {im[0], im[1], im[2]}, ir = im[3], ifr = im[4];
  const float R = ceilf(fmaxf(ir, ifr) * 1.2f) + 1.0f;
  int32_t blo[3], bhi[3];
  for (int a = 0; a < 3; a++) {
    blo[a] = px_clampi(px_f2i(floorf((ic[a] - R) / 8.0f)), 0, e.bd[a] - 1);
    bhi[a] = px_clampi(px_f2i(floorf((ic[a] + R) / 8.0f)), 0, e.bd[a] - 1);
  }
  // impact k's box, from its lane
  int32_t kb[6];
  for (int a = 0; a < 3; a++) {
    kb[a] = __shfl(blo[a], (int)k);
    kb[3 + a] = __shfl(bhi[a], (int)k);
  }
  const uint32_t ex = (uint32_t)(kb[3] - kb[0] + 1), ey = (uint32_t)(kb[4] - kb[1] + 1);
  const uint32_t ez = (uint32_t)(kb[5] - kb[2] + 1);
  const uint32_t total = ex * ey * ez;
  uint32_t jobs = 0;
  for (uint32_t c0 = wave * 64u; c0 < total; c0 += nwaves * 64u) {
    // cell c0 + lane of the box: coordinates, grid word, material terms
    const uint32_t t = c0 + lane;
    const int32_t cx = kb[0] + (int32_t)(t % ex), cy = kb[1] + (int32_t)((t / ex) % ey);
    const int32_t cz = kb[2] + (int32_t)(t / (ex * ey));
    const uint32_t ci = (uint32_t)cx + (uint32_t)e.bd[0] * ((uint32_t)cy + (uint32_t)e.bd[1] * (uint32_t)cz);
    const uint32_t g = t < total ? e.grid[ci] : 0u;
    const bool uni = (g & PX_UNIFORM_BIT) != 0u;
    const float* E = e.mats + (g & 0xffu) * 20u + 16u;
    const uint32_t efrag = uni && E[1] > 0.5f ? 1u : 0u, esoft = uni && E[0] > 0.0f ? 1u : 0u;
    const float em14 = uni ? fminf(1.4f, E[0]) : 0.0f;
    uint64_t my_mask = 0ull;
    bool my_alloc = false;
    for (uint64_t todo = __ballot(g != 0u); todo; todo &= todo - 1ull) {
      const uint32_t c = first_lane(todo);
      const int32_t bx = (int32_t)lane_value((uint32_t)cx, c), by = (int32_t)lane_value((uint32_t)cy, c);
      const int32_t bz = (int32_t)lane_value((uint32_t)cz, c);
      const bool inbox = has && bx >= blo[0] && bx <= bhi[0] && by >= blo[1] && by <= bhi[1] && bz >= blo[2] &&
                         bz <= bhi[2];
      if (__ballot(inbox && lane < k) != 0ull) continue;  // an earlier impact's row decides it
      bool job = false;
      if (inbox) {
        const float lo[3] = {(float)(bx * 8), (float)(by * 8), (float)(bz * 8)};
        float ev[3];
        for (int a = 0; a < 3; a++) ev[a] = fmaxf(fmaxf(lo[a] - ic[a], 0.0f), ic[a] - (lo[a] + 8.0f));
        const float near2 = ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2];
        job = !(near2 > R * R);
        if (job && lane_value(uni ? 1u : 0u, c) != 0u) {
          // reach() of sim.wgsl:242-248 with cell c's material terms
          float rch = lane_value(efrag, c) != 0u ? ifr * 1.2f : 0.0f;
          if (lane_value(esoft, c) != 0u)
            rch = fmaxf(rch, fmaxf(ir * __int_as_float((int)lane_value((uint32_t)__float_as_int(em14), c)),
                                   fminf(ir, 2.5f)) * 1.2f);
          job = !(near2 > rch * rch);
        }
      }
      const uint64_t m = __ballot(job);
      if (lane == c) {
        my_mask = m;
        my_alloc = uni && m != 0ull;
      }
    }
    if (my_mask) e.imask[ci] = my_mask;
    jobs += (uint32_t)__popcll(my_mask);
    const uint64_t am = __ballot(my_alloc), cm = __ballot(my_mask != 0ull);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t abase = 0, cbase = 0;
    if (lane == 0 && am) abase = atomicAdd(&e.ec[EC_ALLOCS], (uint32_t)__popcll(am));
    if (lane == 0 && cm) cbase = atomicAdd(&e.ec[EC_JOBCELLS], (uint32_t)__popcll(cm));
    abase = __shfl(abase, 0);
    cbase = __shfl(cbase, 0);
    if (my_alloc) {
      const uint32_t q = abase + (uint32_t)__popcll(am & below);
      if (q < e.ncell) e.alloc_list[q] = ci;
    }
    if (my_mask) {
      const uint32_t q = cbase + (uint32_t)__popcll(cm & below);
      if (q < e.ncell) e.job_list[q] = ci;
    }
  }
  jobs = wave_sum(jobs);
  if (lane == 0 && jobs) atomicAdd(&e.ec[EC_JOBS], jobs);
}

// allocBricks, sim.wgsl:459-483, one wave per request.  Every request is a
// uniform cell, so request i takes free_list[top - 1 - i] while i < top and
// fails otherwise (popFree :148-156 pops one entry per request too); the top
// moves in k_alloc_commit.
__global__ __launch_bounds__(64) void k_alloc(EditView e) {
  const uint32_t n = min(e.ec[EC_ALLOCS], e.ncell);
  const uint32_t top = e.ctrl[PX_C_FREE_TOP];
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    uint32_t ci = e.alloc_list[i];
    uint32_t g = e.grid[ci];
    if (!(g == 0u || (g & PX_UNIFORM_BIT) != 0u)) continue;
    if (i >= top || top > e.pool_cap) continue;
    uint32_t p = e.free_list[top - 1u - i];
    uint32_t fill = g != 0u ? (g & 0xffffu) : 0u;
    fill |= fill << 16;
    // no fence before publishing: nothing in this kernel reads the new brick,
    // and the kernels after it on the stream see every store (a device-scope
    // release per request cost 38 us of the 64-impact pass)
    reinterpret_cast<uint4*>(e.pool + (size_t)p * 256u)[lane] = make_uint4(fill, fill, fill, fill);
    if (lane == 0) {
      e.grid[ci] = p + 1u;
      if (g == 0u) atomicAdd(&e.sgrid[super_of(e, ci)], 1u);
      set_dirty(e, ci);
    }
  }
}

// The free-list top after k_alloc's requests, and the failed ones (one
// thread).
PX_DEV void alloc_commit(const EditView& e) {
  uint32_t n = min(e.ec[EC_ALLOCS], e.ncell);
  uint32_t top = e.ctrl[PX_C_FREE_TOP];
  uint32_t used = top > e.pool_cap ? 0u : min(n, top);
  e.ctrl[PX_C_FREE_TOP] = top - used;
  e.ctrl[PX_S_ALLOC_FAIL] += n - used;
}

__global__ void k_alloc_commit(EditView e) {
  if (threadIdx.x == 0) alloc_commit(e);
}

// Per-brick labelling.  MODE 0: islands (uniform bricks of a node class are
// one node).  MODE 1: pane voxels (pool bricks only).
template <int MODE>
__global__ __launch_bounds__(64) void k_local_ccl(EditView e) {
  // One wavefront per brick; lane l owns the row of 8 voxels (x = 0..7) at
  // y = l & 7, z = l >> 3 (local indices 8l .. 8l+7).  Same-class runs along
  // x get their label from bit arithmetic; unions happen only between runs of
  // neighbouring rows (y+1: lane l+1, z+1: lane l+8) at the start of every
  // overlap, so a full brick needs ~100 LDS unions instead of ~1300.
  __shared__ uint32_t s_lab[512];
  // per brick-local root: voxel count (bits 0..15) and anchored runs (bits
  // 16..31), one LDS add per run; 4.5 KiB of LDS per block in all
  __shared__ uint32_t s_cnt[512];
  __shared__ uint8_t s_cls[256];
  __shared__ uint8_t s_str[256];
  const uint32_t lane = threadIdx.x;
  const uint32_t ry = lane & 7u, rz = lane >> 3;
  // the island pass's k_alloc_commit, whether the pass runs or not (nothing
  // here reads the free-list top)
  if (MODE == 1 && blockIdx.x == 0u && lane == 0u) alloc_commit(e);
  if (e.ec[MODE == 0 ? EC_ACTIVE : EC_NIMP] == 0u) return;
  load_tables<MODE>(e.mats, s_cls, s_str);
  __syncthreads();
  uint32_t* my_nodes = e.nodes + (size_t)blockIdx.x * e.slots_node;
  uint32_t* my_cells = e.pane_cells + (size_t)blockIdx.x * e.slots_cell;
  uint32_t nn = 0u, nc = 0u;  // wave-uniform list lengths of this block
  uint32_t walk_err = 0u;
  // Lane l reads the grid word of the block's cell of the l-th of the next
  // 64 steps (step_cell) at once, and the brick of the next
  // pool cell is loaded while the current one is labelled, so the loop does
  // not wait on two dependent loads per cell.
  for (uint32_t k0 = 0; k0 * gridDim.x < e.ncell; k0 += 64u) {
    const uint32_t mc = step_cell(k0 + lane);
    const uint32_t mg = mc < e.ncell ? e.grid[mc] : 0u;
    const bool want = mg != 0u && ((mg & PX_UNIFORM_BIT) == 0u || (MODE == 0 && s_cls[mg & 0xffu]));
    uint64_t live = __ballot(want);
    uint4 wn = make_uint4(0u, 0u, 0u, 0u);
    if (live) wn = brick_row(e, lane_value(mg, first_lane(live)), lane);
    while (live) {
      const uint32_t l = first_lane(live);
      live &= live - 1ull;
      const uint32_t ci = lane_value(mc, l);
      const uint32_t g = lane_value(mg, l);
      const uint4 w = wn;
      if (live) wn = brick_row(e, lane_value(mg, first_lane(live)), lane);
      int32_t bx, by, bz;
      cell_coord(e, ci, bx, by, bz);
      uint32_t okey = vkey(e, bx * 8, by * 8, bz * 8);
      if ((g & PX_UNIFORM_BIT) != 0u) {  // a node-class uniform brick (MODE 0)
        if (lane == 0) {
          e.par[okey] = okey;
          e.cnt[okey] = 512u;
          e.flag[okey] = by == 0 ? kFlagAnchor : 0u;
          if (nn < e.slots_node) my_nodes[nn] = okey;
        }
        nn++;
        continue;
      }
      uint32_t p = g - 1u;
      uint32_t ww[4] = {w.x, w.y, w.z, w.w};
      uint32_t cl[8];
      uint32_t m = 0u, sm = 0u;
      for (int x = 0; x < 8; x++) {
        uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
        cl[x] = v ? s_cls[v & 0xffu] : 0u;
        if (cl[x]) m |= 1u << x;
        if (v && s_str[v & 0xffu]) sm |= 1u << x;
      }
      if (MODE == 0) {  // the brick's face masks for the face merge
        uint64_t fm[kFaceMasks];
        face_masks(m, fm);
        face_masks(sm, fm + 6);
        uint64_t mine = 0ull;
        for (uint32_t k = 0; k < kFaceMasks; k++)
          if (lane == k) mine = fm[k];
        if (lane < kFaceMasks) e.bface[(size_t)p * kFaceMasks + lane] = mine;
      }
      if (__ballot(m != 0u) == 0ull) {  // no node voxel in this brick
        if (MODE == 0 && lane == 0u) e.bnode[p] = kNoNode;
        continue;
      }
      uint32_t c_lo = cl[0] | (cl[1] << 8) | (cl[2] << 16) | (cl[3] << 24);
      uint32_t c_hi = cl[4] | (cl[5] << 8) | (cl[6] << 16) | (cl[7] << 24);
      uint32_t lx = 0u;  // x-links: voxel x continues the run of x-1
      for (int x = 1; x < 8; x++)
        if (cl[x] && cl[x] == cl[x - 1]) lx |= 1u << x;
      uint32_t rs = m & ~lx;  // run starts
      uint32_t ylo = __shfl_down(c_lo, 1), yhi = __shfl_down(c_hi, 1), ylx = __shfl_down(lx, 1);
      uint32_t zlo = __shfl_down(c_lo, 8), zhi = __shfl_down(c_hi, 8), zlx = __shfl_down(lx, 8);
      uint32_t sm_yp = __shfl_down(sm, 1), sm_zp = __shfl_down(sm, 8);
      uint32_t sm_ym = __shfl_up(sm, 1), sm_zm = __shfl_up(sm, 8);
      bool has_y = ry < 7u, has_z = rz < 7u;
      uint32_t ey = 0u, ez = 0u;
      for (int x = 0; x < 8; x++) {
        uint32_t cy = ((x < 4 ? ylo : yhi) >> ((x & 3) * 8)) & 0xffu;
        uint32_t cz = ((x < 4 ? zlo : zhi) >> ((x & 3) * 8)) & 0xffu;
        if (has_y && cl[x] && cl[x] == cy) ey |= 1u << x;
        if (has_z && cl[x] && cl[x] == cz) ez |= 1u << x;
      }
      // anchors inside the brick (islands): y == 0 or a structural neighbour
      uint32_t abits = 0u;
      if (MODE == 0) {
        uint32_t adj = ((sm << 1) | (sm >> 1)) & 0xffu;
        if (has_y) adj |= sm_yp;
        if (ry > 0u) adj |= sm_ym;
        if (has_z) adj |= sm_zp;
        if (rz > 0u) adj |= sm_zm;
        if (by == 0 && ry == 0u) adj = 0xffu;
        abits = m & adj;
      }
      // Every row holds at most one run (most bricks): the runs are then
      // joined exactly when their rows' runs touch (ey / ez), and the
      // components are the connected sets of rows, found with bit operations
      // on the wavefront's row masks instead of LDS walks.  A component's
      // root is its lowest row's run start, its minimum local index, and the
      // roots come in increasing order, as from the walks.
      if (__ballot(__popc(rs) > 1) == 0ull) {
        const uint64_t rows = __ballot(m != 0u), yj = __ballot(ey != 0u), zj = __ballot(ez != 0u);
        const uint32_t cnt_l = (uint32_t)__popc(m);
        const bool anc_l = (m & abits) != 0u;
        uint32_t myroot = kLabNone, total = 0u, one = kNoNode;
        for (uint64_t rem = rows; rem != 0ull;) {
          uint64_t comp = rem & (~rem + 1ull);
          for (;;) {
            const uint64_t nxt =
                comp | ((comp & yj) << 1) | ((comp >> 1) & yj) | ((comp & zj) << 8) | ((comp >> 8) & zj);
            if (nxt == comp) break;
            comp = nxt;
          }
          rem &= ~comp;
          const uint32_t rl = first_lane(comp);
          const uint32_t rli = rl * 8u + (uint32_t)__ffs(lane_value(m, rl)) - 1u;
          const bool in = ((comp >> lane) & 1ull) != 0ull;
          const uint32_t cnt = wave_sum(in ? cnt_l : 0u);
          const bool anch = __ballot(in && anc_l) != 0ull;
          if (in) myroot = rli;
          if (total == 0u) one = okey + loff(e, rli);
          if (lane == 0u) {
            const uint32_t node = okey + loff(e, rli);
            e.par[node] = node;
            e.cnt[node] = cnt;
            e.flag[node] = anch ? kFlagAnchor : 0u;
            if (nn + total < e.slots_node) my_nodes[nn + total] = node;
            else atomicOr(&e.ec[EC_OVERFLOW], kErrNodes);
          }
          total++;
        }
        nn += total;
        if (MODE == 0 && lane == 0u) e.bnode[p] = total == 1u ? one : kManyNodes;
        uint32_t rx[8];
        for (int x = 0; x < 8; x++) rx[x] = ((m >> x) & 1u) ? myroot : kLabNone;
        const uint4 lw = make_uint4(rx[0] | (rx[1] << 16), rx[2] | (rx[3] << 16), rx[4] | (rx[5] << 16),
                                    rx[6] | (rx[7] << 16));
        reinterpret_cast<uint4*>(e.lab + (size_t)p * 512u)[lane] = lw;
        if (MODE == 1) {
          if (lane == 0u) my_cells[nc] = ci;
          nc++;
        }
        continue;  // no LDS touched: the next brick needs no barrier
      }
      for (int x = 0; x < 8; x++) {
        uint32_t li = lane * 8u + (uint32_t)x;
        uint32_t start = 31u - (uint32_t)__clz(rs & ((2u << x) - 1u));
        s_lab[li] = ((m >> x) & 1u) ? lane * 8u + start : PX_NONE;
        s_cnt[li] = 0u;
      }
      __syncthreads();
      for (int x = 0; x < 8; x++) {
        uint32_t b = 1u << x, pb = b >> 1;
        if ((ey & b) && !(x > 0 && (ey & pb) && (lx & b) && (ylx & b)))
          lds_union(s_lab, lane * 8u + (uint32_t)x, (lane + 1u) * 8u + (uint32_t)x, walk_err);
        if ((ez & b) && !(x > 0 && (ez & pb) && (lx & b) && (zlx & b)))
          lds_union(s_lab, lane * 8u + (uint32_t)x, (lane + 8u) * 8u + (uint32_t)x, walk_err);
      }
      __syncthreads();
      uint32_t rootx[8];
      uint32_t nroots = 0u;
      for (int x = 0; x < 8; x++) rootx[x] = kLabNone;
      // unrolled, so rootx[] stays in registers
#pragma unroll
      for (int x = 0; x < 8; x++) {
        if (!((rs >> x) & 1u)) continue;
        uint32_t run = 1u << x;
        for (int y = x + 1; y < 8 && ((lx >> y) & 1u); y++) run |= 1u << y;
        uint32_t r = lds_find(s_lab, lane * 8u + (uint32_t)x, walk_err);
        atomicAdd(&s_cnt[r], (uint32_t)__popc(run) + ((run & abits) != 0u ? 0x10000u : 0u));
        for (int y = x; y < 8; y++)
          if ((run >> y) & 1u) rootx[y] = r;
        if (r == lane * 8u + (uint32_t)x) nroots++;
      }
      __syncthreads();
      // roots become union-find nodes in this block's range
      uint32_t incl = wave_scan(nroots);
      uint32_t total = __shfl(incl, 63);
      uint32_t slot = nn + incl - nroots;
      for (int x = 0; x < 8; x++) {
        if (!((rs >> x) & 1u)) continue;
        uint32_t li = lane * 8u + (uint32_t)x;
        if (rootx[x] != li) continue;
        uint32_t node = okey + loff(e, li);
        e.par[node] = node;
        e.cnt[node] = s_cnt[li] & 0xffffu;
        e.flag[node] = (s_cnt[li] >> 16) != 0u ? kFlagAnchor : 0u;
        if (slot < e.slots_node) my_nodes[slot] = node;
        else atomicOr(&e.ec[EC_OVERFLOW], kErrNodes);
        slot++;
      }
      nn += total;
      if (MODE == 0 && lane == 0u) e.bnode[p] = kManyNodes;  // the face merge takes this brick voxel by voxel
      uint4 lw;
      lw.x = rootx[0] | (rootx[1] << 16);
      lw.y = rootx[2] | (rootx[3] << 16);
      lw.z = rootx[4] | (rootx[5] << 16);
      lw.w = rootx[6] | (rootx[7] << 16);
      reinterpret_cast<uint4*>(e.lab + (size_t)p * 512u)[lane] = lw;
      if (MODE == 1 && total) {
        if (lane == 0u) my_cells[nc] = ci;
        nc++;
      }
      __syncthreads();
    }
  }
  if (lane == 0u) {
    e.bcount[blockIdx.x * kCountWords + (MODE == 0 ? BC_NODES : BC_PANENODES)] = min(nn, e.slots_node);
    if (MODE == 1) e.bcount[blockIdx.x * kCountWords + BC_PANECELLS] = nc;
  }
  if (__ballot(walk_err != 0u) != 0ull && lane == 0u) atomicOr(&e.ec[EC_OVERFLOW], kErrWalk);
}

PX_DEV uint32_t cell_voxel(const EditView& e, uint32_t g, uint32_t li) {
  if (g == 0u) return 0u;
  if ((g & PX_UNIFORM_BIT) != 0u) return g & 0xffffu;
  uint32_t idx = (g - 1u) * 512u + li;
  return (e.pool[idx >> 1u] >> ((idx & 1u) * 16u)) & 0xffffu;
}

// Brick label of local index li in the brick grid word g names, read
// whatever the voxel holds (the caller decides whether it is a node), so
// that the read need not wait for the voxel's; 0 for a uniform brick.
PX_DEV uint32_t raw_label(const EditView& e, uint32_t g, uint32_t li) {
  if ((g & PX_UNIFORM_BIT) != 0u) return 0u;
  return e.lab[(size_t)(g - 1u) * 512u + li];
}

// Node of a voxel known to belong to a component, from its brick label r;
// PX_NONE (and the internal error bit in ec[EC_OVERFLOW]) if r is not a
// local index.
PX_DEV uint32_t node_from(const EditView& e, uint32_t okey, uint32_t g, uint32_t r) {
  if ((g & PX_UNIFORM_BIT) != 0u) return okey;
  if (r >= 512u) {
    atomicOr(&e.ec[EC_OVERFLOW], kErrLabel);
    return PX_NONE;
  }
  return okey + loff(e, r);
}

PX_DEV uint32_t node_of(const EditView& e, uint32_t okey, uint32_t g, uint32_t li) {
  return node_from(e, okey, g, raw_label(e, g, li));
}

// grid words of cell ci and of its +x / +y / +z neighbours (0 past the
// world's edge, and all 0 unless `in`)
PX_DEV void cell_and_neighbours(const EditView& e, uint32_t ci, bool in, uint32_t g[4]) {
  g[0] = g[1] = g[2] = g[3] = 0u;
  if (!in) return;
  int32_t bx, by, bz;
  cell_coord(e, ci, bx, by, bz);
  const uint32_t sy = (uint32_t)e.bd[0], sz = (uint32_t)e.bd[0] * (uint32_t)e.bd[1];
  g[0] = e.grid[ci];
  if (bx + 1 < e.bd[0]) g[1] = e.grid[ci + 1u];
  if (by + 1 < e.bd[1]) g[2] = e.grid[ci + sy];
  if (bz + 1 < e.bd[2]) g[3] = e.grid[ci + sz];
}

// Face merges between cell ci (grid word g != 0) and its +x / +y / +z
// neighbours (grid words gn[], 0 past the world's edge): 64 voxel pairs per
// face, one per lane, plus anchors across faces (islands).  The three faces'
// voxel and label reads are issued together before any union, so their
// latencies overlap.
template <int MODE>
PX_DEV void merge_cell(const EditView& e, uint32_t ci, uint32_t g, const uint32_t gn[3], const uint8_t* s_cls,
                       const uint8_t* s_str, uint32_t& walk_err) {
  const uint32_t lane = threadIdx.x;
  const uint32_t u = lane & 7u, v = lane >> 3;
  if (g == 0u || (MODE == 1 && (g & PX_UNIFORM_BIT) != 0u)) return;
  int32_t bc[3];
  cell_coord(e, ci, bc[0], bc[1], bc[2]);
  uint32_t okey = vkey(e, bc[0] * 8, bc[1] * 8, bc[2] * 8);
  uint32_t va[3], vb[3], ra[3], rb[3], na[3], nb[3];
  bool live[3];
  for (int d = 0; d < 3; d++) {
    va[d] = vb[d] = ra[d] = rb[d] = 0u;
    live[d] = gn[d] != 0u && !(MODE == 1 && (gn[d] & PX_UNIFORM_BIT) != 0u);
    if (!live[d]) continue;
    uint32_t la, lb;
    if (d == 0) { la = 7u | (u << 3) | (v << 6); lb = 0u | (u << 3) | (v << 6); }
    else if (d == 1) { la = u | (7u << 3) | (v << 6); lb = u | (0u << 3) | (v << 6); }
    else { la = u | (v << 3) | (7u << 6); lb = u | (v << 3) | (0u << 6); }
    va[d] = cell_voxel(e, g, la);
    vb[d] = cell_voxel(e, gn[d], lb);
    ra[d] = raw_label(e, g, la);
    rb[d] = raw_label(e, gn[d], lb);
  }
  for (int d = 0; d < 3; d++) {
    na[d] = nb[d] = PX_NONE;
    if (!live[d]) continue;
    const uint32_t ca = va[d] ? s_cls[va[d] & 0xffu] : 0u, cb = vb[d] ? s_cls[vb[d] & 0xffu] : 0u;
    const uint32_t okn = okey + 8u * (d == 0 ? 1u : d == 1 ? (uint32_t)e.dims[0] : (uint32_t)e.dims[0] * e.dims[1]);
    na[d] = ca ? node_from(e, okey, g, ra[d]) : PX_NONE;
    nb[d] = cb ? node_from(e, okn, gn[d], rb[d]) : PX_NONE;
  }
  uint32_t ua[3], ub[3], act = 0u;  // the unions of the three faces, run together
  for (int d = 0; d < 3; d++) {
    ua[d] = ub[d] = 0u;
    if (!live[d]) continue;
    const uint32_t ca = va[d] ? s_cls[va[d] & 0xffu] : 0u, cb = vb[d] ? s_cls[vb[d] & 0xffu] : 0u;
    const uint32_t A = na[d], B = nb[d];
    const uint32_t cA = A == PX_NONE ? 0u : ca, cB = B == PX_NONE ? 0u : cb;
    bool join = cA != 0u && cA == cB;
    // skip a pair that the lane to the left or above on the face also has:
    // in raster order the first lane of every such region does the work
    uint32_t jb = join ? 1u : 0u;
    uint32_t pa = __shfl_up(A, 1), pb = __shfl_up(B, 1), pj = __shfl_up(jb, 1);
    uint32_t qa = __shfl_up(A, 8), qb = __shfl_up(B, 8), qj = __shfl_up(jb, 8);
    bool dup = (u > 0u && pj && pa == A && pb == B) || (v > 0u && qj && qa == A && qb == B);
    if (join && !dup) {
      ua[d] = A;
      ub[d] = B;
      act |= 1u << d;
    }
    if (MODE == 0) {
      uint32_t anc_a = (cA && vb[d] && s_str[vb[d] & 0xffu]) ? 1u : 0u;
      uint32_t anc_b = (cB && va[d] && s_str[va[d] & 0xffu]) ? 1u : 0u;
      uint32_t pna = __shfl_up(anc_a, 1), qna = __shfl_up(anc_a, 8);
      uint32_t pnb = __shfl_up(anc_b, 1), qnb = __shfl_up(anc_b, 8);
      if (anc_a && !((u > 0u && pna && pa == A) || (v > 0u && qna && qa == A))) atomicOr(&e.flag[A], kFlagAnchor);
      if (anc_b && !((u > 0u && pnb && pb == B) || (v > 0u && qnb && qb == B))) atomicOr(&e.flag[B], kFlagAnchor);
    }
  }
  uf_union3(e.par, ua, ub, act, e.walk_cap, walk_err);
}

// What the face merge needs of one side of a face: the island node of the
// cell's brick (kNoNode / kManyNodes as in bnode) and the node-class and
// structural voxels of its face `f` (face_masks() order, 0..5).  A uniform
// brick of a node class is one node (its origin key) and fills its faces.
PX_DEV void face_side(const EditView& e, uint32_t g, uint32_t okey, uint32_t f, const uint8_t* s_cls,
                      const uint8_t* s_str, uint32_t& node, uint64_t& cls, uint64_t& str) {
  if ((g & PX_UNIFORM_BIT) != 0u) {
    const uint32_t v = g & 0xffu;
    node = s_cls[v] ? okey : kNoNode;
    cls = s_cls[v] ? ~0ull : 0ull;
    str = s_str[v] ? ~0ull : 0ull;
    return;
  }
  const uint32_t p = g - 1u;
  node = e.bnode[p];
  cls = e.bface[(size_t)p * kFaceMasks + f];
  str = e.bface[(size_t)p * kFaceMasks + 6u + f];
}

// The island face merge of the cell each lane holds (grid words mg: the
// cell, then its +x / +y / +z neighbours), one cell per lane, for faces
// whose two bricks are each one node (or none): the pair joins when the
// faces' node-class masks meet, and a node is anchored when its class voxels
// meet structural ones across the face.  Returns the lanes whose cell has a
// brick of several nodes on some face; merge_cell() takes those voxel by
// voxel.
PX_DEV uint64_t merge_simple(const EditView& e, uint32_t ci, const uint32_t mg[4], const uint8_t* s_cls,
                             const uint8_t* s_str, uint32_t& walk_err) {
  bool hard = false;
  uint32_t ua[3] = {0u, 0u, 0u}, ub[3] = {0u, 0u, 0u}, act = 0u;
  if (mg[0] != 0u && (mg[1] | mg[2] | mg[3]) != 0u) {
    int32_t bx, by, bz;
    cell_coord(e, ci, bx, by, bz);
    const uint32_t okey = vkey(e, bx * 8, by * 8, bz * 8);
    const uint32_t step[3] = {8u, 8u * (uint32_t)e.dims[0], 8u * (uint32_t)e.dims[0] * (uint32_t)e.dims[1]};
    uint32_t na[3], nb[3];
    uint64_t ca[3], sa[3], cb[3], sb[3];
    for (int d = 0; d < 3; d++) {
      na[d] = nb[d] = kNoNode;
      ca[d] = sa[d] = cb[d] = sb[d] = 0ull;
      if (mg[1 + d] == 0u) continue;
      face_side(e, mg[0], okey, 3u + (uint32_t)d, s_cls, s_str, na[d], ca[d], sa[d]);
      face_side(e, mg[1 + d], okey + step[d], (uint32_t)d, s_cls, s_str, nb[d], cb[d], sb[d]);
    }
    for (int d = 0; d < 3; d++) {
      if (mg[1 + d] == 0u) continue;
      if (na[d] == kManyNodes || nb[d] == kManyNodes) {
        hard = true;
        continue;
      }
      if (na[d] != kNoNode && nb[d] != kNoNode && (ca[d] & cb[d]) != 0ull) {
        ua[d] = na[d];
        ub[d] = nb[d];
        act |= 1u << d;
      }
      if (na[d] != kNoNode && (ca[d] & sb[d]) != 0ull) atomicOr(&e.flag[na[d]], kFlagAnchor);
      if (nb[d] != kNoNode && (sa[d] & cb[d]) != 0ull) atomicOr(&e.flag[nb[d]], kFlagAnchor);
    }
  }
  if (hard) act = 0u;
  uf_union3(e.par, ua, ub, act, e.walk_cap, walk_err);
  return __ballot(hard);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_face_merge(EditView e) {
  __shared__ uint8_t s_cls[256];
  __shared__ uint8_t s_str[256];
  if (e.ec[MODE == 0 ? EC_ACTIVE : EC_NIMP] == 0u) return;
  load_tables<MODE>(e.mats, s_cls, s_str);
  __syncthreads();
  uint32_t walk_err = 0u;
  const uint32_t lane = threadIdx.x;
  // Lane l reads the grid words of the l-th of the next 64 cells and of its
  // three neighbours at once; the cells are then merged in order, skipping
  // air and cells whose neighbours are all air without another read.
  if (MODE == 0) {
    for (uint32_t k0 = 0; k0 * gridDim.x < e.ncell; k0 += 64u) {
      // plain round robin here: rotating the steps (step_cell) balances the
      // blocks but made this kernel slower (244 against 213 us on the
      // 512 x 256 x 512 floor scene, with one union walk per face)
      const uint32_t mc = blockIdx.x + (k0 + lane) * gridDim.x;
      uint32_t mg[4];
      cell_and_neighbours(e, mc, mc < e.ncell, mg);
      // one cell per lane where each face's bricks are one node or none,
      // then the rest a cell at a time, voxel pair by voxel pair
      for (uint64_t live = merge_simple(e, mc, mg, s_cls, s_str, walk_err); live; live &= live - 1ull) {
        const uint32_t l = first_lane(live);
        const uint32_t gn[3] = {lane_value(mg[1], l), lane_value(mg[2], l), lane_value(mg[3], l)};
        merge_cell<0>(e, lane_value(mc, l), lane_value(mg[0], l), gn, s_cls, s_str, walk_err);
      }
    }
  } else {
    for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
      const uint32_t n = e.bcount[b * kCountWords + BC_PANECELLS];
      const uint32_t* cells = e.pane_cells + (size_t)b * e.slots_cell;
      for (uint32_t i0 = 0; i0 < n; i0 += 64u) {
        const bool in = i0 + lane < n;
        const uint32_t mc = in ? cells[i0 + lane] : 0u;
        uint32_t mg[4];
        cell_and_neighbours(e, mc, in, mg);
        for (uint64_t live = __ballot(mg[0] != 0u && (mg[1] | mg[2] | mg[3]) != 0u); live; live &= live - 1ull) {
          const uint32_t l = first_lane(live);
          const uint32_t gn[3] = {lane_value(mg[1], l), lane_value(mg[2], l), lane_value(mg[3], l)};
          merge_cell<1>(e, lane_value(mc, l), lane_value(mg[0], l), gn, s_cls, s_str, walk_err);
        }
      }
    }
  }
  if (__ballot(walk_err != 0u) != 0ull && threadIdx.x == 0u) atomicOr(&e.ec[EC_OVERFLOW], kErrWalk);
}

// After the unions: par[x] = root(x) for every node of block ranges `which`.
// The walk only reads (no path halving), so each par[x] is written by x's
// own thread alone and ends as the root.
// The root of node x after the unions, by a read-only walk (same order
// check and step bound as the union walks; kErrWalk into err when either
// fails, stopping on a key below x).
PX_DEV uint32_t root_of(const EditView& e, uint32_t x, uint32_t& err) {
  uint32_t r = x;
  for (uint32_t it = 0; it < e.walk_cap; it++) {
    const uint32_t p = __hip_atomic_load(e.par + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == r) return r;
    if (p > r) break;  // never written by a node of this pass
    r = p;
  }
  err = kErrWalk;
  return r;
}

__global__ __launch_bounds__(64) void k_flatten(EditView e, uint32_t which) {
  if (e.ec[which == BC_NODES ? EC_ACTIVE : EC_NIMP] == 0u) return;
  if (which == BC_PANENODES) {  // cpair of every cell to (0, 0) before the count passes
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < e.ncell; i += gridDim.x * blockDim.x)
      e.cpair[i] = make_uint2(0u, 0u);
  }
  uint32_t walk_err = 0u;
  for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
    uint32_t n = e.bcount[b * kCountWords + which];
    const uint32_t* nodes = e.nodes + (size_t)b * e.slots_node;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t x = nodes[i];
      __hip_atomic_store(e.par + x, root_of(e, x, walk_err), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (__ballot(walk_err != 0u) != 0ull && threadIdx.x == 0u) atomicOr(&e.ec[EC_OVERFLOW], kErrWalk);
}

// maint of one cell (sim.wgsl:1219-1274; the density output is shading only)
// if its dirty bit was set; returns 1 if it was dirty.
PX_DEV uint32_t maint_cell(const EditView& e, uint32_t ci) {
  const uint32_t lane = threadIdx.x;
  uint32_t bit = 1u << (ci & 31u);
  uint32_t was = 0u;
  if (lane == 0) was = (atomicAnd(&e.dirty_flags[ci >> 5], ~bit) & bit) ? 1u : 0u;
  was = __shfl(was, 0);
  if (!was) return 0u;
  uint32_t g = e.grid[ci];
  if (g == 0u || (g & PX_UNIFORM_BIT) != 0u) return 1u;
  uint32_t p = g - 1u;
  uint4 w4 = reinterpret_cast<const uint4*>(e.pool + (size_t)p * 256u)[lane];
  uint32_t first = __shfl(w4.x & 0xffffu, 0);
  uint32_t ww[4] = {w4.x, w4.y, w4.z, w4.w};
  bool nz = false, diff = false;
  for (int j = 0; j < 4; j++) {
    uint32_t lo = ww[j] & 0xffffu, hi = ww[j] >> 16;
    nz |= (lo | hi) != 0u;
    diff |= lo != first || hi != first;
  }
  bool anynz = __ballot(nz) != 0ull, anydiff = __ballot(diff) != 0ull;
  if (lane == 0) {
    if (!anynz) {
      e.grid[ci] = 0u;
      uint32_t k = atomicAdd(&e.ctrl[PX_C_FREE_TOP], 1u);
      e.free_list[k] = p;
      atomicSub(&e.sgrid[super_of(e, ci)], 1u);
    } else if (!anydiff) {
      e.grid[ci] = PX_UNIFORM_BIT | first;
      uint32_t k = atomicAdd(&e.ctrl[PX_C_FREE_TOP], 1u);
      e.free_list[k] = p;
    }
  }
  return 1u;
}

// maint_cell for the cells lane l names (c, when `in`), 64 at a time: every
// lane claims its cell's dirty bit at once, and the claimed pool bricks are
// then taken in turn with the next one's row read ahead.  Returns the number
// of cells claimed (wave-uniform).
PX_DEV uint32_t maint_batch(const EditView& e, uint32_t c, bool in) {
  const uint32_t lane = __lane_id();
  const uint32_t bit = 1u << (c & 31u);
  const bool was = in && (atomicAnd(&e.dirty_flags[c >> 5], ~bit) & bit) != 0u;
  const uint32_t g = was ? e.grid[c] : 0u;
  uint64_t live = __ballot(g != 0u && (g & PX_UNIFORM_BIT) == 0u);
  uint4 wn = make_uint4(0u, 0u, 0u, 0u);
  if (live) wn = brick_row(e, lane_value(g, first_lane(live)), lane);
  while (live) {
    const uint32_t l = first_lane(live);
    live &= live - 1ull;
    const uint32_t ci = lane_value(c, l);
    const uint32_t p = lane_value(g, l) - 1u;
    const uint4 w4 = wn;
    if (live) wn = brick_row(e, lane_value(g, first_lane(live)), lane);
    const uint32_t first = __shfl(w4.x & 0xffffu, 0);
    const uint32_t ww[4] = {w4.x, w4.y, w4.z, w4.w};
    bool nz = false, diff = false;
    for (int j = 0; j < 4; j++) {
      const uint32_t lo = ww[j] & 0xffffu, hi = ww[j] >> 16;
      nz |= (lo | hi) != 0u;
      diff |= lo != first || hi != first;
    }
    const bool anynz = __ballot(nz) != 0ull, anydiff = __ballot(diff) != 0ull;
    if (lane == 0) {
      if (!anynz) {
        e.grid[ci] = 0u;
        const uint32_t k = atomicAdd(&e.ctrl[PX_C_FREE_TOP], 1u);
        e.free_list[k] = p;
        atomicSub(&e.sgrid[super_of(e, ci)], 1u);
      } else if (!anydiff) {
        e.grid[ci] = PX_UNIFORM_BIT | first;
        const uint32_t k = atomicAdd(&e.ctrl[PX_C_FREE_TOP], 1u);
        e.free_list[k] = p;
      }
    }
  }
  return (uint32_t)__popcll(__ballot(was));
}

// Every dirty cell is a job cell (alloc, carve), a pane cell (flood) or a
// cell that dropped rubble lost voxels in (the extraction's dirty list).  The
// job cells' impact masks are cleared here, after the effects read them.
// Each list is walked 64 entries per wave at a time (maint_batch).
__global__ __launch_bounds__(64) void k_maint(EditView e) {
  if (e.ec[EC_NIMP] == 0u) return;
  const uint32_t lane = threadIdx.x;
  // spawn()'s counters after the carve and flood effects (nothing here
  // reads them)
  if (blockIdx.x == 0u && lane == 0u) effects_finish(e);
  uint32_t dirty = 0u;
  const uint32_t n = min(e.ec[EC_JOBCELLS], e.ncell);
  for (uint32_t i0 = blockIdx.x; i0 < n; i0 += 64u * gridDim.x) {
    const uint32_t i = i0 + lane * gridDim.x;
    const uint32_t ci = i < n ? e.job_list[i] : 0u;
    if (i < n) e.imask[ci] = 0ull;
    dirty += maint_batch(e, ci, i < n);
  }
  for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
    const uint32_t nc = e.bcount[b * kCountWords + BC_PANECELLS];
    const uint32_t* cells = e.pane_cells + (size_t)b * e.slots_cell;
    for (uint32_t i0 = 0; i0 < nc; i0 += 64u)
      dirty += maint_batch(e, i0 + lane < nc ? cells[i0 + lane] : 0u, i0 + lane < nc);
  }
  const uint32_t nd = min(e.ec[EC_R_DIRTY], e.ncell);
  for (uint32_t i0 = blockIdx.x; i0 < nd; i0 += 64u * gridDim.x) {
    const uint32_t i = i0 + lane * gridDim.x;
    dirty += maint_batch(e, i < nd ? e.rdirty[i] : 0u, i < nd);
  }
  const uint32_t ns = min(e.shots[1], kShotDirtyCap);
  for (uint32_t i0 = blockIdx.x; i0 < ns; i0 += 64u * gridDim.x) {
    const uint32_t i = i0 + lane * gridDim.x;
    dirty += maint_batch(e, i < ns ? e.shots[kShotDirtyBase + i] : 0u, i < ns);
  }
  if (lane == 0) e.part[blockIdx.x * kPartWords + PT_DIRTY] = dirty;
}

// Per-root totals.  Lanes of a wavefront that share a root are reduced first
// (one loop turn per distinct root), so a component spanning many bricks
// takes one atomic per wavefront instead of one per brick-local node.
// FLATTEN (PHASE 0 only): the kernel also flattens its nodes, as k_flatten
// would have just before it (each node is in one list, so par[x] has one
// writer; other nodes' stores only shorten the walks).
template <int PHASE, bool FLATTEN>
__global__ __launch_bounds__(64) void k_aggregate(EditView e) {
  // PHASE 0: node flags -> root flags; PHASE 1: node counts -> root counts,
  // for roots left unanchored only (anchored components need no count, and
  // the largest of them, the floor, would take an atomic from every block;
  // in seeded mode only searched roots are counted)
  const uint32_t lane = threadIdx.x;
  if (e.ec[EC_ACTIVE] == 0u) return;
  const uint32_t need = e.mode == PX_ISLANDS_SEEDED ? kFlagSeed : 0u;
  uint32_t walk_err = 0u;
  for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
    uint32_t n = e.bcount[b * kCountWords + BC_NODES];
    const uint32_t* nodes = e.nodes + (size_t)b * e.slots_node;
    for (uint32_t base = 0; base < n; base += 64u) {
      uint32_t i = base + lane;
      uint32_t x = i < n ? nodes[i] : 0u;
      uint32_t r = 0u;
      if (i < n) {
        if (PHASE == 0 && FLATTEN) {  // k_flatten's walk and store, done here
          r = root_of(e, x, walk_err);
          __hip_atomic_store(e.par + x, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          r = e.par[x];
        }
      }
      bool pending = i < n && r != x;
      uint32_t v = 0u;
      if (pending) {
        if (PHASE == 0) {
          v = e.flag[x];
          pending = v != 0u;
        } else {
          const uint32_t rf = __hip_atomic_load(e.flag + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pending = (rf & kFlagAnchor) == 0u && (rf & need) == need;
          v = pending ? e.cnt[x] : 0u;
        }
      }
      uint64_t act = __ballot(pending);
      while (act) {
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
        uint32_t r0 = __shfl(r, (int)leader);
        bool mine = pending && r == r0;
        uint64_t peers = __ballot(mine);
        if (PHASE == 0) {
          uint32_t fs = mine ? v : 0u;
          for (int o = 32; o > 0; o >>= 1) fs |= __shfl_xor(fs, o);
          if (lane == leader) {
            uint32_t have = __hip_atomic_load(e.flag + r0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((have & fs) != fs) atomicOr(&e.flag[r0], fs);
          }
        } else {
          uint32_t cs = wave_sum(mine ? v : 0u);
          if (lane == leader) atomicAdd(&e.cnt[r0], cs);
        }
        if (mine) pending = false;
        act &= ~peers;
      }
    }
  }
  if (__ballot(walk_err != 0u) != 0ull && lane == 0u) atomicOr(&e.ec[EC_OVERFLOW], kErrWalk);
}

// Unsupported roots -> (label, count), appended with one atomic per
// wavefront; in seeded mode searched roots only.  REEMIT: the second emit
// after the node budget tripped, keeping labels below the threshold.
template <int REEMIT>
__global__ __launch_bounds__(64) void k_emit(EditView e) {
  const uint32_t lane = threadIdx.x;
  if (e.ec[REEMIT ? EC_REEMIT : EC_ACTIVE] == 0u) return;
  const uint32_t need = e.mode == PX_ISLANDS_SEEDED ? kFlagSeed : 0u;
  const uint32_t below = REEMIT ? e.ec[EC_NB_T] : 0xFFFFFFFFu;
  uint32_t chunks = 0u;
  for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
    uint32_t n = e.bcount[b * kCountWords + BC_NODES];
    const uint32_t* nodes = e.nodes + (size_t)b * e.slots_node;
    for (uint32_t base = 0; base < n; base += 64u) {
      uint32_t i = base + lane;
      uint32_t x = i < n ? nodes[i] : 0u;
      const uint32_t f = i < n ? e.flag[x] : 0u;
      bool lab = i < n && e.par[x] == x && (f & kFlagAnchor) == 0u && (f & need) == need && (!REEMIT || x < below);
      uint32_t c = lab ? e.cnt[x] : 0u;
      uint64_t m = __ballot(lab);
      if (!m) continue;
      uint32_t q0 = 0u;
      if (lane == 0) q0 = atomicAdd(&e.ec[EC_LABELS], (uint32_t)__popcll(m));
      q0 = __shfl(q0, 0);
      if (lab) {
        uint32_t q = q0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (q < e.out_cap) {
          e.out[2u * q] = x;
          e.out[2u * q + 1u] = c;
        }
        if (c >= 48u) chunks++;
      }
    }
  }
  chunks = wave_sum(chunks);
  if (lane == 0 && chunks) atomicAdd(&e.ec[EC_CHUNKS], chunks);
}

// Seeded mode: the reference stops its search once the hash holds more than
// 7/10 of its slots (sim.wgsl:758) and takes the unexplored rest as
// supported.  Here the floating searched components are resolved in label
// order while their voxel total stays within that budget: the smallest
// label T whose prefix total passes it is found by bisection, labels from T
// on are marked anchored, and k_emit<1> emits the rest again.
__global__ __launch_bounds__(1024) void k_node_budget(EditView e) {
  __shared__ unsigned long long s_red[16];
  if (e.ec[EC_ACTIVE] == 0u || e.mode != PX_ISLANDS_SEEDED) return;
  const uint32_t n = min(e.ec[EC_LABELS], e.out_cap);
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  auto prefix = [&](uint64_t t) -> uint64_t {  // voxels of labels below t
    unsigned long long v = 0ull;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
      if ((uint64_t)e.out[2u * i] < t) v += e.out[2u * i + 1u];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    unsigned long long tot = 0ull;
    for (uint32_t k = 0; k < (blockDim.x >> 6); k++) tot += s_red[k];
    __syncthreads();
    return tot;
  };
  const uint64_t budget = e.node_budget;
  if (prefix(1ull << 32) <= budget) return;
  // smallest t with prefix(t + 1) > budget: labels >= t are dropped
  uint64_t lo = 0ull, hi = 0xFFFFFFFFull;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (prefix(mid + 1ull) > budget) hi = mid;
    else lo = mid + 1ull;
  }
  uint32_t dropped = 0u;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = e.out[2u * i];
    if ((uint64_t)key >= lo) {
      e.flag[key] |= kFlagAnchor;
      dropped++;
    }
  }
  dropped = wave_sum(dropped);
  if (lane == 0 && dropped) atomicAdd(&e.ec[EC_NB_DROP], dropped);
  __syncthreads();
  if (threadIdx.x == 0) {
    e.ec[EC_NB_T] = (uint32_t)lo;
    e.ec[EC_REEMIT] = 1u;
    e.ec[EC_LABELS] = 0u;
    e.ec[EC_CHUNKS] = 0u;
  }
}

// counters of the pass into the ctrl words the reference exposes
__global__ __launch_bounds__(256) void k_edit_finish(EditView e) {
  // One partial per thread over kEditFinishBlocks blocks; each block adds its
  // sums to ec[EC_F_SUM ..], and the last block to finish writes the totals
  // (one block reading all 16384 partials was latency-bound at 17 us).
  __shared__ uint32_t s[4][4];
  __shared__ uint32_t s_last;
  if (e.ec[EC_ACTIVE] == 0u) return;
  const bool carved = e.ec[EC_NIMP] != 0u;
  const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
  uint32_t a[4] = {0, 0, 0, 0};
  for (uint32_t b = blockIdx.x * blockDim.x + t; b < kEditBlocks; b += gridDim.x * blockDim.x) {
    if (carved) {
      a[0] += e.part[b * kPartWords + PT_CARVE_D] + e.part[b * kPartWords + PT_FLOOD_D];
      a[1] += e.part[b * kPartWords + PT_CARVE_B] + e.part[b * kPartWords + PT_FLOOD_B];
      a[2] += e.part[b * kPartWords + PT_DIRTY];
    }
    a[3] += e.bcount[b * kCountWords + BC_NODES];
  }
  for (int k = 0; k < 4; k++) {
    a[k] = wave_sum(a[k]);
    if (lane == 0u) s[k][wave] = a[k];
  }
  __syncthreads();
  if (t == 0u) {
    for (int k = 0; k < 4; k++) {
      uint32_t sum = 0u;
      for (uint32_t w = 0; w < blockDim.x / 64u; w++) sum += s[k][w];
      if (sum) atomicAdd(&e.ec[EC_F_SUM + k], sum);
    }
    __threadfence();
    s_last = atomicAdd(&e.ec[EC_F_DONE], 1u) == gridDim.x - 1u ? 1u : 0u;
  }
  __syncthreads();
  if (t != 0u || !s_last) return;
  __threadfence();
  uint32_t tot[4];
  for (int k = 0; k < 4; k++)
    tot[k] = __hip_atomic_load(&e.ec[EC_F_SUM + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  e.ctrl[PX_S_DESTROYED] += tot[0];
  e.ctrl[PX_S_BILL] += tot[1];
  e.ctrl[PX_C_DIRTY] = tot[2];
  e.ctrl[PX_C_NODES] = tot[3];
  e.ctrl[PX_C_JOBS] = e.ec[EC_JOBS];
  e.ctrl[PX_C_ALLOC_N] = e.ec[EC_ALLOCS];
  e.ctrl[PX_C_ALLOCS] = 0u;
  e.ctrl[PX_C_IMPACTS] = 0u;  // postCarve, sim.wgsl:581-584
  if (carved) {               // shoot's flood entries and dirty cells are consumed
    e.shots[0] = 0u;
    e.shots[1] = 0u;
  }
  e.ctrl[PX_S_SEARCHES] += 1u;
  if (e.mode == PX_ISLANDS_SEEDED) e.ctrl[PX_C_SEED_N] = e.ec[EC_SEEDS];
  e.ec[EC_LABELS_OK] = 1u;
}

// ---------------------------------------------------------------------------
// Voxel writes: writeArgs / writeRequest / allocBricks / writeApply
// (sim.wgsl:1178-1203, 94-102, 127-146, 459-483) and maint of the cells they
// touch.  The reference requests bricks and puts voxels through atomics in
// schedule order; here the order is the write list's:
//  * a cell without a brick is requested by its first write in list order
//    (atomicMin of the write index per cell), and the requests are ranked in
//    list order, so the first max_alloc distinct cells are the ones served;
//  * writes only interact inside one world column (a put goes to p, p+y or
//    p+2y), so the list is radix-sorted by (column, index) and one lane per
//    column applies that column's writes in list order: the serial result,
//    with every column in parallel.

PX_DEV void key_pos(const EditView& e, uint32_t key, int32_t& x, int32_t& y, int32_t& z) {
  x = (int32_t)(key % (uint32_t)e.dims[0]);
  y = (int32_t)((key / (uint32_t)e.dims[0]) % (uint32_t)e.dims[1]);
  z = (int32_t)(key / ((uint32_t)e.dims[0] * (uint32_t)e.dims[1]));
}

PX_DEV uint32_t cell_of_pos(const EditView& e, int32_t x, int32_t y, int32_t z) {
  return (uint32_t)(x >> 3) + (uint32_t)e.bd[0] * ((uint32_t)(y >> 3) + (uint32_t)e.bd[1] * (uint32_t)(z >> 3));
}

__global__ void k_write_args(EditView e) {
  if (threadIdx.x != 0) return;
  const uint32_t n = min(e.ctrl[PX_C_WRITES], e.max_writes);
  e.ctrl[PX_C_WRITE_N] = n;
  e.ctrl[PX_C_WRITES] = 0u;
  e.ec[EC_W_N] = n;
  e.ec[EC_W_DIRTY] = 0u;
  e.ec[EC_W_WRITTEN] = 0u;
  e.ec[EC_W_DROP] = 0u;
  e.ec[EC_ALLOCS] = 0u;
}

// First write per brick-less cell, and the sort keys (column << 32 | index;
// all ones past the list).
__global__ __launch_bounds__(256) void k_write_first(EditView e) {
  const uint32_t n = e.ec[EC_W_N];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < e.max_writes; i += gridDim.x * blockDim.x) {
    uint64_t k = ~0ull;
    if (i < n) {
      int32_t x, y, z;
      key_pos(e, e.wlist[2u * i], x, y, z);
      const uint32_t ci = cell_of_pos(e, x, y, z);
      if (e.grid[ci] == 0u) atomicMin(&e.cfirst[ci], i);
      const uint64_t col = (uint64_t)x + (uint64_t)e.dims[0] * (uint64_t)z;
      k = (col << 32) | i;
    }
    e.wkeys[i] = k;
  }
}

// writeRequest: the alloc list in list order (one workgroup, tiles of
// blockDim.x writes, a block prefix sum per tile); resets cfirst.
__global__ __launch_bounds__(1024) void k_write_request(EditView e) {
  __shared__ uint32_t s_wave[16];
  const uint32_t n = e.ec[EC_W_N];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t base = 0u;
  for (uint32_t t0 = 0; t0 < n; t0 += blockDim.x) {
    const uint32_t i = t0 + threadIdx.x;
    uint32_t ci = 0u;
    bool req = false;
    if (i < n) {
      int32_t x, y, z;
      key_pos(e, e.wlist[2u * i], x, y, z);
      ci = cell_of_pos(e, x, y, z);
      req = e.grid[ci] == 0u && e.cfirst[ci] == i;
    }
    const uint64_t m = __ballot(req);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0u, total = 0u;
    for (uint32_t k = 0; k < nw; k++) {
      const uint32_t c = s_wave[k];
      before += k < wave ? c : 0u;
      total += c;
    }
    if (req) {
      const uint32_t r = base + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (r < e.max_alloc && r < e.ncell) e.alloc_list[r] = ci;
      e.cfirst[ci] = 0xFFFFFFFFu;
    }
    base += total;
    __syncthreads();  // s_wave is reused by the next tile
  }
  if (threadIdx.x == 0) {
    const uint32_t served = min(min(base, e.max_alloc), e.ncell);
    e.ec[EC_ALLOCS] = served;
    e.ec[EC_W_DROP] = base - served;
  }
}

// putVoxel (sim.wgsl:127-146): 0 = written, 1 = occupied, 2 = no pool brick.
// The lane owns voxel p for this launch (one lane per column), so the empty
// check and the OR of its half-word cannot race with another write to it; the
// other half of the word may belong to the neighbouring column, hence the
// atomic OR.
PX_DEV uint32_t put_voxel(const EditView& e, int32_t x, int32_t y, int32_t z, uint32_t v) {
  if (x < 0 || y < 0 || z < 0 || x >= e.dims[0] || y >= e.dims[1] || z >= e.dims[2]) return 2u;
  const uint32_t ci = cell_of_pos(e, x, y, z);
  const uint32_t g = e.grid[ci];
  if (g == 0u) return 2u;
  if ((g & PX_UNIFORM_BIT) != 0u) return 1u;
  const uint32_t idx = (g - 1u) * 512u + (uint32_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
  const uint32_t sh = (idx & 1u) * 16u;
  const uint32_t old = __hip_atomic_load(e.pool + (idx >> 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (((old >> sh) & 0xffffu) != 0u) return 1u;
  atomicOr(e.pool + (idx >> 1), (v & 0xffffu) << sh);
  const uint32_t bit = 1u << (ci & 31u);
  if ((atomicOr(&e.dirty_flags[ci >> 5], bit) & bit) == 0u) {
    const uint32_t k = atomicAdd(&e.ec[EC_W_DIRTY], 1u);
    if (k < e.max_writes) e.wdirty[k] = ci;
  }
  return 0u;
}

// writeApply: one lane per column segment of the sorted keys.
__global__ __launch_bounds__(256) void k_write_apply(EditView e) {
  uint32_t written = 0u;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < e.max_writes; t += gridDim.x * blockDim.x) {
    const uint64_t k = e.wsorted[t];
    if (k == ~0ull) continue;
    const uint32_t col = (uint32_t)(k >> 32);
    if (t > 0u && (uint32_t)(e.wsorted[t - 1u] >> 32) == col) continue;  // not the segment head
    for (uint32_t j = t; j < e.max_writes; j++) {
      const uint64_t kj = e.wsorted[j];
      if (kj == ~0ull || (uint32_t)(kj >> 32) != col) break;
      const uint32_t i = (uint32_t)kj;
      int32_t x, y, z;
      key_pos(e, e.wlist[2u * i], x, y, z);
      const uint32_t v = e.wlist[2u * i + 1u];
      // a taken cell stacks the voxel on top, up to two voxels higher
      for (int up = 0; up < 3; up++) {
        const uint32_t r = put_voxel(e, x, y + up, z, v);
        if (r == 0u) written++;
        if (r != 1u) break;
      }
    }
  }
  written = wave_sum(written);
  if (__lane_id() == 0 && written) atomicAdd(&e.ec[EC_W_WRITTEN], written);
}

// maint of the cells the puts dirtied and of the cells alloc gave a brick
__global__ __launch_bounds__(64) void k_write_maint(EditView e) {
  const uint32_t nd = min(e.ec[EC_W_DIRTY], e.max_writes);
  const uint32_t na = min(e.ec[EC_ALLOCS], e.ncell);
  for (uint32_t i = blockIdx.x; i < nd; i += gridDim.x) maint_cell(e, e.wdirty[i]);
  for (uint32_t i = blockIdx.x; i < na; i += gridDim.x) maint_cell(e, e.alloc_list[i]);
}

__global__ void k_write_finish(EditView e) {
  if (threadIdx.x != 0) return;
  e.ctrl[PX_C_DIRTY] += min(e.ec[EC_W_DIRTY], e.max_writes);
  e.ctrl[PX_C_ALLOC_N] = e.ec[EC_ALLOCS];
}

// pushWrite (sim.wgsl:186-192) for n staged pairs: entry C_WRITES + j while
// it is below max_writes; C_WRITES grows by n in k_push_commit.
__global__ __launch_bounds__(256) void k_push_writes(EditView e, const uint32_t* __restrict__ pairs, uint32_t n) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t k0 = e.ctrl[PX_C_WRITES];
  if (k0 >= e.max_writes || j >= e.max_writes - k0) return;
  e.wlist[2u * (k0 + j)] = pairs[2u * j];
  e.wlist[2u * (k0 + j) + 1u] = pairs[2u * j + 1u];
}

__global__ void k_push_commit(EditView e, uint32_t n) {
  if (threadIdx.x == 0) e.ctrl[PX_C_WRITES] += n;
}

__global__ void k_zero64(uint64_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}
__global__ void k_zero32(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}


// ---------------------------------------------------------------------------
// Island extraction: resolveChunks / resolveExpand / resolveExtract
// (sim.wgsl:862-923) with dropRubbleAbove (:274-284) and the crumb debris.
// Order: unsupported components by label, voxels by cell then local index.

PX_DEV float rnd01(uint32_t seed) { return px_rnd(seed); }

PX_DEV bool mat_fragile(const float* mats, uint32_t v) { return mats[(v & 0xffu) * 20u + 17u] > 0.5f; }

// debrisFlags, sim.wgsl:172-178
PX_DEV uint32_t debris_flags(const float* mats, uint32_t v, uint32_t seed) {
  const uint32_t id = v & 0xffu;
  if (mats[id * 20u + 14u] > 0.0f) return 0u;  // liquids splash and vanish
  if (rnd01(seed ^ 0x51ed27u) < (mats[id * 20u + 17u] > 0.5f ? 0.55f : 0.75f)) return PX_P_DEPOSIT;
  if (rnd01(seed ^ 0x9e3779u) < 0.5f) return PX_P_DUST;
  return 0u;
}

// spawn() into the ring for record j of a batch of `total`: stored when j is
// within the budget left at the start of the batch and no later accepted
// record of the batch maps to the same slot.
PX_DEV void spawn_at(const EditView& e, uint32_t j, uint32_t accepted, uint32_t cursor, const float pos[3],
                     const float vel[3], uint32_t value, float life) {
  if (j >= accepted || accepted - j > e.pmask + 1u) return;
  const uint32_t slot = (cursor + j) & e.pmask;
  e.parts[2u * slot] = make_float4(pos[0], pos[1], pos[2], __uint_as_float(value));
  e.parts[2u * slot + 1u] = make_float4(vel[0], vel[1], vel[2], life);
}

PX_DEV uint32_t spawn_accepted(const EditView& e, uint32_t total) {
  const uint32_t spawned = e.ctrl[PX_C_SPAWNED];
  const uint32_t room = spawned >= e.spawn_budget ? 0u : e.spawn_budget - spawned;
  return min(total, room);
}

// decoded voxel (0 outside the world) and whether its cell holds a pool brick
PX_DEV uint32_t voxel_at(const EditView& e, int32_t x, int32_t y, int32_t z, bool* pool) {
  *pool = false;
  if (x < 0 || y < 0 || z < 0 || x >= e.dims[0] || y >= e.dims[1] || z >= e.dims[2]) return 0u;
  const uint32_t g = e.grid[cell_of_pos(e, x, y, z)];
  if (g == 0u) return 0u;
  if ((g & PX_UNIFORM_BIT) != 0u) return g & 0xffffu;
  *pool = true;
  const uint32_t idx = (g - 1u) * 512u + (uint32_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
  return (__hip_atomic_load(e.pool + (idx >> 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> ((idx & 1u) * 16u)) &
         0xffffu;
}

// rubble voxels dropRubbleAbove takes above p: consecutive rubble voxels in
// pool bricks, at most 12
PX_DEV uint32_t rubble_above(const EditView& e, int32_t x, int32_t y, int32_t z) {
  uint32_t n = 0u;
  for (int32_t d = 1; d <= 12; d++) {
    bool pool;
    const uint32_t v = voxel_at(e, x, y + d, z, &pool);
    if (v == 0u || (v & 0x80u) == 0u || !pool) break;
    n++;
  }
  return n;
}

// clear a voxel known to be in a pool brick; the first clear of a cell
// lists it for maint
PX_DEV void clear_voxel(const EditView& e, int32_t x, int32_t y, int32_t z) {
  const uint32_t ci = cell_of_pos(e, x, y, z);
  const uint32_t g = e.grid[ci];
  const uint32_t idx = (g - 1u) * 512u + (uint32_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
  atomicAnd(e.pool + (idx >> 1), ~(0xffffu << ((idx & 1u) * 16u)));
  const uint32_t bit = 1u << (ci & 31u);
  if ((atomicOr(&e.dirty_flags[ci >> 5], bit) & bit) == 0u) {
    const uint32_t k = atomicAdd(&e.ec[EC_R_DIRTY], 1u);
    if (k < e.ncell) e.rdirty[k] = ci;
  }
}

// removes() of sim.wgsl:229-239 for voxel value v at (x, y, z), same FP32
// operation order as the reference
PX_DEV bool removes_voxel(const float* __restrict__ im, const float* __restrict__ mats, int32_t x, int32_t y,
                          int32_t z, uint32_t v) {
  const float* E = mats + (v & 0xffu) * 20u + 16u;
  const float dx = ((float)x + 0.5f) - im[0], dy = ((float)y + 0.5f) - im[1], dz = ((float)z + 0.5f) - im[2];
  const float d2 = dx * dx + dy * dy + dz * dz;
  const float jit = 0.8f + 0.4f * px_rnd(px_hash3i(x, y, z));
  if (E[1] > 0.5f && d2 < im[4] * im[4] * jit * jit) return true;
  if (E[0] <= 0.0f) return false;
  const float rr = fmaxf(im[3] * fminf(1.4f, E[0]), fminf(im[3], 2.5f)) * jit;
  return d2 < rr * rr;
}

// Whether carve removes voxel v at (x, y, z): some impact with a job in its
// cell (the cell's impact mask, kept until maint) removes it.
PX_DEV bool carve_removes(const EditView& e, const float* __restrict__ s_im, int32_t x, int32_t y, int32_t z,
                          uint32_t v) {
  uint64_t m = e.imask[cell_of_pos(e, x, y, z)];
  while (m) {
    const uint32_t k = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    m &= m - 1ull;
    if (removes_voxel(s_im + k * 16u, e.mats, x, y, z, v)) return true;
  }
  return false;
}

// carve_removes() for the 8 voxels x = 0..7 of one row of a pool brick
// (row words ww, at y = vy, z = vz in the column of brick x index bx),
// against the impacts of the row's cell mask: bit x set if some impact
// removes voxel x; rk gets the first such impact (index order, the one whose
// effects the voxel takes) in bits 6x .. 6x + 5.  Only alive fragile or soft
// voxels can go (`can`), so the loop ends once no lane has such a voxel
// left.  A row whose nearest voxel lies beyond an impact's reach bound is
// skipped for that impact (every exact test fails there), and the 8 tests
// run without branches, in the operation order of removes_voxel().
PX_DEV uint32_t removed_row(const EditView& e, const float* __restrict__ s_im, uint64_t mask, int32_t bx, int32_t vy,
                            int32_t vz, const uint32_t ww[4], uint64_t& rk) {
  uint32_t alive = 0u, frag = 0u, soft = 0u;
  float jit[8], m14[8], fx[8];
  for (int x = 0; x < 8; x++) {
    const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
    fx[x] = (float)(bx * 8 + x) + 0.5f;
    jit[x] = 0.0f;
    m14[x] = 0.0f;
    if (!v) continue;
    alive |= 1u << x;
    const float* E = e.mats + (v & 0xffu) * 20u + 16u;
    if (E[1] > 0.5f) frag |= 1u << x;
    if (E[0] > 0.0f) soft |= 1u << x;
    m14[x] = fminf(1.4f, E[0]);
    jit[x] = 0.8f + 0.4f * px_rnd(px_hash3i(bx * 8 + x, vy, vz));
  }
  const float fy = (float)vy + 0.5f, fz = (float)vz + 0.5f;
  const uint32_t can = alive & (frag | soft);
  uint32_t rem = 0u;
  rk = 0ull;
  for (uint64_t m = mask; m != 0ull && __ballot((can & ~rem) != 0u) != 0ull; m &= m - 1ull) {
    const uint32_t k = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    const float* im = s_im + k * 16u;
    const float ix = im[0], r = im[3], fr = im[4];
    // removes() needs d2 < (fr jit)^2 or d2 < (rr jit)^2 with jit < 1.2 and
    // rr <= max(1.4 r, 2.5): nothing beyond b is removed
    const float b = fmaxf(fmaxf(r * 1.4f, 2.5f), fr) * 1.2f * 1.00001f;
    const float dy = fy - im[1], dz = fz - im[2];
    const float dy2 = dy * dy, dz2 = dz * dz;
    const float dxn = fmaxf(fmaxf(fx[0] - ix, ix - fx[7]), 0.0f);
    const uint32_t open = can & ~rem;
    if (__ballot(open != 0u && dxn * dxn + dy2 + dz2 < b * b) == 0ull) continue;
    const float fr2 = fr * fr, rlo = fminf(r, 2.5f);
#pragma unroll
    for (int x = 0; x < 8; x++) {
      const float dx = fx[x] - ix;
      const float d2 = dx * dx + dy2 + dz2;
      const float j = jit[x];
      const float rr = fmaxf(r * m14[x], rlo) * j;
      const uint32_t hit = ((frag >> x) & (uint32_t)(d2 < fr2 * j * j)) | ((soft >> x) & (uint32_t)(d2 < rr * rr));
      const bool take = (hit & (open >> x) & 1u) != 0u;
      rem |= take ? 1u << x : 0u;
      rk |= take ? (uint64_t)k << (6 * x) : 0ull;
    }
  }
  return rem;
}

// dropRubbleAbove (sim.wgsl:274-284) after carve: the consecutive rubble
// voxels of pool bricks above (x, y, z) that carve leaves, at most 12.  A
// voxel carve removes counts as air whether or not its clear has landed yet,
// so the count is the same before and after carve's stores.
PX_DEV uint32_t rubble_chain(const EditView& e, const float* __restrict__ s_im, int32_t x, int32_t y, int32_t z) {
  uint32_t n = 0u;
  for (int32_t d = 1; d <= 12; d++) {
    bool pool;
    const uint32_t v = voxel_at(e, x, y + d, z, &pool);
    if (v == 0u || (v & 0x80u) == 0u || !pool) break;
    if (carve_removes(e, s_im, x, y + d, z, v)) break;
    n++;
  }
  return n;
}

// rubble_chain with its first probe answered from a copy of the row above:
// the voxel above a row's voxel x is in the row of the next lane (y + 1,
// same z), or for the brick's top row in the bottom row of the cell above.
// `up` is word x / 2 of that row and `up_pool` whether it lies in a pool
// brick (rubble_chain stops at once otherwise).  Only a rubble voxel there
// (rare) reads the world.  The copy may predate another wavefront's clear of
// that voxel, and the result is the same either way: no wavefront but this
// one drops the rubble right above a removed voxel (a chain from further
// down stops at the removed voxel), and a rubble voxel that carve removes
// ends the chain whether its clear has landed or not.
PX_DEV uint32_t rubble_chain_row(const EditView& e, const float* __restrict__ s_im, int32_t x, int32_t y, int32_t z,
                                 bool up_pool, uint32_t up) {
  if (!up_pool) return 0u;
  const uint32_t va = (up >> ((uint32_t)(x & 1) * 16u)) & 0xffffu;
  if (va == 0u || (va & 0x80u) == 0u) return 0u;
  return rubble_chain(e, s_im, x, y, z);
}

// grid word of the cell above cell ci (0 on the top layer or unless `in`)
PX_DEV uint32_t grid_above(const EditView& e, uint32_t ci, bool in) {
  if (!in) return 0u;
  const uint32_t sy = (uint32_t)e.bd[0];
  if ((ci / sy) % (uint32_t)e.bd[1] + 1u >= (uint32_t)e.bd[1]) return 0u;
  return e.grid[ci + sy];
}

// For the lanes of a brick's top row (ry = 7, lane = 7 + 8 z): the bottom
// row (y = 0, same z, row 8 z) of the brick that grid word g_up names, if
// it is a pool brick; zero otherwise and on the other lanes.
PX_DEV uint4 top_row_above(const EditView& e, uint32_t g_up, uint32_t lane) {
  if ((lane & 7u) != 7u || g_up == 0u || (g_up & PX_UNIFORM_BIT) != 0u) return make_uint4(0u, 0u, 0u, 0u);
  return reinterpret_cast<const uint4*>(e.pool + (size_t)(g_up - 1u) * 256u)[lane - 7u];
}

// The row above this lane's row of a pool brick (own = the lane's row):
// the next lane's row, or on the brick's top row the row top_row_above()
// read from the brick above (grid word g_up).  `pool` says whether that row
// lies in a pool brick.
PX_DEV void row_above(uint4 own, uint4 above, uint32_t g_up, uint32_t lane, uint32_t up[4], bool& pool) {
  const uint4 n =
      make_uint4(__shfl_down(own.x, 1), __shfl_down(own.y, 1), __shfl_down(own.z, 1), __shfl_down(own.w, 1));
  const bool top = (lane & 7u) == 7u;
  pool = !top || (g_up != 0u && (g_up & PX_UNIFORM_BIT) == 0u);
  up[0] = top ? above.x : n.x;
  up[1] = top ? above.y : n.y;
  up[2] = top ? above.z : n.z;
  up[3] = top ? above.w : n.w;
}

// pushSeed (sim.wgsl:180-184) of the neighbours of a removed voxel of value
// v, as carveEffects (:511-523) and floodStep (:606-616) push them: not air,
// not structural, not rubble, not of v's id when v is a pane.  Whether the
// seed is still a node voxel is checked when the search reads it.
PX_DEV void push_seeds(const EditView& e, int32_t x, int32_t y, int32_t z, uint32_t v) {
  const bool pane = e.mats[(v & 0xffu) * 20u + 17u] > 1.5f;
  for (int k = 0; k < 6; k++) {
    const int32_t qx = x + (k == 0 ? 1 : (k == 1 ? -1 : 0));
    const int32_t qy = y + (k == 2 ? 1 : (k == 3 ? -1 : 0));
    const int32_t qz = z + (k == 4 ? 1 : (k == 5 ? -1 : 0));
    bool pool;
    const uint32_t w = voxel_at(e, qx, qy, qz, &pool);
    if (w == 0u || (w & 0x80u) != 0u || e.mats[(w & 0xffu) * 20u + 18u] > 0.5f) continue;
    if (pane && (w & 0xffu) == (v & 0xffu)) continue;
    const uint32_t key = vkey(e, qx, qy, qz);
    atomicOr(&e.seeds[key >> 5], 1u << (key & 31u));
  }
}

// carveEffects' debris record of voxel v at (x, y, z) removed by impact im
// (sim.wgsl:500-508): spawned when rnd(seed) < E.w * keep.
PX_DEV bool carve_spawns(const EditView& e, const float* __restrict__ im, uint32_t key, uint32_t v) {
  const float r = im[3];
  const float keep = fminf(1.0f, 30000.0f / fmaxf(1.0f, 4.19f * r * r * r));
  return rnd01(key ^ e.fseed) < e.mats[(v & 0xffu) * 20u + 19u] * keep;
}

PX_DEV void carve_record(const EditView& e, const float* __restrict__ im, int32_t x, int32_t y, int32_t z,
                         uint32_t key, uint32_t v, uint32_t j, uint32_t accepted, uint32_t cursor) {
  const uint32_t seed = key ^ e.fseed;
  const float d[3] = {((float)x + 0.5f) - im[0], ((float)y + 0.5f) - im[1], ((float)z + 0.5f) - im[2]};
  const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const float den = len < 1e-3f ? 1e-3f : len;
  const float sp = im[5] * (0.3f + 0.9f * rnd01(seed + 7u));
  const float off[3] = {0.5f, 0.2f, 0.5f}, m[3] = {0.4f, 0.5f, 0.4f};
  float vel[3];
  for (int a = 0; a < 3; a++) {
    const float dirn = d[a] / den;
    const float jitter = ((rnd01(seed + 13u + (uint32_t)a) - off[a]) * im[5]) * m[a];
    vel[a] = ((dirn * 0.7f + im[6 + a] * 0.6f) - im[9 + a] * 0.2f) * sp + jitter;
  }
  const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  spawn_at(e, j, accepted, cursor, pos, vel, v | debris_flags(e.mats, v, seed), 5.0f + 4.0f * rnd01(seed + 11u));
}

// shardEffects' debris record (sim.wgsl:265-270): when rnd(seed) < E.w * 0.6
PX_DEV bool shard_spawns(const EditView& e, uint32_t key, uint32_t v) {
  return rnd01(key ^ e.fseed) < e.mats[(v & 0xffu) * 20u + 19u] * 0.6f;
}

// the particle of that record
PX_DEV void shard_particle(const EditView& e, int32_t x, int32_t y, int32_t z, uint32_t key, uint32_t v, float pos[3],
                           float vel[3], uint32_t& value, float& life) {
  const uint32_t seed = key ^ e.fseed;
  const float s = 1.0f / e.vs;
  const float b[3] = {rnd01(seed + 3u) - 0.5f, rnd01(seed + 4u) * 0.4f, rnd01(seed + 5u) - 0.5f};
  const float m[3] = {1.6f, 1.0f, 1.6f};
  for (int a = 0; a < 3; a++) vel[a] = (b[a] * m[a]) * s;
  pos[0] = (float)x + 0.5f;
  pos[1] = (float)y + 0.5f;
  pos[2] = (float)z + 0.5f;
  value = v | debris_flags(e.mats, v, seed);
  life = 4.0f + 4.0f * rnd01(seed + 6u);
}

PX_DEV void shard_record(const EditView& e, int32_t x, int32_t y, int32_t z, uint32_t key, uint32_t v, uint32_t j,
                         uint32_t accepted, uint32_t cursor) {
  float pos[3], vel[3], life;
  uint32_t value;
  shard_particle(e, x, y, z, key, v, pos, vel, value, life);
  spawn_at(e, j, accepted, cursor, pos, vel, value, life);
}

// dropRubbleAbove's particle for rubble voxel t taken at (x, qy, z)
// (sim.wgsl:281-282)
PX_DEV void rubble_particle(const EditView& e, int32_t x, int32_t qy, int32_t z, uint32_t t, float pos[3],
                            float vel[3], uint32_t& value, float& life) {
  const uint32_t seed = vkey(e, x, qy, z) ^ e.fseed;
  pos[0] = (float)x + 0.5f;
  pos[1] = (float)qy + 0.5f;
  pos[2] = (float)z + 0.5f;
  const float r[3] = {rnd01(seed), rnd01(seed + 1u), rnd01(seed + 2u)};
  const float m[3] = {0.4f, 0.0f, 0.4f};
  for (int k = 0; k < 3; k++) vel[k] = ((r[k] - 0.5f) * m[k]) / e.vs;
  value = t | PX_P_DEPOSIT;
  life = 6.0f + 2.0f * rnd01(seed + 5u);
}

// dropRubbleAbove's records for the n rubble voxels above (x, y, z), bottom
// to top from rank j: each is taken (one clear per voxel, listed for maint)
// and tumbles as DEPOSIT debris.
PX_DEV void drop_rubble(const EditView& e, int32_t x, int32_t y, int32_t z, uint32_t n, uint32_t j,
                        uint32_t accepted, uint32_t cursor) {
  for (uint32_t d = 1; d <= n; d++) {
    const int32_t qy = y + (int32_t)d;
    bool pool;
    const uint32_t t = voxel_at(e, x, qy, z, &pool);
    clear_voxel(e, x, qy, z);
    float pos[3], vel[3], life;
    uint32_t value;
    rubble_particle(e, x, qy, z, t, pos, vel, value, life);
    spawn_at(e, j + d - 1u, accepted, cursor, pos, vel, value, life);
  }
}

PX_DEV void load_impacts(const EditView& e, float* s_im, uint32_t nimp) {
  for (uint32_t i = threadIdx.x; i < nimp * 16u; i += blockDim.x) s_im[i] = e.impacts[i];
  __syncthreads();
}

// carve with carveEffects (sim.wgsl:485-579), one wave per job cell over every
// impact of the cell's mask (a cell has one writer: no CAS loop).  A voxel is
// removed by the lowest impact of the mask that removes it.  Lane l owns the
// row y = l & 7, z = l >> 3 of the brick.
// PASS 0 (count): the pane components the removed voxels hit, their seeds
// (seeded mode) and the number of debris records their effects spawn
// (dropRubbleAbove's rubble, then the carve debris) into cpair[cell].x.
// PASS 1 (apply): the same removals are cleared with atomic ANDs (rubble that
// another cell's voxels drop may share these words), and the records are
// spawned at their rank in (cell, local index) order, carved voxels first.
template <int PASS>
__global__ __launch_bounds__(64) void k_carve(EditView e) {
  __shared__ float s_im[64 * 16];
  const uint32_t lane = threadIdx.x;
  const uint32_t nimp = e.ec[EC_NIMP];
  if (nimp == 0u) return;
  // the last block also marks the pane components shoot's flood entries hit
  // (the count pass only reads the world; flags are only OR-ed in)
  if (PASS == 0 && blockIdx.x == gridDim.x - 1u) shot_flood(e);
  const uint32_t n = min(e.ec[EC_JOBCELLS], e.ncell);
  if (blockIdx.x >= n) {  // no job entry for this block: only the partials
    if (PASS == 1 && lane == 0u) {
      e.part[blockIdx.x * kPartWords + PT_CARVE_D] = 0u;
      e.part[blockIdx.x * kPartWords + PT_CARVE_B] = 0u;
    }
    return;
  }
  load_impacts(e, s_im, nimp);
  const uint32_t total = PASS == 1 ? e.etot[0] + e.etot[1] : 0u;
  const uint32_t accepted = PASS == 1 ? spawn_accepted(e, total) : 0u;
  const uint32_t cursor = PASS == 1 ? e.ctrl[PX_C_PCURSOR] : 0u;
  uint32_t acc_d = 0u, acc_b = 0u;
  // The count pass visits every job cell and writes work[i] = the cell of
  // job entry i if it found removals there, PX_NONE otherwise (an indexed
  // list: an appended one costs a same-address atomic per cell); the apply
  // pass visits the cells so listed.
  // Lane l reads list entry, impact mask and grid word of the block's l-th
  // of the next 64 entries at once.  The count pass also loads the brick of
  // the next pool cell while the current one is carved (the apply pass reads
  // each brick in turn, after its own earlier stores).
  for (uint32_t i0 = blockIdx.x; i0 < n; i0 += 64u * gridDim.x) {
    const uint32_t mi = i0 + lane * gridDim.x;
    const uint32_t mc = mi < n ? (PASS == 0 ? e.job_list[mi] : e.work[mi]) : PX_NONE;
    const bool in = mc != PX_NONE;
    const uint32_t mg = in ? e.grid[mc] : 0u;
    const uint64_t mm = in ? e.imask[mc] : 0ull;
    const uint32_t mu = grid_above(e, mc, in);
    const bool pool = mg != 0u && (mg & PX_UNIFORM_BIT) == 0u;
    if (PASS == 0 && mi < n && !pool) e.work[mi] = PX_NONE;
    uint64_t live = __ballot(pool);
    uint4 wn = make_uint4(0u, 0u, 0u, 0u), an = wn;
    if (PASS == 0 && live) {
      wn = brick_row(e, lane_value(mg, first_lane(live)), lane);
      an = top_row_above(e, lane_value(mu, first_lane(live)), lane);
    }
    while (live) {
      const uint32_t l = first_lane(live);
      live &= live - 1ull;
      const uint32_t ci = lane_value(mc, l);
      const uint32_t g = lane_value(mg, l);
      const uint32_t g_up = lane_value(mu, l);
      const uint64_t mask = lane_value64(mm, l);
      int32_t bx, by, bz;
      cell_coord(e, ci, bx, by, bz);
      const uint32_t okey = vkey(e, bx * 8, by * 8, bz * 8);
      const uint32_t p = g - 1u;
      uint32_t* row = e.pool + (size_t)p * 256u + lane * 4u;
      uint4 w4, a4;
      if (PASS == 0) {
        w4 = wn;
        a4 = an;
        if (live) {
          wn = brick_row(e, lane_value(mg, first_lane(live)), lane);
          an = top_row_above(e, lane_value(mu, first_lane(live)), lane);
        }
      } else {
        w4 = *reinterpret_cast<const uint4*>(row);
        a4 = top_row_above(e, g_up, lane);
      }
      const uint32_t ww[4] = {w4.x, w4.y, w4.z, w4.w};
      uint32_t up[4];
      bool up_pool;
      row_above(w4, a4, g_up, lane, up, up_pool);
      const int32_t vy = by * 8 + (int32_t)(lane & 7u), vz = bz * 8 + (int32_t)(lane >> 3);
      uint64_t rk;  // remover of removed voxel x in bits 6x .. 6x + 5
      uint32_t rem;
      const uint32_t ie = i0 + l * gridDim.x;  // job entry
      if (PASS == 1 && ie < kCarveRowCache) {
        const uint2 c = e.carve_rows[(size_t)ie * 64u + lane];
        rem = c.x & 0xffu;
        rk = ((uint64_t)c.y << 24) | (c.x >> 8);
      } else {
        rem = removed_row(e, s_im, mask, bx, vy, vz, ww, rk);
        if (PASS == 0 && ie < kCarveRowCache)
          e.carve_rows[(size_t)ie * 64u + lane] = make_uint2(rem | ((uint32_t)rk << 8), (uint32_t)(rk >> 24));
      }
      const bool any = __ballot(rem != 0u) != 0ull;
      if (PASS == 0 && lane == 0u) e.work[ie] = any ? ci : PX_NONE;
      if (!any) continue;
      // records of this lane's removed voxels
      uint32_t rub[8];
      uint32_t nrec = 0u, own = 0u;
      for (int x = 0; x < 8; x++) {
        rub[x] = 0u;
        if (!((rem >> x) & 1u)) continue;
        const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
        const int32_t px = bx * 8 + x;
        rub[x] = rubble_chain_row(e, s_im, px, vy, vz, up_pool, up[x >> 1]);
        const uint32_t kx = (uint32_t)(rk >> (6 * x)) & 63u;
        if (carve_spawns(e, s_im + kx * 16u, vkey(e, px, vy, vz), v)) own |= 1u << x;
        nrec += rub[x] + ((own >> x) & 1u);
      }
      if (PASS == 0) {
        uint32_t pane = 0u;
        for (int x = 0; x < 8; x++) {
          if (!((rem >> x) & 1u)) continue;
          const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
          if (e.mode == PX_ISLANDS_SEEDED) push_seeds(e, bx * 8 + x, vy, vz, v);
          if (e.mats[(v & 0xffu) * 20u + 17u] > 1.5f) pane |= 1u << x;
        }
        // the pane components of the removed pane voxels are hit (flooded
        // later): the row's labels in one read, their roots read together
        if (__ballot(pane != 0u) != 0ull) {
          const uint4 l4 = reinterpret_cast<const uint4*>(e.lab + (size_t)p * 512u)[lane];
          const uint32_t ll[4] = {l4.x, l4.y, l4.z, l4.w};
          uint32_t hit_root[8];
          for (int x = 0; x < 8; x++) {
            hit_root[x] = PX_NONE;
            if (!((pane >> x) & 1u)) continue;
            const uint32_t r = (ll[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
            if (r < 512u) hit_root[x] = e.par[okey + loff(e, r)];
            else atomicOr(&e.ec[EC_OVERFLOW], kErrLabel);
          }
          wave_flag_unique8(e.flag, hit_root, kFlagHit);
        }
        const uint32_t t = wave_sum(nrec);
        if (lane == 0u) e.cpair[ci].x = t;
        continue;
      }
      // PASS 1
      uint32_t wmask[4] = {0u, 0u, 0u, 0u};
      uint32_t destroyed = 0u, bill = 0u;
      for (int x = 0; x < 8; x++)
        if ((rem >> x) & 1u) {
          const uint32_t sh = (uint32_t)(x & 1) * 16u;
          wmask[x >> 1] |= 0xffffu << sh;
          destroyed++;
          bill += price_dimes(e.mats, (ww[x >> 1] >> sh) & 0xffffu);
        }
      for (int q = 0; q < 4; q++)
        if (wmask[q]) atomicAnd(row + q, ~wmask[q]);
      uint32_t j = e.coff[ci].x + wave_scan(nrec) - nrec;
      for (int x = 0; x < 8; x++) {
        if (!((rem >> x) & 1u)) continue;
        const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
        const int32_t px = bx * 8 + x;
        drop_rubble(e, px, vy, vz, rub[x], j, accepted, cursor);
        j += rub[x];
        if ((own >> x) & 1u) {
          const uint32_t kx = (uint32_t)(rk >> (6 * x)) & 63u;
          carve_record(e, s_im + kx * 16u, px, vy, vz, vkey(e, px, vy, vz), v, j, accepted, cursor);
          j++;
        }
      }
      destroyed = wave_sum(destroyed);
      bill = wave_sum(bill);
      acc_d += destroyed;
      acc_b += bill;
      if (lane == 0u && destroyed) set_dirty(e, ci);
    }
  }
  if (PASS == 1 && lane == 0u) {
    e.part[blockIdx.x * kPartWords + PT_CARVE_D] = acc_d;
    e.part[blockIdx.x * kPartWords + PT_CARVE_B] = acc_b;
  }
}

// The pane flood's closure with shardEffects (sim.wgsl:250-271, 589-618):
// every pane voxel whose component carve hit and that carve did not remove.
// PASS 0 counts the records (rubble above, shard debris) into cpair[cell].y
// and pushes the seeds; PASS 1 clears the voxels and spawns the records after
// all of carve's, in (cell, local index) order.
template <int PASS>
__global__ __launch_bounds__(64) void k_flood(EditView e) {
  __shared__ float s_im[64 * 16];
  const uint32_t lane = threadIdx.x;
  const uint32_t nimp = e.ec[EC_NIMP];
  if (nimp == 0u) return;
  // a block with no list entry writes its partials and leaves before the
  // impacts are read
  bool work = false;
  if (PASS == 0) {
    for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x)
      work |= e.bcount[b * kCountWords + BC_PANECELLS] != 0u;
  } else {
    work = blockIdx.x < min(e.ec[EC_FLOOD_N], e.ncell);
  }
  if (!work) {
    if (PASS == 1 && lane == 0u) {
      e.part[blockIdx.x * kPartWords + PT_FLOOD_D] = 0u;
      e.part[blockIdx.x * kPartWords + PT_FLOOD_B] = 0u;
    }
    return;
  }
  load_impacts(e, s_im, nimp);
  const uint32_t total = PASS == 1 ? e.etot[0] + e.etot[1] : 0u;
  const uint32_t accepted = PASS == 1 ? spawn_accepted(e, total) : 0u;
  const uint32_t cursor = PASS == 1 ? e.ctrl[PX_C_PCURSOR] : 0u;
  const uint32_t base = PASS == 1 ? e.etot[0] : 0u;
  uint32_t acc_d = 0u, acc_b = 0u;
  // the count pass visits every pane cell (the labelling's per-block lists),
  // the apply pass the cells it found voxels to flood in
  // The count pass walks every block's pane-cell list, the apply pass the
  // list of cells the count pass found voxels to flood in.  Lane l reads the
  // l-th of the next 64 list entries and its grid word at once.
  const uint32_t nb = PASS == 0 ? kEditBlocks : 1u;
  for (uint32_t b = PASS == 0 ? blockIdx.x : 0u; b < nb; b += gridDim.x) {
    const uint32_t nc = PASS == 0 ? e.bcount[b * kCountWords + BC_PANECELLS] : min(e.ec[EC_FLOOD_N], e.ncell);
    const uint32_t* cells = PASS == 0 ? e.pane_cells + (size_t)b * e.slots_cell : e.work + e.ncell;
    const uint32_t first = PASS == 0 ? 0u : blockIdx.x, step = PASS == 0 ? 1u : gridDim.x;
    for (uint32_t i0 = first; i0 < nc; i0 += 64u * step) {
      const uint32_t mi = i0 + lane * step;
      const uint32_t mc = mi < nc ? cells[mi] : 0u;
      const uint32_t mg = mi < nc ? e.grid[mc] : 0u;
      const uint32_t mu = grid_above(e, mc, mi < nc);
      const uint64_t mm = PASS == 0 && mi < nc ? e.imask[mc] : 0ull;
      for (uint64_t live = __ballot(mg != 0u && (mg & PX_UNIFORM_BIT) == 0u); live; live &= live - 1ull) {
        const uint32_t l = first_lane(live);
        const uint32_t ci = lane_value(mc, l);
        const uint32_t g = lane_value(mg, l);
        const uint32_t g_up = lane_value(mu, l);
        int32_t bx, by, bz;
        cell_coord(e, ci, bx, by, bz);
        const uint32_t okey = vkey(e, bx * 8, by * 8, bz * 8);
        const uint32_t p = g - 1u;
        uint32_t* row = e.pool + (size_t)p * 256u + lane * 4u;
        const uint4 w4 = *reinterpret_cast<const uint4*>(row);
        const uint4 l4 = reinterpret_cast<const uint4*>(e.lab + (size_t)p * 512u)[lane];
        const uint4 a4 = top_row_above(e, g_up, lane);
        const uint32_t ww[4] = {w4.x, w4.y, w4.z, w4.w};
        uint32_t up[4];
        bool up_pool;
        row_above(w4, a4, g_up, lane, up, up_pool);
        const uint32_t ll[4] = {l4.x, l4.y, l4.z, l4.w};
        const int32_t vy = by * 8 + (int32_t)(lane & 7u), vz = bz * 8 + (int32_t)(lane >> 3);
        // The voxels of a row mostly share a label: one root lookup per run of
        // a label (bit x of `look`), all issued together, then their flags.
        uint32_t lab[8], look = 0u, valid = 0u, last_r = kLabNone;
        for (int x = 0; x < 8; x++) {
          const uint32_t sh = (uint32_t)(x & 1) * 16u;
          const uint32_t v = (ww[x >> 1] >> sh) & 0xffffu;
          const uint32_t r = (ll[x >> 1] >> sh) & 0xffffu;
          lab[x] = r;
          if (v == 0u || r == kLabNone) continue;
          if (r >= 512u) {
            atomicOr(&e.ec[EC_OVERFLOW], kErrLabel);
            continue;
          }
          valid |= 1u << x;
          if (r != last_r) look |= 1u << x;
          last_r = r;
        }
        uint32_t root[8], hit = 0u;
        for (int x = 0; x < 8; x++) root[x] = ((look >> x) & 1u) ? e.par[okey + loff(e, lab[x])] : 0u;
        for (int x = 0; x < 8; x++)
          if (((look >> x) & 1u) && (e.flag[root[x]] & kFlagHit) != 0u) hit |= 1u << x;
        uint32_t fl = 0u;
        bool last_hit = false;
        for (int x = 0; x < 8; x++) {
          if (!((valid >> x) & 1u)) continue;
          if ((look >> x) & 1u) last_hit = ((hit >> x) & 1u) != 0u;
          if (last_hit) fl |= 1u << x;
        }
        // carve's own voxels (still present before carve's stores) are not
        // flooded: the count pass takes carve's removals of the row out
        if (PASS == 0 && __ballot(fl != 0u) != 0ull) {
          uint64_t rk;
          fl &= ~removed_row(e, s_im, lane_value64(mm, l), bx, vy, vz, ww, rk);
        }
        if (__ballot(fl != 0u) == 0ull) continue;
        if (PASS == 0 && lane == 0u) {
          const uint32_t k = atomicAdd(&e.ec[EC_FLOOD_N], 1u);
          if (k < e.ncell) e.work[e.ncell + k] = ci;
        }
        uint32_t rub[8];
        uint32_t nrec = 0u, own = 0u;
        for (int x = 0; x < 8; x++) {
          rub[x] = 0u;
          if (!((fl >> x) & 1u)) continue;
          const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
          const int32_t px = bx * 8 + x;
          rub[x] = rubble_chain_row(e, s_im, px, vy, vz, up_pool, up[x >> 1]);
          if (shard_spawns(e, vkey(e, px, vy, vz), v)) own |= 1u << x;
          nrec += rub[x] + ((own >> x) & 1u);
          if (PASS == 0 && e.mode == PX_ISLANDS_SEEDED) push_seeds(e, px, vy, vz, v);
        }
        if (PASS == 0) {
          const uint32_t t = wave_sum(nrec);
          if (lane == 0u) e.cpair[ci].y = t;
          continue;
        }
        uint32_t wmask[4] = {0u, 0u, 0u, 0u};
        uint32_t destroyed = 0u, bill = 0u;
        for (int x = 0; x < 8; x++)
          if ((fl >> x) & 1u) {
            const uint32_t sh = (uint32_t)(x & 1) * 16u;
            wmask[x >> 1] |= 0xffffu << sh;
            destroyed++;
            bill += price_dimes(e.mats, (ww[x >> 1] >> sh) & 0xffffu);
          }
        for (int q = 0; q < 4; q++)
          if (wmask[q]) atomicAnd(row + q, ~wmask[q]);
        uint32_t j = base + e.coff[ci].y + wave_scan(nrec) - nrec;
        for (int x = 0; x < 8; x++) {
          if (!((fl >> x) & 1u)) continue;
          const uint32_t v = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
          const int32_t px = bx * 8 + x;
          drop_rubble(e, px, vy, vz, rub[x], j, accepted, cursor);
          j += rub[x];
          if ((own >> x) & 1u) {
            shard_record(e, px, vy, vz, vkey(e, px, vy, vz), v, j, accepted, cursor);
            j++;
          }
        }
        destroyed = wave_sum(destroyed);
        bill = wave_sum(bill);
        acc_d += destroyed;
        acc_b += bill;
        if (lane == 0u && destroyed) set_dirty(e, ci);
      }
    }
  }
  if (PASS == 1 && lane == 0u) {
    e.part[blockIdx.x * kPartWords + PT_FLOOD_D] = acc_d;
    e.part[blockIdx.x * kPartWords + PT_FLOOD_B] = acc_b;
  }
}

// cpair of every cell to (0, 0) before the count passes (a pass with impacts)

// spawn()'s counters after the effects: every call counts, the accepted ones
// advance the cursor (sim.wgsl:166-170)
PX_DEV void effects_finish(const EditView& e) {
  const uint32_t total = e.etot[0] + e.etot[1];
  const uint32_t acc = spawn_accepted(e, total);
  e.ctrl[PX_C_SPAWNED] += total;
  e.ctrl[PX_C_PCURSOR] += acc;
}

// ---------------------------------------------------------------------------
// shoot (sim.wgsl:371-403)

struct ShootArgs {
  float fire[4];   // U.fire: pellets, radius, fragile radius, impulse
  float fire2[2];  // U.fire2: spread, seed
  float time_x;    // U.time.x
};

// spawn() (sim.wgsl:166-170) of one record: every call counts, the ones
// within the frame's budget take the next ring slot.
PX_DEV void spawn_one(const EditView& e, const float pos[3], const float vel[3], uint32_t value, float life) {
  const uint32_t k = e.ctrl[PX_C_SPAWNED]++;
  if (k >= e.spawn_budget) return;
  const uint32_t slot = (e.ctrl[PX_C_PCURSOR]++) & e.pmask;
  e.parts[2u * slot] = make_float4(pos[0], pos[1], pos[2], __uint_as_float(value));
  e.parts[2u * slot + 1u] = make_float4(vel[0], vel[1], vel[2], life);
}

// takeVoxel (sim.wgsl:105-124) by the shoot stage's single thread: clears the
// voxel of a pool brick; the first clear of a cell lists it for the next
// island pass's maint.
PX_DEV void shot_clear(const EditView& e, int32_t x, int32_t y, int32_t z) {
  const uint32_t ci = cell_of_pos(e, x, y, z);
  const uint32_t g = e.grid[ci];
  const uint32_t idx = (g - 1u) * 512u + (uint32_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
  atomicAnd(e.pool + (idx >> 1), ~(0xffffu << ((idx & 1u) * 16u)));
  const uint32_t bit = 1u << (ci & 31u);
  if ((atomicOr(&e.dirty_flags[ci >> 5], bit) & bit) == 0u) {
    const uint32_t k = e.shots[1]++;
    if (k < kShotDirtyCap) e.shots[kShotDirtyBase + k] = ci;
  }
}

// Lane li < pellets traces one pellet (sim.wgsl:373-379) on the world as it
// is before any pellet's effects; lane 0 then applies the hits in pellet
// order (:380-402): pushImpact, the pane take with shardEffects
// (dropRubbleAbove, bill, debris) and its flood entry, pellet 0's flash.
__global__ __launch_bounds__(64) void k_shoot(EditView e, PxWorldView w, const float* __restrict__ cam,
                                              ShootArgs a) {
  __shared__ uint32_t s_key[16], s_face[16], s_val[16];
  __shared__ float s_dir[16][3];
  const uint32_t li = threadIdx.x;
  const uint32_t pellets = min(px::f2u_sat(a.fire[0]), 16u);
  if (li < pellets) {
    float dir[3] = {cam[36], cam[37], cam[38]};
    if (a.fire2[0] > 0.0f) {
      const uint32_t seed = px::f2u_sat(a.fire2[1]) * 7919u + li * 104729u;
      float d[3];
      for (int k = 0; k < 3; k++) d[k] = dir[k] + ((rnd01(seed + (uint32_t)k) - 0.5f) * 2.0f) * a.fire2[0];
      const float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
      for (int k = 0; k < 3; k++) dir[k] = d[k] / len;
    }
    const float ro[3] = {cam[32] / e.vs, cam[33] / e.vs, cam[34] / e.vs};
    const px::MarchOut m = px::trace_solid(w, ro, dir, 80.0f / e.vs);
    s_key[li] = m.key;
    s_face[li] = m.face;
    s_val[li] = m.value;
    for (int k = 0; k < 3; k++) s_dir[li][k] = dir[k];
  }
  __syncthreads();
  if (li != 0u) return;
  for (uint32_t i = 0; i < pellets; i++) {
    const uint32_t key = s_key[i];
    if (key == PX_NONE) continue;
    int32_t x, y, z;
    key_pos(e, key, x, y, z);
    const uint32_t f = s_face[i];
    float n[3] = {0.0f, 0.0f, 0.0f};
    if (f < 6u) n[f >> 1] = (f & 1u) ? -1.0f : 1.0f;
    const float r = a.fire[1];
    const float off = fminf(r * 0.3f, 2.0f);
    const float c[3] = {((float)x + 0.5f) - n[0] * off, ((float)y + 0.5f) - n[1] * off,
                        ((float)z + 0.5f) - n[2] * off};
    // pushImpact (sim.wgsl:203-211)
    const uint32_t k = e.ctrl[PX_C_IMPACTS]++;
    if (k < e.max_imp) {
      float* im = e.impacts + (size_t)k * 16u;
      const float rec[12] = {c[0], c[1], c[2], r, a.fire[2], a.fire[3], n[0], n[1], n[2],
                             s_dir[i][0], s_dir[i][1], s_dir[i][2]};
      for (int q = 0; q < 12; q++) im[q] = rec[q];
    }
    // a pane shatters as a whole: take the voxel, its effects, flood the rest
    const uint32_t id = s_val[i] & 0xffu;
    if (e.mats[id * 20u + 17u] > 1.5f) {
      bool pool;
      const uint32_t v = voxel_at(e, x, y, z, &pool);
      if (pool && v != 0u && (v & 0xffu) == id) {
        shot_clear(e, x, y, z);
        // shardEffects (sim.wgsl:261-271): dropRubbleAbove (:274-284) first
        for (int32_t d = 1; d <= 12; d++) {
          bool qpool;
          const uint32_t t = voxel_at(e, x, y + d, z, &qpool);
          if (t == 0u || (t & 0x80u) == 0u || !qpool) break;
          shot_clear(e, x, y + d, z);
          float pos[3], vel[3], life;
          uint32_t value;
          rubble_particle(e, x, y + d, z, t, pos, vel, value, life);
          spawn_one(e, pos, vel, value, life);
        }
        e.ctrl[PX_S_BILL] += price_dimes(e.mats, v);
        e.ctrl[PX_S_DESTROYED] += 1u;
        if (shard_spawns(e, key, v)) {
          float pos[3], vel[3], life;
          uint32_t value;
          shard_particle(e, x, y, z, key, v, pos, vel, value, life);
          spawn_one(e, pos, vel, value, life);
        }
        // pushFlood (sim.wgsl:250-258)
        const uint32_t q = e.shots[0]++;
        if (q < kShotFloodCap) {
          e.shots[kShotFloodBase + 2u * q] = key;
          e.shots[kShotFloodBase + 2u * q + 1u] = id;
        }
      }
    }
    if (i == 0u) {
      const uint32_t fk = (e.ctrl[PX_C_FLASH]++) % (uint32_t)PX_MAX_FLASHES;
      float* b = e.flashes + fk * 8u;
      for (int q = 0; q < 3; q++) b[q] = (c[q] + n[q] * 4.0f) * e.vs;
      b[3] = r * e.vs;
      b[4] = a.time_x;
    }
  }
}

// The pane flood of shoot's entries (floodStep, sim.wgsl:599-618, on the
// entries shoot pushed): the pane components of each struck voxel's
// same-id neighbours are flooded whole, which is what the breadth-first flood
// from the struck voxel reaches; in seeded mode the struck voxel's other
// neighbours are seeds.  After the pane labelling, before the flood passes.
PX_DEV void shot_flood(const EditView& e) {
  const uint32_t n = min(e.shots[0], kShotFloodCap);
  uint32_t err = 0u;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = e.shots[kShotFloodBase + 2u * i], id = e.shots[kShotFloodBase + 2u * i + 1u];
    int32_t x, y, z;
    key_pos(e, key, x, y, z);
    for (int k = 0; k < 6; k++) {
      const int32_t qx = x + (k == 0 ? 1 : (k == 1 ? -1 : 0));
      const int32_t qy = y + (k == 2 ? 1 : (k == 3 ? -1 : 0));
      const int32_t qz = z + (k == 4 ? 1 : (k == 5 ? -1 : 0));
      bool pool;
      const uint32_t v = voxel_at(e, qx, qy, qz, &pool);
      if (!pool || v == 0u || (v & 0xffu) != id) continue;
      const uint32_t g = e.grid[cell_of_pos(e, qx, qy, qz)];
      const uint32_t li = (uint32_t)((qx & 7) | ((qy & 7) << 3) | ((qz & 7) << 6));
      const uint32_t r = e.lab[(size_t)(g - 1u) * 512u + li];
      if (r >= 512u) {
        err = kErrLabel;
        continue;
      }
      const uint32_t okey = vkey(e, qx & ~7, qy & ~7, qz & ~7);
      atomicOr(&e.flag[e.par[okey + loff(e, r)]], kFlagHit);
    }
    if (e.mode == PX_ISLANDS_SEEDED) push_seeds(e, x, y, z, id);
  }
  if (err) atomicOr(&e.ec[EC_OVERFLOW], err);
}

// Seeded mode (searchSeed, sim.wgsl:722-731): every seed bit whose voxel is
// a node voxel marks its component searched; the bits are cleared as they
// are read.  Seeds are the neighbours of carved and flooded voxels and the
// voxels under landed chunks; they have no capacity limit here (a bit per
// voxel), so nothing is dropped (S_SEED_DROP stays 0).
__global__ __launch_bounds__(256) void k_seed_mark(EditView e) {
  if (e.ec[EC_ACTIVE] == 0u || e.mode != PX_ISLANDS_SEEDED) return;
  uint32_t found = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < e.seed_words; i += gridDim.x * blockDim.x) {
    uint32_t w = e.seeds[i];
    if (w == 0u) continue;
    e.seeds[i] = 0u;
    while (w) {
      const uint32_t b = (uint32_t)__ffs(w) - 1u;
      w &= w - 1u;
      int32_t x, y, z;
      key_pos(e, i * 32u + b, x, y, z);
      if (z >= e.dims[2]) continue;
      const uint32_t ci = cell_of_pos(e, x, y, z);
      const uint32_t g = e.grid[ci];
      bool pool;
      const uint32_t v = voxel_at(e, x, y, z, &pool);
      if (v == 0u || (v & 0x80u) != 0u || e.mats[(v & 0xffu) * 20u + 18u] > 0.5f) continue;
      const uint32_t okey = vkey(e, x & ~7, y & ~7, z & ~7);
      const uint32_t node = node_of(e, okey, g, (uint32_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6)));
      if (node == PX_NONE) continue;
      const uint32_t r = e.par[node];
      found++;
      if ((__hip_atomic_load(e.flag + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFlagSeed) == 0u)
        atomicOr(&e.flag[r], kFlagSeed);
    }
  }
  found = wave_sum(found);
  if (__lane_id() == 0u && found) atomicAdd(&e.ec[EC_SEEDS], found);
}

// R1: chunk ids for the big unsupported components in label order (one
// workgroup).  Labels over the component budget stay (kFlagKeep, S_BUDGET);
// of the big ones (>= 48 voxels) the smallest labels get the free chunk ids,
// found by a binary search for the label threshold; all others crumble.
__global__ __launch_bounds__(1024) void k_resolve_chunks(EditView e) {
  __shared__ uint32_t s_red[3][16];
  __shared__ uint32_t s_roots[512];
  __shared__ uint32_t s_nroots;
  if (e.ec[EC_LABELS_OK] == 0u) return;  // the last island pass did not run
  const uint32_t n = min(e.ec[EC_LABELS], e.out_cap);
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  auto block_sum = [&](uint32_t v, int slot) -> uint32_t {
    v = wave_sum(v);
    if (lane == 0) s_red[slot][wave] = v;
    __syncthreads();
    uint32_t t = 0u;
    for (uint32_t k = 0; k < (blockDim.x >> 6); k++) t += s_red[slot][k];
    __syncthreads();
    return t;
  };
  uint32_t keep = 0u, big = 0u;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = e.out[2u * i], cnt = e.out[2u * i + 1u];
    if (cnt > e.comp_budget) {
      e.flag[key] |= kFlagKeep;
      keep++;
    } else if (cnt >= 48u) {
      big++;
    }
  }
  keep = block_sum(keep, 0);
  big = block_sum(big, 1);
  const uint32_t c0 = e.ctrl[PX_C_CHUNKS];
  const uint32_t room = c0 < e.max_chunks ? e.max_chunks - c0 : 0u;
  const uint32_t K = min(room, big);
  // smallest T with #(big labels <= T) >= K (labels are distinct)
  uint32_t T = 0u;
  bool any = K > 0u;
  if (any && K < big) {
    uint32_t lo = 0u, hi = 0xFFFFFFFFu;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      uint32_t c = 0u;
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t key = e.out[2u * i], cnt = e.out[2u * i + 1u];
        if (cnt >= 48u && cnt <= e.comp_budget && key <= mid) c++;
      }
      c = block_sum(c, 2);
      if (c >= K) hi = mid;
      else lo = mid + 1u;
    }
    T = lo;
  } else if (any) {
    T = 0xFFFFFFFFu;
  }
  if (threadIdx.x == 0) s_nroots = 0u;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = e.out[2u * i], cnt = e.out[2u * i + 1u];
    if (cnt > e.comp_budget) continue;
    if (any && cnt >= 48u && key <= T) {
      const uint32_t k = atomicAdd(&s_nroots, 1u);
      if (k < 512u) s_roots[k] = key;
    } else {
      e.flag[key] |= kFlagCrumb;
    }
  }
  __syncthreads();
  const uint32_t nr = min(s_nroots, 512u);
  for (uint32_t k = threadIdx.x; k < nr; k += blockDim.x) {
    const uint32_t key = s_roots[k];
    uint32_t rank = 0u;
    for (uint32_t j = 0; j < nr; j++) rank += s_roots[j] < key ? 1u : 0u;
    const uint32_t id = c0 + rank;
    uint32_t* b = e.chunks + (size_t)id * 16u;
    for (int w = 0; w < 16; w++) b[w] = 0u;
    b[6] = kNoBlock;
    b[0] = 1u;  // falling
    e.flag[key] |= (id + 1u) << 8;
  }
  if (threadIdx.x == 0) {
    e.ctrl[PX_C_CHUNKS] = c0 + big;
    e.ctrl[PX_S_ISLANDS] += K;
    e.ctrl[PX_S_BUDGET] += keep;
  }
}

// A component stays in the world when it is supported, over the component
// budget, or (seeded mode) was not searched: the reference only resolves
// the nodes its search visited (sim.wgsl:841-923).
PX_DEV bool stays(const EditView& e, uint32_t f) {
  if ((f & (kFlagAnchor | kFlagKeep)) != 0u) return true;
  return e.mode == PX_ISLANDS_SEEDED && (f & kFlagSeed) == 0u;
}

// R2: cells holding voxels of extracted components; uniform ones need a pool
// brick first (resolveExpand)
__global__ __launch_bounds__(64) void k_resolve_mark(EditView e) {
  if (e.ec[EC_LABELS_OK] == 0u) return;
  for (uint32_t b = blockIdx.x; b < kEditBlocks; b += gridDim.x) {
    const uint32_t n = e.bcount[b * kCountWords + BC_NODES];
    const uint32_t* nodes = e.nodes + (size_t)b * e.slots_node;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t x = nodes[i];
      const uint32_t f = e.flag[e.par[x]];
      if (stays(e, f)) continue;
      int32_t px, py, pz;
      key_pos(e, x, px, py, pz);
      const uint32_t ci = cell_of_pos(e, px, py, pz);
      e.cpair[ci] = make_uint2(1u, (e.grid[ci] & PX_UNIFORM_BIT) != 0u ? 1u : 0u);
    }
  }
}

// R3: extract list and alloc list in cell order
__global__ __launch_bounds__(256) void k_resolve_lists(EditView e, const uint32_t* tot) {
  if (e.ec[EC_LABELS_OK] == 0u) return;
  for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < e.ncell; ci += gridDim.x * blockDim.x) {
    const uint2 f = e.cpair[ci];
    if (f.x == 0u) continue;
    const uint2 o = e.coff[ci];
    e.rlist[o.x] = ci | (f.y << 31);  // top bit: the cell was uniform (one node, no brick labels)
    if (f.y && o.y < e.max_alloc && o.y < e.ncell) e.alloc_list[o.y] = ci;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    e.ec[EC_R_CELLS] = tot[0];
    e.ec[EC_ALLOCS] = min(min(tot[1], e.max_alloc), e.ncell);
  }
}

// Class of voxel li of extract cell ci: 0 stays, 1 + chunk id = to the
// chunk, 0xFFFFFFFF = crumbles.  The cell must hold a pool brick now (alloc
// may have failed for a uniform cell: its voxels cannot be taken).
PX_DEV uint32_t extract_class(const EditView& e, uint32_t g, bool was_uniform, uint32_t okey, uint32_t li,
                              uint32_t v) {
  if (v == 0u) return 0u;
  uint32_t node;
  if (was_uniform) {
    node = okey;
  } else {
    const uint32_t r = e.lab[(size_t)(g - 1u) * 512u + li];
    if (r >= 512u) return 0u;
    node = okey + loff(e, r);
  }
  const uint32_t f = e.flag[e.par[node]];
  if (stays(e, f)) return 0u;
  const uint32_t cid = (f >> 8) & 0x1ffu;
  if (cid != 0u && (f & kFlagCrumb) == 0u) return cid;
  return 0xFFFFFFFFu;
}

// Per extract cell (one wavefront; lane = row y|z<<3, 8 voxels in x):
// PASS 0 counts the chunk voxels; PASS 1 counts the debris records (rubble
// above every taken voxel, plus the crumbs, including chunk voxels past the
// arena); PASS 2 takes the voxels, writes the arena and the records.
template <int PASS>
__global__ __launch_bounds__(64) void k_extract(EditView e, const uint32_t* spawn_tot) {
  const uint32_t lane = threadIdx.x;
  const uint32_t n = e.ec[EC_R_CELLS];
  uint32_t bill = 0u, destroyed = 0u;
  uint32_t accepted = 0u, cursor = 0u;
  if (PASS == 2) {
    accepted = spawn_accepted(e, spawn_tot[1]);
    cursor = e.ctrl[PX_C_PCURSOR];
  }
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t ent = e.rlist[i];
    const uint32_t ci = ent & 0x7FFFFFFFu;
    const bool was_uniform = (ent >> 31) != 0u;
    const uint32_t g = e.grid[ci];
    if (g == 0u || (g & PX_UNIFORM_BIT) != 0u) {  // no brick could be given: nothing is taken
      if (PASS < 2 && lane == 0) {
        uint2 c = e.cpair[i];
        if (PASS == 0) c = make_uint2(0u, 0u);
        else c.y = 0u;
        e.cpair[i] = c;
      }
      continue;
    }
    int32_t bx, by, bz;
    cell_coord(e, ci, bx, by, bz);
    const uint32_t okey = vkey(e, bx * 8, by * 8, bz * 8);
    const uint4 w4 = reinterpret_cast<const uint4*>(e.pool + (size_t)(g - 1u) * 256u)[lane];
    const uint32_t ww[4] = {w4.x, w4.y, w4.z, w4.w};
    uint32_t cls[8], val[8];
    uint32_t na = 0u;
    for (int x = 0; x < 8; x++) {
      val[x] = (ww[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
      cls[x] = extract_class(e, g, was_uniform, okey, lane * 8u + (uint32_t)x, val[x]);
      if (cls[x] != 0u && cls[x] != 0xFFFFFFFFu) na++;
    }
    if (PASS == 0) {
      const uint32_t t = wave_sum(na);
      if (lane == 0) e.cpair[i] = make_uint2(t, 0u);
      continue;
    }
    // arena slots of this lane's chunk voxels: cell offset + lanes before + in-lane order
    const uint32_t abase = e.coff[i].x + wave_scan(na) - na;
    const int32_t ry = (int32_t)(lane & 7u), rz = (int32_t)(lane >> 3);
    uint32_t ns = 0u, rub[8];
    uint32_t a = abase;
    for (int x = 0; x < 8; x++) {
      rub[x] = 0u;
      if (cls[x] == 0u) continue;
      if (cls[x] != 0xFFFFFFFFu) {
        if (a >= e.max_arena) cls[x] = 0xFFFFFFFFu;  // the arena is full: crumbles
        a++;
      }
      rub[x] = rubble_above(e, bx * 8 + x, by * 8 + ry, bz * 8 + rz);
      ns += rub[x] + (cls[x] == 0xFFFFFFFFu ? 1u : 0u);
    }
    if (PASS == 1) {
      const uint32_t t = wave_sum(ns);
      if (lane == 0) e.cpair[i].y = t;
      continue;
    }
    // PASS 2: the cell's own voxels are cleared word by word (one atomic AND
    // per pool word: rubble in the same word may be taken by another
    // wavefront) and the cell is listed for maint once
    uint32_t j = e.coff[i].y + wave_scan(ns) - ns;
    a = abase;
    uint32_t wmask[4] = {0u, 0u, 0u, 0u};
    for (int x = 0; x < 8; x++)
      if (cls[x] != 0u) wmask[x >> 1] |= 0xffffu << ((x & 1) * 16);
    uint32_t* row = e.pool + (size_t)(g - 1u) * 256u + lane * 4u;
    for (int q = 0; q < 4; q++)
      if (wmask[q]) atomicAnd(row + q, ~wmask[q]);
    if (__ballot((wmask[0] | wmask[1] | wmask[2] | wmask[3]) != 0u) != 0ull && lane == 0) {
      const uint32_t bit = 1u << (ci & 31u);
      if ((atomicOr(&e.dirty_flags[ci >> 5], bit) & bit) == 0u) {
        const uint32_t k = atomicAdd(&e.ec[EC_R_DIRTY], 1u);
        if (k < e.ncell) e.rdirty[k] = ci;
      }
    }
    for (int x = 0; x < 8; x++) {
      if (cls[x] == 0u) continue;
      const int32_t px = bx * 8 + x, py = by * 8 + ry, pz = bz * 8 + rz;
      const uint32_t v = val[x];
      // dropRubbleAbove: bottom to top, before the voxel's own record
      for (uint32_t d = 1; d <= rub[x]; d++) {
        const int32_t qy = py + (int32_t)d;
        bool pool;
        const uint32_t t = voxel_at(e, px, qy, pz, &pool);
        clear_voxel(e, px, qy, pz);
        const uint32_t seed = vkey(e, px, qy, pz) ^ e.fseed;
        const float pos[3] = {(float)px + 0.5f, (float)qy + 0.5f, (float)pz + 0.5f};
        const float r[3] = {rnd01(seed), rnd01(seed + 1u), rnd01(seed + 2u)};
        const float m[3] = {0.4f, 0.0f, 0.4f};
        float vel[3];
        for (int k = 0; k < 3; k++) vel[k] = ((r[k] - 0.5f) * m[k]) / e.vs;
        spawn_at(e, j, accepted, cursor, pos, vel, t | PX_P_DEPOSIT, 6.0f + 2.0f * rnd01(seed + 5u));
        j++;
      }
      if (cls[x] != 0xFFFFFFFFu) {
        e.arena[2u * a] = vkey(e, px, py, pz);
        e.arena[2u * a + 1u] = v | ((cls[x] - 1u) << 16);
        a++;
        continue;
      }
      // crumbs: straight to debris
      bill += price_dimes(e.mats, v);
      destroyed++;
      const uint32_t seed = vkey(e, px, py, pz) ^ e.fseed;
      const float pos[3] = {(float)px + 0.5f, (float)py + 0.5f, (float)pz + 0.5f};
      const float r[3] = {rnd01(seed), rnd01(seed + 1u), rnd01(seed + 2u)};
      const float off[3] = {0.5f, 0.0f, 0.5f}, m[3] = {0.6f, 0.3f, 0.6f};
      float vel[3];
      for (int k = 0; k < 3; k++) vel[k] = ((r[k] - off[k]) * m[k]) / e.vs;
      spawn_at(e, j, accepted, cursor, pos, vel, v | debris_flags(e.mats, v, seed), 4.0f + 3.0f * rnd01(seed + 9u));
      j++;
    }
  }
  if (PASS == 2) {
    bill = wave_sum(bill);
    destroyed = wave_sum(destroyed);
    if (lane == 0 && (bill | destroyed)) {
      atomicAdd(&e.ec[EC_R_BILL], bill);
      atomicAdd(&e.ec[EC_R_DESTROYED], destroyed);
    }
  }
}

// chunk voxel counts and fragile flags from the arena entries this pass added
// (per-workgroup tallies in LDS, one global add per chunk per workgroup)
__global__ __launch_bounds__(256) void k_arena_tally(EditView e, const uint32_t* tot) {
  __shared__ uint32_t s_cnt[512], s_frag[512];
  for (uint32_t k = threadIdx.x; k < 512u; k += blockDim.x) s_cnt[k] = s_frag[k] = 0u;
  __syncthreads();
  const uint32_t a0 = e.ec[EC_R_ARENA0], a1 = min(a0 + tot[0], e.max_arena);
  for (uint32_t a = a0 + blockIdx.x * blockDim.x + threadIdx.x; a < a1; a += gridDim.x * blockDim.x) {
    const uint32_t vv = e.arena[2u * a + 1u];
    const uint32_t c = vv >> 16;
    atomicAdd(&s_cnt[c], 1u);
    if (mat_fragile(e.mats, vv & 0xffffu)) s_frag[c] = 1u;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < 512u; k += blockDim.x) {
    if (s_cnt[k]) atomicAdd(e.chunks + (size_t)k * 16u + 5u, s_cnt[k]);
    if (s_frag[k]) e.chunks[(size_t)k * 16u + 4u] = 1u;
  }
}

__global__ void k_resolve_begin(EditView e) {
  if (threadIdx.x != 0) return;
  e.ec[EC_R_CELLS] = 0u;
  e.ec[EC_ALLOCS] = 0u;
  e.ec[EC_R_DIRTY] = 0u;
  e.ec[EC_R_BILL] = 0u;
  e.ec[EC_R_DESTROYED] = 0u;
  e.ec[EC_R_ARENA0] = e.ctrl[PX_C_ARENA];  // first arena entry of this pass
}

__global__ void k_resolve_finish(EditView e, const uint32_t* arena_tot, const uint32_t* spawn_tot) {
  if (threadIdx.x != 0) return;
  e.ec[EC_LABELS_OK] = 0u;  // the labels are consumed
  const uint32_t total = spawn_tot[1];
  const uint32_t acc = spawn_accepted(e, total);
  e.ctrl[PX_C_SPAWNED] += total;
  e.ctrl[PX_C_PCURSOR] += acc;
  e.ctrl[PX_C_ARENA] += arena_tot[0];
  e.ctrl[PX_S_BILL] += e.ec[EC_R_BILL];
  e.ctrl[PX_S_DESTROYED] += e.ec[EC_R_DESTROYED];
  e.ctrl[PX_C_DIRTY] += min(e.ec[EC_R_DIRTY], e.ncell);
}

// maint of the cells the extraction dirtied and of the cells alloc expanded
__global__ __launch_bounds__(64) void k_resolve_maint(EditView e) {
  const uint32_t nd = min(e.ec[EC_R_DIRTY], e.ncell);
  const uint32_t na = min(e.ec[EC_ALLOCS], e.ncell);
  for (uint32_t i = blockIdx.x; i < nd; i += gridDim.x) maint_cell(e, e.rdirty[i]);
  for (uint32_t i = blockIdx.x; i < na; i += gridDim.x) maint_cell(e, e.alloc_list[i]);
}

// ---------------------------------------------------------------------------
// Falling chunks: chunkArgs / chunkBegin / chunkCollide / chunkStep /
// chunkLand (sim.wgsl:951-1075).  Arena order decides which voxels of a chunk
// push the (at most 6) impacts of a break-through, the order of the writes
// and of the shatter debris.

PX_DEV float u2f(uint32_t u) { return __uint_as_float(u); }
PX_DEV uint32_t f2bits(float f) { return __float_as_uint(f); }
// WGSL u32(f32): truncation toward zero, saturating, NaN -> 0
PX_DEV uint32_t px_f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// chunkArgs + chunkBegin, one thread per chunk (one workgroup)
__global__ __launch_bounds__(512) void k_chunk_begin(EditView e) {
  __shared__ uint32_t s_act[8];
  const uint32_t c = threadIdx.x;
  const uint32_t n0 = min(e.ctrl[PX_C_CHUNKS], e.max_chunks);
  uint32_t* b = e.chunks + (size_t)c * 16u;
  uint32_t act = 0u;
  if (c < n0) {
    uint32_t st = b[0];
    if (st == 2u) b[0] = 0u;
    if (st == 1u && (b[5] == 0u || b[2] > (uint32_t)e.dims[1])) {
      b[0] = 0u;
      st = 0u;
    }
    act = st == 1u ? 1u : 0u;
  }
  act = wave_sum(act);
  if (__lane_id() == 0) s_act[threadIdx.x >> 6] = act;
  __syncthreads();
  uint32_t active = 0u;
  for (uint32_t k = 0; k < (blockDim.x >> 6); k++) active += s_act[k];
  if (c == 0) {
    if (active == 0u) {
      e.ctrl[PX_C_CHUNKS] = 0u;
      e.ctrl[PX_C_ARENA] = 0u;
    }
    e.ctrl[PX_S_FALLING] = active;
  }
  const uint32_t n = active == 0u ? 0u : n0;
  if (c == 0) e.ec[EC_C_LAND] = active == 0u ? 0u : min(e.ctrl[PX_C_ARENA], e.max_arena);
  if (c < n && b[0] == 1u) {
    const float vy = fminf(u2f(b[1]) + e.grav * e.dt, 25.0f / e.vs);
    const float travel = vy * e.dt + u2f(b[3]);
    const uint32_t steps = min(px_f2u(floorf(travel)), 48u);
    b[1] = f2bits(vy);
    b[3] = f2bits(travel - floorf(travel));
    b[8] = steps;
    b[6] = kNoBlock;
    b[7] = 0u;
  }
}

PX_DEV uint32_t arena_len(const EditView& e) { return min(e.ctrl[PX_C_ARENA], e.max_arena); }

// chunkCollide: the first solid voxel below each falling voxel within the
// frame's steps.  blockedAt is a min (any order); a break-through candidate
// (fragile voxel hit at speed > 1.5) is counted in word 7 and flagged
// (cpair[j].x = 1) for the impact selection.
__global__ __launch_bounds__(256) void k_chunk_collide(EditView e) {
  const uint32_t na = arena_len(e);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < na; j += gridDim.x * blockDim.x) {
    uint32_t cand = 0u;
    const uint32_t vv = e.arena[2u * j + 1u], v = vv & 0xffffu;
    if (v != 0u) {
      uint32_t* b = e.chunks + (size_t)(vv >> 16) * 16u;
      const uint32_t steps = b[0] == 1u ? b[8] : 0u;
      if (steps != 0u) {
        int32_t px, py, pz;
        key_pos(e, e.arena[2u * j], px, py, pz);
        py -= (int32_t)b[2];
        const float speed = u2f(b[1]) * e.vs;
        for (uint32_t s = 1; s <= steps; s++) {
          const int32_t qy = py - (int32_t)s;
          if (qy < 0) {
            atomicMin(b + 6, s);
            break;
          }
          bool pool;
          const uint32_t w = voxel_at(e, px, qy, pz, &pool);
          if (w != 0u) {
            if (mat_fragile(e.mats, w) && speed > 1.5f) {
              atomicAdd(b + 7, 1u);
              cand = 1u;
            }
            atomicMin(b + 6, s);
            break;
          }
        }
      }
    }
    e.cpair[j] = make_uint2(cand, 0u);
  }
}

// One round of the impact selection: the smallest candidate index per chunk
// (a wavefront reduces its lanes per chunk first), marked chosen (2).
__global__ __launch_bounds__(256) void k_cand_min(EditView e, uint32_t* best) {
  const uint32_t na = arena_len(e);
  const uint32_t lane = __lane_id();
  for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < na; j0 += gridDim.x * blockDim.x) {
    const uint32_t j = j0 + threadIdx.x;
    uint32_t c = PX_NONE;
    if (j < na && e.cpair[j].x == 1u) c = e.arena[2u * j + 1u] >> 16;
    for (;;) {
      const uint64_t act = __ballot(c != PX_NONE);
      if (act == 0ull) break;
      const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
      const uint32_t c0 = __shfl(c, (int)leader);
      // lanes are in increasing j: the first lane of the chunk holds its smallest index
      if (lane == leader) atomicMin(best + c0, j);
      if (c == c0) c = PX_NONE;
    }
  }
}

__global__ void k_cand_pick(EditView e, uint32_t* best) {
  const uint32_t c = threadIdx.x;
  if (c >= e.max_chunks) return;
  const uint32_t j = best[c];
  if (j != PX_NONE) {
    e.cpair[j].x = 2u;
    best[c] = PX_NONE;
  }
}

// chosen candidates -> (1, 0) for the ordered compaction of the impacts
__global__ __launch_bounds__(256) void k_cand_flags(EditView e) {
  const uint32_t na = arena_len(e);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < na; j += gridDim.x * blockDim.x)
    e.cpair[j] = make_uint2(e.cpair[j].x == 2u ? 1u : 0u, 0u);
}

// pushImpact (sim.wgsl:203-211) for the chosen candidates in arena order
__global__ __launch_bounds__(256) void k_cand_impacts(EditView e, const uint32_t* tot) {
  const uint32_t na = arena_len(e);
  const uint32_t k0 = e.ctrl[PX_C_IMPACTS];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < na; j += gridDim.x * blockDim.x) {
    if (e.cpair[j].x == 0u) continue;
    const uint32_t k = k0 + e.coff[j].x;
    if (k >= e.max_imp) continue;
    const uint32_t vv = e.arena[2u * j + 1u];
    const uint32_t* b = e.chunks + (size_t)(vv >> 16) * 16u;
    int32_t px, py, pz;
    key_pos(e, e.arena[2u * j], px, py, pz);
    py -= (int32_t)b[2];
    // the candidate's hit is at blockedAt of its own walk: find it again
    int32_t qy = py;
    for (uint32_t s = 1; s <= b[8]; s++) {
      bool pool;
      qy = py - (int32_t)s;
      if (qy >= 0 && voxel_at(e, px, qy, pz, &pool) != 0u) break;
    }
    const float speed = u2f(b[1]) * e.vs;
    float* im = e.impacts + (size_t)k * 16u;
    im[0] = (float)px + 0.5f;
    im[1] = (float)qy + 0.5f;
    im[2] = (float)pz + 0.5f;
    im[3] = 0.02f / e.vs;
    im[4] = 0.12f / e.vs;
    im[5] = speed * 0.4f / e.vs;
    im[6] = 0.0f; im[7] = 1.0f; im[8] = 0.0f;
    im[9] = 0.0f; im[10] = -1.0f; im[11] = 0.0f;
    im[12] = 0.0f; im[13] = 0.0f; im[14] = 0.0f; im[15] = 0.0f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) e.ec[EC_C_CAND] = tot[0];
}

// chunkStep, one thread per chunk
__global__ __launch_bounds__(512) void k_chunk_step(EditView e) {
  const uint32_t c = threadIdx.x;
  if (threadIdx.x == 0) e.ctrl[PX_C_IMPACTS] += e.ec[EC_C_CAND];
  if (c >= min(e.ctrl[PX_C_CHUNKS], e.max_chunks)) return;
  uint32_t* b = e.chunks + (size_t)c * 16u;
  if (b[0] != 1u) return;
  const uint32_t steps = b[8];
  if (steps == 0u) return;
  const uint32_t blocked = b[6];
  uint32_t drop = steps;
  if (blocked <= steps) {
    drop = blocked - 1u;
    if (b[7] > 0u) {
      b[1] = f2bits(u2f(b[1]) * 0.6f);  // broke through something: keep falling
    } else {
      b[9] = f2bits(u2f(b[1]) * e.vs);
      b[0] = 2u;
    }
  }
  b[2] += drop;
}

// chunkLand: per arena entry of a landed chunk, (write, shatter) counts
__global__ __launch_bounds__(256) void k_land_count(EditView e) {
  const uint32_t na = arena_len(e);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < na; j += gridDim.x * blockDim.x) {
    uint2 f = make_uint2(0u, 0u);
    const uint32_t vv = e.arena[2u * j + 1u], v = vv & 0xffffu;
    if (v != 0u) {
      const uint32_t* b = e.chunks + (size_t)(vv >> 16) * 16u;
      if (b[0] == 2u) {
        const bool shatter = b[4] != 0u && u2f(b[9]) > 2.0f && mat_fragile(e.mats, v);
        f = shatter ? make_uint2(0u, 1u) : make_uint2(1u, 0u);
      }
    }
    e.cpair[j] = f;
  }
}

__global__ __launch_bounds__(256) void k_land_emit(EditView e, const uint32_t* tot) {
  const uint32_t na = arena_len(e);
  const uint32_t w0 = e.ctrl[PX_C_WRITES];
  const uint32_t accepted = spawn_accepted(e, tot[1]);
  const uint32_t cursor = e.ctrl[PX_C_PCURSOR];
  uint32_t bill = 0u, destroyed = 0u;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < na; j += gridDim.x * blockDim.x) {
    const uint2 f = e.cpair[j];
    if ((f.x | f.y) == 0u) continue;
    const uint32_t vv = e.arena[2u * j + 1u], v = vv & 0xffffu;
    const uint32_t key = e.arena[2u * j];
    const uint32_t* b = e.chunks + (size_t)(vv >> 16) * 16u;
    e.arena[2u * j + 1u] = 0u;
    int32_t px, py, pz;
    key_pos(e, key, px, py, pz);
    py -= (int32_t)b[2];
    const uint2 o = e.coff[j];
    if (f.y) {
      bill += price_dimes(e.mats, v);
      destroyed++;
      const float speed = u2f(b[9]);
      const uint32_t seed = key ^ e.fseed;
      const float r[3] = {(rnd01(seed) - 0.5f) * 0.9f, rnd01(seed + 1u) * 0.5f, (rnd01(seed + 2u) - 0.5f) * 0.9f};
      float vel[3];
      for (int k = 0; k < 3; k++) vel[k] = (r[k] * speed) / e.vs;
      const float pos[3] = {(float)px + 0.5f, (float)py + 0.5f, (float)pz + 0.5f};
      spawn_at(e, o.y, accepted, cursor, pos, vel, v | debris_flags(e.mats, v, seed), 5.0f + 3.0f * rnd01(seed + 3u));
      continue;
    }
    const uint32_t k = w0 + o.x;  // pushWrite, sim.wgsl:186-192
    if (k < e.max_writes) {
      e.wlist[2u * k] = vkey(e, px, py, pz);
      e.wlist[2u * k + 1u] = v;
    }
    // whatever it landed on is re-checked for support (pushSeed, :1068-1072)
    if (e.mode == PX_ISLANDS_SEEDED && (px_hashu(key) & 15u) == 0u) {
      bool pool;
      const uint32_t w = voxel_at(e, px, py - 1, pz, &pool);
      if (w != 0u && !(e.mats[(w & 0xffu) * 20u + 18u] > 0.5f)) {
        const uint32_t b = vkey(e, px, py - 1, pz);
        atomicOr(&e.seeds[b >> 5], 1u << (b & 31u));
        atomicAdd(&e.ctrl[PX_C_PENDING], 1u);
      }
    }
  }
  bill = wave_sum(bill);
  destroyed = wave_sum(destroyed);
  if (__lane_id() == 0 && (bill | destroyed)) {
    atomicAdd(&e.ec[EC_R_BILL], bill);
    atomicAdd(&e.ec[EC_R_DESTROYED], destroyed);
  }
}

__global__ void k_land_finish(EditView e, const uint32_t* tot) {
  if (threadIdx.x != 0) return;
  const uint32_t acc = spawn_accepted(e, tot[1]);
  e.ctrl[PX_C_WRITES] += tot[0];
  e.ctrl[PX_C_SPAWNED] += tot[1];
  e.ctrl[PX_C_PCURSOR] += acc;
  e.ctrl[PX_S_BILL] += e.ec[EC_R_BILL];
  e.ctrl[PX_S_DESTROYED] += e.ec[EC_R_DESTROYED];
  e.ec[EC_R_BILL] = 0u;
  e.ec[EC_R_DESTROYED] = 0u;
}

EditView make_view(px_ctx* c) {
  EditView e;
  e.grid = c->d_grid;
  e.pool = c->d_pool;
  e.sgrid = c->d_super;
  e.mats = c->d_mats;
  e.ctrl = c->d_ctrl;
  e.free_list = c->d_free;
  e.pool_cap = c->pool_cap;
  for (int a = 0; a < 3; a++) {
    e.dims[a] = (int32_t)c->dims[a];
    e.bd[a] = (int32_t)c->bd[a];
    e.sd[a] = (int32_t)c->sd[a];
  }
  e.ncell = c->ncell;
  e.imask = c->d_cell_imask;
  e.dirty_flags = c->d_dirty_flags;
  e.job_list = c->d_job_list;
  e.alloc_list = c->d_alloc_list;
  e.pane_cells = c->d_pane_cells;
  e.par = c->d_par;
  e.cnt = c->d_ncnt;
  e.flag = c->d_nflag;
  e.lab = c->d_lab;
  e.bnode = c->d_bnode;
  e.bface = c->d_bface;
  e.nodes = c->d_nodes;
  e.slots_cell = (c->ncell + kEditBlocks - 1u) / kEditBlocks;
  e.slots_node = e.slots_cell * 256u;
  // no chain of distinct nodes is longer than the node lists can hold
  e.walk_cap = c->lim.uf_walk_cap ? c->lim.uf_walk_cap : e.slots_node * kEditBlocks;
  e.part = c->d_part;
  e.bcount = c->d_bcount;
  e.out = c->d_out_labels;
  e.out_cap = c->label_cap;
  e.ec = c->d_edit_ctr;
  e.impacts = c->d_impacts;
  e.shots = c->d_shots;
  e.work = c->d_edit_work;
  e.carve_rows = c->d_carve_rows;
  e.flashes = c->d_flashes;
  e.max_imp = std::min<uint32_t>(c->lim.max_impacts, 64u);
  e.seeds = c->d_seeds;
  e.seed_words = (uint32_t)((c->vox_total + 31u) / 32u);
  e.mode = c->lim.island_mode;
  // hash slots the reference's search may fill before it stops (sim.wgsl:758)
  e.node_budget = (uint32_t)(((uint64_t)c->lim.hash_size * 7u) / 10u);
  e.etot = c->d_rtot + 16;
  e.wlist = c->d_write_list;
  e.max_writes = c->lim.max_writes;
  e.max_alloc = c->lim.max_alloc;
  e.cfirst = c->d_cfirst;
  e.wkeys = c->d_wkeys;
  e.wsorted = c->d_wsorted;
  e.wdirty = c->d_wdirty;
  e.chunks = c->d_chunks;
  e.arena = c->d_arena;
  e.max_chunks = std::min<uint32_t>(c->lim.max_chunks, kMaxChunkIds);
  e.max_arena = c->lim.max_arena;
  e.comp_budget = c->lim.component_budget;
  e.max_impacts = c->lim.max_impacts;
  e.cpair = c->d_cpair;
  e.coff = c->d_coff;
  e.tiles = c->d_tiles;
  e.rlist = c->d_rlist;
  e.rdirty = c->d_rdirty;
  e.parts = c->d_particles;
  e.pmask = c->max_particles ? c->max_particles - 1u : 0u;
  e.spawn_budget = c->lim.spawn_budget;
  e.vs = c->h_uniforms[3];
  e.dt = c->h_uniforms[24];
  e.grav = c->h_uniforms[25];
  {
    // frameSeed() = u32(U.time.y) * 0x9E3779B9 (sim.wgsl:69)
    const float fy = c->h_uniforms[29];
    uint32_t f = 0u;
    if (fy > 0.0f) f = fy >= 4294967040.0f ? 0xFFFFFFFFu : (uint32_t)fy;
    e.fseed = f * 0x9E3779B9u;
  }
  return e;
}

}  // namespace

int px_edit_alloc(px_ctx* c) {
  size_t words = (c->ncell + 31u) / 32u;
  size_t slots_cell = (c->ncell + kEditBlocks - 1u) / kEditBlocks;
  PX_HIP_CHECK(c, hipMalloc(&c->d_impacts, 64 * 16 * sizeof(float)));
  PX_HIP_CHECK(c, hipMalloc(&c->d_impact_stage, 64 * 16 * sizeof(float)));
  PX_HIP_CHECK(c, hipMalloc(&c->d_cell_imask, (size_t)c->ncell * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_dirty_flags, words * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_job_list, (size_t)c->ncell * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_alloc_list, (size_t)c->ncell * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_pane_cells, slots_cell * kEditBlocks * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_part, (size_t)kEditBlocks * kPartWords * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_bcount, (size_t)kEditBlocks * kCountWords * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_edit_ctr, 64 * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_shots, kShotWords * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_edit_work, (size_t)c->ncell * 2 * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_carve_rows, (size_t)kCarveRowCache * 64 * sizeof(uint2)));
  PX_HIP_CHECK(c, hipMalloc(&c->d_flashes, PX_MAX_FLASHES * 8 * 4));
  hipLaunchKernelGGL(k_zero32, dim3(1), dim3(256), 0, c->stream, c->d_shots, (size_t)kShotWords);
  hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<uint32_t*>(c->d_flashes),
                     (size_t)PX_MAX_FLASHES * 8);
  hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(256), 0, c->stream, c->d_cell_imask, (size_t)c->ncell);
  hipLaunchKernelGGL(k_zero32, dim3(256), dim3(256), 0, c->stream, c->d_dirty_flags, words);
  hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, c->stream, c->d_edit_ctr, (size_t)64);
  hipLaunchKernelGGL(k_zero32, dim3(64), dim3(256), 0, c->stream, c->d_bcount, (size_t)kEditBlocks * kCountWords);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

void px_edit_free(px_ctx* c) {
  void* bufs[] = {c->d_impacts, c->d_impact_stage, c->d_cell_imask, c->d_dirty_flags,
                  c->d_job_list, c->d_alloc_list, c->d_pane_cells, c->d_part, c->d_bcount, c->d_edit_ctr, c->d_par,
                  c->d_ncnt, c->d_nflag, c->d_lab, c->d_bnode, c->d_bface, c->d_nodes, c->d_out_labels,
                  c->d_seeds, c->d_shots, c->d_flashes, c->d_edit_work, c->d_carve_rows};
  for (void* p : bufs)
    if (p) hipFree(p);
  c->d_impacts = c->d_impact_stage = nullptr;
  c->d_cell_imask = nullptr;
  c->d_par = c->d_ncnt = c->d_nflag = nullptr;
  c->d_lab = nullptr;
  c->d_bnode = nullptr;
  c->d_bface = nullptr;
  c->d_nodes = c->d_out_labels = c->d_seeds = c->d_shots = nullptr;
  c->d_flashes = nullptr;
  c->d_edit_work = nullptr;
  c->d_carve_rows = nullptr;
}

// the labelling arrays are sized by the voxel count and allocated on first use
static int ensure_ccl(px_ctx* c) {
  if (c->d_par) return PX_OK;
  size_t nv = c->vox_total;
  size_t slots_node = (size_t)((c->ncell + kEditBlocks - 1u) / kEditBlocks) * 256u;
  size_t node_cap = slots_node * kEditBlocks;
  size_t label_cap = std::min<size_t>(node_cap, nv);
  PX_HIP_CHECK(c, hipMalloc(&c->d_par, nv * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_ncnt, nv * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_nflag, nv * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_lab, (size_t)c->pool_cap * 1024));
  PX_HIP_CHECK(c, hipMalloc(&c->d_bnode, (size_t)c->pool_cap * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_bface, (size_t)c->pool_cap * kFaceMasks * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_nodes, node_cap * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_out_labels, label_cap * 8));
  c->node_cap = (uint32_t)std::min<size_t>(node_cap, 0xFFFFFFF0u);
  c->label_cap = (uint32_t)std::min<size_t>(label_cap, 0xFFFFFFF0u);
  return PX_OK;
}

// the seed bits (one per voxel key), cleared once; the search clears what it reads
int px_seeds_alloc(px_ctx* c) {
  if (c->d_seeds || c->lim.island_mode != PX_ISLANDS_SEEDED) return PX_OK;
  const size_t words = (c->vox_total + 31u) / 32u;
  PX_HIP_CHECK(c, hipMalloc(&c->d_seeds, words * 4));
  hipLaunchKernelGGL(k_zero32, dim3(1024), dim3(256), 0, c->stream, c->d_seeds, words);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_clear_seeds(px_ctx* c) {
  if (!c->d_seeds) return PX_OK;
  hipLaunchKernelGGL(k_zero32, dim3(1024), dim3(256), 0, c->stream, c->d_seeds, (c->vox_total + 31u) / 32u);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_launch_shoot(px_ctx* c) {
  if (int rc = px_chunks_alloc(c)) return rc;  // the particle ring takes the debris
  EditView e = make_view(c);
  ShootArgs a;
  for (int k = 0; k < 4; k++) a.fire[k] = c->h_uniforms[40 + k];
  a.fire2[0] = c->h_uniforms[44];
  a.fire2[1] = c->h_uniforms[45];
  a.time_x = c->h_uniforms[28];
  hipLaunchKernelGGL(k_shoot, dim3(1), dim3(64), 0, c->stream, e, px_view(c), (const float*)c->d_camera, a);
  PX_HIP_CHECK(c, hipGetLastError());
  c->masks_dirty = c->density_dirty = true;
  return PX_OK;
}

int px_clear_shots(px_ctx* c) {
  hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, c->stream, c->d_shots, (size_t)2);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_launch_push_impacts(px_ctx* c, uint32_t n) {
  EditView e = make_view(c);
  hipLaunchKernelGGL(k_push_impacts, dim3(1), dim3(1024), 0, c->stream, e, (const float*)c->d_impact_stage, n);
  hipLaunchKernelGGL(k_push_impacts_commit, dim3(1), dim3(64), 0, c->stream, e, n);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

// One island pass, every launch recorded whatever the device holds: the
// kernels read the impact count and the run decision k_edit_begin writes
// (force = run the labelling even without impacts or seeds), so the host
// waits for nothing and a frame of the chain can be captured in a graph.
int px_launch_island_pass(px_ctx* c, bool force) {
  if (int rc = ensure_ccl(c)) return rc;
  if (int rc = px_chunks_alloc(c)) return rc;  // effects scan buffers, dirty list, particle ring
  if (int rc = px_seeds_alloc(c)) return rc;
  PxProfScope prof(c, PX_PROF_EDIT);
  EditView e = make_view(c);
  hipStream_t s = c->stream;
  const dim3 G(kEditBlocks), W(64);
  hipLaunchKernelGGL(k_edit_begin, dim3(1), dim3(64), 0, s, e, force ? 1u : 0u);
  hipLaunchKernelGGL(k_impact_jobs, dim3(kJobBlocks, 64), dim3(256), 0, s, e);
  hipLaunchKernelGGL(k_alloc, G, W, 0, s, e);
  // pane components on the world as it is after alloc and before carve
  // (k_local_ccl<1> commits the allocations first)
  hipLaunchKernelGGL(k_local_ccl<1>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_face_merge<1>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_flatten, G, W, 0, s, e, (uint32_t)BC_PANENODES);  // also clears cpair
  // carve + flood: count the effects (k_carve<0> also takes shoot's flood
  // entries), rank them in cell order, apply
  hipLaunchKernelGGL(k_carve<0>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_flood<0>, G, W, 0, s, e);
  scan2_launch(s, e.cpair, e.coff, e.tiles, c->scan_cap, nullptr, c->ncell, nullptr, nullptr, e.etot);
  hipLaunchKernelGGL(k_carve<1>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_flood<1>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_maint, G, W, 0, s, e);  // also spawn()'s counters
  // labelling
  hipLaunchKernelGGL(k_local_ccl<0>, G, W, 0, s, e);
  hipLaunchKernelGGL(k_face_merge<0>, G, W, 0, s, e);
  if (c->lim.island_mode == PX_ISLANDS_SEEDED) {  // the seeds find their roots through flat parents
    hipLaunchKernelGGL(k_flatten, G, W, 0, s, e, (uint32_t)BC_NODES);
    hipLaunchKernelGGL(k_seed_mark, dim3(2048), dim3(256), 0, s, e);
    hipLaunchKernelGGL((k_aggregate<0, false>), G, W, 0, s, e);
  } else {
    hipLaunchKernelGGL((k_aggregate<0, true>), G, W, 0, s, e);
  }
  hipLaunchKernelGGL((k_aggregate<1, false>), G, W, 0, s, e);
  hipLaunchKernelGGL(k_emit<0>, G, W, 0, s, e);
  if (c->lim.island_mode == PX_ISLANDS_SEEDED) {
    hipLaunchKernelGGL(k_node_budget, dim3(1), dim3(1024), 0, s, e);
    hipLaunchKernelGGL(k_emit<1>, G, W, 0, s, e);
  }
  hipLaunchKernelGGL(k_edit_finish, dim3(kEditFinishBlocks), dim3(256), 0, s, e);
  PX_HIP_CHECK(c, hipGetLastError());
  c->labels_valid = true;
  c->masks_dirty = c->density_dirty = true;
  return px_refresh_masks(c);
}

int px_edit_errors(px_ctx* c, const char* what, uint32_t bits) {
  if (bits & kErrWalk)
    return px_fail(c, PX_ERR_STATE, what,
                   "a union-find walk met a parent above its child or passed its step bound; labels not trusted");
  if (bits & kErrLabel) return px_fail(c, PX_ERR_STATE, what, "internal error: a brick label was not a local index");
  return PX_OK;
}

int px_check_island_pass(px_ctx* c, const char* what) {
  if (!c->d_edit_ctr) return PX_OK;
  uint32_t bits = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&bits, c->d_edit_ctr + EC_OVERFLOW, 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return px_edit_errors(c, what, bits);
}

int px_read_island_labels(px_ctx* c, uint32_t* min_key, uint32_t* count, uint32_t cap, uint32_t* n) {
  *n = 0;
  if (!c->labels_valid) return px_fail(c, PX_ERR_STATE, "px_island_labels", "no island pass has run");
  uint32_t ec[EC_WORDS];
  PX_HIP_CHECK(c, hipMemcpyAsync(ec, c->d_edit_ctr, sizeof ec, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  if (int rc = px_edit_errors(c, "px_island_labels", ec[EC_OVERFLOW])) return rc;
  if (ec[EC_OVERFLOW] || ec[EC_LABELS] > c->label_cap)
    return px_fail(c, PX_ERR_CAPACITY, "px_island_labels", "component node list overflowed");
  uint32_t m = ec[EC_LABELS];
  std::vector<uint32_t> pairs((size_t)m * 2);
  if (m) {
    PX_HIP_CHECK(c, hipMemcpyAsync(pairs.data(), c->d_out_labels, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  std::vector<uint64_t> packed(m);
  for (uint32_t i = 0; i < m; i++) packed[i] = ((uint64_t)pairs[2 * i] << 32) | pairs[2 * i + 1];
  std::sort(packed.begin(), packed.end());
  for (uint32_t i = 0; i < m && i < cap; i++) {
    min_key[i] = (uint32_t)(packed[i] >> 32);
    count[i] = (uint32_t)packed[i];
  }
  *n = m;
  return PX_OK;
}

// ---------------------------------------------------------------------------
// voxel writes

int px_writes_alloc(px_ctx* c) {
  if (c->d_write_list) return PX_OK;
  const size_t nw = c->lim.max_writes ? c->lim.max_writes : 1u;
  PX_HIP_CHECK(c, hipMalloc(&c->d_write_list, nw * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_cfirst, (size_t)c->ncell * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_wkeys, nw * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_wsorted, nw * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_wdirty, nw * 4));
  PX_HIP_CHECK(c, hipMemsetAsync(c->d_write_list, 0, nw * 8, c->stream));
  PX_HIP_CHECK(c, hipMemsetAsync(c->d_cfirst, 0xFF, (size_t)c->ncell * 4, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

void px_writes_free(px_ctx* c) {
  void* bufs[] = {c->d_write_list, c->d_cfirst, c->d_wkeys, c->d_wsorted, c->d_wdirty, c->d_sort_tmp};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  c->d_write_list = c->d_cfirst = c->d_wdirty = nullptr;
  c->d_wkeys = c->d_wsorted = nullptr;
  c->d_sort_tmp = nullptr;
  c->sort_tmp_bytes = 0;
}

int px_launch_writes(px_ctx* c) {
  if (int rc = px_writes_alloc(c)) return rc;
  if (c->lim.max_writes == 0) return PX_OK;
  PxProfScope prof(c, PX_PROF_WRITES);
  EditView e = make_view(c);
  hipStream_t s = c->stream;
  const uint32_t nw = c->lim.max_writes;
  const dim3 G(std::min<uint32_t>((nw + 255u) / 256u, 2048u)), W(256);
  hipLaunchKernelGGL(k_write_args, dim3(1), dim3(64), 0, s, e);
  hipLaunchKernelGGL(k_write_first, G, W, 0, s, e);
  hipLaunchKernelGGL(k_write_request, dim3(1), dim3(1024), 0, s, e);
  hipLaunchKernelGGL(k_alloc, dim3(kEditBlocks), dim3(64), 0, s, e);
  hipLaunchKernelGGL(k_alloc_commit, dim3(1), dim3(64), 0, s, e);
  PX_HIP_CHECK(c, hipGetLastError());
  // column bits: x + dx * z < dx * dz
  const uint64_t ncol = (uint64_t)c->dims[0] * c->dims[2];
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < ncol) bits++;
  // from bit 32: the keys are in list order already and the sort is stable
  if (int rc = px_sort_u64(c, c->d_wkeys, c->d_wsorted, nw, 32u, 32u + bits)) return rc;
  hipLaunchKernelGGL(k_write_apply, G, W, 0, s, e);
  hipLaunchKernelGGL(k_write_maint, dim3(kEditBlocks), dim3(64), 0, s, e);
  hipLaunchKernelGGL(k_write_finish, dim3(1), dim3(64), 0, s, e);
  PX_HIP_CHECK(c, hipGetLastError());
  c->masks_dirty = c->density_dirty = true;
  return px_refresh_masks(c);
}

extern "C" int px_push_writes(px_ctx* c, const uint32_t* pairs, uint32_t n) {
  if (!c || (n && !pairs)) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (int rc = px_writes_alloc(c)) return rc;
  if (n == 0) return PX_OK;
  // staged in the sorted-key buffer (max_writes pairs fit; later ones are dropped anyway)
  const uint32_t m = std::min(n, c->lim.max_writes);
  if (m) PX_HIP_CHECK(c, hipMemcpyAsync(c->d_wsorted, pairs, (size_t)m * 8, hipMemcpyHostToDevice, c->stream));
  EditView e = make_view(c);
  if (m)
    hipLaunchKernelGGL(k_push_writes, dim3((m + 255u) / 256u), dim3(256), 0, c->stream, e,
                       reinterpret_cast<const uint32_t*>(c->d_wsorted), m);
  hipLaunchKernelGGL(k_push_commit, dim3(1), dim3(64), 0, c->stream, e, n);
  PX_HIP_CHECK(c, hipGetLastError());
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

// ---------------------------------------------------------------------------
// island extraction and falling chunks

int px_chunks_alloc(px_ctx* c) {
  if (c->d_chunks) return PX_OK;
  if (int rc = px_writes_alloc(c)) return rc;
  if (!c->d_particles) {
    // the debris the extraction and the landings spawn needs the ring
    if (int rc = px_particles_init(c, 0)) return rc;
  }
  const size_t nch = kMaxChunkIds + 1u, na = c->lim.max_arena ? c->lim.max_arena : 1u;
  const size_t cap = std::max<size_t>(c->ncell, na);
  const size_t tiles = (cap + kScanTile - 1u) / kScanTile;
  PX_HIP_CHECK(c, hipMalloc(&c->d_chunks, nch * 16 * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_arena, na * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_cpair, cap * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_coff, cap * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_tiles, tiles * 8));
  PX_HIP_CHECK(c, hipMalloc(&c->d_rlist, (size_t)c->ncell * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_rdirty, (size_t)c->ncell * 4));
  PX_HIP_CHECK(c, hipMalloc(&c->d_rtot, 128));
  PX_HIP_CHECK(c, hipMalloc(&c->d_best, nch * 4));
  PX_HIP_CHECK(c, hipMemsetAsync(c->d_chunks, 0, nch * 16 * 4, c->stream));
  PX_HIP_CHECK(c, hipMemsetAsync(c->d_arena, 0, na * 8, c->stream));
  PX_HIP_CHECK(c, hipMemsetAsync(c->d_best, 0xFF, nch * 4, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  c->scan_cap = cap;
  return PX_OK;
}

void px_chunks_free(px_ctx* c) {
  void* bufs[] = {c->d_chunks, c->d_arena, c->d_cpair, c->d_coff, c->d_tiles, c->d_rlist,
                  c->d_rdirty, c->d_rtot, c->d_best};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  c->d_chunks = c->d_arena = c->d_rlist = c->d_rdirty = c->d_rtot = c->d_best = nullptr;
  c->d_cpair = c->d_coff = c->d_tiles = nullptr;
}

// exclusive (x, y) offsets of e.cpair[0..n) into e.coff, totals into tot[0..1]
static void scan2(px_ctx* c, const EditView& e, const uint32_t* n_dev, uint32_t n_fixed, const uint32_t* base_x,
                  const uint32_t* base_y, uint32_t* tot) {
  scan2_launch(c->stream, e.cpair, e.coff, e.tiles, c->scan_cap, n_dev, n_fixed, base_x, base_y, tot);
}

int px_launch_resolve(px_ctx* c) {
  if (!c->labels_valid) return px_fail(c, PX_ERR_STATE, "px_resolve_islands", "no island pass to resolve");
  if (int rc = px_chunks_alloc(c)) return rc;
  PxProfScope prof(c, PX_PROF_CHUNKS);
  EditView e = make_view(c);
  hipStream_t s = c->stream;
  const dim3 G(kEditBlocks), W(64);
  uint32_t* t1 = c->d_rtot;       // cells: (extract, expand) totals
  uint32_t* t2 = c->d_rtot + 4;   // (arena, -) totals
  uint32_t* t3 = c->d_rtot + 8;   // (arena, spawn) totals
  hipLaunchKernelGGL(k_resolve_begin, dim3(1), dim3(64), 0, s, e);
  hipLaunchKernelGGL(k_zero32, dim3(1024), dim3(256), 0, s, reinterpret_cast<uint32_t*>(c->d_cpair),
                     (size_t)c->ncell * 2);
  hipLaunchKernelGGL(k_resolve_chunks, dim3(1), dim3(1024), 0, s, e);
  hipLaunchKernelGGL(k_resolve_mark, G, W, 0, s, e);
  scan2(c, e, nullptr, c->ncell, nullptr, nullptr, t1);
  hipLaunchKernelGGL(k_resolve_lists, dim3(1024), dim3(256), 0, s, e, t1);
  hipLaunchKernelGGL(k_alloc, G, W, 0, s, e);
  hipLaunchKernelGGL(k_alloc_commit, dim3(1), dim3(64), 0, s, e);
  const uint32_t* ncells = c->d_edit_ctr + EC_R_CELLS;
  const uint32_t* arena0 = c->d_ctrl + PX_C_ARENA;
  hipLaunchKernelGGL(k_extract<0>, G, W, 0, s, e, t3);
  scan2(c, e, ncells, 0, arena0, nullptr, t2);
  hipLaunchKernelGGL(k_extract<1>, G, W, 0, s, e, t3);
  scan2(c, e, ncells, 0, arena0, nullptr, t3);
  hipLaunchKernelGGL(k_extract<2>, G, W, 0, s, e, t3);
  hipLaunchKernelGGL(k_arena_tally, dim3(256), dim3(256), 0, s, e, t2);
  hipLaunchKernelGGL(k_resolve_finish, dim3(1), dim3(64), 0, s, e, t2, t3);
  hipLaunchKernelGGL(k_resolve_maint, G, W, 0, s, e);
  PX_HIP_CHECK(c, hipGetLastError());
  c->labels_valid = false;
  c->masks_dirty = c->density_dirty = true;
  return px_refresh_masks(c);
}

int px_launch_chunk_frame(px_ctx* c) {
  if (int rc = px_chunks_alloc(c)) return rc;
  if (int rc = px_seeds_alloc(c)) return rc;  // landings push seeds in seeded mode
  PxProfScope prof(c, PX_PROF_CHUNKS);
  EditView e = make_view(c);
  hipStream_t s = c->stream;
  const uint32_t ga = (uint32_t)std::min<size_t>((std::max<uint32_t>(c->lim.max_arena, 1u) + 255u) / 256u, 2048u);
  const dim3 GA(ga), WA(256);
  uint32_t* t = c->d_rtot + 12;
  const uint32_t* na = c->d_edit_ctr + EC_C_LAND;
  hipLaunchKernelGGL(k_chunk_begin, dim3(1), dim3(512), 0, s, e);
  hipLaunchKernelGGL(k_chunk_collide, GA, WA, 0, s, e);
  for (int r = 0; r < 6; r++) {
    hipLaunchKernelGGL(k_cand_min, GA, WA, 0, s, e, c->d_best);
    hipLaunchKernelGGL(k_cand_pick, dim3(1), dim3(512), 0, s, e, c->d_best);
  }
  hipLaunchKernelGGL(k_cand_flags, GA, WA, 0, s, e);
  scan2(c, e, na, 0, nullptr, nullptr, t);
  hipLaunchKernelGGL(k_cand_impacts, GA, WA, 0, s, e, t);
  hipLaunchKernelGGL(k_chunk_step, dim3(1), dim3(512), 0, s, e);
  hipLaunchKernelGGL(k_land_count, GA, WA, 0, s, e);
  scan2(c, e, na, 0, nullptr, nullptr, t);
  hipLaunchKernelGGL(k_land_emit, GA, WA, 0, s, e, t);
  hipLaunchKernelGGL(k_land_finish, dim3(1), dim3(64), 0, s, e, t);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

