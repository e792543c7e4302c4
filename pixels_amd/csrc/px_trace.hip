// px_trace.hip - ray queries on the brick map for gfx950.
//
//   march()      src/render/shaders/raymarch.wgsl:44-162 (+ safeDir :22-24,
//                glassSurface :26-38): nested DDA super cell (32^3) -> brick
//                (8^3) -> voxel with the uniform-opaque-brick fast path.
//   frame        fsMain :388-398 primary rays + the shadow rays of shade()
//                :323-356.
//   traceSolid   src/render/shaders/sim.wgsl:296-357.
//   segment DDA  src/render/shaders/sim.wgsl:1110-1145.
//   probe        src/render/shaders/probe.wgsl:12-32.
//
// FP32 arithmetic keeps the WGSL operation order (the library is built with
// -ffp-contract=off), so the hit voxel, face, t bits and step counts are the
// ones the reference's arithmetic defines.  Choices WGSL leaves open are
// fixed as: pow(x,5) = ((x*x)*(x*x))*x, mix(a,b,t) = a*(1-t) + b*t,
// dot = (x*x' + y*y') + z*z', length = sqrt(dot), zero t reported as +0.
#include <cstring>
#include <mutex>

#include "px_internal.h"
#include "px_march.h"

#pragma clang fp contract(off)

namespace {

// Register budget of the march kernels: 6 wavefronts per SIMD (<= 80 VGPRs)
// fits the live state of the three nested DDAs without spilling.
#define PX_MARCH_WAVES __attribute__((amdgpu_waves_per_eu(6)))

using px::MarchOut;
using px::MarchTables;
using px::key_of;
using px::load_ray;
using px::read_voxel;
using px::shadow_ray;
using px::step_face;
using px::store_hit;
using px::store_not_cast;
using px::trace_solid;

// Segment DDA of particleSim, sim.wgsl:1110-1145.  A start inside a solid
// reports face 6; a hit reports the solid cell entered, t of the crossing and
// the face of the step.
PX_DEV MarchOut segment(const PxWorldView& w, const float pos[3], const float mv[3]) {
  MarchOut o;
  o.key = PX_NONE;
  o.t = 0.0f;
  o.face = 0xFFu;
  o.value = 0u;
  o.steps = 0u;
  int32_t cell[3];
  for (int a = 0; a < 3; a++) cell[a] = px_f2i(floorf(pos[a]));
  uint32_t v0 = read_voxel(w, cell[0], cell[1], cell[2]);
  if (v0 != 0u) {
    o.key = key_of(w, cell[0], cell[1], cell[2]);
    o.face = 6u;
    o.value = v0;
    return o;
  }
  float dist = sqrtf(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
  if (!(dist > 1e-6f)) return o;
  float dir[3], tn = 0.0f;
  uint32_t v = 0u, steps = 0u;
  const int axis = px::segment_walk(w, cell, pos, mv, dist, dir, &tn, &v, &steps);
  o.steps = steps;
  if (axis >= 0) {
    o.key = key_of(w, cell[0], cell[1], cell[2]);
    o.t = px_canon0(tn);
    const float da = axis == 0 ? dir[0] : (axis == 1 ? dir[1] : dir[2]);
    o.face = step_face(axis, da > 0.0f ? 1 : -1);
    o.value = v;
  }
  return o;
}

template <int KIND, bool SUP_LDS>
__global__ __launch_bounds__(256) PX_MARCH_WAVES void k_trace(PxWorldView w, const px_ray* __restrict__ rays, uint32_t n,
                                               px_hit* __restrict__ out) {
  extern __shared__ uint32_t s_mem[];
  MarchTables T;
  if (KIND == 0 || KIND == 1 || KIND < 0) T = px::load_march_tables<SUP_LDS>(w, s_mem);
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ro[3], rd[3], maxT;
  uint32_t flags;
  load_ray(rays + i, ro, rd, maxT, flags);
  int kind = KIND >= 0 ? KIND : (int)(flags & 3u);
  MarchOut m;
  if (kind == 0) m = px::march<false>(w, T, ro, rd, maxT);
  else if (kind == 1) m = px::march<true>(w, T, ro, rd, maxT);
  else if (kind == 2) m = trace_solid(w, ro, rd, maxT);
  else m = segment(w, ro, rd);
  store_hit(out + i, m);
}

// k_trace over rays taken in the order of perm (results to their own slots)
template <int KIND, bool SUP_LDS>
__global__ __launch_bounds__(256) PX_MARCH_WAVES void k_trace_perm(PxWorldView w, const px_ray* __restrict__ rays,
                                                                    const uint32_t* __restrict__ perm, uint32_t n,
                                                                    px_hit* __restrict__ out) {
  extern __shared__ uint32_t s_mem[];
  MarchTables T;
  if (KIND == 0 || KIND == 1) T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = perm[k];
  float ro[3], rd[3], maxT;
  uint32_t flags;
  load_ray(rays + i, ro, rd, maxT, flags);
  MarchOut m;
  if (KIND == 0) m = px::march<false>(w, T, ro, rd, maxT);
  else if (KIND == 1) m = px::march<true>(w, T, ro, rd, maxT);
  else if (KIND == 2) m = trace_solid(w, ro, rd, maxT);
  else m = segment(w, ro, rd);
  store_hit(out + i, m);
}

// Primary march + the decision of which shadow rays to cast.  Wavefront `gw`
// owns the queue region [gw * 64 * slots, (gw + 1) * 64 * slots): the output
// slot (pixel * slots + s) of each cast ray, packed there slot by slot (lanes
// placed by ballot prefix, no atomics); its count goes to wave_count[gw].
// The shadow kernel rebuilds each ray from the primary ray and hit, so a
// queue entry is 4 bytes instead of a 32-byte ray.
template <bool SUP_LDS>
__global__ __launch_bounds__(256) PX_MARCH_WAVES void k_frame_primary(PxWorldView w, PxShadowDesc sd,
                                                       const px_ray* __restrict__ rays, uint32_t n,
                                                       px_hit* __restrict__ prim_out,
                                                       uint32_t* __restrict__ queue, uint32_t* __restrict__ wave_count) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  float ro[3] = {0, 0, 0}, rd[3] = {0, 0, 1}, maxT = 0.0f;
  uint32_t flags = 0;
  MarchOut h;
  h.face = 0xFFu;
  h.value = 0u;
  h.t = 0.0f;
  if (live) {
    load_ray(rays + i, ro, rd, maxT, flags);
    h = px::march<false>(w, T, ro, rd, maxT);
    store_hit(prim_out + i, h);
  }
  const uint32_t gwave = i >> 6;
  uint32_t* __restrict__ wq = queue + (size_t)gwave * 64u * sd.slots;
  const uint32_t lane = __lane_id();
  uint32_t wcount = 0u;
  for (uint32_t s = 0; s < sd.slots; s++) {
    float O[3], dir[3], mt;
    const bool cast = live && shadow_ray(w, sd, ro, rd, h.t, h.face, h.value, s, O, dir, mt);
    const uint64_t mask = __ballot(cast);
    if (cast) wq[wcount + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = i * sd.slots + s;
    wcount += (uint32_t)__popcll(mask);
  }
  // the last block may hold wavefronts past the last region (no live lanes)
  if (lane == 0u && gwave < (n + 63u) / 64u) wave_count[gwave] = wcount;
}

// One wavefront per queue region of k_frame_primary.  It writes the "not
// cast" record to the outputs of its pixels that no queued ray covers (a
// bitmap of the queue in LDS), then traces the queued rays into theirs.
template <bool SUP_LDS>
__global__ __launch_bounds__(256) PX_MARCH_WAVES void k_trace_shadow_queue(PxWorldView w, PxShadowDesc sd,
                                                            const px_ray* __restrict__ rays,
                                                            const px_hit* __restrict__ prim,
                                                            const uint32_t* __restrict__ queue,
                                                            const uint32_t* __restrict__ wave_count,
                                                            uint32_t n, uint32_t planar,
                                                            px_hit* __restrict__ sh_out) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  // wavefronts take the regions last first: regions follow the image rows,
  // and the lower rows (near ground, whose shadow rays cross the scene) cost
  // the most, so issuing them first shortens the tail of the launch
  // (the wavefront's region and its derived values are wave-uniform: taken
  // through readfirstlane they live in scalar registers, not in VGPRs that
  // would stay live through the march)
  const uint32_t n_regions = (n + 63u) / 64u;
  const uint32_t lw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (lw >= n_regions) return;
  const uint32_t region = n_regions - 1u - lw;
  const uint32_t slots = sd.slots;
  const uint32_t first = region * 64u * slots, last = min(region * 64u + 64u, n) * slots;
  const uint32_t lane = __lane_id();
  const uint32_t cnt = wave_count[region];
  const uint32_t* __restrict__ rq = queue + (size_t)region * 64u * slots;
  // bitmap of the region's cast slots (64 x slots bits, slots <= 33): only
  // the others get the "not cast" record, so every output is written once
  __shared__ uint32_t s_cast[4][(64u * 33u) / 32u];
  uint32_t* bm = s_cast[threadIdx.x >> 6];
  const uint32_t nwords = (64u * slots + 31u) / 32u;
  for (uint32_t j = lane; j < nwords; j += 64u) bm[j] = 0u;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t k = lane; k < cnt; k += 64u) {
    const uint32_t local = rq[k] - first;
    atomicOr(&bm[local >> 5], 1u << (local & 31u));
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (planar) {
    // slot-planar output (slot s of pixel i at s * n + i): lanes walk the
    // region slot by slot, so each wavefront store covers consecutive pixels
    // (a region has at most 64 pixels: lane j owns pixel j)
    const uint32_t p0 = region * 64u, npix = min(64u, n - p0);
    for (uint32_t s = 0; s < slots; s++) {
      const uint32_t local = lane * slots + s;
      if (lane < npix && ((bm[local >> 5] >> (local & 31u)) & 1u) == 0u)
        store_not_cast(sh_out + (size_t)s * n + p0 + lane);
    }
  } else {
    for (uint32_t k = first + lane; k < last; k += 64u) {
      const uint32_t local = k - first;
      if (((bm[local >> 5] >> (local & 31u)) & 1u) == 0u) store_not_cast(sh_out + k);
    }
  }
  // (a wave-uniform chunk base, so no per-lane counter stays live through
  // the march: the lane's entry is kb + lane)
  for (uint32_t kb = 0; kb < cnt; kb += 64u) {
    const uint32_t k = kb + lane;
    if (k >= cnt) break;
    const uint32_t code = rq[k];
    const uint32_t i = code / slots, s = code - i * slots;
    // output slot as one 32-bit index (codes, i * slots + s, are u32 too),
    // the only value kept through the march
    uint32_t oidx = planar ? s * n + i : code;
    PX_OPAQUE(oidx);
    float ro[3], rd[3], maxT;
    uint32_t flags;
    load_ray(rays + i, ro, rd, maxT, flags);
    const uint4 hv = *reinterpret_cast<const uint4*>(prim + i);
    const uint32_t face = hv.w & 0xffu, value = (hv.w >> 8) & 0xffffu;
    float O[3], dir[3], mt;
    shadow_ray(w, sd, ro, rd, __uint_as_float(hv.y), face, value, s, O, dir, mt);
    MarchOut m = px::march<true>(w, T, O, dir, mt);
    store_hit(sh_out + oidx, m);
  }
}

// ---- batched queries ---------------------------------------------------
// One persistent launch over queries of all four kinds.  The chunk space is
// the kinds laid end to end in `order` (longest-running kinds first, so the
// short ones fill the tail), cut into chunks of `chunk` rays that never
// straddle two kinds.  Chunk c belongs to claim group c % kBatchGroups; a
// workgroup claims from the group of its XCD (workgroups are dealt to the 8
// XCDs round-robin), then from the others once that group is drained, so no
// single counter takes every claim.
constexpr uint32_t kBatchGroups = 8;
constexpr uint32_t kCursorStride = 64;  // words between cursors (256 bytes)

struct BatchArgs {
  const px_ray* rays[4];
  px_hit* out[4];
  const uint32_t* perm[4];  // per kind, NULL = index order
  uint32_t n[4];
  uint32_t first_chunk[5];  // chunk index where each kind of `order` starts
  uint32_t order[4];        // processing order -> kind
  uint32_t chunk;
  uint32_t total_chunks;
};

template <bool SUP_LDS>
__global__ __launch_bounds__(256) PX_MARCH_WAVES void k_trace_batch(PxWorldView w, BatchArgs a,
                                                                     uint32_t* __restrict__ cursors) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t lane = __lane_id();
  const uint32_t home = blockIdx.x % kBatchGroups;
  for (uint32_t gi = 0; gi < kBatchGroups; gi++) {
    const uint32_t g = (home + gi) % kBatchGroups;
    for (;;) {
      uint32_t j = 0u;
      if (lane == 0u) j = atomicAdd(cursors + g * kCursorStride, 1u);
      // (readfirstlane: the chunk and everything derived from it - kind,
      // range, array pointers - are wave-uniform and live in scalar
      // registers instead of VGPRs kept through the traversal)
      j = __builtin_amdgcn_readfirstlane(j);
      const uint32_t c = j * kBatchGroups + g;
      if (c >= a.total_chunks) break;
      uint32_t pos = 0u;
      while (pos < 3u && c >= a.first_chunk[pos + 1u]) pos++;
      const uint32_t kind = a.order[pos];
      const uint32_t start = (c - a.first_chunk[pos]) * a.chunk;
      const uint32_t end = min(start + a.chunk, a.n[kind]);
      const px_ray* __restrict__ rays = a.rays[kind];
      px_hit* __restrict__ out = a.out[kind];
      const uint32_t* __restrict__ perm = a.perm[kind];
      for (uint32_t kb = start; kb < end; kb += 64u) {
        const uint32_t k = kb + lane;
        if (k >= end) break;
        // position k of the kind's order traces ray perm[k] (k without one)
        const uint32_t i = perm ? perm[k] : k;
        float ro[3], rd[3], maxT;
        uint32_t flags;
        load_ray(rays + i, ro, rd, maxT, flags);
        MarchOut m;
        if (kind == PX_RAY_PRIMARY) m = px::march<false>(w, T, ro, rd, maxT);
        else if (kind == PX_RAY_SHADOW) m = px::march<true>(w, T, ro, rd, maxT);
        else if (kind == PX_RAY_SHOOT) m = trace_solid(w, ro, rd, maxT);
        else m = segment(w, ro, rd);
        store_hit(out + i, m);
      }
    }
  }
}

// Sum of the per-wavefront shadow-ray counts (one block).
__global__ __launch_bounds__(1024) void k_sum_counts(const uint32_t* __restrict__ cnt, uint32_t n,
                                                     uint32_t* __restrict__ out) {
  __shared__ uint32_t part[16];
  uint32_t s = 0u;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += cnt[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if (__lane_id() == 0u) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0u) {
    uint32_t t = 0u;
    for (uint32_t k = 0; k < blockDim.x / 64u; k++) t += part[k];
    *out = t;
  }
}

__global__ void k_probe(PxWorldView w, int32_t ox, int32_t oy, int32_t oz, int32_t sx, int32_t sy, int32_t sz,
                        uint32_t* __restrict__ out) {
  size_t n = (size_t)sx * sy * sz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int32_t x = ox + (int32_t)(i % (size_t)sx);
    int32_t y = oy + (int32_t)((i / (size_t)sx) % (size_t)sy);
    int32_t z = oz + (int32_t)(i / ((size_t)sx * sy));
    out[i] = read_voxel(w, x, y, z);
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// one bit per super cell (count != 0) and per brick cell (entry != 0):
// lane l of wave w reads cell 64 w + l (coalesced), and the wave's ballot is
// two mask words; the super cells' words follow the cells' waves
__global__ void k_masks(const uint32_t* __restrict__ grid, uint32_t ncell, const uint32_t* __restrict__ sgrid,
                        uint32_t nsuper, uint32_t* __restrict__ bmask, uint32_t* __restrict__ smask) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t cw = (ncell + 63u) / 64u;  // waves over the cells
  const bool cells = wave < cw;
  const uint32_t base = (cells ? wave : wave - cw) * 64u;
  const uint32_t n = cells ? ncell : nsuper;
  if (base >= n) return;
  const uint32_t i = base + lane;
  const uint64_t m = __ballot(i < n && (cells ? grid[i] : sgrid[i]) != 0u);
  uint32_t* out = cells ? bmask : smask;
  const uint32_t words = (n + 31u) / 32u;
  if (lane < 2u && base / 32u + lane < words) out[base / 32u + lane] = (uint32_t)(m >> (32u * lane));
}

// Per-super-cell 64-bit brick occupancy (bit = lx + 4 (ly + 4 lz)): one
// wave per super cell, lane = the brick's bit.
__global__ void k_super_bm(const uint32_t* __restrict__ grid, int32_t bx, int32_t by, int32_t bz, int32_t sx,
                           int32_t sy, int32_t sz, uint64_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t si = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t ns = (uint32_t)(sx * sy * sz);
  if (si >= ns) return;
  const int32_t s0 = (int32_t)(si % (uint32_t)sx), s1 = (int32_t)((si / (uint32_t)sx) % (uint32_t)sy);
  const int32_t s2 = (int32_t)(si / ((uint32_t)sx * (uint32_t)sy));
  const int32_t x = s0 * 4 + (int32_t)(lane & 3u), y = s1 * 4 + (int32_t)((lane >> 2) & 3u);
  const int32_t z = s2 * 4 + (int32_t)(lane >> 4);
  const bool in = x < bx && y < by && z < bz;
  const uint64_t m = __ballot(in && grid[(uint32_t)(x + bx * (y + by * z))] != 0u);
  if (lane == 0u) out[si] = m;
}

// 2x2x2-block occupancy of a pool brick: bit x2 + 4 y2 + 16 z2 is set when
// the block of voxels (2 x2 .. 2 x2 + 1, 2 y2 .., 2 z2 ..) has a non-zero
// voxel.  w is row l = lane of the brick (y = l & 7, z = l >> 3; 4 words of 2
// voxels); one ballot per word gives the rows in which that x pair has a
// voxel, and OR-ing neighbouring rows in y and z gives the blocks.  The mask
// is stored by grid cell; cells without a pool brick are skipped (march()
// reads the mask of pool-brick cells only).
PX_DEV uint64_t sub_bits(uint4 w) {
  // word k of row (y, z) holds voxels x = 2k, 2k + 1: sub-block column x2 = k
  const uint64_t b[4] = {__ballot(w.x != 0u), __ballot(w.y != 0u), __ballot(w.z != 0u), __ballot(w.w != 0u)};
  uint64_t m = 0ull;
  for (uint32_t x2 = 0; x2 < 4u; x2++) {
    uint64_t t = b[x2] | (b[x2] >> 1);  // rows y, y + 1 (bit at even y)
    t |= t >> 8;                         // rows z, z + 1 (bit at even z)
    // compact the bits at 2 y2 + 16 z2 into bit y2 + 4 z2 ...
    t &= 0x0055005500550055ull;
    t = (t | (t >> 1)) & 0x0033003300330033ull;
    t = (t | (t >> 2)) & 0x000F000F000F000Full;
    t = (t | (t >> 12)) & 0x000000FF000000FFull;
    t = (t | (t >> 24)) & 0xFFFFull;
    // ... then spread bit y2 + 4 z2 to bit 4 (y2 + 4 z2) and add x2
    t = (t | (t << 24)) & 0x000000FF000000FFull;
    t = (t | (t << 12)) & 0x000F000F000F000Full;
    t = (t | (t << 6)) & 0x0303030303030303ull;
    t = (t | (t << 3)) & 0x1111111111111111ull;
    m |= t << x2;
  }
  return m;
}

// Each wavefront owns 64 consecutive cells: one coalesced read of their grid
// words, then their pool bricks four at a time (loads first, so they
// overlap).  Lane l reads row l of a brick (y = l & 7, z = l >> 3; 4 words).
__global__ __launch_bounds__(256) void k_sub_mask(const uint32_t* __restrict__ grid, const uint32_t* __restrict__ pool,
                                                  uint32_t ncell, uint64_t* __restrict__ subs) {
  const uint32_t lane = __lane_id();
  const uint32_t c0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
  if (c0 >= ncell) return;
  const uint32_t gl = c0 + lane < ncell ? grid[c0 + lane] : 0u;
  uint64_t todo = __ballot(gl != 0u && (gl & PX_UNIFORM_BIT) == 0u);
  const uint4* p4 = reinterpret_cast<const uint4*>(pool);
  while (todo) {
    uint32_t k[4];
    uint4 w[4];
    for (int j = 0; j < 4; j++) {
      k[j] = todo ? (uint32_t)(__ffsll((unsigned long long)todo) - 1) : 64u;
      todo &= todo - 1ull;
    }
    for (int j = 0; j < 4; j++)
      if (k[j] < 64u) w[j] = p4[(size_t)(__shfl(gl, (int)k[j]) - 1u) * 64u + lane];
    for (int j = 0; j < 4; j++)
      if (k[j] < 64u) {
        const uint64_t m = sub_bits(w[j]);
        if (lane == 0u) subs[c0 + k[j]] = m;
      }
  }
}

}  // namespace

PxWorldView px_view(const px_ctx* c) {
  PxWorldView w;
  w.grid = c->d_grid;
  w.pool = c->d_pool;
  w.super_grid = c->d_super;
  w.mats = c->d_mats;
  w.super_mask = c->d_super_mask;
  w.brick_mask = c->d_brick_mask;
  w.super_bm = c->d_super_bm;
  w.sub_mask = c->d_sub_mask;
  for (int a = 0; a < 3; a++) {
    w.dims[a] = (int32_t)c->dims[a];
    w.bd[a] = (int32_t)c->bd[a];
    w.sd[a] = (int32_t)c->sd[a];
    w.fdims[a] = (float)c->dims[a];
  }
  return w;
}

int px_launch_masks(px_ctx* c) {
  const uint32_t waves = (c->ncell + 63u) / 64u + (c->nsuper + 63u) / 64u;
  hipLaunchKernelGGL(k_masks, dim3((waves + 3u) / 4u), dim3(256), 0, c->stream, c->d_grid, c->ncell, c->d_super,
                     c->nsuper, c->d_brick_mask, c->d_super_mask);
  PX_HIP_CHECK(c, hipGetLastError());
  const uint32_t ns = c->nsuper;
  hipLaunchKernelGGL(k_super_bm, dim3((ns + 3u) / 4u), dim3(256), 0, c->stream, c->d_grid, (int32_t)c->bd[0],
                     (int32_t)c->bd[1], (int32_t)c->bd[2], (int32_t)c->sd[0], (int32_t)c->sd[1], (int32_t)c->sd[2],
                     c->d_super_bm);
  PX_HIP_CHECK(c, hipGetLastError());
  c->masks_dirty = false;
  c->mask_gen++;
  return PX_OK;
}

int px_refresh_masks(px_ctx* c) { return c->masks_dirty ? px_launch_masks(c) : PX_OK; }

int px_refresh_march(px_ctx* c) {
  int rc = px_refresh_masks(c);
  if (rc) return rc;
  if (c->sub_mask_gen == c->mask_gen) return PX_OK;
  if (c->ncell) {
    hipLaunchKernelGGL(k_sub_mask, dim3((c->ncell + 255u) / 256u), dim3(256), 0, c->stream, c->d_grid, c->d_pool,
                       c->ncell, c->d_sub_mask);
    PX_HIP_CHECK(c, hipGetLastError());
  }
  c->sub_mask_gen = c->mask_gen;
  return PX_OK;
}


template <int KIND>
static void launch_trace_kind(px_ctx* c, const PxWorldView& w, const px_ray* d_rays, uint32_t n, px_hit* d_out) {
  bool sup_lds;
  size_t lds = px_table_lds(c, &sup_lds);
  dim3 grid((n + 255u) / 256u), block(256);
  if (sup_lds) hipLaunchKernelGGL((k_trace<KIND, true>), grid, block, lds, c->stream, w, d_rays, n, d_out);
  else hipLaunchKernelGGL((k_trace<KIND, false>), grid, block, lds, c->stream, w, d_rays, n, d_out);
}

int px_launch_trace(px_ctx* c, uint32_t kind, const px_ray* d_rays, uint32_t n, px_hit* d_out) {
  if (n == 0) return PX_OK;
  int rc = px_refresh_march(c);
  if (rc) return rc;
  PxWorldView w = px_view(c);
  PxProfScope prof(c, PX_PROF_TRACE);
  switch (kind) {
    case PX_RAY_PRIMARY: launch_trace_kind<0>(c, w, d_rays, n, d_out); break;
    case PX_RAY_SHADOW: launch_trace_kind<1>(c, w, d_rays, n, d_out); break;
    case PX_RAY_SHOOT: launch_trace_kind<2>(c, w, d_rays, n, d_out); break;
    case PX_RAY_SEGMENT: launch_trace_kind<3>(c, w, d_rays, n, d_out); break;
    default: launch_trace_kind<-1>(c, w, d_rays, n, d_out); break;
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

template <int KIND>
static void launch_perm_kind(px_ctx* c, const PxWorldView& w, const px_ray* d_rays, const uint32_t* d_perm,
                             uint32_t n, px_hit* d_out) {
  bool sup_lds;
  size_t lds = px_table_lds(c, &sup_lds);
  dim3 grid((n + 255u) / 256u), block(256);
  if (sup_lds) hipLaunchKernelGGL((k_trace_perm<KIND, true>), grid, block, lds, c->stream, w, d_rays, d_perm, n, d_out);
  else hipLaunchKernelGGL((k_trace_perm<KIND, false>), grid, block, lds, c->stream, w, d_rays, d_perm, n, d_out);
}

int px_launch_trace_perm(px_ctx* c, uint32_t kind, const px_ray* d_rays, const uint32_t* d_perm, uint32_t n,
                         px_hit* d_out) {
  if (n == 0) return PX_OK;
  int rc = px_refresh_march(c);
  if (rc) return rc;
  PxWorldView w = px_view(c);
  PxProfScope prof(c, PX_PROF_TRACE);
  switch (kind) {
    case PX_RAY_PRIMARY: launch_perm_kind<0>(c, w, d_rays, d_perm, n, d_out); break;
    case PX_RAY_SHADOW: launch_perm_kind<1>(c, w, d_rays, d_perm, n, d_out); break;
    case PX_RAY_SHOOT: launch_perm_kind<2>(c, w, d_rays, d_perm, n, d_out); break;
    default: launch_perm_kind<3>(c, w, d_rays, d_perm, n, d_out); break;
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

// The frame's two launches over a caller-given queue (n_regions * 64 *
// slots words), per-region counts (n_regions words) and one word for the
// cast total.
static int launch_frame_on(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                           uint32_t* n_cast, uint32_t* queue, uint32_t* wave_count, uint32_t* total) {
  const uint32_t n_regions = (n + 63u) / 64u;
  PxWorldView w = px_view(c);
  bool sup_lds;
  const size_t lds = px_table_lds(c, &sup_lds);
  {
    PxProfScope prof(c, PX_PROF_PRIMARY);
    dim3 grid((n + 255u) / 256u), block(256);
    if (sup_lds)
      hipLaunchKernelGGL(k_frame_primary<true>, grid, block, lds, c->stream, w, c->shadow, d_primary, n, d_prim,
                         queue, wave_count);
    else
      hipLaunchKernelGGL(k_frame_primary<false>, grid, block, lds, c->stream, w, c->shadow, d_primary, n, d_prim,
                         queue, wave_count);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  const uint32_t blocks = (n_regions + 3u) / 4u;
  const uint32_t planar = c->frame_layout == PX_FRAME_SLOT_PLANAR ? 1u : 0u;
  {
    PxProfScope prof(c, PX_PROF_SHADOW);
    if (sup_lds)
      hipLaunchKernelGGL(k_trace_shadow_queue<true>, dim3(blocks), dim3(256), lds, c->stream, w, c->shadow,
                         d_primary, d_prim, queue, wave_count, n, planar, d_shadow);
    else
      hipLaunchKernelGGL(k_trace_shadow_queue<false>, dim3(blocks), dim3(256), lds, c->stream, w, c->shadow,
                         d_primary, d_prim, queue, wave_count, n, planar, d_shadow);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  if (n_cast) {
    hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(1024), 0, c->stream, wave_count, n_regions, total);
    PX_HIP_CHECK(c, hipGetLastError());
    PX_HIP_CHECK(c, hipMemcpyAsync(n_cast, total, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return PX_OK;
}

int px_launch_frame(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                    uint32_t* n_cast) {
  if (n == 0) {
    if (n_cast) *n_cast = 0;
    return PX_OK;
  }
  int rc = px_refresh_march(c);
  if (rc) return rc;
  // shadow-ray queue: one region of 64 * slots entries per wavefront of primaries
  const uint32_t n_regions = (n + 63u) / 64u;
  const size_t need = (size_t)n_regions * 64u * c->shadow.slots;
  if (need > 0xFFFFFFFFull) return px_fail(c, PX_ERR_CAPACITY, "px_trace_frame_device", "too many shadow slots");
  if (need > c->shadow_cap) {
    if (c->d_shadow_queue) (void)hipFree(c->d_shadow_queue);
    c->d_shadow_queue = nullptr;
    c->shadow_cap = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_shadow_queue, need * sizeof(uint32_t)));
    c->shadow_cap = need;
  }
  if (n_regions > c->wave_count_cap) {
    if (c->d_wave_count) hipFree(c->d_wave_count);
    c->d_wave_count = nullptr;
    c->wave_count_cap = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_wave_count, (size_t)n_regions * sizeof(uint32_t)));
    c->wave_count_cap = n_regions;
  }
  return launch_frame_on(c, d_primary, n, d_prim, d_shadow, n_cast, c->d_shadow_queue, c->d_wave_count,
                         c->d_counters);
}

// ---- the frame without allocation -----------------------------------------
// Fixed scratch of px_trace_frame_device_noalloc: the cast total (word 0),
// the per-region counts (from word 64) and the shadow-ray queue, in one
// array that every call on a device shares; each device holds its own copy
// for the life of the process.  Host code reaches it only through the
// address hipGetSymbolAddress gives on that device.
constexpr size_t kFixedScratchWords = (size_t)20 << 20;  // 80 MiB
static __device__ __align__(256) uint32_t scratch[kFixedScratchWords];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kMaxScratchDevices] = {};

int px_fixed_scratch_init(px_ctx* c) {
  if (c->device < 0 || c->device >= kMaxScratchDevices)
    return px_fail(c, PX_ERR_ARG, "fixed scratch", "device ordinal beyond the supported range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return px_fail(c, PX_ERR_HIP, "hipGetSymbolAddress(scratch)", hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<uint32_t*>(p);
  }
  c->d_fixed_scratch = g_scratch_ptr[c->device];
  return PX_OK;
}

int px_launch_frame_noalloc(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                            uint32_t* n_cast) {
  if (n == 0) {
    if (n_cast) *n_cast = 0;
    return PX_OK;
  }
  if (!c->d_fixed_scratch)
    return px_fail(c, PX_ERR_STATE, "px_trace_frame_device_noalloc", "fixed scratch not set up");
  const uint32_t n_regions = (n + 63u) / 64u;
  const size_t count_words = ((size_t)n_regions + 63u) & ~(size_t)63u;
  const size_t need = 64u + count_words + (size_t)n_regions * 64u * c->shadow.slots;
  if (need > kFixedScratchWords)
    return px_fail(c, PX_ERR_CAPACITY, "px_trace_frame_device_noalloc",
                   "64 * ceil(n / 64) * slots + ceil(n / 64) exceeds the fixed scratch of 20,971,520 words");
  int rc = px_refresh_march(c);
  if (rc) return rc;
  uint32_t* base = c->d_fixed_scratch;
  return launch_frame_on(c, d_primary, n, d_prim, d_shadow, n_cast, base + 64 + count_words, base + 64, base);
}

// Kinds in processing order: shoot rays (long, few), primary, shadow, then
// the short debris segments.
static const uint32_t kBatchOrder[4] = {PX_RAY_SHOOT, PX_RAY_PRIMARY, PX_RAY_SHADOW, PX_RAY_SEGMENT};

static BatchArgs batch_args(const px_ctx* c, const px_query_set* q, const uint32_t* const* perm) {
  BatchArgs a;
  memset(&a, 0, sizeof(a));
  a.chunk = c->batch_chunk;
  uint32_t acc = 0u;
  for (int p = 0; p < 4; p++) {
    const uint32_t k = kBatchOrder[p];
    a.order[p] = k;
    a.first_chunk[p] = acc;
    acc += (q->n[k] + a.chunk - 1u) / a.chunk;
  }
  a.first_chunk[4] = acc;
  a.total_chunks = acc;
  for (int k = 0; k < 4; k++) {
    a.rays[k] = q->rays[k];
    a.out[k] = q->out[k];
    a.perm[k] = perm ? perm[k] : nullptr;
    a.n[k] = q->n[k];
  }
  return a;
}

static uint32_t batch_blocks(const px_ctx* c) {
  // default: six wavefronts per SIMD (the march kernels' register budget)
  return c->batch_blocks ? c->batch_blocks : c->n_cu * 6u * 4u * 64u / c->batch_block;
}

// Captures [cursor reset, k_trace_batch] into a graph for this query set and
// shape (kept while they stay the same) and launches it.
int px_launch_batch(px_ctx* c, const px_query_set* q, const uint32_t* const* perm) {
  int rc = px_refresh_march(c);
  if (rc) return rc;
  const BatchArgs a = batch_args(c, q, perm);
  if (a.total_chunks == 0u) return PX_OK;
  unsigned char key[sizeof(c->batch_key)];
  static_assert(sizeof(BatchArgs) + 3 * sizeof(uint32_t) <= sizeof(key), "batch key size");
  memset(key, 0, sizeof(key));
  memcpy(key, &a, sizeof(a));
  const uint32_t shape[3] = {c->batch_block, batch_blocks(c), c->batch_chunk};
  memcpy(key + sizeof(a), shape, sizeof(shape));
  if (!c->batch_exec || memcmp(key, c->batch_key, sizeof(key)) != 0) {
    if (c->batch_exec) {
      hipGraphExecDestroy(c->batch_exec);
      c->batch_exec = nullptr;
    }
    PxWorldView w = px_view(c);
    bool sup_lds;
    const size_t lds = px_table_lds(c, &sup_lds);
    hipGraph_t graph = nullptr;
    PX_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    // a fill kernel, not hipMemsetAsync, inside the graph
    hipLaunchKernelGGL(k_fill_u32, dim3(8), dim3(64), 0, c->stream, c->d_batch_cursors,
                       kBatchGroups * kCursorStride, 0u);
    if (sup_lds)
      hipLaunchKernelGGL(k_trace_batch<true>, dim3(shape[1]), dim3(shape[0]), lds, c->stream, w, a,
                         c->d_batch_cursors);
    else
      hipLaunchKernelGGL(k_trace_batch<false>, dim3(shape[1]), dim3(shape[0]), lds, c->stream, w, a,
                         c->d_batch_cursors);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      return px_fail(c, PX_ERR_HIP, "px_trace_batch", hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipError_t ie = hipGraphInstantiate(&c->batch_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      c->batch_exec = nullptr;
      return px_fail(c, PX_ERR_HIP, "px_trace_batch", hipGetErrorString(ie));
    }
    memcpy(c->batch_key, key, sizeof(key));
  }
  PxProfScope prof(c, PX_PROF_BATCH);
  PX_HIP_CHECK(c, hipGraphLaunch(c->batch_exec, c->stream));
  return PX_OK;
}

int px_autotune_batch_impl(px_ctx* c, const px_query_set* q, const uint32_t* const* perm, uint32_t repeats,
                           uint32_t out[3], float* best_ms) {
  static const uint32_t blocks_opt[3] = {64u, 128u, 256u};
  static const uint32_t waves_opt[3] = {8u, 16u, 24u};  // wavefronts per CU
  static const uint32_t chunk_opt[3] = {64u, 256u, 1024u};
  const uint32_t keep[3] = {c->batch_block, c->batch_blocks, c->batch_chunk};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PX_HIP_CHECK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return px_fail(c, PX_ERR_HIP, "px_autotune_batch", "hipEventCreate");
  }
  float best = 3.0e38f;
  uint32_t bs[3] = {keep[0], keep[1], keep[2]};
  int rc = PX_OK;
  for (uint32_t bi = 0; bi < 3 && rc == PX_OK; bi++)
    for (uint32_t wi = 0; wi < 3 && rc == PX_OK; wi++)
      for (uint32_t ci = 0; ci < 3 && rc == PX_OK; ci++) {
        c->batch_block = blocks_opt[bi];
        c->batch_blocks = c->n_cu * waves_opt[wi] * 64u / blocks_opt[bi];
        c->batch_chunk = chunk_opt[ci];
        rc = px_launch_batch(c, q, perm);  // capture + warm-up
        if (rc) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) rc = px_fail(c, PX_ERR_HIP, "px_autotune_batch", "event");
        for (uint32_t r = 0; r < repeats && rc == PX_OK; r++) rc = px_launch_batch(c, q, perm);
        if (rc) break;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = px_fail(c, PX_ERR_HIP, "px_autotune_batch", "event");
          break;
        }
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)repeats;
        if (ms < best) {
          best = ms;
          bs[0] = c->batch_block;
          bs[1] = c->batch_blocks;
          bs[2] = c->batch_chunk;
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) {
    c->batch_block = keep[0];
    c->batch_blocks = keep[1];
    c->batch_chunk = keep[2];
    return rc;
  }
  c->batch_block = bs[0];
  c->batch_blocks = bs[1];
  c->batch_chunk = bs[2];
  out[0] = bs[0];
  out[1] = bs[1];
  out[2] = bs[2];
  if (best_ms) *best_ms = best;
  return PX_OK;
}

int px_launch_probe(px_ctx* c, const int32_t o[3], const int32_t s[3], uint32_t* d_out) {
  size_t n = (size_t)s[0] * s[1] * s[2];
  if (n == 0) return PX_OK;
  PxWorldView w = px_view(c);
  size_t blocks = (n + 255u) / 256u;
  if (blocks > 65536u) blocks = 65536u;
  hipLaunchKernelGGL(k_probe, dim3((uint32_t)blocks), dim3(256), 0, c->stream, w, o[0], o[1], o[2], s[0], s[1], s[2],
                     d_out);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}
