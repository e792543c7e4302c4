// px_scan.h - ordered exclusive scans of (x, y) counts for gfx950.
//
// The extraction, chunk and scene-build stages need list positions in a
// fixed order (cells, voxels, arena entries) without global atomics: the
// per-element (x, y) counts are scanned in tiles of 1024 (one workgroup
// each), the tile sums in one workgroup, then applied.  Lengths may live on
// the device (n_dev), so grids are sized by capacity and tiles past the
// length exit.  Kernels are in an anonymous namespace: every translation
// unit that includes this header gets its own copy.
#pragma once

#include "px_internal.h"

namespace {

constexpr uint32_t kScanTile = 1024;

// inclusive prefix sum over the wavefront
PX_DEV uint32_t scan_wave_incl(uint32_t v) {
  const uint32_t lane = __lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// Exclusive prefix of (x, y) over the workgroup (up to 1024 threads);
// returns the block total in *tot.  Block-cooperative.
PX_DEV uint2 block_excl2(uint2 v, uint2* tot) {
  __shared__ uint32_t s_x[16], s_y[16];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t ix = scan_wave_incl(v.x), iy = scan_wave_incl(v.y);
  if (lane == 63u) {
    s_x[wave] = ix;
    s_y[wave] = iy;
  }
  __syncthreads();
  uint32_t px = 0u, py = 0u, tx = 0u, ty = 0u;
  const uint32_t nw = blockDim.x >> 6;
  for (uint32_t k = 0; k < nw; k++) {
    px += k < wave ? s_x[k] : 0u;
    py += k < wave ? s_y[k] : 0u;
    tx += s_x[k];
    ty += s_y[k];
  }
  __syncthreads();
  *tot = make_uint2(tx, ty);
  return make_uint2(px + ix - v.x, py + iy - v.y);
}

PX_DEV uint32_t scan_len(const uint32_t* n_dev, uint32_t n_fixed) { return n_dev ? *n_dev : n_fixed; }

__global__ __launch_bounds__(1024) void k_scan2_tiles(const uint2* __restrict__ in, const uint32_t* n_dev,
                                                      uint32_t n_fixed, uint2* __restrict__ tiles) {
  const uint32_t n = scan_len(n_dev, n_fixed);
  const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
  if (blockIdx.x * kScanTile >= n) return;
  const uint2 v = i < n ? in[i] : make_uint2(0u, 0u);
  uint2 tot;
  block_excl2(v, &tot);
  if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// tiles -> exclusive offsets (+ base_x / base_y), totals to tot[0..1]
__global__ __launch_bounds__(1024) void k_scan2_top(uint2* __restrict__ tiles, const uint32_t* n_dev, uint32_t n_fixed,
                                                    const uint32_t* base_x, const uint32_t* base_y,
                                                    uint32_t* __restrict__ tot) {
  const uint32_t n = scan_len(n_dev, n_fixed);
  const uint32_t nt = (n + kScanTile - 1u) / kScanTile;
  uint2 carry = make_uint2(base_x ? *base_x : 0u, base_y ? *base_y : 0u);
  const uint2 base = carry;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < nt; t0 += blockDim.x) {
    const uint32_t t = t0 + threadIdx.x;
    const uint2 v = t < nt ? tiles[t] : make_uint2(0u, 0u);
    uint2 bt;
    const uint2 ex = block_excl2(v, &bt);
    if (t < nt) tiles[t] = make_uint2(carry.x + ex.x, carry.y + ex.y);
    carry.x += bt.x;
    carry.y += bt.y;
  }
  if (threadIdx.x == 0) {
    tot[0] = carry.x - base.x;
    tot[1] = carry.y - base.y;
  }
}

__global__ __launch_bounds__(1024) void k_scan2_apply(const uint2* __restrict__ in, const uint32_t* n_dev,
                                                      uint32_t n_fixed, const uint2* __restrict__ tiles,
                                                      uint2* __restrict__ out) {
  const uint32_t n = scan_len(n_dev, n_fixed);
  const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
  if (blockIdx.x * kScanTile >= n) return;
  const uint2 v = i < n ? in[i] : make_uint2(0u, 0u);
  uint2 tot;
  const uint2 ex = block_excl2(v, &tot);
  const uint2 t = tiles[blockIdx.x];
  if (i < n) out[i] = make_uint2(t.x + ex.x, t.y + ex.y);
}

// k_scan2_apply for at most kScanTile tiles, without k_scan2_top: block b
// adds the totals of tiles 0 .. b-1 itself (one per thread), and the last
// tile's block writes the totals.
__global__ __launch_bounds__(1024) void k_scan2_apply_small(const uint2* __restrict__ in, uint32_t n,
                                                            const uint2* __restrict__ tiles,
                                                            const uint32_t* base_x, const uint32_t* base_y,
                                                            uint2* __restrict__ out, uint32_t* __restrict__ tot) {
  const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
  const uint2 tb = threadIdx.x < blockIdx.x ? tiles[threadIdx.x] : make_uint2(0u, 0u);
  uint2 before;
  block_excl2(tb, &before);
  const uint2 v = i < n ? in[i] : make_uint2(0u, 0u);
  uint2 own;
  const uint2 ex = block_excl2(v, &own);
  const uint32_t bx = base_x ? *base_x : 0u, by = base_y ? *base_y : 0u;
  if (i < n) out[i] = make_uint2(bx + before.x + ex.x, by + before.y + ex.y);
  if (blockIdx.x == gridDim.x - 1u && threadIdx.x == 0u) {
    tot[0] = before.x + own.x;
    tot[1] = before.y + own.y;
  }
}

// Exclusive (x, y) offsets of in[0..n) into out (n = *n_dev, or n_fixed when
// n_dev is null; cap bounds n), totals into tot[0..1]; tiles holds
// ceil(cap / 1024) entries.
inline void scan2_launch(hipStream_t s, const uint2* in, uint2* out, uint2* tiles, size_t cap, const uint32_t* n_dev,
                         uint32_t n_fixed, const uint32_t* base_x, const uint32_t* base_y, uint32_t* tot) {
  if (!n_dev && n_fixed > 0u && n_fixed <= (size_t)kScanTile * kScanTile && n_fixed <= cap) {
    // a length known here, in at most kScanTile tiles: two launches
    const uint32_t nt = (n_fixed + kScanTile - 1u) / kScanTile;
    hipLaunchKernelGGL(k_scan2_tiles, dim3(nt), dim3(kScanTile), 0, s, in, nullptr, n_fixed, tiles);
    hipLaunchKernelGGL(k_scan2_apply_small, dim3(nt), dim3(kScanTile), 0, s, in, n_fixed, tiles, base_x, base_y,
                       out, tot);
    return;
  }
  const uint32_t ntiles = (uint32_t)((cap + kScanTile - 1u) / kScanTile);
  hipLaunchKernelGGL(k_scan2_tiles, dim3(ntiles), dim3(kScanTile), 0, s, in, n_dev, n_fixed, tiles);
  hipLaunchKernelGGL(k_scan2_top, dim3(1), dim3(1024), 0, s, tiles, n_dev, n_fixed, base_x, base_y, tot);
  hipLaunchKernelGGL(k_scan2_apply, dim3(ntiles), dim3(kScanTile), 0, s, in, n_dev, n_fixed, tiles, out);
}

}  // namespace
