// px_scene.hip - the scene rasteriser on the GPU (gfx950): apply() of
// src/world/shapes.ts:334-414 (Op.Add, constant paint) for a list of SDF
// primitives, straight into the device brick map of a new context.
//
// The reference rasterises one primitive after another over the world.  A
// brick's result only depends on the primitives that reach it, in order, so
// here every brick replays its own primitives, all bricks at once:
//   k_prim_pairs<0/1>  one workgroup per primitive: count, then emit the
//                      (brick, primitive) pairs its brick-centre test keeps;
//   radix sort         pairs by (brick, primitive): each brick's primitives
//                      in order (stable LSD radix sort, px_sort.hip);
//   k_raster           one wavefront per brick (launched per pair, the first
//                      pair of each brick does the work), the 512 voxels in registers
//                      (lane = row y | z<<3, 8 voxels in x): the brick test,
//                      the 8 sub-block tests, the voxel tests, compactCell
//                      (VoxelWorld.ts:147-156) after every change;
//   k_pool_flags +     pool bricks numbered in cell order by an ordered scan
//   k_fill_world       (as the CPU scene builder numbers them), grid entries,
//                      brick copies, and the
//                      super-grid counts (GpuWorld.ts:90-100).
// SDFs and tests are in double precision, as the reference's JavaScript
// numbers, with the oracle's operation order (-ffp-contract=off): the world
// is bit-identical to the CPU restatement.
#include <algorithm>
#include <vector>

#include "px_internal.h"
#include "px_scan.h"

#pragma clang fp contract(off)

namespace {

PX_DEV double dmax(double a, double b) { return a < b ? b : a; }
PX_DEV double dmin(double a, double b) { return b < a ? b : a; }
PX_DEV double dlen2(double a, double b) { return sqrt(a * a + b * b); }
PX_DEV double dlen3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

PX_DEV void prim_local(const px_prim& P, double x, double y, double z, double& lx, double& ly, double& lz) {
  const double dx = x - P.p[0], dy = y - P.p[1], dz = z - P.p[2];
  lx = P.cy * dx + P.sy * dz;
  ly = dy;
  lz = -P.sy * dx + P.cy * dz;
}

PX_DEV void prim_axis(int axis, double lx, double ly, double lz, double& a, double& b, double& h) {
  if (axis == 1) { a = lx; b = lz; h = ly; }
  else if (axis == 0) { a = ly; b = lz; h = lx; }
  else { a = lx; b = ly; h = lz; }
}

// SDFs of shapes.ts:24-143 in voxel units, one type at compile time
template <uint32_t T>
__device__ __forceinline__ double sdf_of(const px_prim& P, double x, double y, double z) {
  double lx, ly, lz, a, b, h;
  if constexpr (T == PX_PRIM_BOX) {
    prim_local(P, x, y, z, lx, ly, lz);
    const double qx = fabs(lx) - P.p[3], qy = fabs(ly) - P.p[4], qz = fabs(lz) - P.p[5];
    const double out = dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0));
    return out + dmin(dmax(qx, dmax(qy, qz)), 0.0);
  } else if constexpr (T == PX_PRIM_ROUNDBOX) {
    prim_local(P, x, y, z, lx, ly, lz);
    const double r = P.p[6];
    const double qx = fabs(lx) - (P.p[3] - r), qy = fabs(ly) - (P.p[4] - r), qz = fabs(lz) - (P.p[5] - r);
    return dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0)) + dmin(dmax(qx, dmax(qy, qz)), 0.0) - r;
  } else if constexpr (T == PX_PRIM_SPHERE) {
    return dlen3(x - P.p[0], y - P.p[1], z - P.p[2]) - P.p[3];
  } else if constexpr (T == PX_PRIM_CYLINDER) {
    prim_local(P, x, y, z, lx, ly, lz);
    prim_axis((int)P.p[5], lx, ly, lz, a, b, h);
    const double dr = dlen2(a, b) - P.p[3], dh = fabs(h) - P.p[4];
    return dmin(dmax(dr, dh), 0.0) + dlen2(dmax(dr, 0.0), dmax(dh, 0.0));
  } else if constexpr (T == PX_PRIM_CAPSULE) {
    const double ax = P.p[0], ay = P.p[1], az = P.p[2];
    const double bax = P.p[3] - ax, bay = P.p[4] - ay, baz = P.p[5] - az;
    double bb = bax * bax + bay * bay + baz * baz;
    if (bb == 0.0) bb = 1e-12;
    const double px = x - ax, py = y - ay, pz = z - az;
    const double t = dmin(1.0, dmax(0.0, (px * bax + py * bay + pz * baz) / bb));
    return dlen3(px - bax * t, py - bay * t, pz - baz * t) - P.p[6];
  } else if constexpr (T == PX_PRIM_TORUS) {
    prim_local(P, x, y, z, lx, ly, lz);
    prim_axis((int)P.p[5], lx, ly, lz, a, b, h);
    return dlen2(dlen2(a, b) - P.p[3], h) - P.p[4];
  } else {
    static_assert(T == PX_PRIM_AABB, "primitive type");
    const double cx = 0.5 * (P.p[0] + P.p[3]), cy = 0.5 * (P.p[1] + P.p[4]), cz = 0.5 * (P.p[2] + P.p[5]);
    const double hx = 0.5 * (P.p[3] - P.p[0]), hy = 0.5 * (P.p[4] - P.p[1]), hz = 0.5 * (P.p[5] - P.p[2]);
    const double qx = fabs(x - cx) - hx, qy = fabs(y - cy) - hy, qz = fabs(z - cz) - hz;
    const double out = dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0));
    return out + dmin(dmax(qx, dmax(qy, qz)), 0.0);
  }
}

PX_DEV double prim_sdf(const px_prim& P, double x, double y, double z) {
  switch (P.type) {
    case PX_PRIM_BOX: return sdf_of<PX_PRIM_BOX>(P, x, y, z);
    case PX_PRIM_ROUNDBOX: return sdf_of<PX_PRIM_ROUNDBOX>(P, x, y, z);
    case PX_PRIM_SPHERE: return sdf_of<PX_PRIM_SPHERE>(P, x, y, z);
    case PX_PRIM_CYLINDER: return sdf_of<PX_PRIM_CYLINDER>(P, x, y, z);
    case PX_PRIM_CAPSULE: return sdf_of<PX_PRIM_CAPSULE>(P, x, y, z);
    case PX_PRIM_TORUS: return sdf_of<PX_PRIM_TORUS>(P, x, y, z);
    case PX_PRIM_AABB: return sdf_of<PX_PRIM_AABB>(P, x, y, z);
  }
  return 1e30;
}

// shapes.ts:47-48: the exact containment test of axis-aligned boxes
PX_DEV bool prim_has_contains(const px_prim& P) {
  return P.type == PX_PRIM_AABB || (P.type == PX_PRIM_BOX && P.sy == 0.0 && P.cy == 1.0);
}
PX_DEV bool prim_contains(const px_prim& P, double x0, double y0, double z0, double x1, double y1, double z1) {
  if (P.type == PX_PRIM_AABB)
    return x0 >= P.p[0] && x1 <= P.p[3] && y0 >= P.p[1] && y1 <= P.p[4] && z0 >= P.p[2] && z1 <= P.p[5];
  if (P.type == PX_PRIM_BOX && P.sy == 0.0 && P.cy == 1.0)
    return x0 >= P.p[0] - P.p[3] && x1 <= P.p[0] + P.p[3] && y0 >= P.p[1] - P.p[4] && y1 <= P.p[1] + P.p[4] &&
           z0 >= P.p[2] - P.p[5] && z1 <= P.p[2] + P.p[5];
  return false;
}

struct RasterDims {
  int32_t dims[3], bd[3], sd[3];
  uint32_t ncell;
};

// brick range of apply() (shapes.ts:343-349): voxels [floor(lo - 0.5), ceil(hi + 0.5)] clamped
PX_DEV bool prim_bricks(const px_prim& P, const RasterDims& d, int32_t b0[3], int32_t n[3]) {
  for (int a = 0; a < 3; a++) {
    const double lo = floor(P.lo[a] - 0.5), hi = ceil(P.hi[a] + 0.5);
    const int32_t v0 = lo < 0.0 ? 0 : (lo > 2147483520.0 ? 2147483647 : (int32_t)lo);
    const int32_t v1 = hi > (double)(d.dims[a] - 1) ? d.dims[a] - 1 : (hi < -2147483648.0 ? (int32_t)0x80000000 : (int32_t)hi);
    if (v0 > v1) return false;
    b0[a] = v0 >> 3;
    n[a] = (v1 >> 3) - b0[a] + 1;
  }
  return true;
}

PX_DEV double brick_half_diag() { return 4.0 * sqrt(3.0) + 0.01; }
PX_DEV double sub_half_diag() { return 2.0 * sqrt(3.0) + 0.01; }

// PASS 0: kept bricks per primitive; PASS 1: the (brick << 32 | primitive) pairs
template <int PASS>
__global__ __launch_bounds__(256) void k_prim_pairs(const px_prim* __restrict__ prims, uint32_t n, RasterDims d,
                                                    uint32_t* __restrict__ counts, const uint64_t* __restrict__ offs,
                                                    uint64_t* __restrict__ pairs) {
  __shared__ uint32_t s_cnt;
  for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
    const px_prim P = prims[k];
    if (threadIdx.x == 0) s_cnt = 0u;
    __syncthreads();
    int32_t b0[3], nb[3];
    if (prim_bricks(P, d, b0, nb)) {
      const uint64_t total = (uint64_t)nb[0] * nb[1] * nb[2];
      const bool has = prim_has_contains(P);
      const double hd = brick_half_diag();
      for (uint64_t t = threadIdx.x; t < total; t += blockDim.x) {
        const int32_t bx = b0[0] + (int32_t)(t % (uint64_t)nb[0]);
        const int32_t by = b0[1] + (int32_t)((t / (uint64_t)nb[0]) % (uint64_t)nb[1]);
        const int32_t bz = b0[2] + (int32_t)(t / ((uint64_t)nb[0] * nb[1]));
        if (!has && prim_sdf(P, bx * 8 + 4.0, by * 8 + 4.0, bz * 8 + 4.0) > hd) continue;
        const uint32_t pos = atomicAdd(&s_cnt, 1u);
        if (PASS == 1) {
          const uint32_t ci = (uint32_t)bx + (uint32_t)d.bd[0] * ((uint32_t)by + (uint32_t)d.bd[1] * (uint32_t)bz);
          pairs[offs[k] + pos] = ((uint64_t)ci << 32) | k;
        }
      }
    }
    __syncthreads();
    if (PASS == 0 && threadIdx.x == 0) counts[k] = s_cnt;
    __syncthreads();
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Sets v[base + k] to pv for the k whose bit is set in bits4; true if a voxel
// changed.
__device__ __forceinline__ bool put4(uint32_t (&v)[8], int base, uint32_t bits4, uint32_t pv) {
  bool changed = false;
  for (int k = 0; k < 4; k++) {
    const bool set = ((bits4 >> k) & 1u) != 0u;
    changed |= set && v[base + k] != pv;
    v[base + k] = set ? pv : v[base + k];
  }
  return changed;
}

// The sub-block and voxel tests of one partial brick for one primitive of
// type T (shapes.ts:380-405).  The voxels stay in the row layout (lane = row
// ry | rz << 3, v = its 8 voxels in x); the SDF work is sub-block major:
//   - one wave-wide pass for the 8 sub-block tests (lane & 7 = sub-block
//     s = x half | y half << 1 | z half << 2), reduced to full / partial
//     masks by ballots;
//   - one wave-wide pass per partial sub-block, lane = its voxel
//     (x & 3) | (y & 3) << 2 | (z & 3) << 4, reduced to a 64-bit inside mask
//     that the rows of that sub-block read their 4 bits from.
// Every voxel set writes the same pv, so the sub-block order of the
// reference does not change the result.  No SDF sits under a lane-divergent
// branch (a per-lane early exit around the f64 square roots left partial
// bricks empty with hipcc -O3 for gfx950).
template <uint32_t T>
__device__ __forceinline__ bool raster_partial(const px_prim& P, int32_t ox, int32_t oy, int32_t oz, uint32_t lane,
                                               uint32_t pv, uint32_t (&v)[8]) {
  const double sd = sub_half_diag();
  const uint32_t s8 = lane & 7u;
  const double ds = sdf_of<T>(P, ox + (int32_t)(s8 & 1u) * 4 + 2.0, oy + (int32_t)((s8 >> 1) & 1u) * 4 + 2.0,
                              oz + (int32_t)(s8 >> 2) * 4 + 2.0);
  const uint32_t full_m = (uint32_t)__ballot(ds < -sd) & 0xFFu;
  uint32_t part_m = (uint32_t)__ballot(!(ds > sd) && !(ds < -sd)) & 0xFFu;
  const uint32_t ry = lane & 7u, rz = lane >> 3;
  const uint32_t srow = ((ry >> 2) << 1) | ((rz >> 2) << 2);  // this row's sub-blocks: srow, srow | 1
  const uint32_t rbit = ((ry & 3u) << 2) | ((rz & 3u) << 4);   // its voxel x = 0 in a sub-block's mask
  bool changed = put4(v, 0, ((full_m >> srow) & 1u) ? 0xFu : 0u, pv);
  changed |= put4(v, 4, ((full_m >> (srow | 1u)) & 1u) ? 0xFu : 0u, pv);
  const int32_t lx = (int32_t)(lane & 3u), ly = (int32_t)((lane >> 2) & 3u), lz = (int32_t)(lane >> 4);
  while (part_m != 0u) {
    const uint32_t s = (uint32_t)__builtin_ctz(part_m);
    part_m &= part_m - 1u;
    const int32_t x = ox + (int32_t)(s & 1u) * 4 + lx, y = oy + (int32_t)((s >> 1) & 1u) * 4 + ly,
                  z = oz + (int32_t)(s >> 2) * 4 + lz;
    const double dv = sdf_of<T>(P, x + 0.5, y + 0.5, z + 0.5);
    const uint64_t in = __ballot(!(dv > 0.0));
    const uint32_t bits = (srow == (s & 6u)) ? (uint32_t)(in >> rbit) & 0xFu : 0u;
    if (s & 1u)
      changed |= put4(v, 4, bits, pv);
    else
      changed |= put4(v, 0, bits, pv);
  }
  return changed;
}

// One workgroup (one wavefront) per run of pairs of one brick: the workgroup
// at the first pair j0 of a brick replays the brick's primitives in order; the
// others return.  Cells without pairs keep the state 0 they were filled with.
// state: 0 empty, 1 uniform (val), 2 explicit (voxels staged in temp at
// brick slot j0, so no counter is shared between wavefronts).
__global__ __launch_bounds__(64) void k_raster(const px_prim* __restrict__ prims, const uint64_t* __restrict__ pairs,
                                               uint64_t npairs, RasterDims d, uint8_t* __restrict__ out_state,
                                               uint16_t* __restrict__ out_val, uint32_t* __restrict__ out_slot,
                                               uint32_t* __restrict__ temp) {
  const uint32_t lane = threadIdx.x;
  const double hd = brick_half_diag();
  for (uint64_t j0 = blockIdx.x; j0 < npairs; j0 += gridDim.x) {
    const uint32_t ci = (uint32_t)(pairs[j0] >> 32);
    if (j0 > 0 && (uint32_t)(pairs[j0 - 1] >> 32) == ci) continue;
    const int32_t bx = (int32_t)(ci % (uint32_t)d.bd[0]);
    const int32_t by = (int32_t)((ci / (uint32_t)d.bd[0]) % (uint32_t)d.bd[1]);
    const int32_t bz = (int32_t)(ci / ((uint32_t)d.bd[0] * (uint32_t)d.bd[1]));
    const int32_t ox = bx * 8, oy = by * 8, oz = bz * 8;
    uint32_t state = 0u, val = 0u;
    uint32_t v[8];
    for (int x = 0; x < 8; x++) v[x] = 0u;
    for (uint64_t j = j0; j < npairs; j++) {
      const uint64_t key = pairs[j];
      if ((uint32_t)(key >> 32) != ci) break;
      const px_prim& P = prims[(uint32_t)key];
      const uint32_t pv = P.mat & 0xffffu;
      bool full = prim_contains(P, ox + 0.5, oy + 0.5, oz + 0.5, ox + 7.5, oy + 7.5, oz + 7.5);
      if (!full) {
        const double dc = prim_sdf(P, ox + 4.0, oy + 4.0, oz + 4.0);
        if (dc > hd) continue;
        full = dc < -hd;
      }
      if (full) {  // setBrickUniform
        state = pv != 0u ? 1u : 0u;
        val = pv;
        for (int x = 0; x < 8; x++) v[x] = pv;
        continue;
      }
      bool changed = false;
      switch (P.type) {
        case PX_PRIM_BOX: changed = raster_partial<PX_PRIM_BOX>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_ROUNDBOX: changed = raster_partial<PX_PRIM_ROUNDBOX>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_SPHERE: changed = raster_partial<PX_PRIM_SPHERE>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_CYLINDER: changed = raster_partial<PX_PRIM_CYLINDER>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_CAPSULE: changed = raster_partial<PX_PRIM_CAPSULE>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_TORUS: changed = raster_partial<PX_PRIM_TORUS>(P, ox, oy, oz, lane, pv, v); break;
        case PX_PRIM_AABB: changed = raster_partial<PX_PRIM_AABB>(P, ox, oy, oz, lane, pv, v); break;
        default: break;  // unknown types failed the brick test (SDF 1e30)
      }
      if (__ballot(changed) == 0ull) continue;
      // compactCell
      const uint32_t first = __shfl(v[0], 0);
      bool diff = false;
      for (int x = 0; x < 8; x++) diff |= v[x] != first;
      if (__ballot(diff) == 0ull) {
        state = first != 0u ? 1u : 0u;
        val = first;
      } else {
        state = 2u;
      }
    }
    if (lane == 0) {
      out_state[ci] = (uint8_t)state;
      out_val[ci] = (uint16_t)val;
    }
    if (state == 2u) {
      if (lane == 0) out_slot[ci] = (uint32_t)j0;
      uint4 w;
      w.x = v[0] | (v[1] << 16);
      w.y = v[2] | (v[3] << 16);
      w.z = v[4] | (v[5] << 16);
      w.w = v[6] | (v[7] << 16);
      reinterpret_cast<uint4*>(temp + (size_t)j0 * 256u)[lane] = w;
    }
  }
}

__global__ __launch_bounds__(256) void k_pool_flags(const uint8_t* __restrict__ state, uint32_t ncell,
                                                    uint2* __restrict__ flags) {
  for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < ncell; ci += gridDim.x * blockDim.x)
    flags[ci] = make_uint2(state[ci] == 2u ? 1u : 0u, 0u);
}

// grid entries and brick copies (one wavefront per cell)
__global__ __launch_bounds__(64) void k_fill_world(const uint8_t* __restrict__ state, const uint16_t* __restrict__ val,
                                                   const uint32_t* __restrict__ slot, const uint2* __restrict__ pidx,
                                                   const uint32_t* __restrict__ temp, uint32_t ncell,
                                                   uint32_t* __restrict__ grid, uint32_t* __restrict__ pool) {
  const uint32_t lane = threadIdx.x;
  for (uint32_t ci = blockIdx.x; ci < ncell; ci += gridDim.x) {
    const uint32_t s = state[ci];
    if (s == 2u) {
      const uint32_t p = pidx[ci].x;
      reinterpret_cast<uint4*>(pool + (size_t)p * 256u)[lane] =
          reinterpret_cast<const uint4*>(temp + (size_t)slot[ci] * 256u)[lane];
      if (lane == 0) grid[ci] = p + 1u;
    } else if (lane == 0) {
      grid[ci] = s == 1u ? (PX_UNIFORM_BIT | val[ci]) : 0u;
    }
  }
}

__global__ __launch_bounds__(256) void k_super_counts(const uint32_t* __restrict__ grid, RasterDims d,
                                                      uint32_t* __restrict__ sup) {
  const uint32_t ns = (uint32_t)(d.sd[0] * d.sd[1] * d.sd[2]);
  for (uint32_t si = blockIdx.x * blockDim.x + threadIdx.x; si < ns; si += gridDim.x * blockDim.x) {
    const int32_t sx = (int32_t)(si % (uint32_t)d.sd[0]);
    const int32_t sy = (int32_t)((si / (uint32_t)d.sd[0]) % (uint32_t)d.sd[1]);
    const int32_t sz = (int32_t)(si / ((uint32_t)d.sd[0] * (uint32_t)d.sd[1]));
    uint32_t c = 0u;
    for (int32_t z = sz * 4; z < min(sz * 4 + 4, d.bd[2]); z++)
      for (int32_t y = sy * 4; y < min(sy * 4 + 4, d.bd[1]); y++)
        for (int32_t x = sx * 4; x < min(sx * 4 + 4, d.bd[0]); x++)
          c += grid[(uint32_t)x + (uint32_t)d.bd[0] * ((uint32_t)y + (uint32_t)d.bd[1] * (uint32_t)z)] != 0u ? 1u : 0u;
    sup[si] = c;
  }
}

}  // namespace

static RasterDims raster_dims(const px_ctx* c) {
  RasterDims d;
  for (int a = 0; a < 3; a++) {
    d.dims[a] = (int32_t)c->dims[a];
    d.bd[a] = (int32_t)c->bd[a];
    d.sd[a] = (int32_t)c->sd[a];
  }
  d.ncell = c->ncell;
  return d;
}

int px_raster_begin(px_ctx* c, const px_prim* prims, uint32_t n, PxRaster* R, uint32_t* used) {
  hipStream_t s = c->stream;
  R->stream = s;
  const RasterDims d = raster_dims(c);
  const size_t ncell = c->ncell;
  *used = 0;
  const size_t state_words = (ncell + 3) / 4;
  PX_HIP_CHECK(c, hipMalloc(&R->d_state, state_words * 4));
  PX_HIP_CHECK(c, hipMalloc(&R->d_val, ncell * 2));
  PX_HIP_CHECK(c, hipMalloc(&R->d_slot, ncell * 4));
  hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, reinterpret_cast<uint32_t*>(R->d_state), state_words,
                     0u);
  uint64_t npairs = 0;
  if (n) {
    PX_HIP_CHECK(c, hipMalloc(&R->d_prims, (size_t)n * sizeof(px_prim)));
    PX_HIP_CHECK(c, hipMalloc(&R->d_counts, (size_t)n * 4));
    PX_HIP_CHECK(c, hipMalloc(&R->d_offs, (size_t)n * 8));
    PX_HIP_CHECK(c, hipMemcpyAsync(R->d_prims, prims, (size_t)n * sizeof(px_prim), hipMemcpyHostToDevice, s));
    const uint32_t g = std::min<uint32_t>(n, 65535u);
    hipLaunchKernelGGL(k_prim_pairs<0>, dim3(g), dim3(256), 0, s, R->d_prims, n, d, R->d_counts, nullptr, nullptr);
    PX_HIP_CHECK(c, hipGetLastError());
    std::vector<uint32_t> cnt(n);
    PX_HIP_CHECK(c, hipMemcpyAsync(cnt.data(), R->d_counts, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    PX_HIP_CHECK(c, hipStreamSynchronize(s));
    std::vector<uint64_t> off(n);
    for (uint32_t k = 0; k < n; k++) {
      off[k] = npairs;
      npairs += cnt[k];
    }
    if (npairs > 0xFFFFFFFFull)
      return px_fail(c, PX_ERR_CAPACITY, "px_create_from_prims", "more than 2^32 (brick, primitive) pairs");
    if (npairs) {
      PX_HIP_CHECK(c, hipMemcpyAsync(R->d_offs, off.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
      PX_HIP_CHECK(c, hipMalloc(&R->d_pairs, npairs * 8));
      PX_HIP_CHECK(c, hipMalloc(&R->d_sorted, npairs * 8));
      hipLaunchKernelGGL(k_prim_pairs<1>, dim3(g), dim3(256), 0, s, R->d_prims, n, d, nullptr, R->d_offs,
                         R->d_pairs);
      PX_HIP_CHECK(c, hipGetLastError());
      uint32_t cbits = 1;
      while (cbits < 32 && (1ull << cbits) < ncell) cbits++;
      // by cell only: each primitive's pairs are emitted at its offset, in
      // primitive order, and the sort is stable
      if (int rc = px_sort_u64(c, R->d_pairs, R->d_sorted, (uint32_t)npairs, 32u, 32u + cbits)) return rc;
    }
  }
  // one staging brick per pair: the brick at its first pair's index
  PX_HIP_CHECK(c, hipMalloc(&R->d_temp, (size_t)std::max<uint64_t>(1u, npairs) * 1024));
  if (npairs) {
    hipLaunchKernelGGL(k_raster, dim3((uint32_t)std::min<uint64_t>(npairs, 1u << 24)), dim3(64), 0, s, R->d_prims,
                       R->d_sorted, npairs, d, R->d_state, R->d_val, R->d_slot, R->d_temp);
    PX_HIP_CHECK(c, hipGetLastError());
  }
  // pool bricks numbered in cell order; the total sizes the pool
  PX_HIP_CHECK(c, hipMalloc(&R->d_flags, ncell * 8));
  PX_HIP_CHECK(c, hipMalloc(&R->d_pidx, ncell * 8));
  PX_HIP_CHECK(c, hipMalloc(&R->d_tiles, ((ncell + kScanTile - 1) / kScanTile) * 8 + 8));
  PX_HIP_CHECK(c, hipMalloc(&R->d_tot, 16));
  hipLaunchKernelGGL(k_pool_flags, dim3(1024), dim3(256), 0, s, R->d_state, c->ncell, R->d_flags);
  scan2_launch(s, R->d_flags, R->d_pidx, R->d_tiles, ncell, nullptr, c->ncell, nullptr, nullptr, R->d_tot);
  PX_HIP_CHECK(c, hipGetLastError());
  uint32_t nused = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&nused, R->d_tot, 4, hipMemcpyDeviceToHost, s));
  PX_HIP_CHECK(c, hipStreamSynchronize(s));
  *used = nused;
  return PX_OK;
}

int px_raster_finish(px_ctx* c, PxRaster* R) {
  hipStream_t s = c->stream;
  const RasterDims d = raster_dims(c);
  hipLaunchKernelGGL(k_fill_world, dim3(std::min<uint32_t>(c->ncell, 16384u)), dim3(64), 0, s, R->d_state, R->d_val,
                     R->d_slot, R->d_pidx, R->d_temp, c->ncell, c->d_grid, c->d_pool);
  hipLaunchKernelGGL(k_super_counts, dim3(256), dim3(256), 0, s, c->d_grid, d, c->d_super);
  PX_HIP_CHECK(c, hipGetLastError());
  PX_HIP_CHECK(c, hipStreamSynchronize(s));
  return PX_OK;
}
