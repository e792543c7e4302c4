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
#include "px_internal.h"

#pragma clang fp contract(off)

namespace {

struct MarchOut {
  uint32_t key;
  float t;
  uint32_t steps;
  uint32_t face;   // 0..6, 0xFF miss
  uint32_t value;  // 16-bit voxel value of the hit
};

PX_DEV uint32_t key_of(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  return (uint32_t)x + (uint32_t)w.dims[0] * ((uint32_t)y + (uint32_t)w.dims[1] * (uint32_t)z);
}

// face of the normal vec3f(-stpF[axis]) set when stepping along `axis`
PX_DEV uint32_t step_face(int axis, int32_t stp) { return (uint32_t)(axis * 2) + (stp > 0 ? 1u : 0u); }

PX_DEV void glass_surface(const float* __restrict__ mats, float trans[3], uint32_t id, uint32_t face,
                          const float rd[3], bool shadow) {
  const float* A = mats + id * 20u;
  float g = mats[id * 20u + 13u];
  if (!shadow) {
    // dot(n, rd) with an axis-aligned unit normal (or zero) is +-rd[axis] (or 0)
    float cosT = face < 6u ? fabsf(rd[face >> 1]) : 0.0f;
    float F = 0.04f + 0.96f * px_pow5(1.0f - cosT);
    for (int a = 0; a < 3; a++) trans[a] = trans[a] * (1.0f - F);
  }
  for (int a = 0; a < 3; a++) {
    float m = 1.0f * (1.0f - 0.6f) + A[a] * 0.6f;
    trans[a] = trans[a] * (m * g);
  }
}

// march() of raymarch.wgsl:44-162, one ray per lane.
PX_DEV MarchOut march(const PxWorldView& w, const float ro[3], const float rdIn[3], float maxT, bool shadow) {
  const float* __restrict__ mats = w.mats;
  MarchOut o;
  o.key = PX_NONE;
  o.t = 0.0f;
  o.face = 0xFFu;
  o.value = 0u;
  uint32_t steps = 0;
  float trans[3] = {1.0f, 1.0f, 1.0f};
  float rd[3], inv[3], tmn[3], tmx[3];
  for (int a = 0; a < 3; a++) {
    rd[a] = fabsf(rdIn[a]) < 1e-6f ? px_sign(rdIn[a] + 1e-30f) * 1e-6f : rdIn[a];
    inv[a] = 1.0f / rd[a];
    float t0 = (0.0f - ro[a]) * inv[a];
    float t1 = (w.fdims[a] - ro[a]) * inv[a];
    tmn[a] = fminf(t0, t1);
    tmx[a] = fmaxf(t0, t1);
  }
  float t = fmaxf(fmaxf(tmn[0], tmn[1]), fmaxf(tmn[2], 0.0f));
  float tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
  o.steps = 0;
  if (t >= tExit) return o;

  int32_t stp[3];
  float pos[3];
  for (int a = 0; a < 3; a++) {
    stp[a] = rd[a] > 0.0f ? 1 : (rd[a] < 0.0f ? -1 : 0);
    pos[a] = rd[a] >= 0.0f ? 1.0f : 0.0f;
  }
  uint32_t normal = 6u;
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) normal = step_face(0, stp[0]);
    else if (tmn[1] >= tmn[2]) normal = step_face(1, stp[1]);
    else normal = step_face(2, stp[2]);
  }
  int32_t sc[3];
  float tMaxS[3], tDeltaS[3], tDeltaB[3], tDeltaV[3];
  for (int a = 0; a < 3; a++) {
    float p0 = ro[a] + rd[a] * (t + 1e-3f);
    sc[a] = px_clampi(px_f2i(floorf(p0 / 32.0f)), 0, w.sd[a] - 1);
    tDeltaS[a] = fabsf(inv[a]) * 32.0f;
    tMaxS[a] = (((float)sc[a] + pos[a]) * 32.0f - ro[a]) * inv[a];
    tDeltaB[a] = fabsf(inv[a]) * 8.0f;
    tDeltaV[a] = fabsf(inv[a]);
  }
  bool inGlass = false;
  uint32_t glassId = 0u;

  for (int i = 0; i < 256; i++) {
    steps++;
    if (w.super_grid[(uint32_t)(sc[0] + w.sd[0] * (sc[1] + w.sd[1] * sc[2]))] != 0u) {
      int32_t sLo[3], sHi[3], bc[3];
      float tMaxB[3];
      for (int a = 0; a < 3; a++) {
        sLo[a] = sc[a] * 4;
        sHi[a] = min(sLo[a] + 3, w.bd[a] - 1);
        bc[a] = px_clampi(px_f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), sLo[a], sHi[a]);
        tMaxB[a] = (((float)bc[a] + pos[a]) * 8.0f - ro[a]) * inv[a];
      }
      float tb = t;
      uint32_t nb = normal;
      for (int j = 0; j < 14; j++) {
        steps++;
        uint32_t g = w.grid[(uint32_t)(bc[0] + w.bd[0] * (bc[1] + w.bd[1] * bc[2]))];
        if (g != 0u) {
          int32_t lo[3], vc[3];
          for (int a = 0; a < 3; a++) {
            lo[a] = bc[a] * 8;
            vc[a] = px_clampi(px_f2i(floorf(ro[a] + rd[a] * (tb + 1e-3f))), lo[a], lo[a] + 7);
          }
          if ((g & PX_UNIFORM_BIT) != 0u && !px_is_glass(mats, g & 0xffu)) {
            o.key = key_of(w, vc[0], vc[1], vc[2]);
            o.t = px_canon0(tb);
            o.face = nb;
            o.value = g & 0xffffu;
            o.steps = steps;
            return o;
          }
          float tMaxV[3];
          for (int a = 0; a < 3; a++) tMaxV[a] = ((float)vc[a] + pos[a] - ro[a]) * inv[a];
          float tv = tb;
          uint32_t n = nb;
          for (int k = 0; k < 25; k++) {
            steps++;
            uint32_t v = px_brick_voxel(w.pool, g, vc[0] - lo[0], vc[1] - lo[1], vc[2] - lo[2]);
            if (v != 0u) {
              uint32_t id = v & 0xffu;
              bool hit = true;
              if (px_is_glass(mats, id)) {
                if (!inGlass || id != glassId) glass_surface(mats, trans, id, n, rd, shadow);
                inGlass = true;
                glassId = id;
                hit = fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f;
              }
              if (hit) {
                o.key = key_of(w, vc[0], vc[1], vc[2]);
                o.t = px_canon0(tv);
                o.face = n;
                o.value = v;
                o.steps = steps;
                return o;
              }
            } else {
              inGlass = false;
            }
            if (tMaxV[0] < tMaxV[1] && tMaxV[0] < tMaxV[2]) {
              tv = tMaxV[0]; vc[0] += stp[0]; tMaxV[0] += tDeltaV[0]; n = step_face(0, stp[0]);
            } else if (tMaxV[1] < tMaxV[2]) {
              tv = tMaxV[1]; vc[1] += stp[1]; tMaxV[1] += tDeltaV[1]; n = step_face(1, stp[1]);
            } else {
              tv = tMaxV[2]; vc[2] += stp[2]; tMaxV[2] += tDeltaV[2]; n = step_face(2, stp[2]);
            }
            if (tv > tExit) { o.steps = steps; return o; }
            if (vc[0] < lo[0] || vc[1] < lo[1] || vc[2] < lo[2] || vc[0] > lo[0] + 7 || vc[1] > lo[1] + 7 ||
                vc[2] > lo[2] + 7)
              break;
          }
        } else {
          inGlass = false;
        }
        if (tMaxB[0] < tMaxB[1] && tMaxB[0] < tMaxB[2]) {
          tb = tMaxB[0]; bc[0] += stp[0]; tMaxB[0] += tDeltaB[0]; nb = step_face(0, stp[0]);
        } else if (tMaxB[1] < tMaxB[2]) {
          tb = tMaxB[1]; bc[1] += stp[1]; tMaxB[1] += tDeltaB[1]; nb = step_face(1, stp[1]);
        } else {
          tb = tMaxB[2]; bc[2] += stp[2]; tMaxB[2] += tDeltaB[2]; nb = step_face(2, stp[2]);
        }
        if (tb > tExit) { o.steps = steps; return o; }
        if (bc[0] < sLo[0] || bc[1] < sLo[1] || bc[2] < sLo[2] || bc[0] > sHi[0] || bc[1] > sHi[1] || bc[2] > sHi[2])
          break;
      }
    } else {
      inGlass = false;
    }
    if (tMaxS[0] < tMaxS[1] && tMaxS[0] < tMaxS[2]) {
      t = tMaxS[0]; sc[0] += stp[0]; tMaxS[0] += tDeltaS[0]; normal = step_face(0, stp[0]);
    } else if (tMaxS[1] < tMaxS[2]) {
      t = tMaxS[1]; sc[1] += stp[1]; tMaxS[1] += tDeltaS[1]; normal = step_face(1, stp[1]);
    } else {
      t = tMaxS[2]; sc[2] += stp[2]; tMaxS[2] += tDeltaS[2]; normal = step_face(2, stp[2]);
    }
    if (t > tExit || sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || sc[0] >= w.sd[0] || sc[1] >= w.sd[1] || sc[2] >= w.sd[2])
      break;
  }
  o.steps = steps;
  return o;
}

PX_DEV uint32_t read_voxel(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= w.dims[0] || y >= w.dims[1] || z >= w.dims[2]) return 0u;
  uint32_t g = w.grid[(uint32_t)(x >> 3) + (uint32_t)w.bd[0] * ((uint32_t)(y >> 3) + (uint32_t)w.bd[1] * (uint32_t)(z >> 3))];
  if (g == 0u) return 0u;
  return px_brick_voxel(w.pool, g, x & 7, y & 7, z & 7);
}

// traceSolid() of sim.wgsl:296-357.  steps = brick + voxel iterations, t = the
// voxel-level t of the hit.
PX_DEV MarchOut trace_solid(const PxWorldView& w, const float ro[3], const float rdIn[3], float maxT) {
  MarchOut o;
  o.key = PX_NONE;
  o.t = 0.0f;
  o.face = 0xFFu;
  o.value = 0u;
  o.steps = 0u;
  uint32_t steps = 0;
  float rd[3], inv[3], tmn[3], tmx[3];
  for (int a = 0; a < 3; a++) {
    rd[a] = fabsf(rdIn[a]) < 1e-6f ? px_sign(rdIn[a] + 1e-30f) * 1e-6f : rdIn[a];
    inv[a] = 1.0f / rd[a];
    float t0 = (0.0f - ro[a]) * inv[a];
    float t1 = (w.fdims[a] - ro[a]) * inv[a];
    tmn[a] = fminf(t0, t1);
    tmx[a] = fmaxf(t0, t1);
  }
  float t = fmaxf(fmaxf(tmn[0], tmn[1]), fmaxf(tmn[2], 0.0f));
  float tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
  if (t >= tExit) return o;
  int32_t stp[3];
  float pos[3];
  for (int a = 0; a < 3; a++) {
    stp[a] = rd[a] > 0.0f ? 1 : (rd[a] < 0.0f ? -1 : 0);
    pos[a] = rd[a] >= 0.0f ? 1.0f : 0.0f;
  }
  uint32_t normal = 6u;
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) normal = step_face(0, stp[0]);
    else if (tmn[1] >= tmn[2]) normal = step_face(1, stp[1]);
    else normal = step_face(2, stp[2]);
  }
  int32_t bc[3];
  float tDeltaB[3], tMaxB[3], tDeltaV[3];
  for (int a = 0; a < 3; a++) {
    bc[a] = px_clampi(px_f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), 0, w.bd[a] - 1);
    tDeltaB[a] = fabsf(inv[a]) * 8.0f;
    tMaxB[a] = (((float)bc[a] + pos[a]) * 8.0f - ro[a]) * inv[a];
    tDeltaV[a] = fabsf(inv[a]);
  }
  for (int i = 0; i < 1024; i++) {
    steps++;
    uint32_t g = w.grid[(uint32_t)(bc[0] + w.bd[0] * (bc[1] + w.bd[1] * bc[2]))];
    if (g != 0u) {
      int32_t lo[3], vc[3];
      float tMaxV[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = bc[a] * 8;
        vc[a] = px_clampi(px_f2i(floorf(ro[a] + rd[a] * (t + 1e-3f))), lo[a], lo[a] + 7);
        tMaxV[a] = ((float)vc[a] + pos[a] - ro[a]) * inv[a];
      }
      float tv = t;
      uint32_t n = normal;
      for (int k = 0; k < 25; k++) {
        steps++;
        uint32_t v = read_voxel(w, vc[0], vc[1], vc[2]);
        if (v != 0u) {
          o.key = key_of(w, vc[0], vc[1], vc[2]);
          o.t = px_canon0(tv);
          o.face = n;
          o.value = v;
          o.steps = steps;
          return o;
        }
        if (tMaxV[0] < tMaxV[1] && tMaxV[0] < tMaxV[2]) {
          tv = tMaxV[0]; vc[0] += stp[0]; tMaxV[0] += tDeltaV[0]; n = step_face(0, stp[0]);
        } else if (tMaxV[1] < tMaxV[2]) {
          tv = tMaxV[1]; vc[1] += stp[1]; tMaxV[1] += tDeltaV[1]; n = step_face(1, stp[1]);
        } else {
          tv = tMaxV[2]; vc[2] += stp[2]; tMaxV[2] += tDeltaV[2]; n = step_face(2, stp[2]);
        }
        if (tv > tExit) { o.steps = steps; return o; }
        if (vc[0] < lo[0] || vc[1] < lo[1] || vc[2] < lo[2] || vc[0] > lo[0] + 7 || vc[1] > lo[1] + 7 ||
            vc[2] > lo[2] + 7)
          break;
      }
    }
    if (tMaxB[0] < tMaxB[1] && tMaxB[0] < tMaxB[2]) {
      t = tMaxB[0]; bc[0] += stp[0]; tMaxB[0] += tDeltaB[0]; normal = step_face(0, stp[0]);
    } else if (tMaxB[1] < tMaxB[2]) {
      t = tMaxB[1]; bc[1] += stp[1]; tMaxB[1] += tDeltaB[1]; normal = step_face(1, stp[1]);
    } else {
      t = tMaxB[2]; bc[2] += stp[2]; tMaxB[2] += tDeltaB[2]; normal = step_face(2, stp[2]);
    }
    if (t > tExit || bc[0] < 0 || bc[1] < 0 || bc[2] < 0 || bc[0] >= w.bd[0] || bc[1] >= w.bd[1] || bc[2] >= w.bd[2])
      break;
  }
  o.steps = steps;
  return o;
}

// Segment DDA of particleSim, sim.wgsl:1110-1145.
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
  float dir[3], inv[3], tMax[3], tDelta[3];
  int32_t stp[3];
  for (int a = 0; a < 3; a++) {
    float d = mv[a] / dist;
    dir[a] = fabsf(d) < 1e-6f ? 1e-6f : d;
  }
  for (int a = 0; a < 3; a++) {
    inv[a] = 1.0f / dir[a];
    stp[a] = dir[a] > 0.0f ? 1 : (dir[a] < 0.0f ? -1 : 0);
    tMax[a] = ((float)cell[a] + (dir[a] >= 0.0f ? 1.0f : 0.0f) - pos[a]) * inv[a];
    tDelta[a] = fabsf(inv[a]);
  }
  uint32_t steps = 0;
  for (int k = 0; k < 48; k++) {
    steps++;
    int axis;
    if (tMax[0] < tMax[1] && tMax[0] < tMax[2]) axis = 0;
    else if (tMax[1] < tMax[2]) axis = 1;
    else axis = 2;
    float tn = axis == 0 ? tMax[0] : (axis == 1 ? tMax[1] : tMax[2]);
    if (tn > dist) break;
    int32_t nx = cell[0] + (axis == 0 ? stp[0] : 0);
    int32_t ny = cell[1] + (axis == 1 ? stp[1] : 0);
    int32_t nz = cell[2] + (axis == 2 ? stp[2] : 0);
    uint32_t v = read_voxel(w, nx, ny, nz);
    if (v != 0u) {
      o.key = key_of(w, nx, ny, nz);
      o.t = px_canon0(tn);
      o.face = step_face(axis, axis == 0 ? stp[0] : (axis == 1 ? stp[1] : stp[2]));
      o.value = v;
      o.steps = steps;
      return o;
    }
    cell[0] = nx; cell[1] = ny; cell[2] = nz;
    if (axis == 0) tMax[0] += tDelta[0];
    else if (axis == 1) tMax[1] += tDelta[1];
    else tMax[2] += tDelta[2];
  }
  o.steps = steps;
  return o;
}

PX_DEV void store_hit(px_hit* out, const MarchOut& m) {
  uint4 v;
  v.x = m.key;
  v.y = __float_as_uint(m.t);
  v.z = m.steps;
  v.w = (m.face & 0xffu) | ((m.value & 0xffu) << 8) | (((m.value >> 8) & 0xffu) << 16);
  *reinterpret_cast<uint4*>(out) = v;
}

PX_DEV void store_not_cast(px_hit* out) {
  uint4 v;
  v.x = PX_NONE;
  v.y = 0u;
  v.z = 0u;
  v.w = 0xFEu;
  *reinterpret_cast<uint4*>(out) = v;
}

PX_DEV void load_ray(const px_ray* r, float ro[3], float rd[3], float& maxT, uint32_t& flags) {
  const float4* p = reinterpret_cast<const float4*>(r);
  float4 a = p[0], b = p[1];
  ro[0] = a.x; ro[1] = a.y; ro[2] = a.z;
  rd[0] = a.w; rd[1] = b.x; rd[2] = b.y;
  maxT = b.z;
  flags = __float_as_uint(b.w);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_trace(PxWorldView w, const px_ray* __restrict__ rays, uint32_t n,
                                               px_hit* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ro[3], rd[3], maxT;
  uint32_t flags;
  load_ray(rays + i, ro, rd, maxT, flags);
  int kind = KIND >= 0 ? KIND : (int)(flags & 3u);
  MarchOut m;
  if (kind == 0) m = march(w, ro, rd, maxT, false);
  else if (kind == 1) m = march(w, ro, rd, maxT, true);
  else if (kind == 2) m = trace_solid(w, ro, rd, maxT);
  else m = segment(w, ro, rd);
  store_hit(out + i, m);
}

// Primary march + shadow-ray generation.  Shadow rays are appended to a
// compacted queue (one atomic per wavefront, lanes placed by ballot prefix),
// with the destination slot in `flags`.
__global__ __launch_bounds__(256) void k_frame_primary(PxWorldView w, PxShadowDesc sd,
                                                       const px_ray* __restrict__ rays, uint32_t n,
                                                       px_hit* __restrict__ prim_out, px_hit* __restrict__ sh_out,
                                                       px_ray* __restrict__ queue, uint32_t* __restrict__ counter) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < n;
  float ro[3] = {0, 0, 0}, rd[3] = {0, 0, 1}, maxT = 0.0f;
  uint32_t flags = 0;
  MarchOut h;
  h.face = 0xFFu;
  h.value = 0u;
  if (live) {
    load_ray(rays + i, ro, rd, maxT, flags);
    h = march(w, ro, rd, maxT, false);
    store_hit(prim_out + i, h);
    for (uint32_t s = 0; s < sd.slots; s++) store_not_cast(sh_out + (size_t)i * sd.slots + s);
  }
  bool shade = live && h.face != 0xFFu && !(w.mats[(h.value & 0xffu) * 20u + 3u] > 0.0f);
  float N[3] = {0.0f, 0.0f, 0.0f};
  if (h.face < 6u) N[h.face >> 1] = (h.face & 1u) ? -1.0f : 1.0f;
  float P[3], O[3];
  for (int a = 0; a < 3; a++) {
    P[a] = ro[a] + rd[a] * h.t;
    O[a] = P[a] + N[a] * 0.5f;
  }
  // up to `slots` rays per lane; each slot is compacted across the wave
  for (uint32_t s = 0; s < sd.slots; s++) {
    bool cast = false;
    float dir[3] = {0, 0, 0}, mt = 0.0f;
    if (shade) {
      if (s == 0) {
        float ndl = N[0] * sd.sun[0] + N[1] * sd.sun[1] + N[2] * sd.sun[2];
        if (ndl > 0.0f && sd.sun_w > 0.0f) {
          cast = true;
          dir[0] = sd.sun[0]; dir[1] = sd.sun[1]; dir[2] = sd.sun[2];
          mt = 1e9f;
        }
      } else {
        const float* lp = sd.lp[s - 1];
        float L[3];
        for (int a = 0; a < 3; a++) L[a] = lp[a] / sd.vs - P[a];
        float radius = lp[3] / sd.vs;
        float d = sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
        if (!(d > radius)) {
          for (int a = 0; a < 3; a++) L[a] = L[a] / d;
          float nl = N[0] * L[0] + N[1] * L[1] + N[2] * L[2];
          if (nl > 0.0f && sd.lc_w[s - 1] > 0.5f && sd.lamp_shadows > 0.5f) {
            cast = true;
            dir[0] = L[0]; dir[1] = L[1]; dir[2] = L[2];
            mt = d - 2.0f;
          }
        }
      }
    }
    uint64_t mask = __ballot(cast);
    if (mask == 0) continue;
    uint32_t lane = __lane_id();
    uint32_t base = 0;
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (cast) {
      uint32_t off = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
      float4* q = reinterpret_cast<float4*>(queue + off);
      q[0] = make_float4(O[0], O[1], O[2], dir[0]);
      q[1] = make_float4(dir[1], dir[2], mt, __uint_as_float(i * sd.slots + s));
    }
  }
}

__global__ __launch_bounds__(256) void k_trace_shadow_queue(PxWorldView w, const px_ray* __restrict__ queue,
                                                            const uint32_t* __restrict__ counter,
                                                            px_hit* __restrict__ sh_out) {
  uint32_t n = *counter;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float ro[3], rd[3], maxT;
    uint32_t slot;
    load_ray(queue + i, ro, rd, maxT, slot);
    MarchOut m = march(w, ro, rd, maxT, true);
    store_hit(sh_out + slot, m);
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

// one bit per super cell (count != 0) and per brick cell (entry != 0)
__global__ void k_masks(const uint32_t* __restrict__ grid, uint32_t ncell, const uint32_t* __restrict__ sgrid,
                        uint32_t nsuper, uint32_t* __restrict__ bmask, uint32_t* __restrict__ smask) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t cells = (ncell + 31u) / 32u;
  if (i < cells) {
    uint32_t m = 0;
    for (uint32_t b = 0; b < 32u; b++) {
      uint32_t c = i * 32u + b;
      if (c < ncell && grid[c] != 0u) m |= 1u << b;
    }
    bmask[i] = m;
  }
  uint32_t sw = (nsuper + 31u) / 32u;
  if (i < sw) {
    uint32_t m = 0;
    for (uint32_t b = 0; b < 32u; b++) {
      uint32_t c = i * 32u + b;
      if (c < nsuper && sgrid[c] != 0u) m |= 1u << b;
    }
    smask[i] = m;
  }
  (void)lane;
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
  for (int a = 0; a < 3; a++) {
    w.dims[a] = (int32_t)c->dims[a];
    w.bd[a] = (int32_t)c->bd[a];
    w.sd[a] = (int32_t)c->sd[a];
    w.fdims[a] = (float)c->dims[a];
  }
  return w;
}

int px_launch_masks(px_ctx* c) {
  uint32_t words = (c->ncell + 31u) / 32u;
  uint32_t blocks = (words + 255u) / 256u;
  hipLaunchKernelGGL(k_masks, dim3(blocks), dim3(256), 0, c->stream, c->d_grid, c->ncell, c->d_super, c->nsuper,
                     c->d_brick_mask, c->d_super_mask);
  PX_HIP_CHECK(c, hipGetLastError());
  int rc = px_launch_super_bm(c);
  if (rc) return rc;
  c->masks_dirty = false;
  return PX_OK;
}

int px_refresh_masks(px_ctx* c) { return c->masks_dirty ? px_launch_masks(c) : PX_OK; }

int px_launch_trace(px_ctx* c, uint32_t kind, const px_ray* d_rays, uint32_t n, px_hit* d_out) {
  if (n == 0) return PX_OK;
  int rc = px_refresh_masks(c);
  if (rc) return rc;
  if (c->tune.variant == 1u && (kind == PX_RAY_PRIMARY || kind == PX_RAY_SHADOW) && px_march_fits(c))
    return px_launch_march(c, kind == PX_RAY_SHADOW, d_rays, n, d_out);
  PxWorldView w = px_view(c);
  dim3 grid((n + 255u) / 256u), block(256);
  PxProfScope prof(c, PX_PROF_TRACE);
  switch (kind) {
    case PX_RAY_PRIMARY: hipLaunchKernelGGL(k_trace<0>, grid, block, 0, c->stream, w, d_rays, n, d_out); break;
    case PX_RAY_SHADOW: hipLaunchKernelGGL(k_trace<1>, grid, block, 0, c->stream, w, d_rays, n, d_out); break;
    case PX_RAY_SHOOT: hipLaunchKernelGGL(k_trace<2>, grid, block, 0, c->stream, w, d_rays, n, d_out); break;
    case PX_RAY_SEGMENT: hipLaunchKernelGGL(k_trace<3>, grid, block, 0, c->stream, w, d_rays, n, d_out); break;
    default: hipLaunchKernelGGL(k_trace<-1>, grid, block, 0, c->stream, w, d_rays, n, d_out); break;
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_launch_frame(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow,
                    uint32_t* n_cast) {
  if (n == 0) {
    if (n_cast) *n_cast = 0;
    return PX_OK;
  }
  int rc = px_refresh_masks(c);
  if (rc) return rc;
  size_t need = (size_t)n * c->shadow.slots;
  if (need > 0xFFFFFFFFull) return px_fail(c, PX_ERR_CAPACITY, "px_trace_frame_device", "too many shadow slots");
  if (need > c->shadow_cap) {
    if (c->d_shadow_rays) hipFree(c->d_shadow_rays);
    c->d_shadow_rays = nullptr;
    c->shadow_cap = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_shadow_rays, need * sizeof(px_ray)));
    c->shadow_cap = need;
  }
  if (c->tune.variant == 1u && px_march_fits(c)) {
    rc = px_launch_frame_persistent(c, d_primary, n, d_prim, d_shadow);
    if (rc) return rc;
    if (n_cast) {
      PX_HIP_CHECK(c, hipMemcpyAsync(n_cast, c->d_counters, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    }
    return PX_OK;
  }
  // a fill kernel, not hipMemsetAsync, so the launch sequence can be captured in a graph
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->d_counters, 1u, 0u);
  PxWorldView w = px_view(c);
  {
    PxProfScope prof(c, PX_PROF_PRIMARY);
    hipLaunchKernelGGL(k_frame_primary, dim3((n + 255u) / 256u), dim3(256), 0, c->stream, w, c->shadow, d_primary, n,
                       d_prim, d_shadow, c->d_shadow_rays, c->d_counters);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  uint32_t blocks = (uint32_t)((need + 255u) / 256u);
  if (blocks > 8192u) blocks = 8192u;
  {
    PxProfScope prof(c, PX_PROF_SHADOW);
    hipLaunchKernelGGL(k_trace_shadow_queue, dim3(blocks), dim3(256), 0, c->stream, w, c->d_shadow_rays,
                       c->d_counters, d_shadow);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  if (n_cast) {
    PX_HIP_CHECK(c, hipMemcpyAsync(n_cast, c->d_counters, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
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
