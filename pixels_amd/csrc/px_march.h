// px_march.h - march() of src/render/shaders/raymarch.wgsl:44-162 (+ safeDir
// :22-24, glassSurface :26-38) for gfx950, one ray per lane.
//
// The loops, iteration caps and FP32 operations are the WGSL's: three nested
// DDAs (32^3 super cell, 8^3 brick, voxel), at most 256 / 14 / 25 turns, a
// uniform opaque brick ends the ray at the brick's entry voxel.  What this
// version changes is the cost of a turn on CDNA4:
//
//  * super-cell occupancy is a bit in LDS, brick occupancy a bit of the super
//    cell's 64-bit mask (one 8-byte load on entering the super cell), so only
//    non-empty bricks read the grid and only their voxels read the pool;
//  * every DDA step is branch-free (compare, select), with
//    tMax += |inv| * s written as one fma: |inv| * s is exact for the
//    power-of-two scales s = 1, 8, 32, so the sum rounds once either way;
//  * faces are carried as an axis code (0..2, 3 = zero normal) and turned
//    into the reference's face id only at the hit, brick and voxel
//    coordinates are local to their parent cell, and the glass state is one
//    word - all to keep the live state small enough for more wavefronts per
//    SIMD.
#pragma once

#include "px_internal.h"
#include "px_shade.h"

#pragma clang fp contract(off)

namespace px {

// Shading outputs of a march (the Hit fields of raymarch.wgsl:6-14 the
// traversal does not need): transmittance through glass, the light glass
// surfaces add to camera rays, and the voxel-loop iterations (the debug
// views' inner counter; outer = steps - inner).
struct MarchShade {
  float trans[3];
  float glow[3];
  uint32_t inner;
};

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

// Per-block tables of march(): super-cell occupancy bits (count != 0) and the
// 256-bit glass table (material D.y > 0).  `sup` is LDS when the world's bits
// fit, else the global copy.
struct MarchTables {
  const uint32_t* sup;
  const uint32_t* glass;
};

PX_DEV bool bit_of(const uint32_t* bits, uint32_t i) { return ((bits[i >> 5] >> (i & 31u)) & 1u) != 0u; }

// Block-cooperative: every thread of the block calls it before any early exit.
template <bool SUP_LDS>
PX_DEV MarchTables load_march_tables(const PxWorldView& w, uint32_t* s_mem) {
  uint32_t* s_glass = s_mem;
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < 256u; base += blockDim.x) {
    uint32_t m = base + threadIdx.x;
    uint64_t b = __ballot(m < 256u && w.mats[m * 20u + 13u] > 0.0f);
    uint32_t word = (base + wave * 64u) >> 5;
    if (lane == 0u && word < 8u) {
      s_glass[word] = (uint32_t)b;
      s_glass[word + 1u] = (uint32_t)(b >> 32);
    }
  }
  MarchTables T;
  T.glass = s_glass;
  if (SUP_LDS) {
    uint32_t* s_sup = s_mem + 8;
    const uint32_t sw = ((uint32_t)(w.sd[0] * w.sd[1] * w.sd[2]) + 31u) / 32u;
    for (uint32_t i = threadIdx.x; i < sw; i += blockDim.x) s_sup[i] = w.super_mask[i];
    T.sup = s_sup;
  } else {
    T.sup = w.super_mask;
  }
  __syncthreads();
  return T;
}

// glassSurface (raymarch.wgsl:26-38): Fresnel term for camera rays, then the
// tinted absorption.  dot(n, rd) with an axis-aligned unit normal is
// +-rd[axis] (0 for the zero normal).
PX_DEV void glass_surface(const float* __restrict__ mats, float trans[3], uint32_t id, uint32_t axis,
                          const float rd[3], bool shadow) {
  const float* A = mats + id * 20u;
  float g = mats[id * 20u + 13u];
  if (!shadow) {
    float cosT = axis < 3u ? fabsf(axis == 0u ? rd[0] : (axis == 1u ? rd[1] : rd[2])) : 0.0f;
    float F = 0.04f + 0.96f * px_pow5(1.0f - cosT);
    for (int a = 0; a < 3; a++) trans[a] = trans[a] * (1.0f - F);
  }
  for (int a = 0; a < 3; a++) {
    float m = 1.0f * (1.0f - 0.6f) + A[a] * 0.6f;
    trans[a] = trans[a] * (m * g);
  }
}

// glassSurface with the camera-ray glow (raymarch.wgsl:26-38): the normal is
// -sign(rd) along `axis` (zero for axis 3); the sky and the sun glint seen in
// the reflected direction add to glow before the Fresnel loss.
PX_DEV void glass_surface_glow(const float* __restrict__ mats, const PxSky& sky, float trans[3], float glow[3],
                               uint32_t id, uint32_t axis, const float rd[3], const int32_t stp[3]) {
  const float* A = mats + id * 20u;
  const float* Cm = A + 8;
  const float g = mats[id * 20u + 13u];
  float n[3] = {0.0f, 0.0f, 0.0f};
  if (axis < 3u) n[axis] = -(float)stp[axis];
  const float cosT = fabsf(dot3f(n, rd));
  const float F = 0.04f + 0.96f * px_pow5(1.0f - cosT);
  const float dn = dot3f(n, rd);
  const float r[3] = {rd[0] - 2.0f * dn * n[0], rd[1] - 2.0f * dn * n[1], rd[2] - 2.0f * dn * n[2]};
  const float sr = dot3f(r, sky.sun);
  const float p = fm::powf_(sr < 0.0f ? 0.0f : sr, Cm[3]);
  const float3 skc = sky_color(sky, make_float3(r[0], r[1], r[2]));
  const float sk[3] = {skc.x, skc.y, skc.z};
  for (int a = 0; a < 3; a++) {
    const float glint = p * sky.sun_col[a] * sky.sun[3] * 2.0f;
    glow[a] = glow[a] + trans[a] * (F * sk[a] * 0.35f + glint * Cm[2]);
  }
  for (int a = 0; a < 3; a++) trans[a] = trans[a] * (1.0f - F);
  for (int a = 0; a < 3; a++) {
    const float m = 1.0f * (1.0f - 0.6f) + A[a] * 0.6f;
    trans[a] = trans[a] * (m * g);
  }
}

// face id of an axis code (3 = zero normal)
PX_DEV uint32_t face_of_axis(uint32_t axis, const float rd[3]) {
  if (axis >= 3u) return 6u;
  const float r = axis == 0u ? rd[0] : (axis == 1u ? rd[1] : rd[2]);
  return axis * 2u + (r > 0.0f ? 1u : 0u);
}

// Every DDA step below is the WGSL's if / else-if chain written as selects:
// axis 0 when m0 < m1 && m0 < m2, else axis 1 when m1 < m2, else axis 2
// (ties go to the later axis); that axis's tMax advances by |inv| * scale
// and its cell coordinate by stp.
// With SHADE, `ms` receives the transmittance, glow (camera rays, with `sky`)
// and inner step count of the march.
template <bool SHADOW, bool SHADE = false>
PX_DEV MarchOut march(const PxWorldView& w, const MarchTables& T, const float ro[3], const float rdIn[3],
                      float maxT, const PxSky* sky = nullptr, MarchShade* ms = nullptr) {
  MarchOut o;
  float glow[3] = {0.0f, 0.0f, 0.0f};
  uint32_t inner = 0u;
  float trans[3] = {1.0f, 1.0f, 1.0f};
  auto done = [&]() {
    if constexpr (SHADE) {
      for (int a = 0; a < 3; a++) {
        ms->trans[a] = trans[a];
        ms->glow[a] = glow[a];
      }
      ms->inner = inner;
    }
  };
  o.key = PX_NONE;
  o.t = 0.0f;
  o.face = 0xFFu;
  o.value = 0u;
  o.steps = 0u;
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
  const float tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
  if (t >= tExit) {
    done();
    return o;
  }

  int32_t stp[3];
  float pos[3];
  for (int a = 0; a < 3; a++) {
    stp[a] = rd[a] > 0.0f ? 1 : (rd[a] < 0.0f ? -1 : 0);
    pos[a] = rd[a] >= 0.0f ? 1.0f : 0.0f;
  }
  uint32_t axisS = 3u;  // normal as an axis code
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) axisS = 0u;
    else if (tmn[1] >= tmn[2]) axisS = 1u;
    else axisS = 2u;
  }
  int32_t sc[3];
  float tMaxS[3];
  for (int a = 0; a < 3; a++) {
    float p0 = ro[a] + rd[a] * (t + 1e-3f);
    sc[a] = px_clampi(px_f2i(floorf(p0 / 32.0f)), 0, w.sd[a] - 1);
    tMaxS[a] = (((float)sc[a] + pos[a]) * 32.0f - ro[a]) * inv[a];
  }
  uint32_t glassState = 0u;  // 0x100 | id while inside glass `id`, else 0
  uint32_t steps = 0u;

  for (int i = 0; i < 256; i++) {
    steps++;
    const uint32_t sidx = (uint32_t)(sc[0] + w.sd[0] * (sc[1] + w.sd[1] * sc[2]));
    if (bit_of(T.sup, sidx)) {
      const uint64_t bm = w.super_bm[sidx];
      int32_t hiL[3], lb[3];
      float tMaxB[3];
      for (int a = 0; a < 3; a++) {
        const int32_t sLo = sc[a] * 4;
        const int32_t sHi = min(sLo + 3, w.bd[a] - 1);
        hiL[a] = sHi - sLo;
        const int32_t bc = px_clampi(px_f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), sLo, sHi);
        lb[a] = bc - sLo;
        tMaxB[a] = (((float)bc + pos[a]) * 8.0f - ro[a]) * inv[a];
      }
      float tb = t;
      uint32_t axisB = axisS;
      for (int j = 0; j < 14; j++) {
        steps++;
        if ((bm >> (uint32_t)(lb[0] + 4 * (lb[1] + 4 * lb[2]))) & 1ull) {
          int32_t lv[3];
          float tMaxV[3];
          for (int a = 0; a < 3; a++) {
            const int32_t lo = (sc[a] * 4 + lb[a]) * 8;
            const int32_t vc = px_clampi(px_f2i(floorf(ro[a] + rd[a] * (tb + 1e-3f))), lo, lo + 7);
            lv[a] = vc - lo;
            tMaxV[a] = ((float)vc + pos[a] - ro[a]) * inv[a];
          }
          const uint32_t bcx = (uint32_t)(sc[0] * 4 + lb[0]), bcy = (uint32_t)(sc[1] * 4 + lb[1]),
                         bcz = (uint32_t)(sc[2] * 4 + lb[2]);
          const uint32_t cidx = bcx + (uint32_t)w.bd[0] * (bcy + (uint32_t)w.bd[1] * bcz);
          const uint32_t g = w.grid[cidx];
          // 2x2x2-block occupancy of a pool brick (read beside the grid
          // word): a voxel in an empty block is 0 without a pool read
          const uint64_t subs = w.sub_mask[cidx];
          const bool uni = (g & PX_UNIFORM_BIT) != 0u;
          if (uni && !bit_of(T.glass, g & 0xffu)) {
            o.key = key_of(w, (int32_t)bcx * 8 + lv[0], (int32_t)bcy * 8 + lv[1], (int32_t)bcz * 8 + lv[2]);
            o.t = px_canon0(tb);
            o.face = face_of_axis(axisB, rd);
            o.value = g & 0xffffu;
            o.steps = steps;
            done();
            return o;
          }
          // byte offset of the brick in the pool (pool bricks only; the pool is < 4 GiB)
          const uint32_t bbase = (g - 1u) * 1024u;
          float tv = tb;
          uint32_t axisV = axisB;
          const float d0 = fabsf(inv[0]), d1 = fabsf(inv[1]), d2 = fabsf(inv[2]);
          for (int k = 0; k < 25; k++) {
            steps++;
            if constexpr (SHADE) inner++;
            const uint32_t li = (uint32_t)lv[0] | ((uint32_t)lv[1] << 3) | ((uint32_t)lv[2] << 6);
            uint32_t v = 0u;
            if (uni) v = g & 0xffffu;
            else if ((subs >> (((uint32_t)lv[0] >> 1) | (((uint32_t)lv[1] << 1) & 12u) | (((uint32_t)lv[2] << 3) & 48u))) &
                     1ull)
              v = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(w.pool) + (uint64_t)(bbase + li * 2u));
            if (v != 0u) {
              const uint32_t id = v & 0xffu;
              bool hit = true;
              if (bit_of(T.glass, id)) {
                if (glassState != (0x100u | id)) {
                  if constexpr (SHADE && !SHADOW) glass_surface_glow(w.mats, *sky, trans, glow, id, axisV, rd, stp);
                  else glass_surface(w.mats, trans, id, axisV, rd, SHADOW);
                }
                glassState = 0x100u | id;
                hit = fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f;
              }
              if (hit) {
                o.key = key_of(w, (int32_t)bcx * 8 + lv[0], (int32_t)bcy * 8 + lv[1], (int32_t)bcz * 8 + lv[2]);
                o.t = px_canon0(tv);
                o.face = face_of_axis(axisV, rd);
                o.value = v;
                o.steps = steps;
                done();
                return o;
              }
            } else {
              glassState = 0u;
            }
            const float m0 = tMaxV[0], m1 = tMaxV[1], m2 = tMaxV[2];
            const bool a0 = m0 < m1 && m0 < m2;
            const bool b12 = m1 < m2;
            tv = a0 ? m0 : (b12 ? m1 : m2);
            axisV = a0 ? 0u : (b12 ? 1u : 2u);
            tMaxV[0] = a0 ? m0 + d0 : m0;
            tMaxV[1] = (!a0 && b12) ? m1 + d1 : m1;
            tMaxV[2] = (!a0 && !b12) ? m2 + d2 : m2;
            lv[0] = a0 ? lv[0] + stp[0] : lv[0];
            lv[1] = (!a0 && b12) ? lv[1] + stp[1] : lv[1];
            lv[2] = (!a0 && !b12) ? lv[2] + stp[2] : lv[2];
            if (tv > tExit) {
              o.steps = steps;
              done();
              return o;
            }
            if ((uint32_t)(lv[0] | lv[1] | lv[2]) > 7u) break;
          }
        } else {
          glassState = 0u;
        }
        {
          const float m0 = tMaxB[0], m1 = tMaxB[1], m2 = tMaxB[2];
          const bool a0 = m0 < m1 && m0 < m2;
          const bool b12 = m1 < m2;
          tb = a0 ? m0 : (b12 ? m1 : m2);
          axisB = a0 ? 0u : (b12 ? 1u : 2u);
          tMaxB[0] = a0 ? __builtin_fmaf(fabsf(inv[0]), 8.0f, m0) : m0;
          tMaxB[1] = (!a0 && b12) ? __builtin_fmaf(fabsf(inv[1]), 8.0f, m1) : m1;
          tMaxB[2] = (!a0 && !b12) ? __builtin_fmaf(fabsf(inv[2]), 8.0f, m2) : m2;
          lb[0] = a0 ? lb[0] + stp[0] : lb[0];
          lb[1] = (!a0 && b12) ? lb[1] + stp[1] : lb[1];
          lb[2] = (!a0 && !b12) ? lb[2] + stp[2] : lb[2];
        }
        if (tb > tExit) {
          o.steps = steps;
          done();
          return o;
        }
        if ((uint32_t)lb[0] > (uint32_t)hiL[0] || (uint32_t)lb[1] > (uint32_t)hiL[1] ||
            (uint32_t)lb[2] > (uint32_t)hiL[2])
          break;
      }
    } else {
      glassState = 0u;
    }
    {
      const float m0 = tMaxS[0], m1 = tMaxS[1], m2 = tMaxS[2];
      const bool a0 = m0 < m1 && m0 < m2;
      const bool b12 = m1 < m2;
      t = a0 ? m0 : (b12 ? m1 : m2);
      axisS = a0 ? 0u : (b12 ? 1u : 2u);
      tMaxS[0] = a0 ? __builtin_fmaf(fabsf(inv[0]), 32.0f, m0) : m0;
      tMaxS[1] = (!a0 && b12) ? __builtin_fmaf(fabsf(inv[1]), 32.0f, m1) : m1;
      tMaxS[2] = (!a0 && !b12) ? __builtin_fmaf(fabsf(inv[2]), 32.0f, m2) : m2;
      sc[0] = a0 ? sc[0] + stp[0] : sc[0];
      sc[1] = (!a0 && b12) ? sc[1] + stp[1] : sc[1];
      sc[2] = (!a0 && !b12) ? sc[2] + stp[2] : sc[2];
    }
    if (t > tExit || (uint32_t)sc[0] >= (uint32_t)w.sd[0] || (uint32_t)sc[1] >= (uint32_t)w.sd[1] ||
        (uint32_t)sc[2] >= (uint32_t)w.sd[2])
      break;
  }
  o.steps = steps;
  done();
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

// Shadow ray `s` of shade() (raymarch.wgsl:323-356) for a primary hit at
// t / face: slot 0 the sun, slot 1 + k light k.  Returns whether it is cast;
// origin P + N 0.5, direction and max t are written either way.
PX_DEV bool shadow_ray(const PxWorldView& w, const PxShadowDesc& sd, const float ro[3], const float rd[3], float t,
                       uint32_t face, uint32_t value, uint32_t s, float O[3], float dir[3], float& mt) {
  float N[3] = {0.0f, 0.0f, 0.0f};
  if (face < 6u) N[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  float P[3];
  for (int a = 0; a < 3; a++) {
    P[a] = ro[a] + rd[a] * t;
    O[a] = P[a] + N[a] * 0.5f;
  }
  dir[0] = dir[1] = dir[2] = 0.0f;
  mt = 0.0f;
  if (face == 0xFFu || w.mats[(value & 0xffu) * 20u + 3u] > 0.0f) return false;  // miss or emissive
  if (s == 0u) {
    const float ndl = N[0] * sd.sun[0] + N[1] * sd.sun[1] + N[2] * sd.sun[2];
    dir[0] = sd.sun[0];
    dir[1] = sd.sun[1];
    dir[2] = sd.sun[2];
    mt = 1e9f;
    return ndl > 0.0f && sd.sun_w > 0.0f;
  }
  const float* lp = sd.lp[s - 1u];
  float L[3];
  for (int a = 0; a < 3; a++) L[a] = lp[a] / sd.vs - P[a];
  const float radius = lp[3] / sd.vs;
  const float d = sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  if (d > radius) return false;
  for (int a = 0; a < 3; a++) L[a] = L[a] / d;
  const float nl = N[0] * L[0] + N[1] * L[1] + N[2] * L[2];
  dir[0] = L[0];
  dir[1] = L[1];
  dir[2] = L[2];
  mt = d - 2.0f;
  return nl > 0.0f && sd.lc_w[s - 1u] > 0.5f && sd.lamp_shadows > 0.5f;
}

// readVoxel (sim.wgsl:74-81): decoded voxel at p, 0 outside the world.
PX_DEV uint32_t read_voxel(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= w.dims[0] || y >= w.dims[1] || z >= w.dims[2]) return 0u;
  uint32_t g = w.grid[(uint32_t)(x >> 3) + (uint32_t)w.bd[0] * ((uint32_t)(y >> 3) + (uint32_t)w.bd[1] * (uint32_t)(z >> 3))];
  if (g == 0u) return 0u;
  return px_brick_voxel(w.pool, g, x & 7, y & 7, z & 7);
}

// The debris DDA of particleSim (sim.wgsl:1110-1145) after its inside test:
// from `cell` (the start cell, advanced in place) along the displacement mv
// of length dist > 1e-6, at most 48 cells, stopping at the first solid cell
// entered with t <= dist.  Returns the hit axis (-1 = none); t, the solid
// voxel's value and the loop turns begun go to the out parameters, and dir
// gets the direction the loop used (mv / dist with components under 1e-6 in
// magnitude replaced by 1e-6).
PX_DEV int segment_walk(const PxWorldView& w, int32_t cell[3], const float pos[3], const float mv[3], float dist,
                        float dir[3], float* t_hit, uint32_t* v_hit, uint32_t* steps_out) {
  float inv[3], tMax[3], tDelta[3];
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
    cell[0] = nx; cell[1] = ny; cell[2] = nz;
    if (v != 0u) {
      *t_hit = tn;
      *v_hit = v;
      *steps_out = steps;
      return axis;
    }
    if (axis == 0) tMax[0] += tDelta[0];
    else if (axis == 1) tMax[1] += tDelta[1];
    else tMax[2] += tDelta[2];
  }
  *steps_out = steps;
  return -1;
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

}  // namespace px
