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
// tinted absorption.  The normal is -sign(rd) along `axis` (zero for axis
// 3), and cos = |dot(n, rd)| is taken over all three components, as the
// WGSL does, so a NaN component of rd gives a NaN cosine.
PX_DEV void glass_surface(const float* __restrict__ mats, float trans[3], uint32_t id, uint32_t axis,
                          const float rd[3], const int32_t stp[3], bool shadow) {
  const float* A = mats + id * 20u;
  float g = mats[id * 20u + 13u];
  if (!shadow) {
    // (selects, not n[axis]: a runtime index would put the arrays in scratch)
    const float n[3] = {axis == 0u ? -(float)stp[0] : 0.0f, axis == 1u ? -(float)stp[1] : 0.0f,
                        axis == 2u ? -(float)stp[2] : 0.0f};
    const float cosT = fabsf(dot3f(n, rd));
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
  const float n[3] = {axis == 0u ? -(float)stp[0] : 0.0f, axis == 1u ? -(float)stp[1] : 0.0f,
                      axis == 2u ? -(float)stp[2] : 0.0f};
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

// Cell coordinates inside a super cell (brick level) or a brick (voxel
// level) are packed as x | y << 10 | z << 20, so one DDA step moves them
// with a single add of the taken axis's packed step (+-1, +-1024,
// +-1048576).  A coordinate that leaves its range sets a guard bit of its
// field (a step below 0 borrows, which sets every high bit of the field), so
// one AND tests all three axes.
constexpr uint32_t kPkY = 10u, kPkZ = 20u;
constexpr uint32_t kVoxGuard = 0x3F8u | (0x3F8u << kPkY) | (0x3F8u << kPkZ);    // fields outside 0..7
constexpr uint32_t kBrickGuard = 0x3FCu | (0x3FCu << kPkY) | (0x3FCu << kPkZ);  // fields outside 0..3

// The zero normal (a ray starting inside the world) as a packed step; it is
// only ever read as a face, never added to a coordinate.
constexpr int32_t kPkNoAxis = 1 << 30;

// Axis code (0..2, 3 = zero normal) of a packed step.  A step of 0 (a NaN
// direction component, whose sign is 0) gives the zero normal -sign(rd) = 0.
PX_DEV uint32_t pk_axis(int32_t d) {
  const uint32_t a = (uint32_t)(d < 0 ? -d : d);
  return a == 0u || d == kPkNoAxis ? 3u : (a >= (1u << kPkZ) ? 2u : (a >= (1u << kPkY) ? 1u : 0u));
}

// Bit x + 4 y + 16 z of a brick-level packed coordinate (fields 0..3): one
// 24-bit multiply gathers the three fields at bit 16 without carries.
PX_DEV uint32_t pk_brick_bit(uint32_t pb) {
  return (__umul24(pb & (3u | (3u << kPkY) | (3u << kPkZ)), 0x10101u) >> 16) & 63u;
}
// Bit (x/2) + 4 (y/2) + 16 (z/2) of a voxel-level packed coordinate (fields
// 0..7): the 2x2x2-block index of the sub-block mask, gathered at bit 17.
PX_DEV uint32_t pk_block_bit(uint32_t p) {
  return (__umul24(p & ((6u) | (6u << kPkY) | (6u << kPkZ)), 0x10101u) >> 17) & 63u;
}
// Local voxel index x | y << 3 | z << 6 of a voxel-level packed coordinate,
// gathered at bit 14.
// (The mask keeps the product a 24-bit multiply.)
PX_DEV uint32_t pk_local_index(uint32_t p) {
  return (__umul24(p & (7u | (7u << kPkY) | (7u << kPkZ)), 0x4081u) >> 14) & 511u;
}

// One DDA step: the WGSL's if / else-if chain as selects - axis 0 when
// m0 < m1 && m0 < m2, else axis 1 when m1 < m2, else axis 2 (ties go to the
// later axis).  The taken axis's crossing time advances by its increment
// dl (|inv| scaled by the level: an exact power-of-two scaling, so the add
// rounds once as the WGSL's tMax += tDelta does); t receives the crossing
// time and d the packed step of the axis.
PX_DEV void dda_step(float tm[3], float dl0, float dl1, float dl2, int32_t ps0, int32_t ps1, int32_t ps2, float& t,
                     int32_t& d) {
  const float m0 = tm[0], m1 = tm[1], m2 = tm[2];
  const bool a0 = m0 < m1 && m0 < m2;
  const bool b12 = m1 < m2;
  t = a0 ? m0 : (b12 ? m1 : m2);
  d = a0 ? ps0 : (b12 ? ps1 : ps2);
  tm[0] = a0 ? m0 + dl0 : m0;
  tm[1] = (!a0 && b12) ? m1 + dl1 : m1;
  tm[2] = (!a0 && !b12) ? m2 + dl2 : m2;
}

// dda_step for crossing times that are never NaN, with the increments
// |inv| * scale (scale 1, 8 or 32: the product is exact, so the fma rounds
// once as the WGSL's tMax += tDelta does): the minimum is the time
// the if / else-if chain selects (ties go to the later axis: axis 2 when m2
// is a minimum, else axis 1 when m1 is, else axis 0).  A tie between +0 and
// -0 may return either zero; every use of t (comparisons, t + 1e-3, the
// reported t with +0 for zero) is the same for both.
PX_DEV void dda_min(float tm[3], float dv0, float dv1, float dv2, float scale, int32_t ps0, int32_t ps1, int32_t ps2,
                    float& t, int32_t& d) {
  const float m0 = tm[0], m1 = tm[1], m2 = tm[2];
  const float mn = fminf(m0, fminf(m1, m2));
  const bool e2 = m2 == mn;
  const bool e1 = !e2 && m1 == mn;
  t = mn;
  d = e2 ? ps2 : (e1 ? ps1 : ps0);
  tm[0] = (!e2 && !e1) ? __builtin_fmaf(dv0, scale, m0) : m0;
  tm[1] = e1 ? __builtin_fmaf(dv1, scale, m1) : m1;
  tm[2] = e2 ? __builtin_fmaf(dv2, scale, m2) : m2;
}

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
  // (three scalars, not an array: a select among array elements may be
  // turned into a runtime-indexed scratch load)
  int32_t ps0 = stp[0], ps1 = stp[1] * (1 << kPkY), ps2 = stp[2] * (1 << kPkZ);
  // (opaque, so that the fast path's unit steps are taken from these by a
  // shift instead of keeping stp[] live, and spilled, beside them)
  PX_OPAQUE(ps0);
  PX_OPAQUE(ps1);
  PX_OPAQUE(ps2);
  const float dv0 = fabsf(inv[0]), dv1 = fabsf(inv[1]), dv2 = fabsf(inv[2]);
  // the normal of the entry face as a packed step (only its axis is used)
  int32_t dS = kPkNoAxis;
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) dS = ps0;
    else if (tmn[1] >= tmn[2]) dS = ps1;
    else dS = ps2;
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
  const char* __restrict__ poolb = reinterpret_cast<const char*>(w.pool);

  // The voxel at packed coordinate p of brick (bx, by, bz) as the march's
  // hit: key, t, face of the step that entered it, value.
  auto hit_out = [&](uint32_t bx, uint32_t by, uint32_t bz, uint32_t p, float th, int32_t d, uint32_t v) {
    o.key = key_of(w, (int32_t)(bx * 8u + (p & 1023u)), (int32_t)(by * 8u + ((p >> kPkY) & 1023u)),
                   (int32_t)(bz * 8u + ((p >> kPkZ) & 1023u)));
    o.t = px_canon0(th);
    o.face = face_of_axis(pk_axis(d), rd);
    o.value = v;
    o.steps = steps;
    done();
  };

  if (__builtin_isfinite(ro[0]) && __builtin_isfinite(ro[1]) && __builtin_isfinite(ro[2]) &&
      __builtin_isfinite(rd[0]) && __builtin_isfinite(rd[1]) && __builtin_isfinite(rd[2])) {
    // Finite origin and direction: every crossing time is a number (never
    // NaN) and every axis steps by +-1.  So a brick loop ends within 10 turns
    // and a voxel loop within 22 (a coordinate leaves its range after at most
    // 3 + 3 + 3 + 1 or 7 + 7 + 7 + 1 steps): the WGSL's caps of 14 and 25
    // never bind and are not counted, and each DDA step can take the
    // minimum crossing time as t (dda_min).  The glass state is kept as the
    // value of the last voxel visited (0 after air or an empty brick or
    // super cell): a non-zero one is always glass, since any other solid
    // voxel ends the march.  A cell index clamped into [lo, hi] is taken as
    // lo + med3(floor(x) - lo, 0, hi - lo) in FP32: floor(x) - lo is exact
    // whenever the result is not clamped, and x is a number or +-inf.
    const char* __restrict__ gridb = reinterpret_cast<const char*>(w.grid);
    const char* __restrict__ subb = reinterpret_cast<const char*>(w.sub_mask);
    const char* __restrict__ sbmb = reinterpret_cast<const char*>(w.super_bm);
    uint32_t prev = 0u;
    // the hit: brick, packed voxel, t, entry step, value
    uint32_t hx = 0u, hy = 0u, hz = 0u, hp = 0u, hv = 0u;
    float ht = 0.0f;
    int32_t hd = 0;
    for (int i = 0; i < 256; i++) {
      steps++;
      const uint32_t sidx = px_mad24((uint32_t)sc[2], (uint32_t)w.sd[1], (uint32_t)sc[1]);
      const uint32_t sidx2 = px_mad24(sidx, (uint32_t)w.sd[0], (uint32_t)sc[0]);
      if (bit_of(T.sup, sidx2)) {
        const uint64_t bm = *reinterpret_cast<const uint64_t*>(sbmb + sidx2 * 8u);
        uint32_t pb = 0u, badd = 0u;
        float tMaxB[3];
        const float tt = t + 1e-3f;
        for (int a = 0; a < 3; a++) {
          const int32_t sLo = sc[a] * 4;
          const int32_t hiL = min(3, w.bd[a] - 1 - sLo);
          const float lbf = __builtin_amdgcn_fmed3f(floorf((ro[a] + rd[a] * tt) / 8.0f) - (float)sLo, 0.0f,
                                                    (float)hiL);
          pb |= (uint32_t)(int32_t)lbf << (kPkY * a);
          badd |= (uint32_t)(3 - hiL) << (kPkY * a);
          tMaxB[a] = (((float)sLo + lbf + pos[a]) * 8.0f - ro[a]) * inv[a];
        }
        float tb = t;
        int32_t dB = dS;
        for (;;) {
          steps++;
          if ((bm >> pk_brick_bit(pb)) & 1ull) {
            const uint32_t bcx = (uint32_t)sc[0] * 4u + (pb & 1023u);
            const uint32_t bcy = (uint32_t)sc[1] * 4u + ((pb >> kPkY) & 1023u);
            const uint32_t bcz = (uint32_t)sc[2] * 4u + (pb >> kPkZ);
            const uint32_t cidx = px_mad24(px_mad24(bcz, (uint32_t)w.bd[1], bcy), (uint32_t)w.bd[0], bcx);
            const uint32_t g = *reinterpret_cast<const uint32_t*>(gridb + cidx * 4u);
            const uint64_t subs = *reinterpret_cast<const uint64_t*>(subb + cidx * 8u);
            const uint32_t bco[3] = {bcx, bcy, bcz};
            const float tt2 = tb + 1e-3f;
            uint32_t p = 0u;
            float tMaxV[3];
            for (int a = 0; a < 3; a++) {
              const float lo = (float)(bco[a] * 8u);
              const float lvf = __builtin_amdgcn_fmed3f(floorf(ro[a] + rd[a] * tt2) - lo, 0.0f, 7.0f);
              p |= (uint32_t)(int32_t)lvf << (kPkY * a);
              tMaxV[a] = ((lo + lvf) + pos[a] - ro[a]) * inv[a];
            }
            float tv = tb;
            int32_t dV = dB;
            const bool uni = (g & PX_UNIFORM_BIT) != 0u;
            if (uni && !bit_of(T.glass, g & 0xffu)) {
              hx = bcx, hy = bcy, hz = bcz, hp = p, ht = tv, hd = dV, hv = g & 0xffffu;
              goto hit;
            }
            const uint64_t blocks = uni ? 0ull : subs;
            const uint32_t vuni = uni ? (g & 0xffffu) : 0u;
            const uint32_t bbase = (g - 1u) * 1024u;
            // one exit: a lane that hits keeps its voxel (no step), so after
            // the loop tv > tExit means a miss, a coordinate outside the brick
            // means the next brick, and anything else the hit of voxel p
            uint32_t v;
            for (;;) {
              steps++;
              if constexpr (SHADE) inner++;
              v = vuni;
              if ((blocks >> pk_block_bit(p)) & 1ull)
                v = *reinterpret_cast<const uint16_t*>(poolb + (bbase + pk_local_index(p) * 2u));
              bool hit = false;
              if (v != 0u) {
                const uint32_t id = v & 0xffu;
                hit = true;
                if (bit_of(T.glass, id)) {
                  if (prev == 0u || (prev & 0xffu) != id) {
                    const int32_t st[3] = {ps0, ps1 >> kPkY, ps2 >> kPkZ};
                    if constexpr (SHADE && !SHADOW)
                      glass_surface_glow(w.mats, *sky, trans, glow, id, pk_axis(dV), rd, st);
                    else
                      glass_surface(w.mats, trans, id, pk_axis(dV), rd, st, SHADOW);
                  }
                  hit = fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f;
                }
              }
              if (!hit) {
                prev = v;
                dda_min(tMaxV, dv0, dv1, dv2, 1.0f, ps0, ps1, ps2, tv, dV);
                p += (uint32_t)dV;
              }
              if (hit || tv > tExit || (p & kVoxGuard)) break;
            }
            PX_OPAQUE(tv);
            PX_OPAQUE(p);
            if (tv > tExit) goto miss;
            if (!(p & kVoxGuard)) {
              hx = bcx, hy = bcy, hz = bcz, hp = p, ht = tv, hd = dV, hv = v;
              goto hit;
            }
          } else {
            prev = 0u;
          }
          dda_min(tMaxB, dv0, dv1, dv2, 8.0f, ps0, ps1, ps2, tb, dB);
          pb += (uint32_t)dB;
          if (tb > tExit) goto miss;
          if ((pb | (pb + badd)) & kBrickGuard) break;
        }
      } else {
        prev = 0u;
      }
      {
        const float m0 = tMaxS[0], m1 = tMaxS[1], m2 = tMaxS[2];
        const float mn = fminf(m0, fminf(m1, m2));
        const bool e2 = m2 == mn;
        const bool e1 = !e2 && m1 == mn;
        t = mn;
        dS = e2 ? ps2 : (e1 ? ps1 : ps0);
        tMaxS[0] = (!e2 && !e1) ? __builtin_fmaf(dv0, 32.0f, m0) : m0;
        tMaxS[1] = e1 ? __builtin_fmaf(dv1, 32.0f, m1) : m1;
        tMaxS[2] = e2 ? __builtin_fmaf(dv2, 32.0f, m2) : m2;
        // (the unit steps come from the packed ones, which stay live anyway)
        sc[0] = (!e2 && !e1) ? sc[0] + ps0 : sc[0];
        sc[1] = e1 ? sc[1] + (ps1 >> kPkY) : sc[1];
        sc[2] = e2 ? sc[2] + (ps2 >> kPkZ) : sc[2];
      }
      if (t > tExit || (uint32_t)sc[0] >= (uint32_t)w.sd[0] || (uint32_t)sc[1] >= (uint32_t)w.sd[1] ||
          (uint32_t)sc[2] >= (uint32_t)w.sd[2])
        break;
    }
  miss:
    o.steps = steps;
    done();
    return o;
  hit:
    hit_out(hx, hy, hz, hp, ht, hd, hv);
    return o;
  }

  // A NaN or infinite origin or direction: the WGSL's loops with their caps
  // (a NaN component steps by 0, so only the caps end those loops).  The
  // unit steps are taken back from the packed ones, so that stp[] is not
  // live (and spilled) through the setup both paths share.
  stp[0] = ps0;
  stp[1] = ps1 >> kPkY;
  stp[2] = ps2 >> kPkZ;

  for (int i = 0; i < 256; i++) {
    steps++;
    const uint32_t sidx = (uint32_t)sc[0] + __umul24((uint32_t)w.sd[0], (uint32_t)sc[1] + __umul24((uint32_t)w.sd[1], (uint32_t)sc[2]));
    if (bit_of(T.sup, sidx)) {
      const uint64_t bm = w.super_bm[sidx];
      uint32_t pb = 0u, badd = 0u;  // packed brick coordinate; 3 - (last index) per axis
      float tMaxB[3];
      for (int a = 0; a < 3; a++) {
        const int32_t sLo = sc[a] * 4;
        const int32_t sHi = min(sLo + 3, w.bd[a] - 1);
        const int32_t bc = px_clampi(px_f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), sLo, sHi);
        pb |= (uint32_t)(bc - sLo) << (kPkY * a);
        badd |= (uint32_t)(3 - (sHi - sLo)) << (kPkY * a);
        tMaxB[a] = (((float)bc + pos[a]) * 8.0f - ro[a]) * inv[a];
      }
      float tb = t;
      int32_t dB = dS;
      for (int j = 0; j < 14; j++) {
        steps++;
        if ((bm >> pk_brick_bit(pb)) & 1ull) {
          const uint32_t bcx = (uint32_t)sc[0] * 4u + (pb & 1023u);
          const uint32_t bcy = (uint32_t)sc[1] * 4u + ((pb >> kPkY) & 1023u);
          const uint32_t bcz = (uint32_t)sc[2] * 4u + (pb >> kPkZ);
          const uint32_t cidx = bcx + __umul24((uint32_t)w.bd[0], bcy + __umul24((uint32_t)w.bd[1], bcz));
          const uint32_t g = w.grid[cidx];
          // 2x2x2-block occupancy of a pool brick (read beside the grid
          // word): a voxel in an empty block is 0 without a pool read
          const uint64_t subs = w.sub_mask[cidx];
          const uint32_t bco[3] = {bcx, bcy, bcz};
          uint32_t p = 0u;
          float tMaxV[3];
          for (int a = 0; a < 3; a++) {
            const int32_t lo = (int32_t)bco[a] * 8;
            const int32_t vc = px_clampi(px_f2i(floorf(ro[a] + rd[a] * (tb + 1e-3f))), lo, lo + 7);
            p |= (uint32_t)(vc - lo) << (kPkY * a);
            tMaxV[a] = ((float)vc + pos[a] - ro[a]) * inv[a];
          }
          const bool uni = (g & PX_UNIFORM_BIT) != 0u;
          if (uni && !bit_of(T.glass, g & 0xffu)) {
            o.key = key_of(w, (int32_t)(bcx * 8u + (p & 1023u)), (int32_t)(bcy * 8u + ((p >> kPkY) & 1023u)),
                           (int32_t)(bcz * 8u + (p >> kPkZ)));
            o.t = px_canon0(tb);
            o.face = face_of_axis(pk_axis(dB), rd);
            o.value = g & 0xffffu;
            o.steps = steps;
            done();
            return o;
          }
          // a uniform (glass) brick reads its value from the grid word and
          // no pool voxel; a pool brick starts at byte (g - 1) * 1024 of
          // the pool (the pool is < 4 GiB)
          const uint64_t blocks = uni ? 0ull : subs;
          const uint32_t vuni = uni ? (g & 0xffffu) : 0u;
          const uint32_t bbase = (g - 1u) * 1024u;
          float tv = tb;
          int32_t dV = dB;
          for (int k = 0; k < 25; k++) {
            steps++;
            if constexpr (SHADE) inner++;
            uint32_t v = vuni;
            if ((blocks >> pk_block_bit(p)) & 1ull)
              v = *reinterpret_cast<const uint16_t*>(poolb + (bbase + pk_local_index(p) * 2u));
            if (v != 0u) {
              const uint32_t id = v & 0xffu;
              bool hit = true;
              if (bit_of(T.glass, id)) {
                if (glassState != (0x100u | id)) {
                  if constexpr (SHADE && !SHADOW)
                    glass_surface_glow(w.mats, *sky, trans, glow, id, pk_axis(dV), rd, stp);
                  else
                    glass_surface(w.mats, trans, id, pk_axis(dV), rd, stp, SHADOW);
                }
                glassState = 0x100u | id;
                hit = fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f;
              }
              if (hit) {
                o.key = key_of(w, (int32_t)(bcx * 8u + (p & 1023u)), (int32_t)(bcy * 8u + ((p >> kPkY) & 1023u)),
                               (int32_t)(bcz * 8u + (p >> kPkZ)));
                o.t = px_canon0(tv);
                o.face = face_of_axis(pk_axis(dV), rd);
                o.value = v;
                o.steps = steps;
                done();
                return o;
              }
            } else {
              glassState = 0u;
            }
            dda_step(tMaxV, dv0, dv1, dv2, ps0, ps1, ps2, tv, dV);
            p += (uint32_t)dV;
            if (tv > tExit) {
              o.steps = steps;
              done();
              return o;
            }
            if (p & kVoxGuard) break;
          }
        } else {
          glassState = 0u;
        }
        dda_step(tMaxB, dv0 * 8.0f, dv1 * 8.0f, dv2 * 8.0f, ps0, ps1, ps2, tb, dB);
        pb += (uint32_t)dB;
        if (tb > tExit) {
          o.steps = steps;
          done();
          return o;
        }
        // outside the super cell: a field below 0 or above 3 (pb's own guard
        // bits), or above the last brick of a super cell the world cuts
        if ((pb | (pb + badd)) & kBrickGuard) break;
      }
    } else {
      glassState = 0u;
    }
    {
      const float m0 = tMaxS[0], m1 = tMaxS[1], m2 = tMaxS[2];
      const bool a0 = m0 < m1 && m0 < m2;
      const bool b12 = m1 < m2;
      t = a0 ? m0 : (b12 ? m1 : m2);
      dS = a0 ? ps0 : (b12 ? ps1 : ps2);
      tMaxS[0] = a0 ? __builtin_fmaf(dv0, 32.0f, m0) : m0;
      tMaxS[1] = (!a0 && b12) ? __builtin_fmaf(dv1, 32.0f, m1) : m1;
      tMaxS[2] = (!a0 && !b12) ? __builtin_fmaf(dv2, 32.0f, m2) : m2;
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
  const float* lp = sd.lpv[s - 1u];  // lp / vs
  float L[3];
  for (int a = 0; a < 3; a++) L[a] = lp[a] - P[a];
  const float radius = lp[3];
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
