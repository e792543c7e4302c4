/*
 * px_oracle.c - CPU oracle (TEST INFRASTRUCTURE ONLY, see px_oracle.h).
 *
 * A literal restatement, in plain C with FP32 arithmetic rounded once per
 * operation in source order (-ffp-contract=off), of:
 *   march / safeDir / glassSurface      src/render/shaders/raymarch.wgsl:22-162
 *   brickVoxel                          src/render/shaders/render_common.wgsl:25-31
 *   shade() shadow-ray setup            src/render/shaders/raymarch.wgsl:304-356
 *   hashu / hash3i / rnd                src/render/shaders/shared.wgsl:65-81
 *   readVoxel, traceSolid               src/render/shaders/sim.wgsl:74-81, 296-357
 *   segment DDA of particleSim          src/render/shaders/sim.wgsl:1110-1145
 *   removes / reach / impactJobs        src/render/shaders/sim.wgsl:229-248, 414-443
 *   allocBricks, carve, carveEffects    src/render/shaders/sim.wgsl:459-579
 *   pane flood (floodStep)              src/render/shaders/sim.wgsl:599-618
 *   maint compaction                    src/render/shaders/sim.wgsl:1219-1274
 *   spawn / frameBegin / particleSim    src/render/shaders/sim.wgsl:166-170, 363-369, 1088-1173
 *   voxel writes (request, apply)       src/render/shaders/sim.wgsl:94-102, 127-146, 1178-1203
 *   extraction, falling chunks          src/render/shaders/sim.wgsl:172-178, 274-284, 862-923, 951-1075
 *   scene rasteriser apply()            src/world/shapes.ts:24-143, 334-414
 * and (px_shade.c) the shading of fsMain: sky, material colour, AO, shade(),
 * reflections and the debug views; plus a textbook breadth-first 6-connected labelling for the islands.
 *
 * Choices the WGSL leaves to the implementation, fixed here (the HIP library
 * makes the same choices):
 *   pow(x, 5.0)  = ((x*x)*(x*x))*x
 *   mix(a, b, t) = a*(1-t) + b*t        (WGSL spec formula)
 *   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   length(a)    = sqrt(dot(a, a))
 *   f32 -> i32   = truncation toward zero, saturating
 *   a zero t is reported as +0.
 */
#include "px_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NONE 0xFFFFFFFFu
#define UNIFORM_BIT 0x80000000u

/* ---------------------------------------------------------------------- */
/* scalar helpers */

static inline float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
/* WGSL step(edge, x) = 1 if x >= edge else 0 */
static inline float fstep(float edge, float x) { return x >= edge ? 1.0f : 0.0f; }
static inline int32_t f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
static inline int32_t iclamp(int32_t v, int32_t lo, int32_t hi) {
  int32_t m = v > lo ? v : lo;
  return m < hi ? m : hi;
}
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline float pow5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x;
}
static inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline float canon0(float t) { return t == 0.0f ? 0.0f : t; }

static inline const float* matA(const float* m, uint32_t id) { return m + id * 20u; }
static inline const float* matC(const float* m, uint32_t id) { return m + id * 20u + 8u; }
static inline const float* matD(const float* m, uint32_t id) { return m + id * 20u + 12u; }
static inline const float* matE(const float* m, uint32_t id) { return m + id * 20u + 16u; }
static inline int isGlass(const float* m, uint32_t id) { return matD(m, id)[1] > 0.0f; }

uint32_t o_hashu(uint32_t x) {
  uint32_t h = x;
  h ^= h >> 16u;
  h *= 0x7feb352du;
  h ^= h >> 15u;
  h *= 0x846ca68bu;
  h ^= h >> 16u;
  return h;
}

uint32_t o_hash3i(int32_t x, int32_t y, int32_t z) {
  return o_hashu(((uint32_t)x * 0x9E3779B1u) ^
                 o_hashu(((uint32_t)y * 0x85EBCA77u) ^ o_hashu((uint32_t)z * 0xC2B2AE3Du)));
}

float o_rnd(uint32_t seed) { return (float)(o_hashu(seed) & 0xffffffu) / 16777216.0f; }

/* face id of an axis-aligned normal vector (dirOf order), 6 for zero */
static uint8_t face_of(const float* n) {
  if (n[0] > 0.0f) return 0;
  if (n[0] < 0.0f) return 1;
  if (n[1] > 0.0f) return 2;
  if (n[1] < 0.0f) return 3;
  if (n[2] > 0.0f) return 4;
  if (n[2] < 0.0f) return 5;
  return 6;
}

static inline uint32_t key_of(const o_world* w, const int32_t* p) {
  return (uint32_t)p[0] + (uint32_t)w->dx * ((uint32_t)p[1] + (uint32_t)w->dy * (uint32_t)p[2]);
}

/* render_common.wgsl:25-31 */
static inline uint32_t brickVoxel(const o_world* w, uint32_t g, int32_t lx, int32_t ly, int32_t lz) {
  if ((g & UNIFORM_BIT) != 0u) return g & 0xffffu;
  uint32_t li = (uint32_t)(lx | (ly << 3) | (lz << 6));
  uint32_t idx = (g - 1u) * 512u + li;
  uint32_t word = w->pool[idx >> 1u];
  return (word >> ((idx & 1u) * 16u)) & 0xffffu;
}

/* sim.wgsl:74-81 */
static inline uint32_t readVoxel(const o_world* w, const int32_t* p) {
  if (p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] >= w->dx || p[1] >= w->dy || p[2] >= w->dz) return 0u;
  int32_t bx = w->dx >> 3, by = w->dy >> 3;
  uint32_t g = w->grid[(uint32_t)(p[0] >> 3) + (uint32_t)bx * ((uint32_t)(p[1] >> 3) + (uint32_t)by * (uint32_t)(p[2] >> 3))];
  if (g == 0u) return 0u;
  if ((g & UNIFORM_BIT) != 0u) return g & 0xffffu;
  uint32_t li = (uint32_t)((p[0] & 7) | ((p[1] & 7) << 3) | ((p[2] & 7) << 6));
  uint32_t idx = (g - 1u) * 512u + li;
  return (w->pool[idx >> 1u] >> ((idx & 1u) * 16u)) & 0xffffu;
}

static void set_hit(const o_world* w, o_hit* out, float t, const int32_t* vc, const float* n, uint32_t v,
                    uint32_t steps) {
  out->key = key_of(w, vc);
  out->t = canon0(t);
  out->steps = steps;
  out->face = face_of(n);
  out->mat = (uint8_t)(v & 0xffu);
  out->value_hi = (uint16_t)((v >> 8) & 0xffu);
}

static void set_miss(o_hit* out, uint32_t steps) {
  out->key = NONE;
  out->t = 0.0f;
  out->steps = steps;
  out->face = 0xFF;
  out->mat = 0;
  out->value_hi = 0;
}

/* raymarch.wgsl:26-38.  glow may be NULL (no shading output); with glow the
 * camera-ray branch adds the Fresnel-weighted sky and the sun glint. */
static void glassSurface(const float* mats, const float* u52, float* trans, float* glow, uint32_t id, const float* n,
                         const float* rd, int shadow) {
  const float* tint = matA(mats, id);
  float g = matD(mats, id)[1];
  if (!shadow) {
    float cosT = fabsf(dot3(n, rd));
    float F = 0.04f + 0.96f * pow5(1.0f - cosT);
    if (glow) {
      float r[3], sky[3];
      float dn = dot3(n, rd);
      for (int a = 0; a < 3; a++) r[a] = rd[a] - 2.0f * dn * n[a];
      float sr = dot3(r, u52 + 4);
      float p = o_powf(sr < 0.0f ? 0.0f : sr, matC(mats, id)[3]);
      o_sky_color(u52, r, sky);
      for (int a = 0; a < 3; a++) {
        float glint = p * u52[8 + a] * u52[7] * 2.0f;
        glow[a] = glow[a] + trans[a] * (F * sky[a] * 0.35f + glint * matC(mats, id)[2]);
      }
    }
    for (int a = 0; a < 3; a++) trans[a] = trans[a] * (1.0f - F);
  }
  for (int a = 0; a < 3; a++) {
    float m = 1.0f * (1.0f - 0.6f) + tint[a] * 0.6f;
    trans[a] = trans[a] * (m * g);
  }
}

/* ---------------------------------------------------------------------- */
/* march: raymarch.wgsl:44-162 */

void o_march_ex(const o_world* w, const float* u52, const float ro[3], const float rdIn[3], float maxT, int shadow,
                o_hit* out, float* trans_out, float* glow_out, uint32_t* outer_out, uint32_t* inner_out) {
  const float* mats = w->mats;
  float trans[3] = {1.0f, 1.0f, 1.0f};
  float glow[3] = {0.0f, 0.0f, 0.0f};
  float* gl = (glow_out && u52) ? glow : NULL;
  uint32_t outer = 0, inner = 0;
  float rd[3], inv[3], t0[3], t1[3], tmn[3], tmx[3];
  float dims[3] = {(float)w->dx, (float)w->dy, (float)w->dz};
  for (int a = 0; a < 3; a++) {
    rd[a] = fabsf(rdIn[a]) < 1e-6f ? fsign(rdIn[a] + 1e-30f) * 1e-6f : rdIn[a];
    inv[a] = 1.0f / rd[a];
    t0[a] = (0.0f - ro[a]) * inv[a];
    t1[a] = (dims[a] - ro[a]) * inv[a];
    tmn[a] = fminf(t0[a], t1[a]);
    tmx[a] = fmaxf(t0[a], t1[a]);
  }
  float t = fmaxf(fmaxf(tmn[0], tmn[1]), fmaxf(tmn[2], 0.0f));
  float tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
#define MARCH_OUTPUTS()                                   \
  do {                                                    \
    if (trans_out) memcpy(trans_out, trans, sizeof trans); \
    if (glow_out) memcpy(glow_out, glow, sizeof glow);     \
    if (outer_out) *outer_out = outer;                     \
    if (inner_out) *inner_out = inner;                     \
  } while (0)
#define RETURN_MISS()                                     \
  do {                                                    \
    set_miss(out, outer + inner);                         \
    MARCH_OUTPUTS();                                      \
    return;                                               \
  } while (0)
#define RETURN_HIT(T, VC, N, V)                           \
  do {                                                    \
    set_hit(w, out, (T), (VC), (N), (V), outer + inner);  \
    MARCH_OUTPUTS();                                      \
    return;                                               \
  } while (0)
  if (t >= tExit) RETURN_MISS();

  int32_t stp[3];
  float stpF[3], pos[3];
  for (int a = 0; a < 3; a++) {
    stp[a] = f2i(fsign(rd[a]));
    stpF[a] = (float)stp[a];
    pos[a] = fstep(0.0f, rd[a]);
  }
  float normal[3] = {0.0f, 0.0f, 0.0f};
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) normal[0] = -stpF[0];
    else if (tmn[1] >= tmn[2]) normal[1] = -stpF[1];
    else normal[2] = -stpF[2];
  }

  int32_t bd[3] = {w->dx >> 3, w->dy >> 3, w->dz >> 3};
  int32_t sd[3] = {(bd[0] + 3) >> 2, (bd[1] + 3) >> 2, (bd[2] + 3) >> 2};
  int32_t sc[3];
  float tDeltaS[3], tMaxS[3], tDeltaB[3], tDeltaV[3];
  for (int a = 0; a < 3; a++) {
    float p0 = ro[a] + rd[a] * (t + 1e-3f);
    sc[a] = iclamp(f2i(floorf(p0 / 32.0f)), 0, sd[a] - 1);
    tDeltaS[a] = fabsf(inv[a]) * 32.0f;
    tMaxS[a] = (((float)sc[a] + pos[a]) * 32.0f - ro[a]) * inv[a];
    tDeltaB[a] = fabsf(inv[a]) * 8.0f;
    tDeltaV[a] = fabsf(inv[a]);
  }
  int inGlass = 0;
  uint32_t glassId = 0u;

  for (int i = 0; i < 256; i++) {
    outer += 1u;
    if (w->super_grid[(uint32_t)(sc[0] + sd[0] * (sc[1] + sd[1] * sc[2]))] != 0u) {
      int32_t sLo[3], sHi[3], bc[3];
      float tMaxB[3];
      for (int a = 0; a < 3; a++) {
        sLo[a] = sc[a] * 4;
        sHi[a] = imin(sLo[a] + 3, bd[a] - 1);
        bc[a] = iclamp(f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), sLo[a], sHi[a]);
        tMaxB[a] = (((float)bc[a] + pos[a]) * 8.0f - ro[a]) * inv[a];
      }
      float tb = t;
      float nb[3] = {normal[0], normal[1], normal[2]};
      for (int j = 0; j < 14; j++) {
        outer += 1u;
        uint32_t g = w->grid[(uint32_t)(bc[0] + bd[0] * (bc[1] + bd[1] * bc[2]))];
        if (g != 0u) {
          int32_t lo[3], vc[3];
          for (int a = 0; a < 3; a++) {
            lo[a] = bc[a] * 8;
            vc[a] = iclamp(f2i(floorf(ro[a] + rd[a] * (tb + 1e-3f))), lo[a], lo[a] + 7);
          }
          if ((g & UNIFORM_BIT) != 0u && !isGlass(mats, g & 0xffu)) RETURN_HIT(tb, vc, nb, g & 0xffffu);
          float tMaxV[3];
          for (int a = 0; a < 3; a++) tMaxV[a] = ((float)vc[a] + pos[a] - ro[a]) * inv[a];
          float tv = tb;
          float n[3] = {nb[0], nb[1], nb[2]};
          for (int k = 0; k < 25; k++) {
            inner += 1u;
            uint32_t v = brickVoxel(w, g, vc[0] - lo[0], vc[1] - lo[1], vc[2] - lo[2]);
            if (v != 0u) {
              uint32_t id = v & 0xffu;
              if (isGlass(mats, id)) {
                if (!inGlass || id != glassId) glassSurface(mats, u52, trans, gl, id, n, rd, shadow);
                inGlass = 1;
                glassId = id;
                if (fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f) RETURN_HIT(tv, vc, n, v);
              } else {
                RETURN_HIT(tv, vc, n, v);
              }
            } else {
              inGlass = 0;
            }
            if (tMaxV[0] < tMaxV[1] && tMaxV[0] < tMaxV[2]) {
              tv = tMaxV[0]; vc[0] += stp[0]; tMaxV[0] += tDeltaV[0];
              n[0] = -stpF[0]; n[1] = 0.0f; n[2] = 0.0f;
            } else if (tMaxV[1] < tMaxV[2]) {
              tv = tMaxV[1]; vc[1] += stp[1]; tMaxV[1] += tDeltaV[1];
              n[0] = 0.0f; n[1] = -stpF[1]; n[2] = 0.0f;
            } else {
              tv = tMaxV[2]; vc[2] += stp[2]; tMaxV[2] += tDeltaV[2];
              n[0] = 0.0f; n[1] = 0.0f; n[2] = -stpF[2];
            }
            if (tv > tExit) RETURN_MISS();
            if (vc[0] < lo[0] || vc[1] < lo[1] || vc[2] < lo[2] || vc[0] > lo[0] + 7 || vc[1] > lo[1] + 7 ||
                vc[2] > lo[2] + 7)
              break;
          }
        } else {
          inGlass = 0;
        }
        if (tMaxB[0] < tMaxB[1] && tMaxB[0] < tMaxB[2]) {
          tb = tMaxB[0]; bc[0] += stp[0]; tMaxB[0] += tDeltaB[0];
          nb[0] = -stpF[0]; nb[1] = 0.0f; nb[2] = 0.0f;
        } else if (tMaxB[1] < tMaxB[2]) {
          tb = tMaxB[1]; bc[1] += stp[1]; tMaxB[1] += tDeltaB[1];
          nb[0] = 0.0f; nb[1] = -stpF[1]; nb[2] = 0.0f;
        } else {
          tb = tMaxB[2]; bc[2] += stp[2]; tMaxB[2] += tDeltaB[2];
          nb[0] = 0.0f; nb[1] = 0.0f; nb[2] = -stpF[2];
        }
        if (tb > tExit) RETURN_MISS();
        if (bc[0] < sLo[0] || bc[1] < sLo[1] || bc[2] < sLo[2] || bc[0] > sHi[0] || bc[1] > sHi[1] || bc[2] > sHi[2])
          break;
      }
    } else {
      inGlass = 0;
    }
    if (tMaxS[0] < tMaxS[1] && tMaxS[0] < tMaxS[2]) {
      t = tMaxS[0]; sc[0] += stp[0]; tMaxS[0] += tDeltaS[0];
      normal[0] = -stpF[0]; normal[1] = 0.0f; normal[2] = 0.0f;
    } else if (tMaxS[1] < tMaxS[2]) {
      t = tMaxS[1]; sc[1] += stp[1]; tMaxS[1] += tDeltaS[1];
      normal[0] = 0.0f; normal[1] = -stpF[1]; normal[2] = 0.0f;
    } else {
      t = tMaxS[2]; sc[2] += stp[2]; tMaxS[2] += tDeltaS[2];
      normal[0] = 0.0f; normal[1] = 0.0f; normal[2] = -stpF[2];
    }
    if (t > tExit || sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || sc[0] >= sd[0] || sc[1] >= sd[1] || sc[2] >= sd[2]) break;
  }
  RETURN_MISS();
#undef RETURN_MISS
#undef RETURN_HIT
#undef MARCH_OUTPUTS
}

void o_march(const o_world* w, const float ro[3], const float rdIn[3], float maxT, int shadow, o_hit* out,
             float* trans_out) {
  o_march_ex(w, NULL, ro, rdIn, maxT, shadow, out, trans_out, NULL, NULL, NULL);
}

/* march() that also returns the hit normal as a vector (for the frame). */
static void march_n(const o_world* w, const float ro[3], const float rd[3], float maxT, int shadow, o_hit* h,
                    float* n_out) {
  o_march(w, ro, rd, maxT, shadow, h, NULL);
  n_out[0] = n_out[1] = n_out[2] = 0.0f;
  if (h->face < 6) n_out[h->face >> 1] = (h->face & 1) ? -1.0f : 1.0f;
}

/* ---------------------------------------------------------------------- */
/* traceSolid: sim.wgsl:296-357.  steps = brick iterations + voxel
 * iterations; t = the voxel-level t of the hit (traceSolid itself returns no t). */

/* The world as traceSolid reads it: the brick-cell word (0 = empty) and the
 * voxel value at p (0 outside), over either the brick map or a dense copy. */
typedef struct o_src {
  int32_t dx, dy, dz;
  const void* ctx;
  uint32_t (*cell)(const void* ctx, uint32_t ci);
  uint32_t (*voxel)(const void* ctx, const int32_t* p);
} o_src;

static void trace_solid_src(const o_src* S, const float ro[3], const float rdIn[3], float maxT, o_hit* out) {
  uint32_t steps = 0;
  float rd[3], inv[3], t0[3], t1[3], tmn[3], tmx[3];
  float dims[3] = {(float)S->dx, (float)S->dy, (float)S->dz};
  for (int a = 0; a < 3; a++) {
    rd[a] = fabsf(rdIn[a]) < 1e-6f ? fsign(rdIn[a] + 1e-30f) * 1e-6f : rdIn[a];
    inv[a] = 1.0f / rd[a];
    t0[a] = (0.0f - ro[a]) * inv[a];
    t1[a] = (dims[a] - ro[a]) * inv[a];
    tmn[a] = fminf(t0[a], t1[a]);
    tmx[a] = fmaxf(t0[a], t1[a]);
  }
  float t = fmaxf(fmaxf(tmn[0], tmn[1]), fmaxf(tmn[2], 0.0f));
  float tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
  if (t >= tExit) { set_miss(out, 0); return; }
  int32_t stp[3];
  float stpF[3], pos[3];
  for (int a = 0; a < 3; a++) {
    stp[a] = f2i(fsign(rd[a]));
    stpF[a] = (float)stp[a];
    pos[a] = fstep(0.0f, rd[a]);
  }
  float normal[3] = {0.0f, 0.0f, 0.0f};
  if (t > 0.0f) {
    if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) normal[0] = -stpF[0];
    else if (tmn[1] >= tmn[2]) normal[1] = -stpF[1];
    else normal[2] = -stpF[2];
  }
  int32_t bd[3] = {S->dx >> 3, S->dy >> 3, S->dz >> 3};
  int32_t bc[3];
  float tDeltaB[3], tMaxB[3], tDeltaV[3];
  for (int a = 0; a < 3; a++) {
    bc[a] = iclamp(f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), 0, bd[a] - 1);
    tDeltaB[a] = fabsf(inv[a]) * 8.0f;
    tMaxB[a] = (((float)bc[a] + pos[a]) * 8.0f - ro[a]) * inv[a];
    tDeltaV[a] = fabsf(inv[a]);
  }
  for (int i = 0; i < 1024; i++) {
    steps++;
    uint32_t g = S->cell(S->ctx, (uint32_t)(bc[0] + bd[0] * (bc[1] + bd[1] * bc[2])));
    if (g != 0u) {
      int32_t lo[3], vc[3];
      float tMaxV[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = bc[a] * 8;
        vc[a] = iclamp(f2i(floorf(ro[a] + rd[a] * (t + 1e-3f))), lo[a], lo[a] + 7);
        tMaxV[a] = ((float)vc[a] + pos[a] - ro[a]) * inv[a];
      }
      float tv = t;
      float n[3] = {normal[0], normal[1], normal[2]};
      for (int k = 0; k < 25; k++) {
        steps++;
        uint32_t v = S->voxel(S->ctx, vc);
        if (v != 0u) {
          o_world kw;
          kw.dx = S->dx;
          kw.dy = S->dy;
          set_hit(&kw, out, tv, vc, n, v, steps);
          return;
        }
        if (tMaxV[0] < tMaxV[1] && tMaxV[0] < tMaxV[2]) {
          tv = tMaxV[0]; vc[0] += stp[0]; tMaxV[0] += tDeltaV[0];
          n[0] = -stpF[0]; n[1] = 0.0f; n[2] = 0.0f;
        } else if (tMaxV[1] < tMaxV[2]) {
          tv = tMaxV[1]; vc[1] += stp[1]; tMaxV[1] += tDeltaV[1];
          n[0] = 0.0f; n[1] = -stpF[1]; n[2] = 0.0f;
        } else {
          tv = tMaxV[2]; vc[2] += stp[2]; tMaxV[2] += tDeltaV[2];
          n[0] = 0.0f; n[1] = 0.0f; n[2] = -stpF[2];
        }
        if (tv > tExit) { set_miss(out, steps); return; }
        if (vc[0] < lo[0] || vc[1] < lo[1] || vc[2] < lo[2] || vc[0] > lo[0] + 7 || vc[1] > lo[1] + 7 ||
            vc[2] > lo[2] + 7)
          break;
      }
    }
    if (tMaxB[0] < tMaxB[1] && tMaxB[0] < tMaxB[2]) {
      t = tMaxB[0]; bc[0] += stp[0]; tMaxB[0] += tDeltaB[0];
      normal[0] = -stpF[0]; normal[1] = 0.0f; normal[2] = 0.0f;
    } else if (tMaxB[1] < tMaxB[2]) {
      t = tMaxB[1]; bc[1] += stp[1]; tMaxB[1] += tDeltaB[1];
      normal[0] = 0.0f; normal[1] = -stpF[1]; normal[2] = 0.0f;
    } else {
      t = tMaxB[2]; bc[2] += stp[2]; tMaxB[2] += tDeltaB[2];
      normal[0] = 0.0f; normal[1] = 0.0f; normal[2] = -stpF[2];
    }
    if (t > tExit || bc[0] < 0 || bc[1] < 0 || bc[2] < 0 || bc[0] >= bd[0] || bc[1] >= bd[1] || bc[2] >= bd[2]) break;
  }
  set_miss(out, steps);
}

static uint32_t src_world_cell(const void* ctx, uint32_t ci) { return ((const o_world*)ctx)->grid[ci]; }
static uint32_t src_world_voxel(const void* ctx, const int32_t* p) { return readVoxel((const o_world*)ctx, p); }

void o_trace_solid(const o_world* w, const float ro[3], const float rdIn[3], float maxT, o_hit* out) {
  o_src S = {w->dx, w->dy, w->dz, w, src_world_cell, src_world_voxel};
  trace_solid_src(&S, ro, rdIn, maxT, out);
}

/* ---------------------------------------------------------------------- */
/* Segment DDA: sim.wgsl:1110-1145.  Inside case -> face 6, t 0, steps 0.
 * Hit -> key of the solid cell entered, face = -step along the hit axis,
 * t = tn, steps = loop iterations begun. */

void o_segment(const o_world* w, const float pos[3], const float mv[3], o_hit* out) {
  int32_t cell[3];
  for (int a = 0; a < 3; a++) cell[a] = f2i(floorf(pos[a]));
  uint32_t v0 = readVoxel(w, cell);
  if (v0 != 0u) {
    float zero[3] = {0.0f, 0.0f, 0.0f};
    set_hit(w, out, 0.0f, cell, zero, v0, 0);
    return;
  }
  float dist = sqrtf(dot3(mv, mv));
  if (!(dist > 1e-6f)) { set_miss(out, 0); return; }
  float dir[3], inv[3], tMax[3], tDelta[3];
  int32_t stp[3];
  for (int a = 0; a < 3; a++) {
    float d = mv[a] / dist;
    dir[a] = fabsf(d) < 1e-6f ? 1e-6f : d;
  }
  for (int a = 0; a < 3; a++) {
    inv[a] = 1.0f / dir[a];
    stp[a] = f2i(fsign(dir[a]));
    tMax[a] = ((float)cell[a] + fstep(0.0f, dir[a]) - pos[a]) * inv[a];
    tDelta[a] = fabsf(inv[a]);
  }
  uint32_t steps = 0;
  for (int k = 0; k < 48; k++) {
    steps++;
    int axis;
    if (tMax[0] < tMax[1] && tMax[0] < tMax[2]) axis = 0;
    else if (tMax[1] < tMax[2]) axis = 1;
    else axis = 2;
    float tn = tMax[axis];
    if (tn > dist) break;
    int32_t next[3] = {cell[0], cell[1], cell[2]};
    next[axis] += stp[axis];
    uint32_t v = readVoxel(w, next);
    if (v != 0u) {
      float n[3] = {0.0f, 0.0f, 0.0f};
      n[axis] = -(float)stp[axis];
      set_hit(w, out, tn, next, n, v, steps);
      return;
    }
    cell[0] = next[0]; cell[1] = next[1]; cell[2] = next[2];
    tMax[axis] += tDelta[axis];
  }
  set_miss(out, steps);
}

/* ---------------------------------------------------------------------- */
/* threaded batches */

typedef struct batch_job {
  const o_world* w;
  int kind;
  const float* rays;
  uint32_t n;
  o_hit* out;
  /* frame */
  const float* u52;
  const float* lights;
  uint32_t n_vec4;
  o_hit* shadow_out;
  uint32_t slots;
  volatile uint32_t next;
  uint32_t cast;
  pthread_mutex_t mu;
  int frame;
} batch_job;

static void trace_one(const o_world* w, int kind, const float* r, o_hit* out) {
  int k = kind >= 0 ? kind : (int)(((const uint32_t*)r)[7] & 3u);
  if (k == 0 || k == 1) o_march(w, r, r + 3, r[6], k == 1, out, NULL);
  else if (k == 2) o_trace_solid(w, r, r + 3, r[6], out);
  else o_segment(w, r, r + 3, out);
}

static void not_cast(o_hit* h) {
  h->key = NONE;
  h->t = 0.0f;
  h->steps = 0;
  h->face = 0xFE;
  h->mat = 0;
  h->value_hi = 0;
}

/* fsMain primary march + shade() shadow rays (quality 1). Returns rays cast. */
static uint32_t frame_one(const batch_job* J, uint32_t i) {
  const o_world* w = J->w;
  const float* u = J->u52;
  const float* r = J->rays + (size_t)i * 8u;
  o_hit* h = J->out + i;
  o_hit* sh = J->shadow_out + (size_t)i * J->slots;
  float N[3];
  march_n(w, r, r + 3, r[6], 0, h, N);
  for (uint32_t s = 0; s < J->slots; s++) not_cast(&sh[s]);
  if (h->face == 0xFF) return 0;
  uint32_t id = h->mat;
  if (matA(w->mats, id)[3] > 0.0f) return 0; /* emissive: shade() returns early (:309-311) */
  uint32_t cast = 0;
  float P[3], O[3];
  for (int a = 0; a < 3; a++) P[a] = r[a] + r[3 + a] * h->t;
  for (int a = 0; a < 3; a++) O[a] = P[a] + N[a] * 0.5f;
  const float* sunL = u + 4;
  float ndl = dot3(N, sunL);
  if (ndl > 0.0f && u[7] > 0.0f) {
    o_march(w, O, sunL, 1e9f, 1, &sh[0], NULL);
    cast++;
  }
  int count = f2i(J->lights[0]);
  float vs = u[3];
  for (int li = 0; li < count && (uint32_t)li + 1u < J->slots; li++) {
    const float* lp = J->lights + (1 + 2 * li) * 4;
    const float* lc = J->lights + (2 + 2 * li) * 4;
    float Lp[3] = {lp[0] / vs, lp[1] / vs, lp[2] / vs};
    float radius = lp[3] / vs;
    float L[3] = {Lp[0] - P[0], Lp[1] - P[1], Lp[2] - P[2]};
    float d = sqrtf(dot3(L, L));
    if (d > radius) continue;
    L[0] = L[0] / d; L[1] = L[1] / d; L[2] = L[2] / d;
    float nl = dot3(N, L);
    if (nl <= 0.0f) continue;
    if (lc[3] > 0.5f && u[21] > 0.5f) {
      o_march(w, O, L, d - 2.0f, 1, &sh[1 + li], NULL);
      cast++;
    }
  }
  return cast;
}

static void* batch_worker(void* arg) {
  batch_job* J = (batch_job*)arg;
  const uint32_t chunk = 256;
  uint32_t cast = 0;
  for (;;) {
    uint32_t b = __atomic_fetch_add(&J->next, chunk, __ATOMIC_RELAXED);
    if (b >= J->n) break;
    uint32_t e = b + chunk < J->n ? b + chunk : J->n;
    for (uint32_t i = b; i < e; i++) {
      if (J->frame) cast += frame_one(J, i);
      else trace_one(J->w, J->kind, J->rays + (size_t)i * 8u, J->out + i);
    }
  }
  __atomic_fetch_add(&J->cast, cast, __ATOMIC_RELAXED);
  return NULL;
}

static void run_batch(batch_job* J, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  J->next = 0;
  J->cast = 0;
  if (threads == 1) {
    batch_worker(J);
    return;
  }
  pthread_t th[256];
  int started = 0;
  for (int i = 0; i < threads; i++) {
    if (pthread_create(&th[i], NULL, batch_worker, J) != 0) break;
    started++;
  }
  if (started == 0) batch_worker(J);
  for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}

void o_trace_batch(const o_world* w, int kind, const float* rays, uint32_t n, o_hit* out, int threads) {
  batch_job J;
  memset(&J, 0, sizeof J);
  J.w = w;
  J.kind = kind;
  J.rays = rays;
  J.n = n;
  J.out = out;
  run_batch(&J, threads);
}

void o_frame(const o_world* w, const float* u52, const float* lights_vec4, uint32_t n_vec4, const float* rays,
             uint32_t n, o_hit* prim_out, o_hit* shadow_out, uint32_t* n_cast, int threads) {
  batch_job J;
  memset(&J, 0, sizeof J);
  J.w = w;
  J.rays = rays;
  J.n = n;
  J.out = prim_out;
  J.u52 = u52;
  J.lights = lights_vec4;
  J.n_vec4 = n_vec4;
  J.shadow_out = shadow_out;
  int count = n_vec4 > 0 ? f2i(lights_vec4[0]) : 0;
  if (count < 0) count = 0;
  if ((uint32_t)(1 + 2 * count) > n_vec4) count = n_vec4 > 0 ? (int)((n_vec4 - 1) / 2) : 0;
  J.slots = 1u + (uint32_t)count;
  J.frame = 1;
  run_batch(&J, threads);
  if (n_cast) *n_cast = J.cast;
}

/* ---------------------------------------------------------------------- */
/* Edit chain */

int o_removes(const float* im, const float* mats, int32_t x, int32_t y, int32_t z, uint32_t v) {
  uint32_t id = v & 0xffu;
  float d[3] = {((float)x + 0.5f) - im[0], ((float)y + 0.5f) - im[1], ((float)z + 0.5f) - im[2]};
  float d2 = dot3(d, d);
  float jit = 0.8f + 0.4f * o_rnd(o_hash3i(x, y, z));
  const float* E = matE(mats, id);
  if (E[1] > 0.5f && d2 < im[4] * im[4] * jit * jit) return 1;
  if (E[0] <= 0.0f) return 0;
  float rr = fmaxf(im[3] * fminf(1.4f, E[0]), fminf(im[3], 2.5f)) * jit;
  return d2 < rr * rr;
}

/* sim.wgsl:242-248 */
static float reach(const float* im, const float* mats, uint32_t id) {
  const float* E = matE(mats, id);
  float r = 0.0f;
#ifndef O_PLANT_REACH_NO_FR
  if (E[1] > 0.5f) r = im[4] * 1.2f;
#endif
  if (E[0] > 0.0f) r = fmaxf(r, fmaxf(im[3] * fminf(1.4f, E[0]), fminf(im[3], 2.5f)) * 1.2f);
  return r;
}

static inline uint32_t price_dimes(const float* mats, uint32_t v) {
  float p = matD(mats, v & 0xffu)[3];
  return (uint32_t)f2i(p * 10.0f + 0.5f);
}

typedef struct ed_world {
  int32_t dx, dy, dz, bx, by, bz;
  uint16_t* vox;
  uint8_t* kind;
  uint8_t* dirty;
  const float* mats;
} ed_world;

static inline size_t vkey(const ed_world* e, int32_t x, int32_t y, int32_t z) {
  return (size_t)x + (size_t)e->dx * ((size_t)y + (size_t)e->dy * (size_t)z);
}
static inline int in_world(const ed_world* e, int32_t x, int32_t y, int32_t z) {
  return x >= 0 && y >= 0 && z >= 0 && x < e->dx && y < e->dy && z < e->dz;
}
static inline size_t cell_of(const ed_world* e, int32_t x, int32_t y, int32_t z) {
  return (size_t)(x >> 3) + (size_t)e->bx * ((size_t)(y >> 3) + (size_t)e->by * (size_t)(z >> 3));
}

static int structural(const float* mats, uint32_t v) { return matE(mats, v & 0xffu)[2] > 0.5f; }
static int is_rubble(uint32_t v) { return (v & 0x80u) != 0u; }

static const int32_t DIRS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

int o_island_pass(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                  const float* impacts16, uint32_t n_impacts, uint32_t* lab_key, uint32_t* lab_count,
                  uint32_t cap, uint32_t* n_labels, uint64_t* stats) {
  o_edit_params P;
  P.mode = 0u;
  P.node_budget = 0xFFFFFFFFu;
  P.vs = 1.0f / 80.0f;
  P.frame_seed = 0u;
  uint64_t st[8];
  int rc = o_edit_pass(dx, dy, dz, vox, kind, mats, &P, impacts16, n_impacts, NULL, 0u, NULL, 0u, NULL, lab_key,
                       lab_count, cap, n_labels, NULL, 0u, NULL, st);
  if (stats)
    for (int k = 0; k < 5; k++) stats[k] = st[k];
  return rc;
}

/* Breadth-first labelling of the components of non-air, non-rubble,
 * non-structural voxels.  Scanning keys in increasing order makes the first
 * voxel of every component its minimum key. */
int o_islands(int32_t dx, int32_t dy, int32_t dz, const uint16_t* vox, const float* mats, uint32_t* lab_key,
              uint32_t* lab_count, uint32_t cap, uint32_t* n_labels) {
  size_t N = (size_t)dx * dy * dz;
  uint8_t* seen = (uint8_t*)calloc((N + 7) / 8, 1);
  uint32_t* q = (uint32_t*)malloc(N * sizeof(uint32_t) + 4);
  if (!seen || !q) { free(seen); free(q); return -1; }
  uint32_t nl = 0;
  size_t plane = (size_t)dx * dy;
  for (size_t s = 0; s < N; s++) {
    uint32_t v = vox[s];
    if (v == 0u || is_rubble(v) || structural(mats, v)) continue;
    if (seen[s >> 3] & (1u << (s & 7))) continue;
    seen[s >> 3] |= (uint8_t)(1u << (s & 7));
    size_t qh = 0, qt = 0;
    q[qt++] = (uint32_t)s;
    int anchored = 0;
    while (qh < qt) {
      size_t k = q[qh++];
      int32_t x = (int32_t)(k % (size_t)dx), y = (int32_t)((k / (size_t)dx) % (size_t)dy), z = (int32_t)(k / plane);
      if (y == 0) anchored = 1;
      for (int d = 0; d < 6; d++) {
        int32_t nx = x + DIRS[d][0], ny = y + DIRS[d][1], nz = z + DIRS[d][2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= dx || ny >= dy || nz >= dz) continue;
        size_t nk = (size_t)nx + (size_t)dx * ((size_t)ny + (size_t)dy * (size_t)nz);
        uint32_t w = vox[nk];
        if (w == 0u || is_rubble(w)) continue;
        if (structural(mats, w)) { anchored = 1; continue; }
        if (seen[nk >> 3] & (1u << (nk & 7))) continue;
        seen[nk >> 3] |= (uint8_t)(1u << (nk & 7));
        q[qt++] = (uint32_t)nk;
      }
    }
    if (!anchored) {
      if (nl < cap) {
        lab_key[nl] = (uint32_t)s;
        lab_count[nl] = (uint32_t)qt;
      }
      nl++;
    }
  }
  free(seen);
  free(q);
  *n_labels = nl;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Debris particles: spawn() sim.wgsl:166-170, frameBegin :363-369 (its
 * particle words), particleSim :1088-1173.  The reference appends to the
 * alive and write lists with atomics, so their order is whatever the GPU
 * schedule gives; this restatement is the serial schedule in particle-index
 * order, which the HIP library reproduces exactly: list entries in index
 * order and, when write requests exceed max_writes, the lowest-indexed
 * requests win. */

#define P_SETTLED 0x10000u
#define P_DEPOSIT 0x20000u
#define P_DUST 0x40000u

void o_frame_begin(uint32_t* ctr) {
  ctr[O_P_SPAWNED] = 0u;
  ctr[O_P_WRITES] = 0u;
  ctr[O_P_ALIVE] = 0u;
}

void o_spawn(o_particle* parts, uint32_t max_particles, uint32_t budget, uint32_t* ctr, const o_particle* src,
             uint32_t n) {
  for (uint32_t j = 0; j < n; j++) {
    uint32_t s = ctr[O_P_SPAWNED]++;
    if (s >= budget) continue;
    uint32_t i = (ctr[O_P_PCURSOR]++) & (max_particles - 1u);
    parts[i] = src[j];
  }
}

static inline int solid_at(const o_world* w, const int32_t* p) { return readVoxel(w, p) != 0u; }

/* pushAlive (sim.wgsl:1082-1085); entries past the list's capacity are dropped */
#define PUSH_ALIVE(i)                                        \
  do {                                                       \
    uint32_t k_ = ctr[O_P_ALIVE]++;                          \
    if (k_ < max_particles) alive[k_] = (i);                 \
  } while (0)

void o_particle_step(const o_world* w, o_particle* parts, uint32_t max_particles, float dt, float gravity, float vs,
                     uint32_t max_writes, uint32_t* ctr, uint32_t* write_list, uint32_t* alive) {
  for (uint32_t i = 0; i < max_particles; i++) {
    o_particle p = parts[i];
    if (p.life <= 0.0f) continue;
    p.life -= dt;
    if (p.life <= 0.0f) {
      parts[i].life = 0.0f;
      continue;
    }
    if ((p.value & P_SETTLED) != 0u) {
      parts[i].life = p.life;
      PUSH_ALIVE(i);
      continue;
    }
    int dust = (p.value & P_DUST) != 0u;
    float g = dust ? gravity * 0.15f : gravity;
    p.vel[1] -= g * dt;
    float drag = 1.0f - (dust ? 2.5f : 0.2f) * dt;
    for (int a = 0; a < 3; a++) p.vel[a] *= drag;
    int32_t cell[3];
    for (int a = 0; a < 3; a++) cell[a] = f2i(floorf(p.pos[a]));
    if (solid_at(w, cell)) {
      p.pos[1] += 1.0f;
      parts[i] = p;
      PUSH_ALIVE(i);
      continue;
    }
    float mv[3] = {p.vel[0] * dt, p.vel[1] * dt, p.vel[2] * dt};
    float dist = sqrtf(dot3(mv, mv));
    if (dist > 1e-6f) {
      float dir[3], inv[3], tMax[3], tDelta[3];
      int32_t stp[3];
      for (int a = 0; a < 3; a++) {
        float d = mv[a] / dist;
        dir[a] = fabsf(d) < 1e-6f ? 1e-6f : d;
      }
      for (int a = 0; a < 3; a++) {
        inv[a] = 1.0f / dir[a];
        stp[a] = f2i(fsign(dir[a]));
        tMax[a] = ((float)cell[a] + fstep(0.0f, dir[a]) - p.pos[a]) * inv[a];
        tDelta[a] = fabsf(inv[a]);
      }
      float t = 0.0f;
      int hit_axis = -1;
      for (int k = 0; k < 48; k++) {
        int axis;
        if (tMax[0] < tMax[1] && tMax[0] < tMax[2]) axis = 0;
        else if (tMax[1] < tMax[2]) axis = 1;
        else axis = 2;
        float tn = tMax[axis];
        if (tn > dist) break;
        int32_t next[3] = {cell[0], cell[1], cell[2]};
        next[axis] += stp[axis];
        if (solid_at(w, next)) {
          t = tn;
          hit_axis = axis;
          break;
        }
        cell[0] = next[0]; cell[1] = next[1]; cell[2] = next[2];
        tMax[axis] += tDelta[axis];
      }
      if (hit_axis < 0) {
        for (int a = 0; a < 3; a++) p.pos[a] += mv[a];
      } else {
        float s = fmaxf(t - 1e-3f, 0.0f);
        for (int a = 0; a < 3; a++) p.pos[a] += dir[a] * s;
        float vin = p.vel[hit_axis];
        p.vel[hit_axis] = -vin * 0.25f;
        float fr = hit_axis == 1 ? 0.55f : 0.75f;
        for (int a = 0; a < 3; a++)
          if (a != hit_axis) p.vel[a] *= fr;
      }
    }
    /* rest check: slow and supported from below */
    int32_t c[3];
    for (int a = 0; a < 3; a++) c[a] = f2i(floorf(p.pos[a]));
    int32_t below[3] = {c[0], c[1] - 1, c[2]};
    if (sqrtf(dot3(p.vel, p.vel)) < 0.6f / vs && solid_at(w, below)) {
      for (int a = 0; a < 3; a++) {
        p.pos[a] = (float)c[a] + 0.5f;
        p.vel[a] = 0.0f;
      }
      p.value |= P_SETTLED;
      int wrote = 0;
      if ((p.value & P_DEPOSIT) != 0u && c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < w->dx && c[1] < w->dy &&
          c[2] < w->dz) {
        uint32_t k = ctr[O_P_WRITES]++;
        if (k < max_writes) {
          write_list[2 * k] = key_of(w, c);
          write_list[2 * k + 1] = (p.value & 0xffffu) | 0x80u;
          wrote = 1;
        }
      }
      if (wrote) p.life = fminf(p.life, 0.5f);
      else p.life = fminf(p.life, dust ? 0.8f : 4.0f);
    }
    if (p.pos[1] < -200.0f) p.life = 0.0f;
    parts[i] = p;
    if (p.life > 0.0f) PUSH_ALIVE(i);
  }
}

/* ---------------------------------------------------------------------- */
/* Voxel writes: writeArgs / writeRequest / allocBricks / writeApply
 * (sim.wgsl:1178-1203, requestBrick :94-102, putVoxel :127-146, allocBricks
 * :459-483) and maint on the cells they touched (:1219-1274), on the dense
 * world of o_island_pass.  Serial in list order: the reference's requests
 * and puts race through atomics; the HIP library reproduces this order. */

/* putVoxel: 0 = written, 1 = cell occupied, 2 = no pool brick there */
static uint32_t put_voxel(ed_world* e, int32_t x, int32_t y, int32_t z, uint32_t v) {
  if (!in_world(e, x, y, z)) return 2u;
  size_t ci = cell_of(e, x, y, z);
  if (e->kind[ci] == 0) return 2u;
  if (e->kind[ci] == 1) return 1u;
  size_t k = vkey(e, x, y, z);
  if (e->vox[k] != 0u) return 1u;
  e->vox[k] = (uint16_t)v;
  e->dirty[ci] = 1;
  return 0u;
}

static void maint_dirty(ed_world* e, size_t ncell) {
  for (size_t ci = 0; ci < ncell; ci++) {
    if (!e->dirty[ci]) continue;
    e->dirty[ci] = 0;
    if (e->kind[ci] != 2) continue;
    int32_t ox = (int32_t)(ci % (size_t)e->bx) * 8;
    int32_t oy = (int32_t)((ci / (size_t)e->bx) % (size_t)e->by) * 8;
    int32_t oz = (int32_t)(ci / ((size_t)e->bx * e->by)) * 8;
    uint32_t first = e->vox[vkey(e, ox, oy, oz)];
    int nz = 0, diff = 0;
    for (uint32_t l = 0; l < 512u; l++) {
      uint32_t v = e->vox[vkey(e, ox + (int32_t)(l & 7u), oy + (int32_t)((l >> 3) & 7u), oz + (int32_t)(l >> 6))];
      if (v != 0u) nz = 1;
      if (v != first) diff = 1;
    }
    if (!nz) e->kind[ci] = 0;
    else if (!diff) e->kind[ci] = 1;
  }
}

int o_apply_writes(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const uint32_t* writes,
                   uint32_t n, uint32_t max_alloc, uint32_t free_top, uint64_t* stats) {
  ed_world E;
  E.dx = dx; E.dy = dy; E.dz = dz;
  E.bx = dx >> 3; E.by = dy >> 3; E.bz = dz >> 3;
  E.vox = vox;
  E.kind = kind;
  E.mats = NULL;
  size_t ncell = (size_t)E.bx * E.by * E.bz;
  E.dirty = (uint8_t*)calloc(ncell, 1);
  uint8_t* req = (uint8_t*)calloc(ncell, 1);
  if (!E.dirty || !req) { free(E.dirty); free(req); return -1; }
  uint64_t written = 0, allocated = 0, dropped = 0, dirty = 0;
  /* writeRequest: the first cell of each write, if it has no brick; the
   * alloc list keeps the first max_alloc distinct cells in list order */
  uint32_t nreq = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t key = writes[2 * i];
    int32_t x = (int32_t)(key % (uint32_t)dx), y = (int32_t)((key / (uint32_t)dx) % (uint32_t)dy);
    int32_t z = (int32_t)(key / ((uint32_t)dx * (uint32_t)dy));
    size_t ci = cell_of(&E, x, y, z);
    if (kind[ci] != 0 || req[ci]) continue;
    req[ci] = 1;
    if (nreq < max_alloc) {
      /* allocBricks: request r gets a brick while free ones remain */
      if (nreq < free_top) {
        kind[ci] = 2;
        E.dirty[ci] = 1;
        allocated++;
      } else {
        dropped++;
      }
      nreq++;
    } else {
      dropped++;
    }
  }
  /* writeApply: stack up to 2 voxels above a taken cell */
  for (uint32_t i = 0; i < n; i++) {
    uint32_t key = writes[2 * i], v = writes[2 * i + 1];
    int32_t x = (int32_t)(key % (uint32_t)dx), y = (int32_t)((key / (uint32_t)dx) % (uint32_t)dy);
    int32_t z = (int32_t)(key / ((uint32_t)dx * (uint32_t)dy));
    for (int up = 0; up < 3; up++) {
      uint32_t r = put_voxel(&E, x, y + up, z, v);
      if (r == 0u) written++;
      if (r != 1u) break;
    }
  }
  for (size_t ci = 0; ci < ncell; ci++) dirty += E.dirty[ci];
  maint_dirty(&E, ncell);
  free(E.dirty);
  free(req);
  if (stats) {
    stats[0] = written;
    stats[1] = allocated;
    stats[2] = dropped;
    stats[3] = dirty;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Island extraction and falling chunks: resolveChunks / resolveExpand /
 * resolveExtract (sim.wgsl:862-923) with dropRubbleAbove (:274-284) and the
 * crumb debris; chunkArgs / chunkBegin / chunkCollide / chunkStep / chunkLand
 * (:951-1075).  Serial restatement in a fixed order, which the HIP library
 * reproduces: unsupported components in increasing label order, voxels in
 * cell order then local index order (x | y<<3 | z<<6), arena entries in
 * arena order.  The reference takes these orders from its atomics. */

#define ANY_NONE 0xFFFFFFFFu

static inline float rnd01(uint32_t seed) { return (float)(o_hashu(seed) & 0xffffffu) / 16777216.0f; }

static uint32_t debris_flags(const float* mats, uint32_t v, uint32_t seed) {
  uint32_t id = v & 0xffu;
  if (matD(mats, id)[2] > 0.0f) return 0u; /* liquids splash and vanish */
  int fragile = matE(mats, id)[1] > 0.5f;
  if (rnd01(seed ^ 0x51ed27u) < (fragile ? 0.55f : 0.75f)) return P_DEPOSIT;
  if (rnd01(seed ^ 0x9e3779u) < 0.5f) return P_DUST;
  return 0u;
}

static void push_spawn(o_particle* out, uint32_t cap, uint32_t* n, const float pos[3], const float vel[3],
                       uint32_t value, float life) {
  if (*n < cap) {
    o_particle* p = &out[*n];
    for (int a = 0; a < 3; a++) {
      p->pos[a] = pos[a];
      p->vel[a] = vel[a];
    }
    p->value = value;
    p->life = life;
  }
  (*n)++;
}

/* takeVoxel (sim.wgsl:105-124): pool bricks only; mat = material id or 256 (any) */
static uint32_t take_voxel(ed_world* e, int32_t x, int32_t y, int32_t z, uint32_t mat) {
  if (!in_world(e, x, y, z)) return 0u;
  size_t ci = cell_of(e, x, y, z);
  if (e->kind[ci] != 2) return 0u;
  size_t k = vkey(e, x, y, z);
  uint32_t v = e->vox[k];
  if (v == 0u || (mat < 256u && (v & 0xffu) != mat)) return 0u;
  e->vox[k] = 0;
  e->dirty[ci] = 1;
  return v;
}

static uint32_t read_dense(const ed_world* e, int32_t x, int32_t y, int32_t z) {
  if (!in_world(e, x, y, z)) return 0u;
  return e->vox[vkey(e, x, y, z)];
}

/* dropRubbleAbove (sim.wgsl:274-284) */
static void drop_rubble_above(ed_world* e, int32_t x, int32_t y, int32_t z, float vs, uint32_t fs, o_particle* sp,
                              uint32_t cap, uint32_t* ns) {
  for (int32_t d = 1; d <= 12; d++) {
    int32_t qy = y + d;
    uint32_t v = read_dense(e, x, qy, z);
    if (v == 0u || !is_rubble(v)) return;
    uint32_t t = take_voxel(e, x, qy, z, v & 0xffu);
    if (t == 0u) return;
    uint32_t seed = (uint32_t)vkey(e, x, qy, z) ^ fs;
    float pos[3] = {(float)x + 0.5f, (float)qy + 0.5f, (float)z + 0.5f};
    float r[3] = {rnd01(seed), rnd01(seed + 1u), rnd01(seed + 2u)};
    float m[3] = {0.4f, 0.0f, 0.4f};
    float vel[3];
    for (int a = 0; a < 3; a++) vel[a] = ((r[a] - 0.5f) * m[a]) / vs;
    push_spawn(sp, cap, ns, pos, vel, t | P_DEPOSIT, 6.0f + 2.0f * rnd01(seed + 5u));
  }
}

/* ---------------------------------------------------------------------- */
/* The edit pass with its side effects: impactJobs, allocBricks, carve with
 * carveEffects (sim.wgsl:414-524), the pane flood with shardEffects
 * (:250-271, 589-618), maint (:1219-1274) and the island search
 * (:620-882), restated serially in the order the HIP library fixes:
 *  - a voxel several impacts remove is removed by the lowest-index one;
 *  - effects run for the carved voxels in (cell, local index) order, then for
 *    the flooded ones in the same order; a voxel's effects are its
 *    dropRubbleAbove records (bottom to top), then its own debris record;
 *  - seeds are the 6-neighbours of carved and flooded voxels that pushSeed
 *    keeps (not structural, not rubble, not a same-id neighbour of a pane
 *    voxel), plus the pending seeds of landed chunks; a seed counts when its
 *    voxel is a node voxel once the edits are done.
 * In seeded mode only components that hold a seed are searched, and the
 * floating ones are resolved in label order while their voxel total stays
 * within node_budget (the rest are assumed supported, as the reference
 * assumes the unexplored frontier is when its node budget trips). */

static uint32_t bit_get(const uint8_t* b, size_t i) { return (b[i >> 3] >> (i & 7)) & 1u; }
static void bit_set(uint8_t* b, size_t i) { b[i >> 3] |= (uint8_t)(1u << (i & 7)); }

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* (cell << 9 | local index) of a voxel, the effects order */
static uint64_t cell_order(const ed_world* e, int32_t x, int32_t y, int32_t z) {
  return ((uint64_t)cell_of(e, x, y, z) << 9) | (uint64_t)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
}

/* carveEffects' debris record (sim.wgsl:500-508) */
static void carve_debris(const float* mats, const float* im, int32_t x, int32_t y, int32_t z, uint32_t key, uint32_t v,
                         uint32_t fs, o_particle* sp, uint32_t cap, uint32_t* ns) {
  const float r = im[3];
  const float keep = fminf(1.0f, 30000.0f / fmaxf(1.0f, 4.19f * r * r * r));
  const uint32_t seed = key ^ fs;
  if (!(rnd01(seed) < matE(mats, v & 0xffu)[3] * keep)) return;
  float d[3] = {((float)x + 0.5f) - im[0], ((float)y + 0.5f) - im[1], ((float)z + 0.5f) - im[2]};
  const float len = sqrtf(dot3(d, d));
  const float den = len < 1e-3f ? 1e-3f : len;
  const float s = im[5] * (0.3f + 0.9f * rnd01(seed + 7u));
  const float off[3] = {0.5f, 0.2f, 0.5f}, m[3] = {0.4f, 0.5f, 0.4f};
  float vel[3];
  for (int a = 0; a < 3; a++) {
    const float dirn = d[a] / den;
    const float jitter = ((rnd01(seed + 13u + (uint32_t)a) - off[a]) * im[5]) * m[a];
    vel[a] = ((dirn * 0.7f + im[6 + a] * 0.6f) - im[9 + a] * 0.2f) * s + jitter;
  }
  const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  push_spawn(sp, cap, ns, pos, vel, v | debris_flags(mats, v, seed), 5.0f + 4.0f * rnd01(seed + 11u));
}

/* shardEffects' debris record (sim.wgsl:265-270) */
static void shard_debris(const float* mats, int32_t x, int32_t y, int32_t z, uint32_t key, uint32_t v, float vs,
                         uint32_t fs, o_particle* sp, uint32_t cap, uint32_t* ns) {
  const uint32_t seed = key ^ fs;
  if (!(rnd01(seed) < matE(mats, v & 0xffu)[3] * 0.6f)) return;
  const float s = 1.0f / vs;
  const float b[3] = {rnd01(seed + 3u) - 0.5f, rnd01(seed + 4u) * 0.4f, rnd01(seed + 5u) - 0.5f};
  const float m[3] = {1.6f, 1.0f, 1.6f};
  float vel[3];
  for (int a = 0; a < 3; a++) vel[a] = (b[a] * m[a]) * s;
  const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  push_spawn(sp, cap, ns, pos, vel, v | debris_flags(mats, v, seed), 4.0f + 4.0f * rnd01(seed + 6u));
}

/* pushSeed candidates around a removed voxel of value v (carveEffects
 * :511-523, floodStep :606-616): neighbours that are not air, structural or
 * rubble, and not of v's id when v is a pane; checked on the final world. */
static void mark_seeds(const ed_world* e, const float* mats, int32_t x, int32_t y, int32_t z, uint32_t v, uint8_t* seed) {
  const int pane = matE(mats, v & 0xffu)[1] > 1.5f;
  for (int k = 0; k < 6; k++) {
    int32_t qx = x + DIRS[k][0], qy = y + DIRS[k][1], qz = z + DIRS[k][2];
    if (!in_world(e, qx, qy, qz)) continue;
    uint32_t w = e->vox[vkey(e, qx, qy, qz)];
    if (w == 0u || structural(mats, w) || is_rubble(w)) continue;
    if (pane && (w & 0xffu) == (v & 0xffu)) continue;
    bit_set(seed, vkey(e, qx, qy, qz));
  }
}

/* components of node voxels (non-air, non-rubble, non-structural) holding a
 * seed, in key order; floating ones are kept while their running voxel total
 * stays within node_budget */
static int seeded_islands(int32_t dx, int32_t dy, int32_t dz, const uint16_t* vox, const float* mats,
                          const uint8_t* seed, uint32_t node_budget, uint32_t* lab_key, uint32_t* lab_count,
                          uint32_t cap, uint32_t* n_labels, uint64_t* dropped) {
  size_t N = (size_t)dx * dy * dz;
  uint8_t* seen = (uint8_t*)calloc((N + 7) / 8, 1);
  uint32_t* q = (uint32_t*)malloc(N * sizeof(uint32_t) + 4);
  if (!seen || !q) { free(seen); free(q); return -1; }
  uint32_t nl = 0;
  uint64_t total = 0, drop = 0;
  int over = 0;
  size_t plane = (size_t)dx * dy;
  for (size_t s = 0; s < N; s++) {
    uint32_t v = vox[s];
    if (v == 0u || is_rubble(v) || structural(mats, v)) continue;
    if (bit_get(seen, s)) continue;
    bit_set(seen, s);
    size_t qh = 0, qt = 0;
    q[qt++] = (uint32_t)s;
    int anchored = 0, seeded = 0;
    while (qh < qt) {
      size_t k = q[qh++];
      if (bit_get(seed, k)) seeded = 1;
      int32_t x = (int32_t)(k % (size_t)dx), y = (int32_t)((k / (size_t)dx) % (size_t)dy), z = (int32_t)(k / plane);
      if (y == 0) anchored = 1;
      for (int d = 0; d < 6; d++) {
        int32_t nx = x + DIRS[d][0], ny = y + DIRS[d][1], nz = z + DIRS[d][2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= dx || ny >= dy || nz >= dz) continue;
        size_t nk = (size_t)nx + (size_t)dx * ((size_t)ny + (size_t)dy * (size_t)nz);
        uint32_t w = vox[nk];
        if (w == 0u || is_rubble(w)) continue;
        if (structural(mats, w)) { anchored = 1; continue; }
        if (bit_get(seen, nk)) continue;
        bit_set(seen, nk);
        q[qt++] = (uint32_t)nk;
      }
    }
    if (anchored || !seeded) continue;
    if (!over && total + qt <= node_budget) {
      total += qt;
      if (nl < cap) {
        lab_key[nl] = (uint32_t)s;
        lab_count[nl] = (uint32_t)qt;
      }
      nl++;
    } else {
      over = 1;
      drop++;
    }
  }
  free(seen);
  free(q);
  *n_labels = nl;
  if (dropped) *dropped = drop;
  return 0;
}

int o_edit_pass(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                const o_edit_params* P, const float* impacts16, uint32_t n_impacts, const uint32_t* seed_keys,
                uint32_t n_seed_keys, const uint32_t* flood_in, uint32_t n_flood_in, const uint8_t* dirty_in,
                uint32_t* lab_key, uint32_t* lab_count, uint32_t cap, uint32_t* n_labels, o_particle* spawns,
                uint32_t spawn_cap, uint32_t* n_spawns, uint64_t* stats) {
  ed_world E;
  E.dx = dx; E.dy = dy; E.dz = dz;
  E.bx = dx >> 3; E.by = dy >> 3; E.bz = dz >> 3;
  E.vox = vox;
  E.kind = kind;
  E.mats = mats;
  const size_t ncell = (size_t)E.bx * E.by * E.bz;
  const size_t N = (size_t)dx * dy * dz;
  E.dirty = (uint8_t*)calloc(ncell, 1);
  uint8_t* req = (uint8_t*)calloc(ncell, 1);
  uint64_t* cmask = (uint64_t*)calloc(ncell, sizeof(uint64_t));
  uint8_t* seed = (uint8_t*)calloc((N + 7) / 8, 1);
  uint8_t* flooded = (uint8_t*)calloc((N + 7) / 8, 1);
  if (!E.dirty || !req || !cmask || !seed || !flooded) {
    free(E.dirty); free(req); free(cmask); free(seed); free(flooded);
    return -1;
  }
  uint64_t destroyed = 0, bill = 0, allocs = 0, njobs = 0, nflood = 0, ndrop = 0;
  uint32_t ns = 0;
  if (n_impacts > 64u) n_impacts = 64u;
  /* cells shoot's takes dirtied are maintained by this pass */
  if (dirty_in)
    for (size_t ci = 0; ci < ncell; ci++)
      if (dirty_in[ci]) E.dirty[ci] = 1;
  /* impactJobs (sim.wgsl:414-443) on the grid before any alloc: per cell,
   * the mask of impacts with a job there */
  const int32_t bd[3] = {E.bx, E.by, E.bz};
  for (uint32_t k = 0; k < n_impacts; k++) {
    const float* im = impacts16 + (size_t)k * 16u;
#ifdef O_PLANT_AABB_NO_1_2
    float R = ceilf(fmaxf(im[3], im[4])) + 1.0f;
#else
    float R = ceilf(fmaxf(im[3], im[4]) * 1.2f) + 1.0f;
#endif
    int32_t bmin[3], bmax[3];
    for (int a = 0; a < 3; a++) {
      bmin[a] = iclamp(f2i(floorf((im[a] - R) / 8.0f)), 0, bd[a] - 1);
      bmax[a] = iclamp(f2i(floorf((im[a] + R) / 8.0f)), 0, bd[a] - 1);
    }
    for (int32_t bz = bmin[2]; bz <= bmax[2]; bz++)
      for (int32_t by = bmin[1]; by <= bmax[1]; by++)
        for (int32_t bx = bmin[0]; bx <= bmax[0]; bx++) {
          size_t ci = (size_t)bx + (size_t)E.bx * ((size_t)by + (size_t)E.by * (size_t)bz);
          if (kind[ci] == 0) continue;
          float lo[3] = {(float)((uint32_t)bx * 8u), (float)((uint32_t)by * 8u), (float)((uint32_t)bz * 8u)};
          float ev[3];
          for (int a = 0; a < 3; a++) ev[a] = fmaxf(fmaxf(lo[a] - im[a], 0.0f), im[a] - (lo[a] + 8.0f));
          float near2 = dot3(ev, ev);
          if (near2 > R * R) continue;
          if (kind[ci] == 1) {
            uint32_t uval = vox[vkey(&E, bx * 8, by * 8, bz * 8)];
            float rch = reach(im, mats, uval & 0xffu);
            if (near2 > rch * rch) continue;
            req[ci] = 1;
          }
          cmask[ci] |= 1ull << k;
          njobs++;
        }
  }
  /* allocBricks (sim.wgsl:459-483): uniform cells in reach become pool bricks */
  for (size_t ci = 0; ci < ncell; ci++)
    if (req[ci] && kind[ci] == 1) {
      kind[ci] = 2;
      E.dirty[ci] = 1;
      allocs++;
    }
  /* carve (sim.wgsl:529-579), cell by cell: (order key, remover, value) of
   * every carved voxel */
  size_t ccap = 1024, cn = 0;
  uint64_t* carved = (uint64_t*)malloc(ccap * 2 * sizeof(uint64_t));
  for (size_t ci = 0; ci < ncell; ci++) {
    if (!cmask[ci] || kind[ci] != 2) continue;
    int32_t ox = (int32_t)(ci % (size_t)E.bx) * 8;
    int32_t oy = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8;
    int32_t oz = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8;
    int any = 0;
    for (uint32_t l = 0; l < 512u; l++) {
      int32_t x = ox + (int32_t)(l & 7u), y = oy + (int32_t)((l >> 3) & 7u), z = oz + (int32_t)(l >> 6);
      size_t kk = vkey(&E, x, y, z);
      uint32_t v = vox[kk];
      if (v == 0u) continue;
      uint64_t m = cmask[ci];
      int rem = -1;
      while (m) {
        int k = __builtin_ctzll(m);
        m &= m - 1ull;
        if (o_removes(impacts16 + (size_t)k * 16u, mats, x, y, z, v)) {
          rem = k;
          break;
        }
      }
      if (rem < 0) continue;
      vox[kk] = 0;
      any = 1;
      destroyed++;
      bill += price_dimes(mats, v);
      if (cn == ccap) {
        ccap *= 2;
        carved = (uint64_t*)realloc(carved, ccap * 2 * sizeof(uint64_t));
      }
      carved[cn * 2] = ((uint64_t)ci << 9) | l;
      carved[cn * 2 + 1] = ((uint64_t)rem << 32) | v;
      cn++;
    }
    if (any) E.dirty[ci] = 1;
  }
  /* carveEffects of every carved voxel, in (cell, local index) order:
   * dropRubbleAbove, then the carve debris */
  for (size_t i = 0; i < cn; i++) {
    size_t ci = (size_t)(carved[i * 2] >> 9);
    uint32_t l = (uint32_t)(carved[i * 2] & 511u);
    int32_t x = (int32_t)(ci % (size_t)E.bx) * 8 + (int32_t)(l & 7u);
    int32_t y = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8 + (int32_t)((l >> 3) & 7u);
    int32_t z = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8 + (int32_t)(l >> 6);
    uint32_t k = (uint32_t)(carved[i * 2 + 1] >> 32), v = (uint32_t)(carved[i * 2 + 1] & 0xffffu);
    uint32_t before = ns;
    drop_rubble_above(&E, x, y, z, P->vs, P->frame_seed, spawns, spawn_cap, &ns);
    ndrop += ns - before;
    carve_debris(mats, impacts16 + (size_t)k * 16u, x, y, z, (uint32_t)vkey(&E, x, y, z), v, P->frame_seed, spawns,
                 spawn_cap, &ns);
  }
  /* the pane flood to completion (carveEffects :511-523 + floodStep
   * :599-618): breadth-first over same-id voxels in pool bricks (takeVoxel
   * removes voxels of pool bricks only), from every carved pane voxel */
  size_t qcap = 1024, qh = 0, qt = 0;
  uint32_t* q = (uint32_t*)malloc(qcap * 4 * sizeof(uint32_t));
  size_t fcap = 1024, fn = 0;
  uint64_t* fl = (uint64_t*)malloc(fcap * 2 * sizeof(uint64_t));
#define TAKE_PUSH(X, Y, Z, M)                                                        \
  do {                                                                               \
    int32_t X_ = (X), Y_ = (Y), Z_ = (Z);                                            \
    if (in_world(&E, X_, Y_, Z_) && kind[cell_of(&E, X_, Y_, Z_)] == 2) {             \
      size_t kk_ = vkey(&E, X_, Y_, Z_);                                             \
      uint32_t v_ = vox[kk_];                                                        \
      if (v_ != 0u && (v_ & 0xffu) == (M) FLOOD_ID_TEST) {                            \
        vox[kk_] = 0;                                                                \
        E.dirty[cell_of(&E, X_, Y_, Z_)] = 1;                                        \
        destroyed++;                                                                 \
        nflood++;                                                                    \
        bill += price_dimes(mats, v_);                                               \
        bit_set(flooded, kk_);                                                       \
        if (qt == qcap) {                                                            \
          qcap *= 2;                                                                 \
          q = (uint32_t*)realloc(q, qcap * 4 * sizeof(uint32_t));                    \
        }                                                                            \
        q[qt * 4] = (uint32_t)X_;                                                    \
        q[qt * 4 + 1] = (uint32_t)Y_;                                                \
        q[qt * 4 + 2] = (uint32_t)Z_;                                                \
        q[qt * 4 + 3] = (M);                                                         \
        qt++;                                                                        \
        if (fn == fcap) {                                                            \
          fcap *= 2;                                                                 \
          fl = (uint64_t*)realloc(fl, fcap * 2 * sizeof(uint64_t));                  \
        }                                                                            \
        fl[fn * 2] = cell_order(&E, X_, Y_, Z_);                                     \
        fl[fn * 2 + 1] = v_;                                                         \
        fn++;                                                                        \
      }                                                                              \
    }                                                                                \
  } while (0)
#ifdef O_PLANT_FLOOD_ANY_PANE
#define FLOOD_ID_TEST || (matE(mats, v_ & 0xffu)[1] > 1.5f)
#else
#define FLOOD_ID_TEST
#endif
  for (size_t i = 0; i < cn; i++) {
    uint32_t v = (uint32_t)(carved[i * 2 + 1] & 0xffffu);
    if (!(matE(mats, v & 0xffu)[1] > 1.5f)) continue;
    size_t ci = (size_t)(carved[i * 2] >> 9);
    uint32_t l = (uint32_t)(carved[i * 2] & 511u);
    int32_t x = (int32_t)(ci % (size_t)E.bx) * 8 + (int32_t)(l & 7u);
    int32_t y = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8 + (int32_t)((l >> 3) & 7u);
    int32_t z = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8 + (int32_t)(l >> 6);
    uint32_t m = v & 0xffu;
    for (int k = 0; k < 6; k++) TAKE_PUSH(x + DIRS[k][0], y + DIRS[k][1], z + DIRS[k][2], m);
  }
  /* the entries shoot pushed (pushFlood of a struck pane voxel): floodStep
   * takes that voxel's same-id neighbours */
#ifdef O_PLANT_SHOT_FLOOD_SKIPPED
  n_flood_in = 0u;
#endif
  for (uint32_t i = 0; i < n_flood_in; i++) {
    uint32_t key = flood_in[2 * i], m = flood_in[2 * i + 1];
    int32_t x = (int32_t)(key % (uint32_t)dx), y = (int32_t)((key / (uint32_t)dx) % (uint32_t)dy);
    int32_t z = (int32_t)(key / ((uint32_t)dx * (uint32_t)dy));
    for (int k = 0; k < 6; k++) TAKE_PUSH(x + DIRS[k][0], y + DIRS[k][1], z + DIRS[k][2], m);
  }
  while (qh < qt) {
    int32_t x = (int32_t)q[qh * 4], y = (int32_t)q[qh * 4 + 1], z = (int32_t)q[qh * 4 + 2];
    uint32_t m = q[qh * 4 + 3];
    qh++;
    for (int k = 0; k < 6; k++) TAKE_PUSH(x + DIRS[k][0], y + DIRS[k][1], z + DIRS[k][2], m);
  }
#undef TAKE_PUSH
#undef FLOOD_ID_TEST
  free(q);
  /* shardEffects of every flooded voxel, in (cell, local index) order */
  qsort(fl, fn, 2 * sizeof(uint64_t), cmp_u64);
  for (size_t i = 0; i < fn; i++) {
    size_t ci = (size_t)(fl[i * 2] >> 9);
    uint32_t l = (uint32_t)(fl[i * 2] & 511u);
    int32_t x = (int32_t)(ci % (size_t)E.bx) * 8 + (int32_t)(l & 7u);
    int32_t y = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8 + (int32_t)((l >> 3) & 7u);
    int32_t z = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8 + (int32_t)(l >> 6);
    uint32_t v = (uint32_t)fl[i * 2 + 1];
    uint32_t before = ns;
    drop_rubble_above(&E, x, y, z, P->vs, P->frame_seed, spawns, spawn_cap, &ns);
    ndrop += ns - before;
    shard_debris(mats, x, y, z, (uint32_t)vkey(&E, x, y, z), v, P->vs, P->frame_seed, spawns, spawn_cap, &ns);
  }
  /* maint (sim.wgsl:1219-1274) on every dirty cell */
  for (size_t ci = 0; ci < ncell; ci++) {
    if (!E.dirty[ci] || kind[ci] != 2) continue;
    int32_t ox = (int32_t)(ci % (size_t)E.bx) * 8;
    int32_t oy = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8;
    int32_t oz = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8;
    uint32_t first = vox[vkey(&E, ox, oy, oz)];
    int nz = 0, diff = 0;
    for (uint32_t l = 0; l < 512u; l++) {
      uint32_t v = vox[vkey(&E, ox + (int32_t)(l & 7u), oy + (int32_t)((l >> 3) & 7u), oz + (int32_t)(l >> 6))];
      if (v != 0u) nz = 1;
      if (v != first) diff = 1;
    }
    if (!nz) kind[ci] = 0;
    else if (!diff) kind[ci] = 1;
  }
  /* seeds on the final world */
  uint64_t nseed = 0;
  if (P->mode == 1u) {
    for (size_t i = 0; i < cn; i++) {
      size_t ci = (size_t)(carved[i * 2] >> 9);
      uint32_t l = (uint32_t)(carved[i * 2] & 511u);
      mark_seeds(&E, mats, (int32_t)(ci % (size_t)E.bx) * 8 + (int32_t)(l & 7u),
                 (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8 + (int32_t)((l >> 3) & 7u),
                 (int32_t)(ci / ((size_t)E.bx * E.by)) * 8 + (int32_t)(l >> 6),
                 (uint32_t)(carved[i * 2 + 1] & 0xffffu), seed);
    }
    for (size_t i = 0; i < fn; i++) {
      size_t ci = (size_t)(fl[i * 2] >> 9);
      uint32_t l = (uint32_t)(fl[i * 2] & 511u);
      mark_seeds(&E, mats, (int32_t)(ci % (size_t)E.bx) * 8 + (int32_t)(l & 7u),
                 (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8 + (int32_t)((l >> 3) & 7u),
                 (int32_t)(ci / ((size_t)E.bx * E.by)) * 8 + (int32_t)(l >> 6), (uint32_t)fl[i * 2 + 1], seed);
    }
    for (uint32_t i = 0; i < n_flood_in; i++) {
      uint32_t key = flood_in[2 * i];
      mark_seeds(&E, mats, (int32_t)(key % (uint32_t)dx), (int32_t)((key / (uint32_t)dx) % (uint32_t)dy),
                 (int32_t)(key / ((uint32_t)dx * (uint32_t)dy)), flood_in[2 * i + 1], seed);
    }
    for (uint32_t i = 0; i < n_seed_keys; i++)
      if (seed_keys[i] < N) bit_set(seed, seed_keys[i]);
    for (size_t i = 0; i < N; i++)
      if (bit_get(seed, i)) {
        uint32_t w = vox[i];
        if (w != 0u && !is_rubble(w) && !structural(mats, w)) nseed++;
      }
  }
  free(carved);
  free(fl);
  free(E.dirty);
  free(req);
  free(cmask);
  free(flooded);
  uint64_t dropped = 0;
  int rc;
  if (P->mode == 1u)
    rc = seeded_islands(dx, dy, dz, vox, mats, seed, P->node_budget, lab_key, lab_count, cap, n_labels, &dropped);
  else
    rc = o_islands(dx, dy, dz, vox, mats, lab_key, lab_count, cap, n_labels);
  free(seed);
  if (n_spawns) *n_spawns = ns;
  if (stats) {
    stats[0] = destroyed;
    stats[1] = bill;
    stats[2] = allocs;
    stats[3] = njobs;
    stats[4] = nflood;
    stats[5] = nseed;
    stats[6] = dropped;
    stats[7] = ndrop;
  }
  return rc;
}

static uint32_t src_dense_cell(const void* ctx, uint32_t ci) { return ((const ed_world*)ctx)->kind[ci]; }
static uint32_t src_dense_voxel(const void* ctx, const int32_t* p) {
  return read_dense((const ed_world*)ctx, p[0], p[1], p[2]);
}

static inline uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

int o_shoot(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
            const o_shoot_params* P, float* impacts16, uint32_t* n_impacts, uint32_t* flood, uint32_t* n_flood,
            uint8_t* dirty, float* flashes, uint32_t* flash_ctr, o_particle* spawns, uint32_t spawn_cap,
            uint32_t* n_spawns, uint64_t* stats) {
  ed_world E;
  E.dx = dx; E.dy = dy; E.dz = dz;
  E.bx = dx >> 3; E.by = dy >> 3; E.bz = dz >> 3;
  E.vox = vox;
  E.kind = kind;
  E.mats = mats;
  E.dirty = dirty;
  o_src S = {dx, dy, dz, &E, src_dense_cell, src_dense_voxel};
  uint32_t pellets = f2u_sat(P->fire[0]);
  if (pellets > 16u) pellets = 16u;
  /* every pellet is traced on the world before any pellet's effects */
  o_hit hits[16];
  float dirs[16][3];
  for (uint32_t li = 0; li < pellets; li++) {
    float dir[3] = {P->look[0], P->look[1], P->look[2]};
    if (P->fire2[0] > 0.0f) {
      uint32_t seed = f2u_sat(P->fire2[1]) * 7919u + li * 104729u;
      float d[3];
      for (int k = 0; k < 3; k++) d[k] = dir[k] + ((rnd01(seed + (uint32_t)k) - 0.5f) * 2.0f) * P->fire2[0];
      float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
      for (int k = 0; k < 3; k++) dir[k] = d[k] / len;
    }
    float ro[3] = {P->eye[0] / P->vs, P->eye[1] / P->vs, P->eye[2] / P->vs};
    trace_solid_src(&S, ro, dir, 80.0f / P->vs, &hits[li]);
    for (int k = 0; k < 3; k++) dirs[li][k] = dir[k];
  }
  uint32_t ns = 0;
  uint64_t destroyed = 0, bill = 0, nhit = 0;
  for (uint32_t i = 0; i < pellets; i++) {
    const o_hit* h = &hits[i];
    if (h->key == NONE) continue;
    nhit++;
    int32_t x = (int32_t)(h->key % (uint32_t)dx), y = (int32_t)((h->key / (uint32_t)dx) % (uint32_t)dy);
    int32_t z = (int32_t)(h->key / ((uint32_t)dx * (uint32_t)dy));
    float n[3] = {0.0f, 0.0f, 0.0f};
    if (h->face < 6) n[h->face >> 1] = (h->face & 1) ? -1.0f : 1.0f;
    float r = P->fire[1];
    float off = fminf(r * 0.3f, 2.0f);
#ifdef O_PLANT_SHOOT_CENTRE_SIGN
    float c[3] = {((float)x + 0.5f) + n[0] * off, ((float)y + 0.5f) + n[1] * off, ((float)z + 0.5f) + n[2] * off};
#else
    float c[3] = {((float)x + 0.5f) - n[0] * off, ((float)y + 0.5f) - n[1] * off, ((float)z + 0.5f) - n[2] * off};
#endif
    /* pushImpact (sim.wgsl:203-211) */
    uint32_t k = (*n_impacts)++;
    if (k < P->max_impacts && k < 64u) {
      float* im = impacts16 + (size_t)k * 16u;
      float rec[12] = {c[0], c[1], c[2], r, P->fire[2], P->fire[3], n[0], n[1], n[2],
                       dirs[i][0], dirs[i][1], dirs[i][2]};
      for (int q = 0; q < 12; q++) im[q] = rec[q];
    }
    /* a struck pane voxel: takeVoxel, shardEffects, pushFlood (:386-394) */
    uint32_t id = h->mat;
    if (matE(mats, id)[1] > 1.5f) {
      uint32_t v = take_voxel(&E, x, y, z, id);
      if (v != 0u) {
        drop_rubble_above(&E, x, y, z, P->vs, P->frame_seed, spawns, spawn_cap, &ns);
        bill += price_dimes(mats, v);
        destroyed++;
        shard_debris(mats, x, y, z, h->key, v, P->vs, P->frame_seed, spawns, spawn_cap, &ns);
        uint32_t q = (*n_flood)++;
        if (q < 64u) {
          flood[2 * q] = h->key;
          flood[2 * q + 1] = id;
        }
      }
    }
    /* pellet 0's flash (:395-401) */
    if (i == 0u) {
      float* b = flashes + ((*flash_ctr)++ % 8u) * 8u;
      for (int q = 0; q < 3; q++) b[q] = (c[q] + n[q] * 4.0f) * P->vs;
      b[3] = r * P->vs;
      b[4] = P->time_x;
    }
  }
  *n_spawns = ns;
  if (stats) {
    stats[0] = destroyed;
    stats[1] = bill;
    stats[2] = nhit;
  }
  return 0;
}

int o_resolve(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
              const uint32_t* lab_key, const uint32_t* lab_count, uint32_t n_labels, const o_resolve_params* prm,
              uint32_t* chunks, uint32_t* arena, uint32_t* ctr, o_particle* spawns, uint32_t spawn_cap,
              uint32_t* n_spawns) {
  ed_world E;
  E.dx = dx; E.dy = dy; E.dz = dz;
  E.bx = dx >> 3; E.by = dy >> 3; E.bz = dz >> 3;
  E.vox = vox;
  E.kind = kind;
  E.mats = mats;
  size_t N = (size_t)dx * dy * dz, ncell = (size_t)E.bx * E.by * E.bz;
  E.dirty = (uint8_t*)calloc(ncell, 1);
  uint32_t* comp = (uint32_t*)calloc(N, sizeof(uint32_t)); /* 0 = not extracted, else tag */
  uint32_t* q = (uint32_t*)malloc(N * sizeof(uint32_t) + 4);
  uint8_t* req = (uint8_t*)calloc(ncell, 1);
  if (!E.dirty || !comp || !q || !req) {
    free(E.dirty); free(comp); free(q); free(req);
    return -1;
  }
  *n_spawns = 0;
  size_t plane = (size_t)dx * dy;
  /* resolveChunks: tag = (cid + 1) for chunk voxels, 0x8000 | anything for crumbs */
  for (uint32_t i = 0; i < n_labels; i++) {
    uint32_t cnt = lab_count[i];
    if (cnt > prm->component_budget) {
      ctr[O_R_S_BUDGET]++;
      continue; /* assumed supported */
    }
    uint32_t tag = 0x10000u; /* crumble */
    if (cnt >= 48u) {
      uint32_t id = ctr[O_R_C_CHUNKS]++;
      if (id < prm->max_chunks) {
        uint32_t* b = chunks + (size_t)id * 16u;
        for (int k = 0; k < 16; k++) b[k] = 0u;
        b[6] = ANY_NONE;
        b[0] = 1u;
        ctr[O_R_S_ISLANDS]++;
        tag = id + 1u;
      }
    }
    /* flood the component from its minimum voxel */
    size_t qh = 0, qt = 0;
    size_t s0 = lab_key[i];
    comp[s0] = tag;
    q[qt++] = (uint32_t)s0;
    while (qh < qt) {
      size_t k = q[qh++];
      int32_t x = (int32_t)(k % (size_t)dx), y = (int32_t)((k / (size_t)dx) % (size_t)dy), z = (int32_t)(k / plane);
      for (int d = 0; d < 6; d++) {
        int32_t nx = x + DIRS[d][0], ny = y + DIRS[d][1], nz = z + DIRS[d][2];
        if (!in_world(&E, nx, ny, nz)) continue;
        size_t nk = vkey(&E, nx, ny, nz);
        uint32_t w = vox[nk];
        if (w == 0u || is_rubble(w) || structural(mats, w) || comp[nk]) continue;
        comp[nk] = tag;
        q[qt++] = (uint32_t)nk;
      }
    }
  }
  /* resolveExpand + allocBricks: uniform cells holding extracted voxels, in cell order */
  uint32_t nreq = 0;
  for (size_t ci = 0; ci < ncell; ci++) {
    if (kind[ci] != 1) continue;
    int32_t ox = (int32_t)(ci % (size_t)E.bx) * 8, oy = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8;
    int32_t oz = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8;
    if (!comp[vkey(&E, ox, oy, oz)]) continue; /* a uniform brick is one component */
    if (nreq < prm->max_alloc) {
      if (nreq < prm->free_top) {
        kind[ci] = 2;
        E.dirty[ci] = 1;
        ctr[O_R_ALLOCS]++;
      } else {
        ctr[O_R_ALLOC_FAIL]++;
      }
      nreq++;
    }
  }
  /* resolveExtract in cell order, local index order */
  for (size_t ci = 0; ci < ncell; ci++) {
    if (kind[ci] != 2) continue;
    int32_t ox = (int32_t)(ci % (size_t)E.bx) * 8, oy = (int32_t)((ci / (size_t)E.bx) % (size_t)E.by) * 8;
    int32_t oz = (int32_t)(ci / ((size_t)E.bx * E.by)) * 8;
    for (uint32_t l = 0; l < 512u; l++) {
      int32_t x = ox + (int32_t)(l & 7u), y = oy + (int32_t)((l >> 3) & 7u), z = oz + (int32_t)(l >> 6);
      size_t k = vkey(&E, x, y, z);
      uint32_t tag = comp[k];
      if (!tag) continue;
      comp[k] = 0;
      uint32_t v = take_voxel(&E, x, y, z, 256u);
      if (v == 0u) continue;
      drop_rubble_above(&E, x, y, z, prm->vs, prm->frame_seed, spawns, spawn_cap, n_spawns);
      if (tag < 0x10000u) {
        uint32_t a = ctr[O_R_C_ARENA]++;
        if (a < prm->max_arena) {
          arena[2u * a] = (uint32_t)k;
          arena[2u * a + 1u] = v | ((tag - 1u) << 16);
          uint32_t* b = chunks + (size_t)(tag - 1u) * 16u;
          if (matE(mats, v & 0xffu)[1] > 0.5f) b[4] = 1u;
          b[5]++;
          continue;
        }
      }
      /* crumbs: straight to debris */
      ctr[O_R_S_BILL] += price_dimes(mats, v);
      ctr[O_R_S_DESTROYED]++;
      uint32_t seed = (uint32_t)k ^ prm->frame_seed;
      float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
      float r[3] = {rnd01(seed), rnd01(seed + 1u), rnd01(seed + 2u)};
      float off[3] = {0.5f, 0.0f, 0.5f}, m[3] = {0.6f, 0.3f, 0.6f};
      float vel[3];
      for (int a = 0; a < 3; a++) vel[a] = ((r[a] - off[a]) * m[a]) / prm->vs;
      push_spawn(spawns, spawn_cap, n_spawns, pos, vel, v | debris_flags(mats, v, seed), 4.0f + 3.0f * rnd01(seed + 9u));
    }
  }
  maint_dirty(&E, ncell);
  free(E.dirty);
  free(comp);
  free(q);
  free(req);
  return 0;
}

static inline float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static inline uint32_t f2u_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
/* WGSL u32(f32): truncation toward zero, saturating, NaN -> 0 */
static inline uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

int o_chunk_frame(int32_t dx, int32_t dy, int32_t dz, uint16_t* vox, uint8_t* kind, const float* mats,
                  uint32_t* chunks, uint32_t* arena, uint32_t* ctr, const o_resolve_params* prm, float dt,
                  float gravity, uint32_t max_writes, uint32_t* writes, uint32_t max_impacts, float* impacts,
                  o_particle* spawns, uint32_t spawn_cap, uint32_t* n_spawns, uint32_t* seeds, uint32_t seed_cap,
                  uint32_t* n_seeds) {
  ed_world E;
  E.dx = dx; E.dy = dy; E.dz = dz;
  E.bx = dx >> 3; E.by = dy >> 3; E.bz = dz >> 3;
  E.vox = vox;
  E.kind = kind;
  E.mats = mats;
  E.dirty = NULL;
  *n_spawns = 0;
  const float vs = prm->vs;
  /* chunkArgs (:951-975) */
  uint32_t n = ctr[O_R_C_CHUNKS] < prm->max_chunks ? ctr[O_R_C_CHUNKS] : prm->max_chunks;
  uint32_t active = 0;
  for (uint32_t c = 0; c < n; c++) {
    uint32_t* b = chunks + (size_t)c * 16u;
    uint32_t st = b[0];
    if (st == 2u) b[0] = 0u;
    if (st == 1u && (b[5] == 0u || b[2] > (uint32_t)dy)) {
      b[0] = 0u;
      st = 0u;
    }
    if (st == 1u) active++;
  }
  if (active == 0u) {
    ctr[O_R_C_CHUNKS] = 0u;
    ctr[O_R_C_ARENA] = 0u;
  }
  ctr[O_R_S_FALLING] = active;
  n = ctr[O_R_C_CHUNKS] < prm->max_chunks ? ctr[O_R_C_CHUNKS] : prm->max_chunks;
  uint32_t na = ctr[O_R_C_ARENA] < prm->max_arena ? ctr[O_R_C_ARENA] : prm->max_arena;
  /* chunkBegin (:977-991) */
  for (uint32_t c = 0; c < n; c++) {
    uint32_t* b = chunks + (size_t)c * 16u;
    if (b[0] != 1u) continue;
    float vy = fminf(u2f(b[1]) + gravity * dt, 25.0f / vs);
    float travel = vy * dt + u2f(b[3]);
    uint32_t steps = f2u(floorf(travel));
    if (steps > 48u) steps = 48u;
    b[1] = f2u_bits(vy);
    b[3] = f2u_bits(travel - floorf(travel));
    b[8] = steps;
    b[6] = ANY_NONE;
    b[7] = 0u;
  }
  /* chunkCollide (:993-1021), arena order */
  for (uint32_t j = 0; j < na; j++) {
    uint32_t vv = arena[2u * j + 1u], v = vv & 0xffffu;
    if (v == 0u) continue;
    uint32_t* b = chunks + (size_t)(vv >> 16) * 16u;
    if (b[0] != 1u) continue;
    uint32_t steps = b[8];
    if (steps == 0u) continue;
    uint32_t key = arena[2u * j];
    int32_t px = (int32_t)(key % (uint32_t)dx), py = (int32_t)((key / (uint32_t)dx) % (uint32_t)dy) - (int32_t)b[2];
    int32_t pz = (int32_t)(key / ((uint32_t)dx * (uint32_t)dy));
    float speed = u2f(b[1]) * vs;
    for (uint32_t s = 1; s <= steps; s++) {
      int32_t qy = py - (int32_t)s;
      if (qy < 0) {
        if (s < b[6]) b[6] = s;
        break;
      }
      uint32_t w = read_dense(&E, px, qy, pz);
      if (w != 0u) {
        if (matE(mats, w & 0xffu)[1] > 0.5f && speed > 1.5f && (b[7]++) < 6u) {
          uint32_t k = ctr[O_R_C_IMPACTS]++;
          if (k < max_impacts) {
            float* im = impacts + (size_t)k * 16u;
            for (int t = 0; t < 16; t++) im[t] = 0.0f;
            im[0] = (float)px + 0.5f; im[1] = (float)qy + 0.5f; im[2] = (float)pz + 0.5f;
            im[3] = 0.02f / vs; im[4] = 0.12f / vs; im[5] = speed * 0.4f / vs;
            im[6] = 0.0f; im[7] = 1.0f; im[8] = 0.0f;
            im[9] = 0.0f; im[10] = -1.0f; im[11] = 0.0f;
          }
        }
        if (s < b[6]) b[6] = s;
        break;
      }
    }
  }
  /* chunkStep (:1023-1042) */
  for (uint32_t c = 0; c < n; c++) {
    uint32_t* b = chunks + (size_t)c * 16u;
    if (b[0] != 1u) continue;
    uint32_t steps = b[8];
    if (steps == 0u) continue;
    uint32_t blocked = b[6], drop = steps;
    if (blocked <= steps) {
      drop = blocked - 1u;
      if (b[7] > 0u) {
        b[1] = f2u_bits(u2f(b[1]) * 0.6f);
      } else {
        b[9] = f2u_bits(u2f(b[1]) * vs);
        b[0] = 2u;
      }
    }
    b[2] += drop;
  }
  /* chunkLand (:1044-1075), arena order */
  for (uint32_t j = 0; j < na; j++) {
    uint32_t vv = arena[2u * j + 1u], v = vv & 0xffffu;
    if (v == 0u) continue;
    uint32_t* b = chunks + (size_t)(vv >> 16) * 16u;
    if (b[0] != 2u) continue;
    arena[2u * j + 1u] = 0u;
    uint32_t key = arena[2u * j];
    int32_t px = (int32_t)(key % (uint32_t)dx), py = (int32_t)((key / (uint32_t)dx) % (uint32_t)dy) - (int32_t)b[2];
    int32_t pz = (int32_t)(key / ((uint32_t)dx * (uint32_t)dy));
    float speed = u2f(b[9]);
    if (b[4] != 0u && speed > 2.0f && matE(mats, v & 0xffu)[1] > 0.5f) {
      ctr[O_R_S_BILL] += price_dimes(mats, v);
      ctr[O_R_S_DESTROYED]++;
      uint32_t seed = key ^ prm->frame_seed;
      float r[3] = {(rnd01(seed) - 0.5f) * 0.9f, rnd01(seed + 1u) * 0.5f, (rnd01(seed + 2u) - 0.5f) * 0.9f};
      float vel[3];
      for (int a = 0; a < 3; a++) vel[a] = (r[a] * speed) / vs;
      float pos[3] = {(float)px + 0.5f, (float)py + 0.5f, (float)pz + 0.5f};
      push_spawn(spawns, spawn_cap, n_spawns, pos, vel, v | debris_flags(mats, v, seed), 5.0f + 3.0f * rnd01(seed + 3u));
      continue;
    }
    uint32_t k = ctr[O_R_C_WRITES]++;
    if (k < max_writes) {
      int32_t p[3] = {px, py, pz};
      writes[2u * k] = (uint32_t)vkey(&E, p[0], p[1], p[2]);
      writes[2u * k + 1u] = v;
    }
    /* whatever it landed on is re-checked for support (pushSeed, :1068-1072) */
    if ((o_hashu(key) & 15u) == 0u) {
      uint32_t w = read_dense(&E, px, py - 1, pz);
      if (w != 0u && !structural(mats, w)) {
        if (n_seeds && *n_seeds < seed_cap) seeds[*n_seeds] = (uint32_t)vkey(&E, px, py - 1, pz);
        if (n_seeds) (*n_seeds)++;
      }
    }
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Scene rasteriser: apply() of src/world/shapes.ts:334-414 (Op.Add with a
 * constant paint), one primitive after another over the whole world as the
 * reference runs it, on a dense voxel array with a per-cell kind (0 empty,
 * 1 uniform, 2 pool brick) that follows setBrickUniform / compactCell
 * (src/world/VoxelWorld.ts:147-156).  SDFs (shapes.ts:24-143): box, rounded
 * box, sphere, capped cylinder, capsule, torus, axis-aligned box, all in
 * voxel units; a yaw turns the primitive about the vertical axis through its
 * centre (local = (c·dx + s·dz, dy, −s·dx + c·dz)).  Double precision as the
 * reference's JavaScript numbers. */

enum { OP_BOX = 0, OP_ROUNDBOX, OP_SPHERE, OP_CYLINDER, OP_CAPSULE, OP_TORUS, OP_AABB };

static inline double dmax(double a, double b) { return a < b ? b : a; }
static inline double dmin(double a, double b) { return b < a ? b : a; }
static inline double dlen2(double a, double b) { return sqrt(a * a + b * b); }
static inline double dlen3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

static void prim_local(const o_prim* P, double x, double y, double z, double* lx, double* ly, double* lz) {
  double dx = x - P->p[0], dy = y - P->p[1], dz = z - P->p[2];
  *lx = P->cy * dx + P->sy * dz;
  *ly = dy;
  *lz = -P->sy * dx + P->cy * dz;
}

static void prim_axis(int axis, double lx, double ly, double lz, double* a, double* b, double* h) {
  if (axis == 1) { *a = lx; *b = lz; *h = ly; }
  else if (axis == 0) { *a = ly; *b = lz; *h = lx; }
  else { *a = lx; *b = ly; *h = lz; }
}

double o_prim_sdf(const o_prim* P, double x, double y, double z) {
  double lx, ly, lz, a, b, h;
  switch (P->type) {
    case OP_BOX: {
      prim_local(P, x, y, z, &lx, &ly, &lz);
      double qx = fabs(lx) - P->p[3], qy = fabs(ly) - P->p[4], qz = fabs(lz) - P->p[5];
      double out = dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0));
      return out + dmin(dmax(qx, dmax(qy, qz)), 0.0);
    }
    case OP_ROUNDBOX: {
      prim_local(P, x, y, z, &lx, &ly, &lz);
      double r = P->p[6];
      double qx = fabs(lx) - (P->p[3] - r), qy = fabs(ly) - (P->p[4] - r), qz = fabs(lz) - (P->p[5] - r);
      return dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0)) + dmin(dmax(qx, dmax(qy, qz)), 0.0) - r;
    }
    case OP_SPHERE:
      return dlen3(x - P->p[0], y - P->p[1], z - P->p[2]) - P->p[3];
    case OP_CYLINDER: {
      prim_local(P, x, y, z, &lx, &ly, &lz);
      prim_axis((int)P->p[5], lx, ly, lz, &a, &b, &h);
      double dr = dlen2(a, b) - P->p[3], dh = fabs(h) - P->p[4];
      return dmin(dmax(dr, dh), 0.0) + dlen2(dmax(dr, 0.0), dmax(dh, 0.0));
    }
    case OP_CAPSULE: {
      double ax = P->p[0], ay = P->p[1], az = P->p[2];
      double bax = P->p[3] - ax, bay = P->p[4] - ay, baz = P->p[5] - az;
      double bb = bax * bax + bay * bay + baz * baz;
      if (bb == 0.0) bb = 1e-12;
      double px = x - ax, py = y - ay, pz = z - az;
      double t = dmin(1.0, dmax(0.0, (px * bax + py * bay + pz * baz) / bb));
      return dlen3(px - bax * t, py - bay * t, pz - baz * t) - P->p[6];
    }
    case OP_TORUS: {
      prim_local(P, x, y, z, &lx, &ly, &lz);
      prim_axis((int)P->p[5], lx, ly, lz, &a, &b, &h);
      return dlen2(dlen2(a, b) - P->p[3], h) - P->p[4];
    }
    case OP_AABB: {
      double cx = 0.5 * (P->p[0] + P->p[3]), cy = 0.5 * (P->p[1] + P->p[4]), cz = 0.5 * (P->p[2] + P->p[5]);
      double hx = 0.5 * (P->p[3] - P->p[0]), hy = 0.5 * (P->p[4] - P->p[1]), hz = 0.5 * (P->p[5] - P->p[2]);
      double qx = fabs(x - cx) - hx, qy = fabs(y - cy) - hy, qz = fabs(z - cz) - hz;
      double out = dlen3(dmax(qx, 0.0), dmax(qy, 0.0), dmax(qz, 0.0));
      return out + dmin(dmax(qx, dmax(qy, qz)), 0.0);
    }
  }
  return 1e30;
}

/* shapes.ts:47-48: the exact containment test of an axis-aligned box */
static int prim_contains(const o_prim* P, int* has, double x0, double y0, double z0, double x1, double y1, double z1) {
  *has = 0;
  if (P->type == OP_AABB) {
    *has = 1;
    return x0 >= P->p[0] && x1 <= P->p[3] && y0 >= P->p[1] && y1 <= P->p[4] && z0 >= P->p[2] && z1 <= P->p[5];
  }
  if (P->type == OP_BOX && P->sy == 0.0 && P->cy == 1.0) {
    *has = 1;
    return x0 >= P->p[0] - P->p[3] && x1 <= P->p[0] + P->p[3] && y0 >= P->p[1] - P->p[4] &&
           y1 <= P->p[1] + P->p[4] && z0 >= P->p[2] - P->p[5] && z1 <= P->p[2] + P->p[5];
  }
  return 0;
}

void o_raster(int32_t dx, int32_t dy, int32_t dz, const o_prim* prims, uint32_t n, uint16_t* vox, uint8_t* kind) {
  const double brickHalfDiag = 4.0 * sqrt(3.0) + 0.01;
  const double subHalfDiag = 2.0 * sqrt(3.0) + 0.01;
  const int32_t bx_n = dx >> 3, by_n = dy >> 3;
  const int32_t dims[3] = {dx, dy, dz};
  uint16_t scratch[512];
  for (uint32_t k = 0; k < n; k++) {
    const o_prim* P = &prims[k];
    const uint16_t pv = (uint16_t)P->mat;
    int32_t v0[3], v1[3];
    for (int a = 0; a < 3; a++) {
      double lo = floor(P->lo[a] - 0.5), hi = ceil(P->hi[a] + 0.5);
      v0[a] = lo < 0.0 ? 0 : (int32_t)lo;
      v1[a] = hi > (double)(dims[a] - 1) ? dims[a] - 1 : (int32_t)hi;
    }
    if (v0[0] > v1[0] || v0[1] > v1[1] || v0[2] > v1[2]) continue;
    for (int32_t bz = v0[2] >> 3; bz <= v1[2] >> 3; bz++)
      for (int32_t by = v0[1] >> 3; by <= v1[1] >> 3; by++)
        for (int32_t bx = v0[0] >> 3; bx <= v1[0] >> 3; bx++) {
          size_t ci = (size_t)bx + (size_t)bx_n * ((size_t)by + (size_t)by_n * (size_t)bz);
          int32_t ox = bx * 8, oy = by * 8, oz = bz * 8;
          int has = 0;
          int full = prim_contains(P, &has, ox + 0.5, oy + 0.5, oz + 0.5, ox + 7.5, oy + 7.5, oz + 7.5);
          if (!full) {
            double d = o_prim_sdf(P, ox + 4.0, oy + 4.0, oz + 4.0);
            if (d > brickHalfDiag) continue;
            full = d < -brickHalfDiag;
          }
          if (full) { /* setBrickUniform */
            for (uint32_t l = 0; l < 512u; l++)
              vox[(size_t)(ox + (l & 7u)) + (size_t)dx * ((size_t)(oy + ((l >> 3) & 7u)) + (size_t)dy * (size_t)(oz + (l >> 6)))] = pv;
            kind[ci] = pv ? 1 : 0;
            continue;
          }
          for (uint32_t l = 0; l < 512u; l++)
            scratch[l] = vox[(size_t)(ox + (l & 7u)) + (size_t)dx * ((size_t)(oy + ((l >> 3) & 7u)) + (size_t)dy * (size_t)(oz + (l >> 6)))];
          int changed = 0;
          for (int s = 0; s < 8; s++) {
            int32_t sx = ox + (s & 1) * 4, sy = oy + ((s >> 1) & 1) * 4, sz = oz + ((s >> 2) & 1) * 4;
            double d = o_prim_sdf(P, sx + 2.0, sy + 2.0, sz + 2.0);
            if (d > subHalfDiag) continue;
            int subFull = d < -subHalfDiag;
            for (int32_t z = sz; z < sz + 4; z++)
              for (int32_t y = sy; y < sy + 4; y++)
                for (int32_t x = sx; x < sx + 4; x++) {
                  if (!subFull && o_prim_sdf(P, x + 0.5, y + 0.5, z + 0.5) > 0) continue;
                  int li = (x & 7) | ((y & 7) << 3) | ((z & 7) << 6);
                  if (scratch[li] != pv) {
                    scratch[li] = pv;
                    changed = 1;
                  }
                }
          }
          if (!changed) continue;
          int same = 1;
          for (uint32_t l = 0; l < 512u; l++) {
            vox[(size_t)(ox + (l & 7u)) + (size_t)dx * ((size_t)(oy + ((l >> 3) & 7u)) + (size_t)dy * (size_t)(oz + (l >> 6)))] = scratch[l];
            if (scratch[l] != scratch[0]) same = 0;
          }
          kind[ci] = same ? (scratch[0] ? 1 : 0) : 2; /* compactCell */
        }
  }
}
