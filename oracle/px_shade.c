/*
 * px_shade.c - CPU oracle of the shading of fsMain (TEST INFRASTRUCTURE ONLY,
 * see px_oracle.h).
 *
 * A restatement, in FP32 rounded once per operation in WGSL source order
 * (-ffp-contract=off), of:
 *   skyline / skyColor            src/render/shaders/raymarch.wgsl:167-264
 *   fineAO / coarseAO / shade     src/render/shaders/raymarch.wgsl:269-366
 *   fsMain (colour, debug views)  src/render/shaders/raymarch.wgsl:387-450
 *   voxelAt / opaqueAt            src/render/shaders/render_common.wgsl:33-44
 *   marble / materialColor        src/render/shaders/render_common.wgsl:49-140
 *   hashf / hash2f / vnoise / fbm / hsv   src/render/shaders/shared.wgsl:76-106
 *   cell density of the AO texture        src/world/VoxelWorld.ts:159-170
 *
 * Choices WGSL leaves to the implementation, fixed here (the HIP library
 * makes the same ones):
 *   sin, cos, exp, exp2, log2, pow, atan2 are evaluated in double precision
 *   (argument reduction + Taylor series, below) and rounded once to FP32;
 *   pow(x, y) = exp2(y log2 x), pow(0, y > 0) = 0; pow(x, 5.0) in fsMain's
 *   Fresnel term = ((x*x)*(x*x))*x as in glassSurface;
 *   max(a, b) = a < b ? b : a, min(a, b) = b < a ? b : a, clamp = min(max());
 *   fract(x) = x - floor(x); round = round half to even;
 *   f32 -> i32 truncates and saturates, f32 -> u32 truncates, clamps to 0;
 *   textureSampleLevel on the rgba8unorm density texture with the linear,
 *   clamp-to-edge sampler = trilinear weights in FP32: s = uvw * size - 0.5,
 *   i0 = floor(s), f = s - i0, texels clamped to the edge, value / 255;
 *   the step counters of the debug views count super + brick iterations
 *   (outer) and voxel iterations (inner) of every march of the pixel.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "px_oracle.h"

/* ---------------------------------------------------------------------- */
/* Elementary functions: double-precision evaluation, one rounding to FP32. */

/* pi/2 = PIO2_HI + PIO2_LO, PIO2_HI with 33 significant bits (k * PIO2_HI is
 * exact for |k| < 2^20); the other constants are the doubles nearest to
 * 2/pi, pi, pi/2, pi/4, ln 2, log2 e, 2 / ln 2, sqrt 2 and tan(pi/8). */
static const double PIO2_HI = 0x1.921fb54400000p+0;
static const double PIO2_LO = 0x1.0b4611a626331p-34;
static const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double D_PI = 0x1.921fb54442d18p+1;
static const double D_PIO2 = 0x1.921fb54442d18p+0;
static const double D_PIO4 = 0x1.921fb54442d18p-1;
static const double D_LN2 = 0x1.62e42fefa39efp-1;
static const double D_LOG2E = 0x1.71547652b82fep+0;
static const double D_TWO_OVER_LN2 = 0x1.71547652b82fep+1;
static const double D_SQRT2 = 0x1.6a09e667f3bcdp+0;
static const double D_TAN_PI_8 = 0x1.a827999fcef34p-2;

/* Taylor coefficients: sin r = r * sum S[k] r^2k, cos r = sum C[k] r^2k
 * (|r| <= pi/4, truncation below 1e-19), e^t = sum E[k] t^k (|t| <= ln2 / 2),
 * atanh s / s = sum L[k] s^2k (|s| <= 0.172), atan u / u = sum A[k] u^2k
 * (|u| <= tan(pi/8)). */
static const double S_[9] = {1.0, -1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0,
                             1.0 / 6227020800.0, -1.0 / 1307674368000.0, 1.0 / 355687428096000.0};
static const double C_[10] = {1.0, -1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0,
                              1.0 / 479001600.0, -1.0 / 87178291200.0, 1.0 / 20922789888000.0,
                              -1.0 / 6402373705728000.0};
static const double E_[14] = {1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
                              1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0,
                              1.0 / 6227020800.0};

static double sin_r(double r) {
  double r2 = r * r, p = S_[8];
  for (int k = 7; k >= 0; k--) p = p * r2 + S_[k];
  return r * p;
}
static double cos_r(double r) {
  double r2 = r * r, p = C_[9];
  for (int k = 8; k >= 0; k--) p = p * r2 + C_[k];
  return p;
}

/* x = k pi/2 + r (Cody-Waite), quadrant q = k mod 4 */
static double sincos_d(float xf, int cosine) {
  double x = (double)xf;
  if (x != x || x - x != 0.0) return x - x; /* NaN for NaN and +-inf */
  double k = rint(x * TWO_OVER_PI);
  double r = (x - k * PIO2_HI) - k * PIO2_LO;
  int q = (int)(k - 4.0 * floor(k * 0.25)) + cosine;
  switch (q & 3) {
    case 0: return sin_r(r);
    case 1: return cos_r(r);
    case 2: return -sin_r(r);
    default: return -cos_r(r);
  }
}

static double exp2_d(double x) {
  if (x != x) return x;
  if (x >= 128.0) return INFINITY;
  if (x <= -150.0) return 0.0;
  double n = rint(x);
  double t = (x - n) * D_LN2;
  double p = E_[13];
  for (int k = 12; k >= 0; k--) p = p * t + E_[k];
  return ldexp(p, (int)n);
}

/* FP32 result; magnitudes below 2^-126 flush to zero */
static float to_f32(double r) {
  if (fabs(r) < 0x1p-126) r = 0.0 * r;
  return (float)r;
}

static uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float bitsf(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static double log2_d(float xf) {
  if (xf != xf || xf < 0.0f) return NAN;
  if (xf == 0.0f) return -INFINITY;
  if (xf - xf != 0.0f) return INFINITY;
  int e = 0;
  if (xf < 0x1p-126f) {
    xf = xf * 0x1p24f;
    e = -24;
  }
  uint32_t b = fbits(xf);
  e += (int)(b >> 23) - 127;
  double m = (double)bitsf((b & 0x7fffffu) | 0x3f800000u);
  if (m > D_SQRT2) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s, p = 1.0 / 25.0;
  for (int k = 11; k >= 0; k--) p = p * s2 + 1.0 / (double)(2 * k + 1);
  return (double)e + D_TWO_OVER_LN2 * (s * p);
}

float o_sinf(float x) { return (float)sincos_d(x, 0); }
float o_cosf(float x) { return (float)sincos_d(x, 1); }
float o_exp2f(float x) { return to_f32(exp2_d((double)x)); }
float o_expf(float x) { return to_f32(exp2_d((double)x * D_LOG2E)); }
float o_log2f(float x) { return (float)log2_d(x); }
float o_powf(float x, float y) {
  if (x != x || y != y || x < 0.0f) return NAN;
  if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
  return to_f32(exp2_d((double)y * log2_d(x)));
}
float o_atan2f(float yf, float xf) {
  double y = (double)yf, x = (double)xf;
  if (x != x || y != y) return NAN;
  double ax = fabs(x), ay = fabs(y), r;
  if (ay == 0.0) {
    r = signbit(x) ? D_PI : 0.0;
  } else if (ax - ax != 0.0 || ay - ay != 0.0) { /* an infinity */
    if (ax == ay) r = x > 0.0 ? D_PIO4 : 3.0 * D_PIO4;
    else if (ay > ax) r = D_PIO2;
    else r = x > 0.0 ? 0.0 : D_PI;
  } else {
    int swap = ay > ax;
    double t = swap ? ax / ay : ay / ax, u = t, base = 0.0;
    if (t > D_TAN_PI_8) {
      u = (t - 1.0) / (t + 1.0);
      base = D_PIO4;
    }
    double u2 = u * u, p = 1.0 / 41.0;
    for (int k = 19; k >= 0; k--) p = p * u2 + ((k & 1) ? -1.0 : 1.0) / (double)(2 * k + 1);
    r = base + u * p;
    if (swap) r = D_PIO2 - r;
    if (x < 0.0) r = D_PI - r;
  }
  return (float)copysign(r, y);
}

/* ---------------------------------------------------------------------- */
/* WGSL builtins and helpers */

static inline float maxf(float a, float b) { return a < b ? b : a; }
static inline float minf(float a, float b) { return b < a ? b : a; }
static inline float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }
static inline float fractf(float x) { return x - floorf(x); }
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float stepf(float edge, float x) { return x >= edge ? 1.0f : 0.0f; }
static inline float smoothstepf(float lo, float hi, float x) {
  float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
static inline float dot3f(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline float len3f(const float* a) { return sqrtf(dot3f(a, a)); }
static inline float len2f(float a, float b) { return sqrtf(a * a + b * b); }
static inline int32_t f2i_(float x) {
  if (x != x) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
static inline uint32_t f2u_(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
static inline int32_t iclamp_(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline const float* mat_(const float* m, uint32_t id, int k) { return m + id * 20u + 4u * (uint32_t)k; }

/* shared.wgsl:76-106 */
static float hashf3(int32_t x, int32_t y, int32_t z) {
  return (float)(o_hash3i(x, y, z) & 0xffffffu) / 16777215.0f;
}
static float hash2f(float x, float y) {
  uint32_t a = (uint32_t)f2i_(floorf(x)) + 100000u;
  uint32_t b = (uint32_t)f2i_(floorf(y)) + 7777u;
  return (float)(o_hashu((a * 0x27d4eb2du) ^ o_hashu(b)) & 0xffffu) / 65535.0f;
}
static float vnoise(const float p[3]) {
  int32_t i[3];
  float u[3];
  for (int a = 0; a < 3; a++) {
    float fl = floorf(p[a]);
    float f = p[a] - fl;
    i[a] = f2i_(fl);
    u[a] = f * f * (3.0f - 2.0f * f);
  }
  uint32_t x0 = (uint32_t)i[0], y0 = (uint32_t)i[1], z0 = (uint32_t)i[2];
  float a = hashf3((int32_t)x0, (int32_t)y0, (int32_t)z0);
  float b = hashf3((int32_t)(x0 + 1u), (int32_t)y0, (int32_t)z0);
  float c = hashf3((int32_t)x0, (int32_t)(y0 + 1u), (int32_t)z0);
  float d = hashf3((int32_t)(x0 + 1u), (int32_t)(y0 + 1u), (int32_t)z0);
  float e = hashf3((int32_t)x0, (int32_t)y0, (int32_t)(z0 + 1u));
  float f1 = hashf3((int32_t)(x0 + 1u), (int32_t)y0, (int32_t)(z0 + 1u));
  float g = hashf3((int32_t)x0, (int32_t)(y0 + 1u), (int32_t)(z0 + 1u));
  float h = hashf3((int32_t)(x0 + 1u), (int32_t)(y0 + 1u), (int32_t)(z0 + 1u));
  return mixf(mixf(mixf(a, b, u[0]), mixf(c, d, u[0]), u[1]), mixf(mixf(e, f1, u[0]), mixf(g, h, u[0]), u[1]), u[2]);
}
static float fbm(const float p[3]) {
  float q2[3], q3[3];
  for (int a = 0; a < 3; a++) {
    q2[a] = p[a] * 2.03f + 17.0f;
    q3[a] = p[a] * 4.1f + 41.0f;
  }
  return vnoise(p) * 0.5f + vnoise(q2) * 0.3f + vnoise(q3) * 0.2f;
}
static void hsv(float h, float s, float v, float out[3]) {
  const float k[3] = {1.0f, 2.0f / 3.0f, 1.0f / 3.0f};
  for (int c = 0; c < 3; c++) {
    float p = fabsf(fractf(h + k[c]) * 6.0f - 3.0f);
    out[c] = v * mixf(1.0f, clampf(p - 1.0f, 0.0f, 1.0f), s);
  }
}

/* ---------------------------------------------------------------------- */
/* Sky: raymarch.wgsl:167-264 (uniform words: sunDir 4-7, sunColor 8-10, night 27) */

static int skyline(float az, float e, float layer, float night, float out[3]) {
  float cellW = layer < 0.5f ? 0.075f : (layer < 1.5f ? 0.03f : 0.012f);
  float cid = floorf(az / cellW);
  float present = hash2f(cid, layer * 13.0f + 1.0f);
  if (layer < 0.5f && present < 0.55f) return 0;
  float hseed = hash2f(cid, layer * 31.0f + 3.0f);
  float lo = -0.045f, hi = -0.008f;
  if (layer < 1.5f) {
    lo = -0.07f;
    hi = 0.025f;
  }
  if (layer < 0.5f) {
    lo = -0.04f;
    hi = 0.13f;
  }
  float top = mixf(lo, hi, o_powf(hseed, 1.6f));
  float fx = fractf(az / cellW);
  float inset = layer < 0.5f ? 0.22f : 0.1f;
  if (fx < inset || fx > 1.0f - inset) {
    if (hash2f(cid, 9.0f + layer) > 0.3f) return 0;
  }
  float base = layer < 0.5f ? -0.22f : (layer < 1.5f ? -0.095f : -0.05f);
  if (e > top || e < base) return 0;
  float depth = 1.0f - layer * 0.3f;
  const float k0[3] = {0.018f, 0.016f, 0.032f};
  float hs = 0.7f + 0.6f * hash2f(cid, 5.0f);
  float c[3];
  for (int a = 0; a < 3; a++) c[a] = k0[a] * hs * depth;
  float cols = layer < 0.5f ? 14.0f : 6.0f;
  float rowsPerRad = layer < 0.5f ? 520.0f : 260.0f;
  float wx = (fx - inset) / (1.0f - 2.0f * inset) * cols;
  float wy = (top - e) * rowsPerRad;
  float cellx = floorf(wx), celly = floorf(wy);
  float lit = hash2f(cellx + cid * 31.0f, celly + layer * 101.0f);
  float fwx = fractf(wx), fwy = fractf(wy);
  float thresh = mixf(0.72f, 0.5f, night);
  if (lit > thresh && fwx > 0.2f && fwx < 0.8f && fwy > 0.35f && fwy < 0.85f && wx >= 0.0f) {
    float warm = hash2f(cid, 77.0f);
    const float wc0[3] = {1.0f, 0.72f, 0.42f}, wc1[3] = {0.55f, 0.8f, 1.0f};
    const float* wc = warm > 0.8f ? wc1 : wc0;
    float s = 0.12f + 0.25f * (lit - thresh);
    for (int a = 0; a < 3; a++) c[a] = c[a] + wc[a] * s * depth * (0.7f + night);
  }
  float hue = hash2f(cid, 41.0f);
  if (layer < 0.5f && hue > 0.45f) {
    float neon[3];
    hsv(hue * 2.3f, 0.75f, 1.0f, neon);
    float nk = 0.8f + night * 1.2f;
    for (int a = 0; a < 3; a++) neon[a] = neon[a] * nk;
    if (fabsf(fx - inset) < 0.015f || fabsf(fx - (1.0f - inset)) < 0.015f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + neon[a] * 0.6f;
    if (top - e < 0.0035f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + neon[a] * 1.4f;
  }
  memcpy(out, c, sizeof c);
  return 1;
}

void o_sky_color(const float* u, const float rd[3], float out[3]) {
  float e = rd[1];
  float night = u[27];
  const float h0[3] = {1.25f, 0.5f, 0.22f}, h1[3] = {0.12f, 0.06f, 0.16f};
  const float m0[3] = {0.55f, 0.22f, 0.38f}, m1[3] = {0.04f, 0.03f, 0.09f};
  const float z0[3] = {0.07f, 0.07f, 0.25f}, z1[3] = {0.005f, 0.006f, 0.02f};
  float s1 = smoothstepf(0.0f, 0.18f, e), s2 = smoothstepf(0.12f, 0.7f, e);
  float c[3];
  for (int a = 0; a < 3; a++) {
    float horizon = mixf(h0[a], h1[a], night);
    float mid = mixf(m0[a], m1[a], night);
    float zenith = mixf(z0[a], z1[a], night);
    c[a] = mixf(horizon, mid, s1);
    c[a] = mixf(c[a], zenith, s2);
  }
  float sd = maxf(dot3f(rd, u + 4), 0.0f);
  float sk = o_powf(sd, 1500.0f) * 40.0f + o_powf(sd, 12.0f) * 0.35f;
  for (int a = 0; a < 3; a++) c[a] = c[a] + u[8 + a] * sk * (1.0f - night);
  if (e > 0.05f) {
    int32_t sp[3];
    for (int a = 0; a < 3; a++) sp[a] = f2i_(floorf(rd[a] * 400.0f));
    if (hashf3(sp[0], sp[1], sp[2]) > 0.9985f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + 0.6f * (0.3f + night);
  }
  float az = o_atan2f(rd[2], rd[0]);
  float beamAz = 0.35f;
  float bw = fabsf(az - beamAz);
  if (e > -0.02f) {
    const float bc[3] = {0.6f, 0.75f, 1.0f};
    float ex = o_expf(-bw * 2500.0f);
    float ss = smoothstepf(0.9f, 0.0f, e);
    for (int a = 0; a < 3; a++) c[a] = c[a] + bc[a] * ex * (0.5f + 1.5f * night) * ss;
  }
  float pyr = -0.02f - bw * 1.6f;
  if (e < pyr && e > -0.12f) {
    const float p0[3] = {0.03f, 0.025f, 0.04f}, p1[3] = {0.4f, 0.3f, 0.15f};
    float st = stepf(0.97f, fractf((pyr - e) * 140.0f));
    for (int a = 0; a < 3; a++) out[a] = p0[a] + p1[a] * st * 0.3f;
    return;
  }
  float towerAz = 1.05f;
  float tw = fabsf(az - towerAz);
  if ((tw < 0.0025f && e < 0.16f) || (tw < 0.008f && fabsf(e - 0.15f) < 0.008f)) {
    float tc[3] = {0.03f, 0.03f, 0.05f};
    if (fabsf(e - 0.15f) < 0.008f) {
      const float k[3] = {0.8f, 0.3f, 0.6f};
      float st = stepf(0.5f, fractf(az * 900.0f));
      for (int a = 0; a < 3; a++) tc[a] = tc[a] + k[a] * st;
    }
    memcpy(out, tc, sizeof tc);
    return;
  }
  for (int l = 0; l < 3; l++) {
    float s[3];
    if (skyline(az + (float)l * 1.7f, e, (float)l, night, s)) {
      float hz = 0.03f + 0.05f * (float)l;
      for (int a = 0; a < 3; a++) out[a] = s[a] + c[a] * hz;
      return;
    }
  }
  if (e < 0.0f) {
    float g = 1.5f / maxf(-e, 1e-3f);
    float gp[2] = {rd[0] * g, rd[2] * g};
    float city[3] = {0.012f, 0.01f, 0.02f};
    float fx = fabsf(fractf(gp[0]) - 0.5f), fy = fabsf(fractf(gp[1]) - 0.5f);
    float street = maxf(fx, fy);
    float w = 0.012f / maxf(-e * 8.0f, 0.05f);
    float lights = mixf(0.25f, 0.55f, night);
    if (street > 0.5f - w) {
      const float k[3] = {1.0f, 0.55f, 0.2f};
      for (int a = 0; a < 3; a++) city[a] = city[a] + k[a] * lights;
    }
    float lotx = floorf(gp[0] * 4.0f), loty = floorf(gp[1] * 4.0f);
    if (hash2f(lotx, loty) > 0.9f) {
      float hc[3];
      hsv(hash2f(lotx + 3.0f, loty), 0.5f, 1.0f, hc);
      for (int a = 0; a < 3; a++) city[a] = city[a] + hc[a] * 0.25f * (0.5f + night);
    }
    float m = o_expf(e * 25.0f);
    for (int a = 0; a < 3; a++) out[a] = mixf(city[a], c[a], m);
    return;
  }
  memcpy(out, c, sizeof c);
}

/* ---------------------------------------------------------------------- */
/* Material colour: render_common.wgsl:49-140 */

static void marble(const float q[3], const float* A, const float* B, float out[3]) {
  float q13[3], q3[3];
  for (int a = 0; a < 3; a++) {
    q13[a] = q[a] * 1.3f;
    q3[a] = q[a] * 3.0f + 5.0f;
  }
  float f = fbm(q13);
  float v = o_sinf((q[0] + q[1] * 0.6f + q[2] * 0.8f) * 2.2f + f * 7.0f);
  float vein = o_powf(1.0f - fabsf(v), 10.0f);
  float cloud = fbm(q3);
  float ck = 0.93f + 0.07f * cloud;
  for (int a = 0; a < 3; a++) out[a] = mixf(A[a] * ck, B[a], vein * 0.85f);
}

void o_material_color(const float* mats, const float* u, uint32_t v, const int32_t vp[3], float out[3]) {
  uint32_t id = v & 0xffu;
  uint32_t shade = (v >> 8) & 0xffu;
  const float* A = mat_(mats, id, 0);
  const float* B = mat_(mats, id, 1);
  const float* Cm = mat_(mats, id, 2);
  float s = maxf(Cm[0], 1e-4f);
  float vs = u[3];
  float wp[3];
  for (int a = 0; a < 3; a++) wp[a] = ((float)vp[a] + 0.5f) * vs;
  float c[3] = {A[0], A[1], A[2]};
  if (shade > 0u) {
    float t = (float)(shade - 1u) / 254.0f;
    for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], t);
  } else {
    uint32_t pat = f2u_(B[3] + 0.5f);
    switch (pat) {
      case 1u: {
        float q[3] = {wp[0] / s, wp[1] / s, wp[2] / s};
        marble(q, A, B, c);
        break;
      }
      case 2u:
      case 3u:
      case 4u: {
        float r = len2f(wp[1], wp[2]);
        float along = wp[0];
        if (pat == 3u) {
          r = len2f(wp[0], wp[1]);
          along = wp[2];
        }
        if (pat == 4u) {
          r = len2f(wp[0], wp[2]);
          along = wp[1];
        }
        float p1[3] = {along * 3.0f, r * 20.0f, 0.0f};
        float p2[3] = {wp[0] * 40.0f, wp[1] * 40.0f, wp[2] * 40.0f};
        float n = vnoise(p1) * 0.8f + vnoise(p2) * 0.15f;
        float ring = 0.5f + 0.5f * o_sinf((r / s + n * 2.0f) * 6.2831f);
        float t = smoothstepf(0.2f, 0.95f, ring);
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], t);
        break;
      }
      case 5u: {
        int32_t tv = f2i_(rintf(s / vs));
        if (tv < 1) tv = 1;
        int32_t tx = vp[0] / tv, tz = vp[2] / tv;
        int32_t lx = vp[0] - tx * tv, lz = vp[2] - tz * tv;
        float seed[3] = {(float)tx * 7.31f, 0.0f, (float)tz * 3.77f};
        float sc = s * 0.45f;
        float q[3];
        for (int a = 0; a < 3; a++) q[a] = wp[a] / sc + seed[a];
        marble(q, A, B, c);
        if (((tx + tz) & 1) == 1) {
          const float k[3] = {0.94f, 0.93f, 0.92f};
          for (int a = 0; a < 3; a++) c[a] = c[a] * k[a];
        }
        if (lx == 0 || lz == 0) {
          c[0] = 0.12f;
          c[1] = 0.1f;
          c[2] = 0.09f;
        }
        break;
      }
      case 6u: {
        float qx = wp[0] / s, qy = wp[2] / s;
        float cx = floorf(qx), cy = floorf(qy);
        float fx = fractf(qx) - 0.5f, fy = fractf(qy) - 0.5f;
        float r = len2f(fx, fy);
        float ang = o_atan2f(fy, fx);
        float petals = 0.22f + 0.07f * o_cosf(ang * 6.0f + floorf(hash2f(cx, cy) * 2.0f) * 0.5236f);
        for (int a = 0; a < 3; a++) c[a] = A[a];
        const float teal[3] = {0.02f, 0.18f, 0.2f}, purple[3] = {0.12f, 0.02f, 0.18f};
        if (fabsf(fabsf(fx) + fabsf(fy) - 0.46f) < 0.035f) memcpy(c, teal, sizeof c);
        if (fabsf(r - petals) < 0.03f)
          for (int a = 0; a < 3; a++) c[a] = B[a];
        if (r < 0.07f)
          for (int a = 0; a < 3; a++) c[a] = B[a] * 1.1f;
        float gx = fractf(qx * 2.0f + 0.25f) - 0.5f, gy = fractf(qy * 2.0f + 0.25f) - 0.5f;
        if (len2f(gx, gy) < 0.06f && r > 0.3f) memcpy(c, purple, sizeof c);
        float sw = o_sinf(qx * 9.0f + o_sinf(qy * 7.0f) * 1.5f);
        if (fabsf(sw) < 0.08f && r > 0.32f && r < 0.42f)
          for (int a = 0; a < 3; a++) c[a] = B[a] * 0.8f;
        break;
      }
      case 7u: {
        float t = fractf((wp[0] + wp[2]) / s);
        float m = smoothstepf(0.45f, 0.5f, t) * (1.0f - smoothstepf(0.95f, 1.0f, t));
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 8u: {
        float weave = (float)((uint32_t)(vp[0] + vp[1] + vp[2]) & 1u);
        float s20 = s * 20.0f;
        float q[3] = {wp[0] / s20, wp[1] / s20, wp[2] / s20};
        float m = weave * 0.35f + vnoise(q) * 0.4f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 9u: {
        float m = stepf(0.75f, hashf3(vp[0], vp[1], vp[2]));
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 10u: {
        float q[3] = {wp[0] / s * 0.3f, wp[1] / s * 6.0f, wp[2] / s * 0.3f};
        float m = vnoise(q) * 0.6f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 11u: {
        float tx = u[28];
        float row = floorf(wp[1] / s);
        float col = floorf((wp[0] + wp[2]) / s);
        float spin = floorf(tx * 6.0f + col * 1.7f);
        float sym = hash2f(row + spin, col);
        hsv(fractf(sym * 3.7f + tx * 0.05f), 0.85f, 0.6f + 0.4f * sym, c);
        float edge = fractf(wp[1] / s);
        if (edge < 0.08f)
          for (int a = 0; a < 3; a++) c[a] = c[a] * 0.15f;
        break;
      }
      case 12u: {
        float tx = u[28];
        float t = fractf(wp[1] / s + tx * 0.03f);
        float wave = 0.5f + 0.5f * o_sinf((wp[0] + wp[2]) * 3.0f + tx * 1.5f + wp[1] * 5.0f);
        float m = t * 0.6f + wave * 0.4f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        if ((vp[1] & 3) == 0)
          for (int a = 0; a < 3; a++) c[a] = c[a] * 0.75f;
        break;
      }
      case 13u: {
        float q[3] = {wp[0] / s, wp[1] / s, wp[2] / s};
        float m = vnoise(q);
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      default:
        break;
    }
  }
  float n = hashf3(vp[0], vp[1], vp[2]);
  float k = 1.0f + Cm[1] * (n - 0.5f) * 2.0f;
  for (int a = 0; a < 3; a++) out[a] = maxf(c[a] * k, 0.0f);
}

/* ---------------------------------------------------------------------- */
/* Voxel reads and AO */

static uint32_t voxel_at(const o_world* w, int32_t x, int32_t y, int32_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= w->dx || y >= w->dy || z >= w->dz) return 0u;
  int32_t bx = w->dx >> 3, by = w->dy >> 3;
  uint32_t g = w->grid[(uint32_t)(x >> 3) + (uint32_t)bx * ((uint32_t)(y >> 3) + (uint32_t)by * (uint32_t)(z >> 3))];
  if (g == 0u) return 0u;
  if (g & 0x80000000u) return g & 0xffffu;
  uint32_t li = (uint32_t)(x & 7) | ((uint32_t)(y & 7) << 3) | ((uint32_t)(z & 7) << 6);
  uint32_t idx = (g - 1u) * 512u + li;
  return (w->pool[idx >> 1] >> ((idx & 1u) * 16u)) & 0xffffu;
}
static float opaque_at(const o_world* w, int32_t x, int32_t y, int32_t z) {
  uint32_t v = voxel_at(w, x, y, z);
  if (v == 0u || mat_(w->mats, v & 0xffu, 3)[1] > 0.0f) return 0.0f;
  return 1.0f;
}

static float fine_ao(const o_world* w, const int32_t vox[3], const float N[3], const float P[3]) {
  int32_t n[3], u[3] = {1, 0, 0}, wv[3] = {0, 0, 1};
  for (int a = 0; a < 3; a++) n[a] = f2i_(rintf(N[a]));
  if (fabsf(N[0]) > 0.5f) {
    u[0] = 0; u[1] = 1; u[2] = 0;
    wv[0] = 0; wv[1] = 0; wv[2] = 1;
  } else if (fabsf(N[2]) > 0.5f) {
    u[0] = 1; u[1] = 0; u[2] = 0;
    wv[0] = 0; wv[1] = 1; wv[2] = 0;
  }
  int32_t b[3];
  for (int a = 0; a < 3; a++) b[a] = vox[a] + n[a];
#define OQ(su, sw) opaque_at(w, b[0] + (su) * u[0] + (sw) * wv[0], b[1] + (su) * u[1] + (sw) * wv[1], \
                             b[2] + (su) * u[2] + (sw) * wv[2])
  float su0 = OQ(-1, 0), su1 = OQ(1, 0), sw0 = OQ(0, -1), sw1 = OQ(0, 1);
  float c00 = OQ(-1, -1), c10 = OQ(1, -1), c01 = OQ(-1, 1), c11 = OQ(1, 1);
#undef OQ
  float a00 = (su0 + sw0 > 1.5f ? 0.0f : 3.0f - (su0 + sw0 + c00)) / 3.0f;
  float a10 = (su1 + sw0 > 1.5f ? 0.0f : 3.0f - (su1 + sw0 + c10)) / 3.0f;
  float a01 = (su0 + sw1 > 1.5f ? 0.0f : 3.0f - (su0 + sw1 + c01)) / 3.0f;
  float a11 = (su1 + sw1 > 1.5f ? 0.0f : 3.0f - (su1 + sw1 + c11)) / 3.0f;
  float f[3];
  for (int a = 0; a < 3; a++) f[a] = clampf(P[a] - (float)vox[a], 0.0f, 1.0f);
  float uf[3] = {(float)u[0], (float)u[1], (float)u[2]}, wf[3] = {(float)wv[0], (float)wv[1], (float)wv[2]};
  float fu = dot3f(f, uf), fw = dot3f(f, wf);
  float ao = mixf(mixf(a00, a10, fu), mixf(a01, a11, fu), fw);
  return mixf(0.35f, 1.0f, ao);
}

/* density texel of cell (x, y, z) as the linear sampler reads it */
static float texel(const uint8_t* dens, const int32_t bd[3], int32_t x, int32_t y, int32_t z) {
  return (float)dens[(uint32_t)x + (uint32_t)bd[0] * ((uint32_t)y + (uint32_t)bd[1] * (uint32_t)z)] / 255.0f;
}
static float sample_density(const uint8_t* dens, const int32_t bd[3], const float uvw[3]) {
  int32_t i0[3], i1[3];
  float f[3];
  for (int a = 0; a < 3; a++) {
    float s = uvw[a] * (float)bd[a] - 0.5f;
    float fl = floorf(s);
    f[a] = s - fl;
    int32_t i = iclamp_(f2i_(fl), -1, bd[a]);
    i0[a] = iclamp_(i, 0, bd[a] - 1);
    i1[a] = iclamp_(i + 1, 0, bd[a] - 1);
  }
  float c00 = mixf(texel(dens, bd, i0[0], i0[1], i0[2]), texel(dens, bd, i1[0], i0[1], i0[2]), f[0]);
  float c10 = mixf(texel(dens, bd, i0[0], i1[1], i0[2]), texel(dens, bd, i1[0], i1[1], i0[2]), f[0]);
  float c01 = mixf(texel(dens, bd, i0[0], i0[1], i1[2]), texel(dens, bd, i1[0], i0[1], i1[2]), f[0]);
  float c11 = mixf(texel(dens, bd, i0[0], i1[1], i1[2]), texel(dens, bd, i1[0], i1[1], i1[2]), f[0]);
  return mixf(mixf(c00, c10, f[1]), mixf(c01, c11, f[1]), f[2]);
}

static float coarse_ao(const o_world* w, const uint8_t* dens, const float* u, const float P[3], const float N[3]) {
  int32_t bd[3] = {w->dx >> 3, w->dy >> 3, w->dz >> 3};
  float inv[3] = {1.0f / u[0], 1.0f / u[1], 1.0f / u[2]};
  const float dist[3] = {10.0f, 22.0f, 44.0f}, wt[3] = {0.55f, 0.35f, 0.3f};
  float occ = 0.0f;
  for (int k = 0; k < 3; k++) {
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = (P[a] + N[a] * dist[k]) * inv[a];
    occ = occ + sample_density(dens, bd, q) * wt[k];
  }
  return clampf(1.0f - occ * u[15], 0.12f, 1.0f);
}

void o_density(const o_world* w, uint8_t* out) {
  int32_t bd[3] = {w->dx >> 3, w->dy >> 3, w->dz >> 3};
  size_t ncell = (size_t)bd[0] * bd[1] * bd[2];
  for (size_t ci = 0; ci < ncell; ci++) {
    uint32_t g = w->grid[ci];
    if (g == 0u) {
      out[ci] = 0;
    } else if (g & 0x80000000u) {
      out[ci] = mat_(w->mats, g & 0xffu, 3)[1] > 0.0f ? 0 : 255;
    } else {
      uint32_t n = 0;
      for (uint32_t i = 0; i < 512u; i++) {
        uint32_t idx = (g - 1u) * 512u + i;
        uint32_t v = (w->pool[idx >> 1] >> ((idx & 1u) * 16u)) & 0xffffu;
        if (v != 0u && !(mat_(w->mats, v & 0xffu, 3)[1] > 0.0f)) n++;
      }
      out[ci] = (uint8_t)((n * 255u + 256u) >> 9); /* Math.round(n / 512 * 255) */
    }
  }
}

/* ---------------------------------------------------------------------- */
/* shade(): raymarch.wgsl:304-366 */

typedef struct o_shade_job {
  const o_world* w;
  const float* u;
  const float* lights;
  int n_lights;
  const uint8_t* dens;
  const float* rays;
  uint32_t n;
  float* rgba;
  volatile uint32_t next;
} o_shade_job;

/* a march() result as shade() and fsMain read it */
typedef struct o_mhit {
  int hit;
  float t;
  int32_t vox[3];
  float N[3];
  uint32_t value;
  float trans[3], glow[3];
  uint32_t outer, inner;
} o_mhit;

static void march_h(const o_shade_job* J, const float ro[3], const float rd[3], float max_t, int shadow, o_mhit* h) {
  o_hit o;
  o_march_ex(J->w, J->u, ro, rd, max_t, shadow, &o, h->trans, h->glow, &h->outer, &h->inner);
  h->hit = o.face != 0xFF;
  h->t = o.t;
  h->value = (uint32_t)o.mat | ((uint32_t)o.value_hi << 8);
  h->N[0] = h->N[1] = h->N[2] = 0.0f;
  h->vox[0] = h->vox[1] = h->vox[2] = 0;
  if (h->hit) {
    uint32_t dx = (uint32_t)J->w->dx, dy = (uint32_t)J->w->dy;
    h->vox[0] = (int32_t)(o.key % dx);
    h->vox[1] = (int32_t)((o.key / dx) % dy);
    h->vox[2] = (int32_t)(o.key / (dx * dy));
    if (o.face < 6) h->N[o.face >> 1] = (o.face & 1) ? -1.0f : 1.0f;
  }
}

/* colour of a hit; *outer / *inner accumulate the steps of its shadow marches */
static void shade(const o_shade_job* J, const o_mhit* h, const float ro[3], const float rd[3], int quality,
                  float out[3], uint32_t* outer, uint32_t* inner) {
  const float* u = J->u;
  const float* mats = J->w->mats;
  uint32_t id = h->value & 0xffu;
  float albedo[3];
  o_material_color(mats, u, h->value, h->vox, albedo);
  const float* A = mat_(mats, id, 0);
  const float* Cm = mat_(mats, id, 2);
  if (A[3] > 0.0f) {
    for (int a = 0; a < 3; a++) out[a] = albedo[a] * A[3];
    return;
  }
  const float* N = h->N;
  float P[3], O[3], V[3];
  for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h->t;
  for (int a = 0; a < 3; a++) O[a] = P[a] + N[a] * 0.5f;
  for (int a = 0; a < 3; a++) V[a] = -rd[a];
  float vs = u[3];
  float ao = coarse_ao(J->w, J->dens, u, P, N);
  if (quality > 0) ao = ao * fine_ao(J->w, h->vox, N, P);
  float diffuse[3] = {0.0f, 0.0f, 0.0f}, specular[3] = {0.0f, 0.0f, 0.0f};
  const float* sunL = u + 4;
  float ndl = dot3f(N, sunL);
  if (ndl > 0.0f && u[7] > 0.0f) {
    float sh[3] = {1.0f, 1.0f, 1.0f};
    if (quality > 0) {
      o_mhit s;
      march_h(J, O, sunL, 1e9f, 1, &s);
      *outer += s.outer;
      *inner += s.inner;
      for (int a = 0; a < 3; a++) sh[a] = s.hit ? 0.0f : s.trans[a];
    } else {
      sh[0] = sh[1] = sh[2] = 0.35f;
    }
    float Hv[3];
    for (int a = 0; a < 3; a++) Hv[a] = sunL[a] + V[a];
    float hl = len3f(Hv);
    for (int a = 0; a < 3; a++) Hv[a] = Hv[a] / hl;
    float sp = o_powf(maxf(dot3f(N, Hv), 0.0f), Cm[3]);
    for (int a = 0; a < 3; a++) {
      float sc = u[8 + a] * u[7] * sh[a];
      diffuse[a] = diffuse[a] + sc * ndl;
      specular[a] = specular[a] + sc * sp * Cm[2];
    }
  }
  for (int i = 0; i < J->n_lights; i++) {
    const float* lp = J->lights + (1 + 2 * i) * 4;
    const float* lc = J->lights + (2 + 2 * i) * 4;
    float L[3];
    for (int a = 0; a < 3; a++) L[a] = lp[a] / vs - P[a];
    float radius = lp[3] / vs;
    float d = len3f(L);
    if (d > radius) continue;
    for (int a = 0; a < 3; a++) L[a] = L[a] / d;
    float nl = dot3f(N, L);
    if (nl <= 0.0f) continue;
    float x = d / radius;
    float dm = d * vs;
    float atten = (1.0f - x * x) * (1.0f - x * x) / (1.0f + dm * dm * 1.5f);
    float sh[3] = {1.0f, 1.0f, 1.0f};
    if (quality > 0 && lc[3] > 0.5f && u[21] > 0.5f) {
      o_mhit s;
      march_h(J, O, L, d - 2.0f, 1, &s);
      *outer += s.outer;
      *inner += s.inner;
      for (int a = 0; a < 3; a++) sh[a] = s.hit ? 0.0f : s.trans[a];
    }
    float Hv[3];
    for (int a = 0; a < 3; a++) Hv[a] = L[a] + V[a];
    float hl = len3f(Hv);
    for (int a = 0; a < 3; a++) Hv[a] = Hv[a] / hl;
    float sp = o_powf(maxf(dot3f(N, Hv), 0.0f), Cm[3]);
    for (int a = 0; a < 3; a++) {
      float c = lc[a] * atten * sh[a];
      diffuse[a] = diffuse[a] + c * nl;
      specular[a] = specular[a] + c * sp * Cm[2];
    }
  }
  const float hemi0[3] = {0.35f, 0.25f, 0.22f};
  float ht = N[1] * 0.5f + 0.5f;
  float am = mixf(1.0f, ao, 0.5f);
  for (int a = 0; a < 3; a++) {
    float hemi = mixf(hemi0[a], u[12 + a], ht);
    float ambient = hemi * ao;
    out[a] = albedo[a] * (diffuse[a] * am + ambient) + specular[a];
  }
}

/* ---------------------------------------------------------------------- */
/* fsMain: raymarch.wgsl:387-450 (the colour; depth is not produced) */

static void pixel(const o_shade_job* J, uint32_t i) {
  const float* u = J->u;
  const float* r = J->rays + (size_t)i * 8u;
  const float* ro = r;
  const float* rd = r + 3;
  float* out = J->rgba + (size_t)i * 4u;
  o_mhit h;
  march_h(J, ro, rd, r[6], 0, &h);
  uint32_t outer = h.outer, inner = h.inner;
  int32_t dbg = f2i_(u[23]);
  float c[3] = {0.0f, 0.0f, 0.0f};
  if (dbg > 0) {
    if (h.hit) {
      if (dbg == 1)
        for (int a = 0; a < 3; a++) c[a] = h.N[a] * 0.5f + 0.5f;
      if (dbg == 2) o_material_color(J->w->mats, u, h.value, h.vox, c);
      if (dbg == 3) c[0] = c[1] = c[2] = fractf(h.t / 100.0f);
      if (dbg == 4) {
        float id = (float)(h.value & 0xffu);
        c[0] = id / 64.0f;
        c[1] = fractf(id * 0.37f);
        c[2] = 0.5f;
      }
      if (dbg == 5) shade(J, &h, ro, rd, 0, c, &outer, &inner);
      if (dbg == 8 || dbg == 9) {
        if (dbg == 9) {
          float s2[3];
          shade(J, &h, ro, rd, 1, s2, &outer, &inner);
        }
        c[0] = (float)outer / 300.0f;
        c[1] = (float)inner / 300.0f;
        c[2] = 0.0f;
      }
      if (dbg == 6) {
        float P[3];
        for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
        float ao = coarse_ao(J->w, J->dens, u, P, h.N) * fine_ao(J->w, h.vox, h.N, P);
        c[0] = c[1] = c[2] = ao;
      }
      if (dbg == 7) {
        float P[3], O[3];
        for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
        for (int a = 0; a < 3; a++) O[a] = P[a] + h.N[a] * 0.5f;
        o_mhit s;
        march_h(J, O, u + 4, 1e9f, 1, &s);
        float k = maxf(dot3f(h.N, u + 4), 0.0f);
        for (int a = 0; a < 3; a++) c[a] = (s.hit ? 0.0f : s.trans[a]) * k;
      }
    } else {
      c[0] = 0.2f;
      c[1] = 0.0f;
      c[2] = 0.3f;
    }
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
    out[3] = 1.0f;
    return;
  }
  if (!h.hit) {
    float sky[3];
    o_sky_color(u, rd, sky);
    for (int a = 0; a < 3; a++) c[a] = h.trans[a] * sky[a] + h.glow[a];
  } else {
    float base[3];
    shade(J, &h, ro, rd, 1, base, &outer, &inner);
    uint32_t id = h.value & 0xffu;
    float refl = mat_(J->w->mats, id, 3)[0];
    if (refl > 0.0f && u[20] > 0.5f) {
      float P[3], O[3], R[3], nrd[3];
      for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
      float dn = dot3f(h.N, rd);
      for (int a = 0; a < 3; a++) R[a] = rd[a] - 2.0f * dn * h.N[a];
      for (int a = 0; a < 3; a++) nrd[a] = -rd[a];
      float cosT = maxf(dot3f(nrd, h.N), 0.0f);
      float om = 1.0f - cosT;
      float p5 = om * om * (om * om) * om;
      float F = refl + (1.0f - refl) * p5 * refl * 2.0f;
      for (int a = 0; a < 3; a++) O[a] = P[a] + h.N[a] * 0.5f;
      o_mhit rh;
      march_h(J, O, R, 1e9f, 0, &rh);
      float rc[3];
      if (rh.hit) {
        float rs[3];
        shade(J, &rh, O, R, 0, rs, &outer, &inner);
        for (int a = 0; a < 3; a++) rc[a] = rh.trans[a] * rs[a] + rh.glow[a];
      } else {
        float sky[3];
        o_sky_color(u, R, sky);
        for (int a = 0; a < 3; a++) rc[a] = rh.trans[a] * sky[a] + rh.glow[a];
      }
      float fm = clampf(F, 0.0f, 0.9f);
      for (int a = 0; a < 3; a++) base[a] = mixf(base[a], rc[a], fm);
    }
    for (int a = 0; a < 3; a++) c[a] = h.trans[a] * base[a] + h.glow[a];
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = 1.0f;
}

static void* render_worker(void* arg) {
  o_shade_job* J = (o_shade_job*)arg;
  const uint32_t chunk = 64;
  for (;;) {
    uint32_t b = __atomic_fetch_add(&J->next, chunk, __ATOMIC_RELAXED);
    if (b >= J->n) break;
    uint32_t e = b + chunk < J->n ? b + chunk : J->n;
    for (uint32_t i = b; i < e; i++) pixel(J, i);
  }
  return NULL;
}

void o_render(const o_world* w, const float* u52, const float* lights_vec4, uint32_t n_vec4, const uint8_t* density,
              const float* rays, uint32_t n, float* rgba, int threads) {
  o_shade_job J;
  memset(&J, 0, sizeof J);
  J.w = w;
  J.u = u52;
  J.lights = lights_vec4;
  int count = n_vec4 > 0 ? f2i_(lights_vec4[0]) : 0;
  if (count < 0) count = 0;
  if ((uint32_t)(1 + 2 * count) > n_vec4) count = n_vec4 > 0 ? (int)((n_vec4 - 1) / 2) : 0;
  J.n_lights = count;
  J.dens = density;
  J.rays = rays;
  J.n = n;
  J.rgba = rgba;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  if (threads == 1) {
    render_worker(&J);
    return;
  }
  pthread_t th[256];
  int started = 0;
  for (int i = 0; i < threads; i++) {
    if (pthread_create(&th[i], NULL, render_worker, &J) != 0) break;
    started++;
  }
  if (started == 0) render_worker(&J);
  for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}
