// px_shade.h - the shading of fsMain for gfx950, one pixel per lane:
//   skyline / skyColor            src/render/shaders/raymarch.wgsl:167-264
//   fineAO / coarseAO / shade     src/render/shaders/raymarch.wgsl:269-366
//   marble / materialColor        src/render/shaders/render_common.wgsl:49-140
//   hashf / hash2f / vnoise / fbm / hsv   src/render/shaders/shared.wgsl:76-106
//
// FP32 in WGSL source order (-ffp-contract=off).  Choices WGSL leaves to the
// implementation are fixed as follows:
//   sin, cos, exp, exp2, log2, pow and atan2 are evaluated in double precision
//   (Cody-Waite reduction by pi/2 with a 33-bit head, Taylor series of sin,
//   cos, e^t, atanh and atan on the reduced ranges) and rounded once to FP32,
//   so they are correctly rounded almost everywhere; pow(x, y) =
//   exp2(y log2 x) with pow(0, y > 0) = 0; FP32 results below 2^-126 flush
//   to zero;
//   max(a, b) = a < b ? b : a and min(a, b) = b < a ? b : a (a defined sign
//   of zero), clamp = min(max()), fract(x) = x - floor(x), round = half to
//   even, f32 -> i32 truncates and saturates;
//   the density texture (rgba8unorm, linear clamp-to-edge sampler) is read
//   with trilinear weights in FP32: s = uvw * size - 0.5, i0 = floor(s),
//   f = s - i0, texels clamped to the edge, each texel byte / 255.
#pragma once

#include "px_internal.h"

#pragma clang fp contract(off)

// What skyColor reads of the uniforms.
struct PxSky {
  float sun[4];        // U.sunDir: direction to the sun, intensity
  float sun_col[3];    // U.sunColor.rgb
  float night;         // U.sim.w
};

// Per-frame shading inputs decoded from the uniforms (shared.wgsl:5-19) and
// the light list (render_common.wgsl:8-18).
struct PxShadeDesc {
  PxSky sky;
  float ambient[4];    // U.ambient: sky ambient, coarse AO strength
  float flags[4];      // U.flags: reflections, lamp shadows, exposure, debug view
  float time;          // U.time.x
  float vs;            // U.dims.w, metres per voxel
  float fdims[3];      // U.dims.xyz
  int32_t bd[3];       // density texture size (brick cells)
  uint32_t n_lights;
  float lp[32][4];     // light position (m), radius (m)
  float lc[32][4];     // light colour, shadow flag
};

namespace px {
namespace fm {

// pi/2 = PIO2_HI + PIO2_LO (33-bit head: k * PIO2_HI is exact for |k| < 2^20),
// and the doubles nearest to 2/pi, pi, pi/2, pi/4, ln 2, log2 e, 2 / ln 2,
// sqrt 2, tan(pi/8).
constexpr double PIO2_HI = 0x1.921fb54400000p+0;
constexpr double PIO2_LO = 0x1.0b4611a626331p-34;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double D_PI = 0x1.921fb54442d18p+1;
constexpr double D_PIO2 = 0x1.921fb54442d18p+0;
constexpr double D_PIO4 = 0x1.921fb54442d18p-1;
constexpr double D_LN2 = 0x1.62e42fefa39efp-1;
constexpr double D_LOG2E = 0x1.71547652b82fep+0;
constexpr double D_TWO_OVER_LN2 = 0x1.71547652b82fep+1;
constexpr double D_SQRT2 = 0x1.6a09e667f3bcdp+0;
constexpr double D_TAN_PI_8 = 0x1.a827999fcef34p-2;

// Taylor coefficients 1/k! with signs: sin r = r sum S[k] r^2k, cos r = sum
// C[k] r^2k on |r| <= pi/4; e^t = sum E[k] t^k on |t| <= ln2/2.
PX_DEV double sin_r(double r) {
  const double r2 = r * r;
  double p = 1.0 / 355687428096000.0;
  p = p * r2 + -1.0 / 1307674368000.0;
  p = p * r2 + 1.0 / 6227020800.0;
  p = p * r2 + -1.0 / 39916800.0;
  p = p * r2 + 1.0 / 362880.0;
  p = p * r2 + -1.0 / 5040.0;
  p = p * r2 + 1.0 / 120.0;
  p = p * r2 + -1.0 / 6.0;
  p = p * r2 + 1.0;
  return r * p;
}
PX_DEV double cos_r(double r) {
  const double r2 = r * r;
  double p = -1.0 / 6402373705728000.0;
  p = p * r2 + 1.0 / 20922789888000.0;
  p = p * r2 + -1.0 / 87178291200.0;
  p = p * r2 + 1.0 / 479001600.0;
  p = p * r2 + -1.0 / 3628800.0;
  p = p * r2 + 1.0 / 40320.0;
  p = p * r2 + -1.0 / 720.0;
  p = p * r2 + 1.0 / 24.0;
  p = p * r2 + -1.0 / 2.0;
  p = p * r2 + 1.0;
  return p;
}
PX_DEV double sincos_d(float xf, int cosine) {
  const double x = (double)xf;
  if (x != x || x - x != 0.0) return x - x;
  const double k = rint(x * TWO_OVER_PI);
  const double r = (x - k * PIO2_HI) - k * PIO2_LO;
  const int q = ((int)(k - 4.0 * floor(k * 0.25)) + cosine) & 3;
  const double s = (q & 1) ? cos_r(r) : sin_r(r);
  return (q & 2) ? -s : s;
}
PX_DEV double exp2_d(double x) {
  if (x != x) return x;
  if (x >= 128.0) return __builtin_inf();
  if (x <= -150.0) return 0.0;
  const double n = rint(x);
  const double t = (x - n) * D_LN2;
  double p = 1.0 / 6227020800.0;
  p = p * t + 1.0 / 479001600.0;
  p = p * t + 1.0 / 39916800.0;
  p = p * t + 1.0 / 3628800.0;
  p = p * t + 1.0 / 362880.0;
  p = p * t + 1.0 / 40320.0;
  p = p * t + 1.0 / 5040.0;
  p = p * t + 1.0 / 720.0;
  p = p * t + 1.0 / 120.0;
  p = p * t + 1.0 / 24.0;
  p = p * t + 1.0 / 6.0;
  p = p * t + 1.0 / 2.0;
  p = p * t + 1.0;
  p = p * t + 1.0;
  return ldexp(p, (int)n);
}
PX_DEV float to_f32(double r) {
  if (fabs(r) < 0x1p-126) r = 0.0 * r;
  return (float)r;
}
PX_DEV double log2_d(float xf) {
  if (xf != xf || xf < 0.0f) return __builtin_nan("");
  if (xf == 0.0f) return -__builtin_inf();
  if (xf - xf != 0.0f) return __builtin_inf();
  int e = 0;
  if (xf < 0x1p-126f) {
    xf = xf * 0x1p24f;
    e = -24;
  }
  const uint32_t b = __float_as_uint(xf);
  e += (int)(b >> 23) - 127;
  double m = (double)__uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m > D_SQRT2) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  for (int k = 11; k >= 0; k--) p = p * s2 + 1.0 / (double)(2 * k + 1);
  return (double)e + D_TWO_OVER_LN2 * (s * p);
}

PX_DEV float sinf_(float x) { return (float)sincos_d(x, 0); }
PX_DEV float cosf_(float x) { return (float)sincos_d(x, 1); }
PX_DEV float expf_(float x) { return to_f32(exp2_d((double)x * D_LOG2E)); }
PX_DEV float powf_(float x, float y) {
  if (x != x || y != y || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_inff());
  return to_f32(exp2_d((double)y * log2_d(x)));
}
PX_DEV float atan2f_(float yf, float xf) {
  const double y = (double)yf, x = (double)xf;
  if (x != x || y != y) return __builtin_nanf("");
  const double ax = fabs(x), ay = fabs(y);
  double r;
  if (ay == 0.0) {
    r = signbit(x) ? D_PI : 0.0;
  } else if (ax - ax != 0.0 || ay - ay != 0.0) {
    if (ax == ay) r = x > 0.0 ? D_PIO4 : 3.0 * D_PIO4;
    else if (ay > ax) r = D_PIO2;
    else r = x > 0.0 ? 0.0 : D_PI;
  } else {
    const bool swap = ay > ax;
    const double t = swap ? ax / ay : ay / ax;
    double u = t, base = 0.0;
    if (t > D_TAN_PI_8) {
      u = (t - 1.0) / (t + 1.0);
      base = D_PIO4;
    }
    const double u2 = u * u;
    double p = 1.0 / 41.0;
    for (int k = 19; k >= 0; k--) p = p * u2 + ((k & 1) ? -1.0 : 1.0) / (double)(2 * k + 1);
    r = base + u * p;
    if (swap) r = D_PIO2 - r;
    if (x < 0.0) r = D_PI - r;
  }
  return (float)copysign(r, y);
}

}  // namespace fm

// ---- WGSL builtins -------------------------------------------------------
PX_DEV float maxf(float a, float b) { return a < b ? b : a; }
PX_DEV float minf(float a, float b) { return b < a ? b : a; }
PX_DEV float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }
PX_DEV float fractf(float x) { return x - floorf(x); }
PX_DEV float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
PX_DEV float stepf(float edge, float x) { return x >= edge ? 1.0f : 0.0f; }
PX_DEV float smoothstepf(float lo, float hi, float x) {
  const float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
PX_DEV float dot3f(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
PX_DEV float len3f(const float* a) { return sqrtf(dot3f(a, a)); }
PX_DEV float len2f(float a, float b) { return sqrtf(a * a + b * b); }
PX_DEV int32_t f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
PX_DEV uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// ---- hashes and noise (shared.wgsl:65-106) --------------------------------
PX_DEV uint32_t hashu(uint32_t x) {
  uint32_t h = x;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
PX_DEV uint32_t hash3i(int32_t x, int32_t y, int32_t z) {
  return hashu(((uint32_t)x * 0x9E3779B1u) ^ hashu(((uint32_t)y * 0x85EBCA77u) ^ hashu((uint32_t)z * 0xC2B2AE3Du)));
}
PX_DEV float hashf3(int32_t x, int32_t y, int32_t z) {
  return (float)(hash3i(x, y, z) & 0xffffffu) / 16777215.0f;
}
PX_DEV float hash2f(float x, float y) {
  const uint32_t a = (uint32_t)f2i_sat(floorf(x)) + 100000u;
  const uint32_t b = (uint32_t)f2i_sat(floorf(y)) + 7777u;
  return (float)(hashu((a * 0x27d4eb2du) ^ hashu(b)) & 0xffffu) / 65535.0f;
}
PX_DEV float vnoise(const float p[3]) {
  int32_t i[3];
  float u[3];
  for (int a = 0; a < 3; a++) {
    const float fl = floorf(p[a]);
    const float f = p[a] - fl;
    i[a] = f2i_sat(fl);
    u[a] = f * f * (3.0f - 2.0f * f);
  }
  const uint32_t x0 = (uint32_t)i[0], y0 = (uint32_t)i[1], z0 = (uint32_t)i[2];
  const float a = hashf3((int32_t)x0, (int32_t)y0, (int32_t)z0);
  const float b = hashf3((int32_t)(x0 + 1u), (int32_t)y0, (int32_t)z0);
  const float c = hashf3((int32_t)x0, (int32_t)(y0 + 1u), (int32_t)z0);
  const float d = hashf3((int32_t)(x0 + 1u), (int32_t)(y0 + 1u), (int32_t)z0);
  const float e = hashf3((int32_t)x0, (int32_t)y0, (int32_t)(z0 + 1u));
  const float f1 = hashf3((int32_t)(x0 + 1u), (int32_t)y0, (int32_t)(z0 + 1u));
  const float g = hashf3((int32_t)x0, (int32_t)(y0 + 1u), (int32_t)(z0 + 1u));
  const float h = hashf3((int32_t)(x0 + 1u), (int32_t)(y0 + 1u), (int32_t)(z0 + 1u));
  return mixf(mixf(mixf(a, b, u[0]), mixf(c, d, u[0]), u[1]), mixf(mixf(e, f1, u[0]), mixf(g, h, u[0]), u[1]), u[2]);
}
PX_DEV float fbm(const float p[3]) {
  float q2[3], q3[3];
  for (int a = 0; a < 3; a++) {
    q2[a] = p[a] * 2.03f + 17.0f;
    q3[a] = p[a] * 4.1f + 41.0f;
  }
  return vnoise(p) * 0.5f + vnoise(q2) * 0.3f + vnoise(q3) * 0.2f;
}
PX_DEV void hsv(float h, float s, float v, float out[3]) {
  const float k[3] = {1.0f, 2.0f / 3.0f, 1.0f / 3.0f};
  for (int c = 0; c < 3; c++) {
    const float p = fabsf(fractf(h + k[c]) * 6.0f - 3.0f);
    out[c] = v * mixf(1.0f, clampf(p - 1.0f, 0.0f, 1.0f), s);
  }
}

// ---- sky (raymarch.wgsl:167-264) ------------------------------------------
PX_DEV bool skyline(float az, float e, float layer, float night, float out[3]) {
  const float cellW = layer < 0.5f ? 0.075f : (layer < 1.5f ? 0.03f : 0.012f);
  const float cid = floorf(az / cellW);
  const float present = hash2f(cid, layer * 13.0f + 1.0f);
  if (layer < 0.5f && present < 0.55f) return false;
  const float hseed = hash2f(cid, layer * 31.0f + 3.0f);
  float lo = -0.045f, hi = -0.008f;
  if (layer < 1.5f) {
    lo = -0.07f;
    hi = 0.025f;
  }
  if (layer < 0.5f) {
    lo = -0.04f;
    hi = 0.13f;
  }
  const float top = mixf(lo, hi, fm::powf_(hseed, 1.6f));
  const float fx = fractf(az / cellW);
  const float inset = layer < 0.5f ? 0.22f : 0.1f;
  if (fx < inset || fx > 1.0f - inset) {
    if (hash2f(cid, 9.0f + layer) > 0.3f) return false;
  }
  const float base = layer < 0.5f ? -0.22f : (layer < 1.5f ? -0.095f : -0.05f);
  if (e > top || e < base) return false;
  const float depth = 1.0f - layer * 0.3f;
  const float k0[3] = {0.018f, 0.016f, 0.032f};
  const float hs = 0.7f + 0.6f * hash2f(cid, 5.0f);
  float c[3];
  for (int a = 0; a < 3; a++) c[a] = k0[a] * hs * depth;
  const float cols = layer < 0.5f ? 14.0f : 6.0f;
  const float rowsPerRad = layer < 0.5f ? 520.0f : 260.0f;
  const float wx = (fx - inset) / (1.0f - 2.0f * inset) * cols;
  const float wy = (top - e) * rowsPerRad;
  const float cellx = floorf(wx), celly = floorf(wy);
  const float lit = hash2f(cellx + cid * 31.0f, celly + layer * 101.0f);
  const float fwx = fractf(wx), fwy = fractf(wy);
  const float thresh = mixf(0.72f, 0.5f, night);
  if (lit > thresh && fwx > 0.2f && fwx < 0.8f && fwy > 0.35f && fwy < 0.85f && wx >= 0.0f) {
    const float warm = hash2f(cid, 77.0f);
    const float wc[3] = {warm > 0.8f ? 0.55f : 1.0f, warm > 0.8f ? 0.8f : 0.72f, warm > 0.8f ? 1.0f : 0.42f};
    const float s = 0.12f + 0.25f * (lit - thresh);
    for (int a = 0; a < 3; a++) c[a] = c[a] + wc[a] * s * depth * (0.7f + night);
  }
  const float hue = hash2f(cid, 41.0f);
  if (layer < 0.5f && hue > 0.45f) {
    float neon[3];
    hsv(hue * 2.3f, 0.75f, 1.0f, neon);
    const float nk = 0.8f + night * 1.2f;
    for (int a = 0; a < 3; a++) neon[a] = neon[a] * nk;
    if (fabsf(fx - inset) < 0.015f || fabsf(fx - (1.0f - inset)) < 0.015f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + neon[a] * 0.6f;
    if (top - e < 0.0035f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + neon[a] * 1.4f;
  }
  for (int a = 0; a < 3; a++) out[a] = c[a];
  return true;
}

// Not inlined: it is reached from the glass branch of the innermost DDA loop
// and from a few call sites of the shading kernels.
__device__ __noinline__ float3 sky_color(PxSky S, float3 rdv) {
  const float rd[3] = {rdv.x, rdv.y, rdv.z};
  float out[3];
  const float e = rd[1];
  const float night = S.night;
  const float h0[3] = {1.25f, 0.5f, 0.22f}, h1[3] = {0.12f, 0.06f, 0.16f};
  const float m0[3] = {0.55f, 0.22f, 0.38f}, m1[3] = {0.04f, 0.03f, 0.09f};
  const float z0[3] = {0.07f, 0.07f, 0.25f}, z1[3] = {0.005f, 0.006f, 0.02f};
  const float s1 = smoothstepf(0.0f, 0.18f, e), s2 = smoothstepf(0.12f, 0.7f, e);
  float c[3];
  for (int a = 0; a < 3; a++) {
    const float horizon = mixf(h0[a], h1[a], night);
    const float mid = mixf(m0[a], m1[a], night);
    const float zenith = mixf(z0[a], z1[a], night);
    c[a] = mixf(horizon, mid, s1);
    c[a] = mixf(c[a], zenith, s2);
  }
  const float sd = maxf(dot3f(rd, S.sun), 0.0f);
  const float sk = fm::powf_(sd, 1500.0f) * 40.0f + fm::powf_(sd, 12.0f) * 0.35f;
  for (int a = 0; a < 3; a++) c[a] = c[a] + S.sun_col[a] * sk * (1.0f - night);
  if (e > 0.05f) {
    int32_t sp[3];
    for (int a = 0; a < 3; a++) sp[a] = f2i_sat(floorf(rd[a] * 400.0f));
    if (hashf3(sp[0], sp[1], sp[2]) > 0.9985f)
      for (int a = 0; a < 3; a++) c[a] = c[a] + 0.6f * (0.3f + night);
  }
  const float az = fm::atan2f_(rd[2], rd[0]);
  const float beamAz = 0.35f;
  const float bw = fabsf(az - beamAz);
  if (e > -0.02f) {
    const float bc[3] = {0.6f, 0.75f, 1.0f};
    const float ex = fm::expf_(-bw * 2500.0f);
    const float ss = smoothstepf(0.9f, 0.0f, e);
    for (int a = 0; a < 3; a++) c[a] = c[a] + bc[a] * ex * (0.5f + 1.5f * night) * ss;
  }
  const float pyr = -0.02f - bw * 1.6f;
  if (e < pyr && e > -0.12f) {
    const float p0[3] = {0.03f, 0.025f, 0.04f}, p1[3] = {0.4f, 0.3f, 0.15f};
    const float st = stepf(0.97f, fractf((pyr - e) * 140.0f));
    for (int a = 0; a < 3; a++) out[a] = p0[a] + p1[a] * st * 0.3f;
    return make_float3(out[0], out[1], out[2]);
  }
  const float towerAz = 1.05f;
  const float tw = fabsf(az - towerAz);
  if ((tw < 0.0025f && e < 0.16f) || (tw < 0.008f && fabsf(e - 0.15f) < 0.008f)) {
    float tc[3] = {0.03f, 0.03f, 0.05f};
    if (fabsf(e - 0.15f) < 0.008f) {
      const float k[3] = {0.8f, 0.3f, 0.6f};
      const float st = stepf(0.5f, fractf(az * 900.0f));
      for (int a = 0; a < 3; a++) tc[a] = tc[a] + k[a] * st;
    }
    return make_float3(tc[0], tc[1], tc[2]);
  }
  for (int l = 0; l < 3; l++) {
    float s[3];
    if (skyline(az + (float)l * 1.7f, e, (float)l, night, s)) {
      const float hz = 0.03f + 0.05f * (float)l;
      for (int a = 0; a < 3; a++) out[a] = s[a] + c[a] * hz;
      return make_float3(out[0], out[1], out[2]);
    }
  }
  if (e < 0.0f) {
    const float g = 1.5f / maxf(-e, 1e-3f);
    const float gp[2] = {rd[0] * g, rd[2] * g};
    float city[3] = {0.012f, 0.01f, 0.02f};
    const float fx = fabsf(fractf(gp[0]) - 0.5f), fy = fabsf(fractf(gp[1]) - 0.5f);
    const float street = maxf(fx, fy);
    const float w = 0.012f / maxf(-e * 8.0f, 0.05f);
    const float lights = mixf(0.25f, 0.55f, night);
    if (street > 0.5f - w) {
      const float k[3] = {1.0f, 0.55f, 0.2f};
      for (int a = 0; a < 3; a++) city[a] = city[a] + k[a] * lights;
    }
    const float lotx = floorf(gp[0] * 4.0f), loty = floorf(gp[1] * 4.0f);
    if (hash2f(lotx, loty) > 0.9f) {
      float hc[3];
      hsv(hash2f(lotx + 3.0f, loty), 0.5f, 1.0f, hc);
      for (int a = 0; a < 3; a++) city[a] = city[a] + hc[a] * 0.25f * (0.5f + night);
    }
    const float m = fm::expf_(e * 25.0f);
    for (int a = 0; a < 3; a++) out[a] = mixf(city[a], c[a], m);
    return make_float3(out[0], out[1], out[2]);
  }
  return make_float3(c[0], c[1], c[2]);
}

// ---- material colour (render_common.wgsl:49-140) --------------------------
PX_DEV void marble(const float q[3], const float* A, const float* B, float out[3]) {
  float q13[3], q3[3];
  for (int a = 0; a < 3; a++) {
    q13[a] = q[a] * 1.3f;
    q3[a] = q[a] * 3.0f + 5.0f;
  }
  const float f = fbm(q13);
  const float v = fm::sinf_((q[0] + q[1] * 0.6f + q[2] * 0.8f) * 2.2f + f * 7.0f);
  const float vein = fm::powf_(1.0f - fabsf(v), 10.0f);
  const float cloud = fbm(q3);
  const float ck = 0.93f + 0.07f * cloud;
  for (int a = 0; a < 3; a++) out[a] = mixf(A[a] * ck, B[a], vein * 0.85f);
}

__device__ __noinline__ float3 material_color(const float* __restrict__ mats, float vs, float time, uint32_t v,
                                              int3 vpi) {
  const int32_t vp[3] = {vpi.x, vpi.y, vpi.z};
  const uint32_t id = v & 0xffu;
  const uint32_t shade = (v >> 8) & 0xffu;
  const float* A = mats + id * 20u;
  const float* B = A + 4;
  const float* Cm = A + 8;
  const float s = maxf(Cm[0], 1e-4f);
  float wp[3];
  for (int a = 0; a < 3; a++) wp[a] = ((float)vp[a] + 0.5f) * vs;
  float c[3] = {A[0], A[1], A[2]};
  if (shade > 0u) {
    const float t = (float)(shade - 1u) / 254.0f;
    for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], t);
  } else {
    const uint32_t pat = f2u_sat(B[3] + 0.5f);
    switch (pat) {
      case 1u: {
        const float q[3] = {wp[0] / s, wp[1] / s, wp[2] / s};
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
        const float p1[3] = {along * 3.0f, r * 20.0f, 0.0f};
        const float p2[3] = {wp[0] * 40.0f, wp[1] * 40.0f, wp[2] * 40.0f};
        const float n = vnoise(p1) * 0.8f + vnoise(p2) * 0.15f;
        const float ring = 0.5f + 0.5f * fm::sinf_((r / s + n * 2.0f) * 6.2831f);
        const float t = smoothstepf(0.2f, 0.95f, ring);
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], t);
        break;
      }
      case 5u: {
        int32_t tv = f2i_sat(rintf(s / vs));
        if (tv < 1) tv = 1;
        const int32_t tx = vp[0] / tv, tz = vp[2] / tv;
        const int32_t lx = vp[0] - tx * tv, lz = vp[2] - tz * tv;
        const float seed[3] = {(float)tx * 7.31f, 0.0f, (float)tz * 3.77f};
        const float sc = s * 0.45f;
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
        const float qx = wp[0] / s, qy = wp[2] / s;
        const float cx = floorf(qx), cy = floorf(qy);
        const float fx = fractf(qx) - 0.5f, fy = fractf(qy) - 0.5f;
        const float r = len2f(fx, fy);
        const float ang = fm::atan2f_(fy, fx);
        const float petals = 0.22f + 0.07f * fm::cosf_(ang * 6.0f + floorf(hash2f(cx, cy) * 2.0f) * 0.5236f);
        for (int a = 0; a < 3; a++) c[a] = A[a];
        if (fabsf(fabsf(fx) + fabsf(fy) - 0.46f) < 0.035f) {
          c[0] = 0.02f;
          c[1] = 0.18f;
          c[2] = 0.2f;
        }
        if (fabsf(r - petals) < 0.03f)
          for (int a = 0; a < 3; a++) c[a] = B[a];
        if (r < 0.07f)
          for (int a = 0; a < 3; a++) c[a] = B[a] * 1.1f;
        const float gx = fractf(qx * 2.0f + 0.25f) - 0.5f, gy = fractf(qy * 2.0f + 0.25f) - 0.5f;
        if (len2f(gx, gy) < 0.06f && r > 0.3f) {
          c[0] = 0.12f;
          c[1] = 0.02f;
          c[2] = 0.18f;
        }
        const float sw = fm::sinf_(qx * 9.0f + fm::sinf_(qy * 7.0f) * 1.5f);
        if (fabsf(sw) < 0.08f && r > 0.32f && r < 0.42f)
          for (int a = 0; a < 3; a++) c[a] = B[a] * 0.8f;
        break;
      }
      case 7u: {
        const float t = fractf((wp[0] + wp[2]) / s);
        const float m = smoothstepf(0.45f, 0.5f, t) * (1.0f - smoothstepf(0.95f, 1.0f, t));
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 8u: {
        const float weave = (float)((uint32_t)(vp[0] + vp[1] + vp[2]) & 1u);
        const float s20 = s * 20.0f;
        const float q[3] = {wp[0] / s20, wp[1] / s20, wp[2] / s20};
        const float m = weave * 0.35f + vnoise(q) * 0.4f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 9u: {
        const float m = stepf(0.75f, hashf3(vp[0], vp[1], vp[2]));
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 10u: {
        const float q[3] = {wp[0] / s * 0.3f, wp[1] / s * 6.0f, wp[2] / s * 0.3f};
        const float m = vnoise(q) * 0.6f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      case 11u: {
        const float tx = time;
        const float row = floorf(wp[1] / s);
        const float col = floorf((wp[0] + wp[2]) / s);
        const float spin = floorf(tx * 6.0f + col * 1.7f);
        const float sym = hash2f(row + spin, col);
        hsv(fractf(sym * 3.7f + tx * 0.05f), 0.85f, 0.6f + 0.4f * sym, c);
        const float edge = fractf(wp[1] / s);
        if (edge < 0.08f)
          for (int a = 0; a < 3; a++) c[a] = c[a] * 0.15f;
        break;
      }
      case 12u: {
        const float tx = time;
        const float t = fractf(wp[1] / s + tx * 0.03f);
        const float wave = 0.5f + 0.5f * fm::sinf_((wp[0] + wp[2]) * 3.0f + tx * 1.5f + wp[1] * 5.0f);
        const float m = t * 0.6f + wave * 0.4f;
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        if ((vp[1] & 3) == 0)
          for (int a = 0; a < 3; a++) c[a] = c[a] * 0.75f;
        break;
      }
      case 13u: {
        const float q[3] = {wp[0] / s, wp[1] / s, wp[2] / s};
        const float m = vnoise(q);
        for (int a = 0; a < 3; a++) c[a] = mixf(A[a], B[a], m);
        break;
      }
      default:
        break;
    }
  }
  const float n = hashf3(vp[0], vp[1], vp[2]);
  const float k = 1.0f + Cm[1] * (n - 0.5f) * 2.0f;
  return make_float3(maxf(c[0] * k, 0.0f), maxf(c[1] * k, 0.0f), maxf(c[2] * k, 0.0f));
}

// ---- AO (raymarch.wgsl:269-302) ---------------------------------------------
PX_DEV uint32_t voxel_at(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= w.dims[0] || y >= w.dims[1] || z >= w.dims[2]) return 0u;
  const uint32_t g = w.grid[(uint32_t)(x >> 3) +
                            (uint32_t)w.bd[0] * ((uint32_t)(y >> 3) + (uint32_t)w.bd[1] * (uint32_t)(z >> 3))];
  if (g == 0u) return 0u;
  return px_brick_voxel(w.pool, g, x & 7, y & 7, z & 7);
}
PX_DEV float opaque_at(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  const uint32_t v = voxel_at(w, x, y, z);
  if (v == 0u || w.mats[(v & 0xffu) * 20u + 13u] > 0.0f) return 0.0f;
  return 1.0f;
}

PX_DEV float fine_ao(const PxWorldView& w, const int32_t vox[3], const float N[3], const float P[3]) {
  int32_t n[3], u[3] = {1, 0, 0}, wv[3] = {0, 0, 1};
  for (int a = 0; a < 3; a++) n[a] = f2i_sat(rintf(N[a]));
  if (fabsf(N[0]) > 0.5f) {
    u[0] = 0; u[1] = 1; u[2] = 0;
    wv[0] = 0; wv[1] = 0; wv[2] = 1;
  } else if (fabsf(N[2]) > 0.5f) {
    u[0] = 1; u[1] = 0; u[2] = 0;
    wv[0] = 0; wv[1] = 1; wv[2] = 0;
  }
  int32_t b[3];
  for (int a = 0; a < 3; a++) b[a] = vox[a] + n[a];
  auto oq = [&](int su, int sw) {
    return opaque_at(w, b[0] + su * u[0] + sw * wv[0], b[1] + su * u[1] + sw * wv[1], b[2] + su * u[2] + sw * wv[2]);
  };
  const float su0 = oq(-1, 0), su1 = oq(1, 0), sw0 = oq(0, -1), sw1 = oq(0, 1);
  const float c00 = oq(-1, -1), c10 = oq(1, -1), c01 = oq(-1, 1), c11 = oq(1, 1);
  const float a00 = (su0 + sw0 > 1.5f ? 0.0f : 3.0f - (su0 + sw0 + c00)) / 3.0f;
  const float a10 = (su1 + sw0 > 1.5f ? 0.0f : 3.0f - (su1 + sw0 + c10)) / 3.0f;
  const float a01 = (su0 + sw1 > 1.5f ? 0.0f : 3.0f - (su0 + sw1 + c01)) / 3.0f;
  const float a11 = (su1 + sw1 > 1.5f ? 0.0f : 3.0f - (su1 + sw1 + c11)) / 3.0f;
  float f[3];
  for (int a = 0; a < 3; a++) f[a] = clampf(P[a] - (float)vox[a], 0.0f, 1.0f);
  const float uf[3] = {(float)u[0], (float)u[1], (float)u[2]}, wf[3] = {(float)wv[0], (float)wv[1], (float)wv[2]};
  const float fu = dot3f(f, uf), fw = dot3f(f, wf);
  const float ao = mixf(mixf(a00, a10, fu), mixf(a01, a11, fu), fw);
  return mixf(0.35f, 1.0f, ao);
}

PX_DEV float texel(const uint8_t* __restrict__ dens, const int32_t bd[3], int32_t x, int32_t y, int32_t z) {
  return (float)dens[(uint32_t)x + (uint32_t)bd[0] * ((uint32_t)y + (uint32_t)bd[1] * (uint32_t)z)] / 255.0f;
}
PX_DEV float sample_density(const uint8_t* __restrict__ dens, const int32_t bd[3], const float uvw[3]) {
  int32_t i0[3], i1[3];
  float f[3];
  for (int a = 0; a < 3; a++) {
    const float s = uvw[a] * (float)bd[a] - 0.5f;
    const float fl = floorf(s);
    f[a] = s - fl;
    const int32_t i = min(max(f2i_sat(fl), -1), bd[a]);
    i0[a] = min(max(i, 0), bd[a] - 1);
    i1[a] = min(max(i + 1, 0), bd[a] - 1);
  }
  const float c00 = mixf(texel(dens, bd, i0[0], i0[1], i0[2]), texel(dens, bd, i1[0], i0[1], i0[2]), f[0]);
  const float c10 = mixf(texel(dens, bd, i0[0], i1[1], i0[2]), texel(dens, bd, i1[0], i1[1], i0[2]), f[0]);
  const float c01 = mixf(texel(dens, bd, i0[0], i0[1], i1[2]), texel(dens, bd, i1[0], i0[1], i1[2]), f[0]);
  const float c11 = mixf(texel(dens, bd, i0[0], i1[1], i1[2]), texel(dens, bd, i1[0], i1[1], i1[2]), f[0]);
  return mixf(mixf(c00, c10, f[1]), mixf(c01, c11, f[1]), f[2]);
}
PX_DEV float coarse_ao(const PxShadeDesc& S, const uint8_t* __restrict__ dens, const float P[3], const float N[3]) {
  const float inv[3] = {1.0f / S.fdims[0], 1.0f / S.fdims[1], 1.0f / S.fdims[2]};
  const float dist[3] = {10.0f, 22.0f, 44.0f}, wt[3] = {0.55f, 0.35f, 0.3f};
  float occ = 0.0f;
  for (int k = 0; k < 3; k++) {
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = (P[a] + N[a] * dist[k]) * inv[a];
    occ = occ + sample_density(dens, S.bd, q) * wt[k];
  }
  return clampf(1.0f - occ * S.ambient[3], 0.12f, 1.0f);
}

}  // namespace px
