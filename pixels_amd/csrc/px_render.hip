// px_render.hip - fsMain to an image (src/render/shaders/raymarch.wgsl:387-450)
// for gfx950: the colour of every primary ray, with shade() (:304-366), the
// sky (:167-264), glass transmittance and glow, glossy reflections and the
// debug views of U.flags.w.
//
// The reference runs one fragment invocation per pixel that marches the
// primary ray, then every shadow ray of shade() in turn, then the reflection
// ray.  Here the frame is split so each march kernel runs coherent work:
//   k_density         AO density per brick cell (VoxelWorld.ts:159-170, the
//                     maint rule of sim.wgsl:1255-1271), rebuilt after edits;
//   k_render_primary  primary march with transmittance and glow, the
//                     per-wavefront queue of the shadow rays shade() casts
//                     (the same queue as the traversal frame) and of the
//                     reflection rays;
//   k_render_shadow   the queued shadow rays: transmittance and hit per slot;
//   k_render_reflect  the reflection rays of reflective pixels, compacted in
//                     tile order into dense wavefronts (ordered scan of the
//                     per-region counts), marched with glow and shaded at
//                     quality 0;
//   k_shade           per pixel: material colour, coarse + fine AO, sun and
//                     lights with the queued shadow results, hemisphere
//                     ambient, the reflection mixed in by its Fresnel term,
//                     the sky for misses;
//   k_render_debug    debug views 1-9, each pixel marched in the reference's
//                     order (their step counters count every march).
#include <algorithm>
#include <cstring>

#include "px_internal.h"
#include "px_march.h"
#include "px_scan.h"
#include "px_shade.h"

#pragma clang fp contract(off)

namespace {

using px::MarchOut;
using px::MarchShade;
using px::MarchTables;

constexpr float kShMiss = 0.0f, kShHit = 1.0f, kShNotCast = 2.0f;

// ---- density ------------------------------------------------------------
// One wavefront per cell: byte = round(opaque voxels / 512 * 255); uniform
// cells 255 unless glass; empty cells 0.
__global__ __launch_bounds__(256) void k_density(PxWorldView w, uint32_t ncell, uint8_t* __restrict__ dens) {
  __shared__ uint32_t s_glass[8];
  if (threadIdx.x < 8u) s_glass[threadIdx.x] = 0u;
  __syncthreads();
  if (threadIdx.x < 256u && w.mats[threadIdx.x * 20u + 13u] > 0.0f)
    atomicOr(&s_glass[threadIdx.x >> 5], 1u << (threadIdx.x & 31u));
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t ci = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); ci < ncell; ci += waves) {
    const uint32_t g = w.grid[ci];
    uint32_t d;
    if (g == 0u) {
      d = 0u;
    } else if (g & PX_UNIFORM_BIT) {
      d = px::bit_of(s_glass, g & 0xffu) ? 0u : 255u;
    } else {
      const uint4 q = reinterpret_cast<const uint4*>(w.pool + (size_t)(g - 1u) * 256u)[lane];
      const uint32_t words[4] = {q.x, q.y, q.z, q.w};
      uint32_t cnt = 0u;
      for (int k = 0; k < 4; k++)
        for (int h = 0; h < 2; h++) {
          const uint32_t v = (words[k] >> (16 * h)) & 0xffffu;
          cnt += (v != 0u && !px::bit_of(s_glass, v & 0xffu)) ? 1u : 0u;
        }
      for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
      d = (cnt * 255u + 256u) >> 9;
    }
    if (lane == 0u) dens[ci] = (uint8_t)d;
  }
}

// ---- a march result as shade() reads it -----------------------------------
struct HitS {
  bool hit;
  float t;
  int32_t vox[3];
  float N[3];
  uint32_t value;
};

PX_DEV HitS decode_hit(const PxWorldView& w, uint32_t key, float t, uint32_t face, uint32_t value) {
  HitS h;
  h.hit = face != 0xFFu;
  h.t = t;
  h.value = value;
  h.N[0] = h.N[1] = h.N[2] = 0.0f;
  h.vox[0] = h.vox[1] = h.vox[2] = 0;
  if (h.hit) {
    const uint32_t dx = (uint32_t)w.dims[0], dy = (uint32_t)w.dims[1];
    h.vox[0] = (int32_t)(key % dx);
    h.vox[1] = (int32_t)((key / dx) % dy);
    h.vox[2] = (int32_t)(key / (dx * dy));
    if (face < 6u) h.N[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  }
  return h;
}

PX_DEV HitS decode_march(const PxWorldView& w, const MarchOut& m) {
  return decode_hit(w, m.key, m.t, m.face, m.value);
}

// shade() of raymarch.wgsl:304-366.  Q = 0: quality 0 (no shadow rays, no
// fine AO); Q = 1: quality 1 with the shadow results of the queue (`shq`,
// slot 0 the sun, 1 + i light i); Q = 2: quality 1 marching the shadow rays
// here, adding their steps to *outer / *inner.
template <int Q>
PX_DEV float3 shade_px(const PxWorldView& w, const MarchTables& T, const PxShadeDesc& S,
                       const uint8_t* __restrict__ dens, const HitS& h, const float ro[3], const float rd[3],
                       const float4* __restrict__ shq, uint32_t* outer, uint32_t* inner) {
  const uint32_t id = h.value & 0xffu;
  const float3 alb = px::material_color(w.mats, S.vs, S.time, h.value, make_int3(h.vox[0], h.vox[1], h.vox[2]));
  const float albedo[3] = {alb.x, alb.y, alb.z};
  const float* A = w.mats + id * 20u;
  const float* Cm = A + 8;
  if (A[3] > 0.0f) return make_float3(albedo[0] * A[3], albedo[1] * A[3], albedo[2] * A[3]);
  const float* N = h.N;
  float P[3], O[3], V[3];
  for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
  for (int a = 0; a < 3; a++) O[a] = P[a] + N[a] * 0.5f;
  for (int a = 0; a < 3; a++) V[a] = -rd[a];
  const float vs = S.vs;
  float ao = px::coarse_ao(S, dens, P, N);
  if (Q > 0) ao = ao * px::fine_ao(w, h.vox, N, P);
  float diffuse[3] = {0.0f, 0.0f, 0.0f}, specular[3] = {0.0f, 0.0f, 0.0f};
  const float* sunL = S.sky.sun;
  const float ndl = px::dot3f(N, sunL);
  if (ndl > 0.0f && S.sky.sun[3] > 0.0f) {
    float sh[3] = {1.0f, 1.0f, 1.0f};
    if (Q == 1) {
      const float4 r = shq[0];
      if (r.w == kShHit) sh[0] = sh[1] = sh[2] = 0.0f;
      else if (r.w == kShMiss) {
        sh[0] = r.x;
        sh[1] = r.y;
        sh[2] = r.z;
      }
    } else if (Q == 2) {
      MarchShade ms;
      const MarchOut m = px::march<true, true>(w, T, O, sunL, 1e9f, nullptr, &ms);
      *outer += m.steps - ms.inner;
      *inner += ms.inner;
      for (int a = 0; a < 3; a++) sh[a] = m.face != 0xFFu ? 0.0f : ms.trans[a];
    } else {
      sh[0] = sh[1] = sh[2] = 0.35f;
    }
    float Hv[3];
    for (int a = 0; a < 3; a++) Hv[a] = sunL[a] + V[a];
    const float hl = px::len3f(Hv);
    for (int a = 0; a < 3; a++) Hv[a] = Hv[a] / hl;
    const float sp = px::fm::powf_(px::maxf(px::dot3f(N, Hv), 0.0f), Cm[3]);
    for (int a = 0; a < 3; a++) {
      const float sc = S.sky.sun_col[a] * S.sky.sun[3] * sh[a];
      diffuse[a] = diffuse[a] + sc * ndl;
      specular[a] = specular[a] + sc * sp * Cm[2];
    }
  }
  for (uint32_t i = 0; i < S.n_lights; i++) {
    const float* lp = S.lp[i];
    const float* lc = S.lc[i];
    float L[3];
    for (int a = 0; a < 3; a++) L[a] = lp[a] / vs - P[a];
    const float radius = lp[3] / vs;
    const float d = px::len3f(L);
    if (d > radius) continue;
    for (int a = 0; a < 3; a++) L[a] = L[a] / d;
    const float nl = px::dot3f(N, L);
    if (nl <= 0.0f) continue;
    const float x = d / radius;
    const float dm = d * vs;
    const float atten = (1.0f - x * x) * (1.0f - x * x) / (1.0f + dm * dm * 1.5f);
    float sh[3] = {1.0f, 1.0f, 1.0f};
    if (Q > 0 && lc[3] > 0.5f && S.flags[1] > 0.5f) {
      if (Q == 1) {
        const float4 r = shq[1u + i];
        if (r.w == kShHit) sh[0] = sh[1] = sh[2] = 0.0f;
        else if (r.w == kShMiss) {
          sh[0] = r.x;
          sh[1] = r.y;
          sh[2] = r.z;
        }
      } else {
        MarchShade ms;
        const MarchOut m = px::march<true, true>(w, T, O, L, d - 2.0f, nullptr, &ms);
        *outer += m.steps - ms.inner;
        *inner += ms.inner;
        for (int a = 0; a < 3; a++) sh[a] = m.face != 0xFFu ? 0.0f : ms.trans[a];
      }
    }
    float Hv[3];
    for (int a = 0; a < 3; a++) Hv[a] = L[a] + V[a];
    const float hl = px::len3f(Hv);
    for (int a = 0; a < 3; a++) Hv[a] = Hv[a] / hl;
    const float sp = px::fm::powf_(px::maxf(px::dot3f(N, Hv), 0.0f), Cm[3]);
    for (int a = 0; a < 3; a++) {
      const float c = lc[a] * atten * sh[a];
      diffuse[a] = diffuse[a] + c * nl;
      specular[a] = specular[a] + c * sp * Cm[2];
    }
  }
  const float hemi0[3] = {0.35f, 0.25f, 0.22f};
  const float ht = N[1] * 0.5f + 0.5f;
  const float am = px::mixf(1.0f, ao, 0.5f);
  float out[3];
  for (int a = 0; a < 3; a++) {
    const float hemi = px::mixf(hemi0[a], S.ambient[a], ht);
    const float ambient = hemi * ao;
    out[a] = albedo[a] * (diffuse[a] * am + ambient) + specular[a];
  }
  return make_float3(out[0], out[1], out[2]);
}

// The reflection ray of fsMain (raymarch.wgsl:429-440) for a hit on a
// reflective material: origin P + N 0.5 and direction reflect(rd, N).
PX_DEV void reflection_ray(const HitS& h, const float ro[3], const float rd[3], float O[3], float R[3]) {
  float P[3];
  for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
  const float dn = px::dot3f(h.N, rd);
  for (int a = 0; a < 3; a++) R[a] = rd[a] - 2.0f * dn * h.N[a];
  for (int a = 0; a < 3; a++) O[a] = P[a] + h.N[a] * 0.5f;
}

PX_DEV bool reflects(const PxWorldView& w, const PxShadeDesc& S, const HitS& h) {
  return h.hit && w.mats[(h.value & 0xffu) * 20u + 12u] > 0.0f && S.flags[0] > 0.5f;
}

PX_DEV void store_rgba(float* __restrict__ rgba, uint32_t i, const float c[3]) {
  reinterpret_cast<float4*>(rgba)[i] = make_float4(c[0], c[1], c[2], 1.0f);
}

// ---- the frame ------------------------------------------------------------
// Primary march with transmittance and glow (tg: 2 float4 per pixel) and the
// shadow-ray queue of the traversal frame (per-wavefront regions).
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_render_primary(PxWorldView w, PxShadowDesc sd, PxSky sky, float refl_on,
                                                        const px_ray* __restrict__ rays, uint32_t n,
                                                        px_hit* __restrict__ prim_out, float4* __restrict__ tg,
                                                        uint32_t* __restrict__ queue,
                                                        uint32_t* __restrict__ wave_count,
                                                        uint32_t* __restrict__ rqueue,
                                                        uint2* __restrict__ rcount) {
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
    px::load_ray(rays + i, ro, rd, maxT, flags);
    MarchShade ms;
    h = px::march<false, true>(w, T, ro, rd, maxT, &sky, &ms);
    px::store_hit(prim_out + i, h);
    tg[2u * i] = make_float4(ms.trans[0], ms.trans[1], ms.trans[2], 0.0f);
    tg[2u * i + 1u] = make_float4(ms.glow[0], ms.glow[1], ms.glow[2], 0.0f);
  }
  const uint32_t gwave = i >> 6;
  uint32_t* __restrict__ wq = queue + (size_t)gwave * 64u * sd.slots;
  const uint32_t lane = __lane_id();
  uint32_t wcount = 0u;
  for (uint32_t s = 0; s < sd.slots; s++) {
    float O[3], dir[3], mt;
    const bool cast = live && px::shadow_ray(w, sd, ro, rd, h.t, h.face, h.value, s, O, dir, mt);
    const uint64_t mask = __ballot(cast);
    if (cast) wq[wcount + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = i * sd.slots + s;
    wcount += (uint32_t)__popcll(mask);
  }
  // reflection rays of the region's pixels, packed by ballot prefix
  const bool refl = live && h.face != 0xFFu && w.mats[(h.value & 0xffu) * 20u + 12u] > 0.0f && refl_on > 0.5f;
  const uint64_t rmask = __ballot(refl);
  if (refl) rqueue[(size_t)gwave * 64u + (uint32_t)__popcll(rmask & ((1ull << lane) - 1ull))] = i;
  // the last block may hold wavefronts past the last region (no live lanes)
  if (lane == 0u && gwave < (n + 63u) / 64u) {
    wave_count[gwave] = wcount;
    rcount[gwave] = make_uint2((uint32_t)__popcll(rmask), 0u);
  }
}

// The reflection rays of all regions in region order: region r's entries go
// to offset roff[r] (ordered scan of the counts).
__global__ __launch_bounds__(256) void k_refl_compact(const uint32_t* __restrict__ rqueue,
                                                      const uint2* __restrict__ rcount,
                                                      const uint2* __restrict__ roff, uint32_t n_regions,
                                                      uint32_t* __restrict__ dense) {
  const uint32_t region = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (region >= n_regions) return;
  const uint32_t lane = __lane_id();
  if (lane < rcount[region].x) dense[roff[region].x + lane] = rqueue[(size_t)region * 64u + lane];
}

// One lane per queued reflection ray (dense list, *total of them): marched
// with glow and shaded at quality 0; rc[pixel] = trans * colour + glow (or
// the sky).
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_render_reflect(PxWorldView w, PxShadeDesc S, const px_ray* __restrict__ rays,
                                                        const px_hit* __restrict__ prim,
                                                        const uint32_t* __restrict__ dense,
                                                        const uint32_t* __restrict__ total,
                                                        const uint8_t* __restrict__ dens, float4* __restrict__ rc) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  // the dense list last first, as the shadow regions (costly rows first)
  const uint32_t tot = *total;
  const uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
  if (jj >= tot) return;
  const uint32_t i = dense[tot - 1u - jj];
  float ro[3], rd[3], maxT;
  uint32_t flags;
  px::load_ray(rays + i, ro, rd, maxT, flags);
  const uint4 hv = *reinterpret_cast<const uint4*>(prim + i);
  const HitS h = decode_hit(w, hv.x, __uint_as_float(hv.y), hv.w & 0xffu, (hv.w >> 8) & 0xffffu);
  float O[3], R[3];
  reflection_ray(h, ro, rd, O, R);
  MarchShade ms;
  const MarchOut m = px::march<false, true>(w, T, O, R, 1e9f, &S.sky, &ms);
  float3 c;
  if (m.face != 0xFFu) {
    const HitS rh = decode_march(w, m);
    uint32_t o = 0, in = 0;
    c = shade_px<0>(w, T, S, dens, rh, O, R, nullptr, &o, &in);
  } else {
    c = px::sky_color(S.sky, make_float3(R[0], R[1], R[2]));
  }
  rc[i] = make_float4(ms.trans[0] * c.x + ms.glow[0], ms.trans[1] * c.y + ms.glow[1], ms.trans[2] * c.z + ms.glow[2],
                      0.0f);
}

// One wavefront per queue region: the slots no queued ray covers read "not
// cast" (a bitmap of the queue in LDS), the queued rays write
// (transmittance, hit).
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_render_shadow(PxWorldView w, PxShadowDesc sd, const px_ray* __restrict__ rays,
                                                       const px_hit* __restrict__ prim,
                                                       const uint32_t* __restrict__ queue,
                                                       const uint32_t* __restrict__ wave_count, uint32_t n,
                                                       float4* __restrict__ sh_out) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  // regions last first, as in k_trace_shadow_queue (costly rows first)
  const uint32_t n_regions = (n + 63u) / 64u;
  const uint32_t lw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (lw >= n_regions) return;
  const uint32_t region = n_regions - 1u - lw;
  const uint32_t slots = sd.slots;
  const uint32_t first = region * 64u * slots, last = min(region * 64u + 64u, n) * slots;
  const uint32_t lane = __lane_id();
  const uint32_t cnt = wave_count[region];
  const uint32_t* __restrict__ rq = queue + (size_t)region * 64u * slots;
  // bitmap of the region's cast slots (slots <= 33): only the others get the
  // "not cast" record, so every output is written once
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
  for (uint32_t k = first + lane; k < last; k += 64u) {
    const uint32_t local = k - first;
    if (((bm[local >> 5] >> (local & 31u)) & 1u) == 0u) sh_out[k] = make_float4(1.0f, 1.0f, 1.0f, kShNotCast);
  }
  for (uint32_t k = lane; k < cnt; k += 64u) {
    const uint32_t code = rq[k];
    const uint32_t i = code / slots, s = code - i * slots;
    float ro[3], rd[3], maxT;
    uint32_t flags;
    px::load_ray(rays + i, ro, rd, maxT, flags);
    const uint4 hv = *reinterpret_cast<const uint4*>(prim + i);
    const uint32_t face = hv.w & 0xffu, value = (hv.w >> 8) & 0xffffu;
    float O[3], dir[3], mt;
    px::shadow_ray(w, sd, ro, rd, __uint_as_float(hv.y), face, value, s, O, dir, mt);
    MarchShade ms;
    const MarchOut m = px::march<true, true>(w, T, O, dir, mt, nullptr, &ms);
    sh_out[code] = make_float4(ms.trans[0], ms.trans[1], ms.trans[2], m.face != 0xFFu ? kShHit : kShMiss);
  }
}

// The colour of every pixel from its primary hit, transmittance / glow and
// shadow results.
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_shade(PxWorldView w, PxShadeDesc S, uint32_t slots,
                                               const px_ray* __restrict__ rays, uint32_t n,
                                               const px_hit* __restrict__ prim, const float4* __restrict__ tg,
                                               const float4* __restrict__ sh, const float4* __restrict__ rc,
                                               const uint8_t* __restrict__ dens, float* __restrict__ rgba) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ro[3], rd[3], maxT;
  uint32_t flags;
  px::load_ray(rays + i, ro, rd, maxT, flags);
  const uint4 hv = *reinterpret_cast<const uint4*>(prim + i);
  const HitS h = decode_hit(w, hv.x, __uint_as_float(hv.y), hv.w & 0xffu, (hv.w >> 8) & 0xffffu);
  const float4 tr = tg[2u * i], gl = tg[2u * i + 1u];
  const float trans[3] = {tr.x, tr.y, tr.z}, glow[3] = {gl.x, gl.y, gl.z};
  float c[3];
  if (!h.hit) {
    const float3 sk = px::sky_color(S.sky, make_float3(rd[0], rd[1], rd[2]));
    c[0] = trans[0] * sk.x + glow[0];
    c[1] = trans[1] * sk.y + glow[1];
    c[2] = trans[2] * sk.z + glow[2];
  } else {
    uint32_t o = 0, in = 0;
    const float3 b = shade_px<1>(w, T, S, dens, h, ro, rd, sh + (size_t)i * slots, &o, &in);
    float base[3] = {b.x, b.y, b.z};
    if (reflects(w, S, h)) {  // mix(base, rc, clamp(F, 0, 0.9)), raymarch.wgsl:431-440
      const float refl = w.mats[(h.value & 0xffu) * 20u + 12u];
      const float nrd[3] = {-rd[0], -rd[1], -rd[2]};
      const float cosT = px::maxf(px::dot3f(nrd, h.N), 0.0f);
      const float F = refl + (1.0f - refl) * px_pow5(1.0f - cosT) * refl * 2.0f;
      const float fm = px::clampf(F, 0.0f, 0.9f);
      const float4 r = rc[i];
      const float rv[3] = {r.x, r.y, r.z};
      for (int a = 0; a < 3; a++) base[a] = px::mixf(base[a], rv[a], fm);
    }
    for (int a = 0; a < 3; a++) c[a] = trans[a] * base[a] + glow[a];
  }
  store_rgba(rgba, i, c);
}

// Debug views (raymarch.wgsl:399-423), every march of the pixel made here.
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_render_debug(PxWorldView w, PxShadeDesc S, const px_ray* __restrict__ rays,
                                                      uint32_t n, const uint8_t* __restrict__ dens,
                                                      float* __restrict__ rgba) {
  extern __shared__ uint32_t s_mem[];
  const MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ro[3], rd[3], maxT;
  uint32_t flags;
  px::load_ray(rays + i, ro, rd, maxT, flags);
  MarchShade ms;
  const MarchOut m = px::march<false, true>(w, T, ro, rd, maxT, &S.sky, &ms);
  const HitS h = decode_march(w, m);
  uint32_t outer = m.steps - ms.inner, inner = ms.inner;
  const int32_t dbg = px::f2i_sat(S.flags[3]);
  float c[3] = {0.0f, 0.0f, 0.0f};
  if (h.hit) {
    if (dbg == 1)
      for (int a = 0; a < 3; a++) c[a] = h.N[a] * 0.5f + 0.5f;
    if (dbg == 2) {
      const float3 mc = px::material_color(w.mats, S.vs, S.time, h.value, make_int3(h.vox[0], h.vox[1], h.vox[2]));
      c[0] = mc.x;
      c[1] = mc.y;
      c[2] = mc.z;
    }
    if (dbg == 3) c[0] = c[1] = c[2] = px::fractf(h.t / 100.0f);
    if (dbg == 4) {
      const float id = (float)(h.value & 0xffu);
      c[0] = id / 64.0f;
      c[1] = px::fractf(id * 0.37f);
      c[2] = 0.5f;
    }
    if (dbg == 5) {
      const float3 s = shade_px<0>(w, T, S, dens, h, ro, rd, nullptr, &outer, &inner);
      c[0] = s.x;
      c[1] = s.y;
      c[2] = s.z;
    }
    if (dbg == 8 || dbg == 9) {
      if (dbg == 9) (void)shade_px<2>(w, T, S, dens, h, ro, rd, nullptr, &outer, &inner);
      c[0] = (float)outer / 300.0f;
      c[1] = (float)inner / 300.0f;
      c[2] = 0.0f;
    }
    if (dbg == 6) {
      float P[3];
      for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
      const float ao = px::coarse_ao(S, dens, P, h.N) * px::fine_ao(w, h.vox, h.N, P);
      c[0] = c[1] = c[2] = ao;
    }
    if (dbg == 7) {
      float P[3], O[3];
      for (int a = 0; a < 3; a++) P[a] = ro[a] + rd[a] * h.t;
      for (int a = 0; a < 3; a++) O[a] = P[a] + h.N[a] * 0.5f;
      MarchShade ss;
      const MarchOut sm = px::march<true, true>(w, T, O, S.sky.sun, 1e9f, nullptr, &ss);
      const float k = px::maxf(px::dot3f(h.N, S.sky.sun), 0.0f);
      for (int a = 0; a < 3; a++) c[a] = (sm.face != 0xFFu ? 0.0f : ss.trans[a]) * k;
    }
  } else {
    c[0] = 0.2f;
    c[1] = 0.0f;
    c[2] = 0.3f;
  }
  store_rgba(rgba, i, c);
}

int32_t f2i_host(float x) {
  if (x != x) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}


PxShadeDesc shade_desc(const px_ctx* c) {
  PxShadeDesc S;
  std::memset(&S, 0, sizeof S);
  const float* u = c->h_uniforms;
  for (int a = 0; a < 4; a++) S.sky.sun[a] = u[4 + a];
  for (int a = 0; a < 3; a++) S.sky.sun_col[a] = u[8 + a];
  S.sky.night = u[27];
  for (int a = 0; a < 4; a++) S.ambient[a] = u[12 + a];
  for (int a = 0; a < 4; a++) S.flags[a] = u[20 + a];
  S.time = u[28];
  S.vs = u[3];
  for (int a = 0; a < 3; a++) S.fdims[a] = u[a];
  for (int a = 0; a < 3; a++) S.bd[a] = (int32_t)c->bd[a];
  S.n_lights = c->shadow.n_lights;
  for (uint32_t i = 0; i < S.n_lights; i++)
    for (int a = 0; a < 4; a++) {
      S.lp[i][a] = c->h_lights[(1 + 2 * i) * 4 + a];
      S.lc[i][a] = c->h_lights[(2 + 2 * i) * 4 + a];
    }
  return S;
}

template <typename T>
int grow(px_ctx* c, T** p, size_t* cap, size_t need, const char* what) {
  if (need <= *cap) return PX_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), need * sizeof(T));
  if (e != hipSuccess) return px_fail(c, PX_ERR_HIP, what, hipGetErrorString(e));
  *cap = need;
  return PX_OK;
}

}  // namespace

int px_refresh_density(px_ctx* c) {
  if (!c->density_dirty && c->d_density) return PX_OK;
  if (!c->d_density) PX_HIP_CHECK(c, hipMalloc(&c->d_density, (size_t)c->ncell + 4u));
  PxWorldView w = px_view(c);
  PxProfScope prof(c, PX_PROF_DENSITY);
  const uint32_t blocks = std::min<uint32_t>((c->ncell + 3u) / 4u, 65536u);
  hipLaunchKernelGGL(k_density, dim3(blocks), dim3(256), 0, c->stream, w, c->ncell, c->d_density);
  PX_HIP_CHECK(c, hipGetLastError());
  c->density_dirty = false;
  return PX_OK;
}

int px_launch_render(px_ctx* c, const px_ray* d_rays, uint32_t n, float* d_rgba) {
  if (n == 0) return PX_OK;
  int rc = px_refresh_march(c);
  if (rc) return rc;
  if ((rc = px_refresh_density(c))) return rc;
  const PxShadeDesc S = shade_desc(c);
  const PxWorldView w = px_view(c);
  bool sup_lds;
  const size_t lds = px_table_lds(c, &sup_lds);
  const dim3 grid((n + 255u) / 256u), block(256);
  if (f2i_host(S.flags[3]) > 0) {
    PxProfScope prof(c, PX_PROF_SHADE);
    if (sup_lds)
      hipLaunchKernelGGL(k_render_debug<true>, grid, block, lds, c->stream, w, S, d_rays, n, c->d_density, d_rgba);
    else
      hipLaunchKernelGGL(k_render_debug<false>, grid, block, lds, c->stream, w, S, d_rays, n, c->d_density, d_rgba);
    PX_HIP_CHECK(c, hipGetLastError());
    return PX_OK;
  }
  const uint32_t slots = c->shadow.slots;
  const uint32_t n_regions = (n + 63u) / 64u;
  const size_t qneed = (size_t)n_regions * 64u * slots;
  if (qneed > 0xFFFFFFFFull) return px_fail(c, PX_ERR_CAPACITY, "px_render_frame_device", "too many shadow slots");
  if ((rc = grow(c, &c->d_shadow_queue, &c->shadow_cap, qneed, "shadow queue"))) return rc;
  size_t wcap = c->wave_count_cap;
  if ((rc = grow(c, &c->d_wave_count, &wcap, n_regions, "wave counts"))) return rc;
  c->wave_count_cap = (uint32_t)wcap;
  if ((rc = grow(c, &c->d_rprim, &c->rprim_cap, (size_t)n, "render hits"))) return rc;
  if ((rc = grow(c, &c->d_rtg, &c->rtg_cap, (size_t)n * 2u, "render transmittance"))) return rc;
  if ((rc = grow(c, &c->d_rsh, &c->rsh_cap, qneed, "render shadow results"))) return rc;
  // reflection queue (u32 words): 64 entries per region, per-region counts
  // and their offsets (uint2), scan tiles, the total, the dense list
  const size_t ntile = ((size_t)n_regions + kScanTile - 1u) / kScanTile + 1u;
  const size_t o_cnt = (size_t)n_regions * 64u, o_off = o_cnt + 2u * n_regions, o_tiles = o_off + 2u * n_regions;
  const size_t o_tot = o_tiles + 2u * ntile, o_dense = o_tot + 4u;
  if ((rc = grow(c, &c->d_rrq, &c->rrq_cap, o_dense + (size_t)n, "reflection queue"))) return rc;
  if ((rc = grow(c, &c->d_rrc, &c->rrc_cap, (size_t)n, "reflection colours"))) return rc;
  uint32_t* rqueue = c->d_rrq;
  uint2* rcount = reinterpret_cast<uint2*>(c->d_rrq + o_cnt);
  uint2* roff = reinterpret_cast<uint2*>(c->d_rrq + o_off);
  uint2* rtiles = reinterpret_cast<uint2*>(c->d_rrq + o_tiles);
  uint32_t* rtot = c->d_rrq + o_tot;
  uint32_t* rdense = c->d_rrq + o_dense;
  {
    PxProfScope prof(c, PX_PROF_RENDER_PRIMARY);
    if (sup_lds)
      hipLaunchKernelGGL(k_render_primary<true>, grid, block, lds, c->stream, w, c->shadow, S.sky, S.flags[0],
                         d_rays, n, c->d_rprim, c->d_rtg, c->d_shadow_queue, c->d_wave_count, rqueue, rcount);
    else
      hipLaunchKernelGGL(k_render_primary<false>, grid, block, lds, c->stream, w, c->shadow, S.sky, S.flags[0],
                         d_rays, n, c->d_rprim, c->d_rtg, c->d_shadow_queue, c->d_wave_count, rqueue, rcount);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  {
    PxProfScope prof(c, PX_PROF_RENDER_SHADOW);
    const uint32_t blocks = (n_regions + 3u) / 4u;
    if (sup_lds)
      hipLaunchKernelGGL(k_render_shadow<true>, dim3(blocks), dim3(256), lds, c->stream, w, c->shadow, d_rays,
                         c->d_rprim, c->d_shadow_queue, c->d_wave_count, n, c->d_rsh);
    else
      hipLaunchKernelGGL(k_render_shadow<false>, dim3(blocks), dim3(256), lds, c->stream, w, c->shadow, d_rays,
                         c->d_rprim, c->d_shadow_queue, c->d_wave_count, n, c->d_rsh);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  {
    PxProfScope prof(c, PX_PROF_REFLECT);
    scan2_launch(c->stream, rcount, roff, rtiles, n_regions, nullptr, n_regions, nullptr, nullptr, rtot);
    hipLaunchKernelGGL(k_refl_compact, dim3((n_regions + 3u) / 4u), dim3(256), 0, c->stream, rqueue, rcount, roff,
                       n_regions, rdense);
    if (sup_lds)
      hipLaunchKernelGGL(k_render_reflect<true>, grid, block, lds, c->stream, w, S, d_rays, c->d_rprim, rdense, rtot,
                         c->d_density, c->d_rrc);
    else
      hipLaunchKernelGGL(k_render_reflect<false>, grid, block, lds, c->stream, w, S, d_rays, c->d_rprim, rdense, rtot,
                         c->d_density, c->d_rrc);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  {
    PxProfScope prof(c, PX_PROF_SHADE);
    if (sup_lds)
      hipLaunchKernelGGL(k_shade<true>, grid, block, lds, c->stream, w, S, slots, d_rays, n, c->d_rprim, c->d_rtg,
                         c->d_rsh, c->d_rrc, c->d_density, d_rgba);
    else
      hipLaunchKernelGGL(k_shade<false>, grid, block, lds, c->stream, w, S, slots, d_rays, n, c->d_rprim, c->d_rtg,
                         c->d_rsh, c->d_rrc, c->d_density, d_rgba);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}
