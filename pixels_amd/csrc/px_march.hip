// px_march.hip - persistent wavefront traversal of march() for gfx950.
//
// The arithmetic is that of march() (src/render/shaders/raymarch.wgsl:44-162)
// and of the frame's shadow rays (shade(), :323-356), in the same FP32
// operation order as px_trace.hip's nested-loop kernel, so the outputs are
// bit-identical to it.  What changes is how a wavefront spends its time:
//
//  * The three nested loops (super cell -> brick -> voxel) become one loop
//    whose every turn is one iteration of one of the three original loops
//    (a per-lane `level`); leaving a level continues the parent loop's DDA
//    step exactly where the nested code would.
//  * Lanes whose ray has ended take the next ray from a chunk the wavefront
//    claimed from a global cursor (one atomic per chunk), so a wavefront is
//    not held by its longest ray.
//  * Super-cell occupancy (1 bit per 32^3 region) and a 256-bit glass table
//    sit in LDS; entering a non-empty super cell loads its 64-bit brick mask
//    (bit = brick non-empty) once, so empty bricks cost a register bit test
//    and only non-empty bricks read the grid.
#include "px_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kLvlSuper = 0, kLvlBrick = 1, kLvlVoxel = 2;

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

PX_DEV void store_hit4(px_hit* out, uint32_t key, float t, uint32_t steps, uint32_t face, uint32_t value) {
  uint4 v;
  v.x = key;
  v.y = __float_as_uint(t);
  v.z = steps;
  v.w = (face & 0xffu) | ((value & 0xffu) << 8) | (((value >> 8) & 0xffu) << 16);
  *reinterpret_cast<uint4*>(out) = v;
}

struct TuneParams {
  uint32_t chunk;       // rays a wavefront claims per atomic
  uint32_t refill_min;  // refill when at least this many lanes are idle (or all)
};

// SRC 0: rays[i] -> out[i]; SRC 1: queue entries carry the output slot in flags.
template <int SRC, bool SHADOW>
__global__ __launch_bounds__(256) void k_march_persistent(PxWorldView w, const px_ray* __restrict__ rays, uint32_t n_arg,
                                                          const uint32_t* __restrict__ n_ptr, px_hit* __restrict__ out,
                                                          uint32_t* __restrict__ cursor, TuneParams tp) {
  extern __shared__ uint32_t s_mem[];
  uint32_t* s_glass = s_mem;      // 8 words
  uint32_t* s_sup = s_mem + 8;    // (nsuper + 31) / 32 words
  const uint32_t nsup = (uint32_t)(w.sd[0] * w.sd[1] * w.sd[2]);
  const uint32_t sw = (nsup + 31u) / 32u;
  for (uint32_t i = threadIdx.x; i < sw; i += blockDim.x) s_sup[i] = w.super_mask[i];
  if (threadIdx.x < 8u) {
    uint32_t m = 0u;
    for (uint32_t b = 0; b < 32u; b++)
      if (w.mats[(threadIdx.x * 32u + b) * 20u + 13u] > 0.0f) m |= 1u << b;
    s_glass[threadIdx.x] = m;
  }
  __syncthreads();
  const uint32_t n = SRC == 1 ? *n_ptr : n_arg;
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const float* __restrict__ mats = w.mats;

  // wave-uniform ray chunk [cur, end)
  uint32_t cur = 0u, end = 0u;
  bool exhausted = false;

  // per-lane ray state
  bool active = false;
  uint32_t slot = 0u, steps = 0u, iters = 0u;  // iters: super (bits 0-8), brick (9-13), voxel (14-18)
  float ro[3], rd[3], inv[3], tExit = 0.0f;
  int32_t stp[3];
  int level = kLvlSuper;
  int32_t sc[3], bc[3], vc[3];
  float tMaxS[3], tMaxB[3], tMaxV[3];
  float t = 0.0f, tb = 0.0f, tv = 0.0f;
  uint32_t normal = 6u, nb = 6u, nv = 6u, g = 0u;
  uint64_t bmask = 0ull;
  float trans[3];
  bool inGlass = false;
  uint32_t glassId = 0u;
  for (int a = 0; a < 3; a++) {
    ro[a] = rd[a] = inv[a] = tMaxS[a] = tMaxB[a] = tMaxV[a] = trans[a] = 0.0f;
    stp[a] = sc[a] = bc[a] = vc[a] = 0;
  }

  for (;;) {
    // ---- refill idle lanes from the wavefront's chunk ----
    uint64_t idle = __ballot(!active);
    uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle && (nidle >= tp.refill_min || nidle == 64u)) {
      if (cur >= end) {
        uint32_t base = 0u;
        if (lane == 0u) base = atomicAdd(cursor, tp.chunk);
        base = __shfl(base, 0);
        if (base >= n) {
          exhausted = true;
        } else {
          cur = base;
          end = min(base + tp.chunk, n);
        }
      }
      if (!exhausted) {
        uint32_t rank = (uint32_t)__popcll(idle & below);
        uint32_t take = min(nidle, end - cur);
        if (!active && rank < take) {
          uint32_t ri = cur + rank;
          const float4* rp = reinterpret_cast<const float4*>(rays + ri);
          float4 r0 = rp[0], r1 = rp[1];
          float rdIn[3] = {r0.w, r1.x, r1.y};
          float maxT = r1.z;
          ro[0] = r0.x; ro[1] = r0.y; ro[2] = r0.z;
          slot = SRC == 1 ? __float_as_uint(r1.w) : ri;
          steps = 0u;
          iters = 0u;
          trans[0] = trans[1] = trans[2] = 1.0f;
          inGlass = false;
          glassId = 0u;
          float tmn[3], tmx[3];
          for (int a = 0; a < 3; a++) {
            rd[a] = fabsf(rdIn[a]) < 1e-6f ? px_sign(rdIn[a] + 1e-30f) * 1e-6f : rdIn[a];
            inv[a] = 1.0f / rd[a];
            float t0 = (0.0f - ro[a]) * inv[a];
            float t1 = (w.fdims[a] - ro[a]) * inv[a];
            tmn[a] = fminf(t0, t1);
            tmx[a] = fmaxf(t0, t1);
          }
          t = fmaxf(fmaxf(tmn[0], tmn[1]), fmaxf(tmn[2], 0.0f));
          tExit = fminf(fminf(tmx[0], tmx[1]), fminf(tmx[2], maxT));
          if (t >= tExit) {
            store_hit4(out + slot, PX_NONE, 0.0f, 0u, 0xFFu, 0u);
          } else {
            for (int a = 0; a < 3; a++) stp[a] = rd[a] > 0.0f ? 1 : (rd[a] < 0.0f ? -1 : 0);
            normal = 6u;
            if (t > 0.0f) {
              if (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) normal = step_face(0, stp[0]);
              else if (tmn[1] >= tmn[2]) normal = step_face(1, stp[1]);
              else normal = step_face(2, stp[2]);
            }
            for (int a = 0; a < 3; a++) {
              float pos = rd[a] >= 0.0f ? 1.0f : 0.0f;
              float p0 = ro[a] + rd[a] * (t + 1e-3f);
              sc[a] = px_clampi(px_f2i(floorf(p0 / 32.0f)), 0, w.sd[a] - 1);
              tMaxS[a] = (((float)sc[a] + pos) * 32.0f - ro[a]) * inv[a];
            }
            level = kLvlSuper;
            active = true;
          }
        }
        cur += take;
      }
    }
    if (__ballot(active) == 0ull) {
      if (exhausted) break;
      continue;
    }
    if (!active) continue;

    // ---- one iteration of the current level's loop ----
    bool done = false;
    bool hit = false;
    uint32_t hkey = 0u, hface = 0u, hval = 0u;
    float ht = 0.0f;
    // The three advance steps share this code: `adv` is the level whose DDA
    // steps next (cascading to the parent level when the child range is left).
    int adv = -1;
    if (level == kLvlSuper) {
      if ((iters & 0x1FFu) >= 256u) {
        done = true;
      } else {
        iters += 1u;
        steps++;
        uint32_t si = (uint32_t)(sc[0] + w.sd[0] * (sc[1] + w.sd[1] * sc[2]));
        if ((s_sup[si >> 5] >> (si & 31u)) & 1u) {
          int32_t sLo[3], sHi[3];
          for (int a = 0; a < 3; a++) {
            float pos = rd[a] >= 0.0f ? 1.0f : 0.0f;
            sLo[a] = sc[a] * 4;
            sHi[a] = min(sLo[a] + 3, w.bd[a] - 1);
            bc[a] = px_clampi(px_f2i(floorf((ro[a] + rd[a] * (t + 1e-3f)) / 8.0f)), sLo[a], sHi[a]);
            tMaxB[a] = (((float)bc[a] + pos) * 8.0f - ro[a]) * inv[a];
          }
          tb = t;
          nb = normal;
          iters &= 0x1FFu;  // brick and voxel counters restart
          bmask = w.super_bm[si];
          level = kLvlBrick;
        } else {
          inGlass = false;
          adv = kLvlSuper;
        }
      }
    } else if (level == kLvlBrick) {
      if (((iters >> 9) & 0x1Fu) >= 14u) {
        adv = kLvlSuper;
      } else {
        iters += 1u << 9;
        steps++;
        uint32_t lb = (uint32_t)((bc[0] & 3) + 4 * ((bc[1] & 3) + 4 * (bc[2] & 3)));
        g = ((bmask >> lb) & 1ull) ? w.grid[(uint32_t)(bc[0] + w.bd[0] * (bc[1] + w.bd[1] * bc[2]))] : 0u;
        if (g != 0u) {
          int32_t lo[3];
          for (int a = 0; a < 3; a++) {
            lo[a] = bc[a] * 8;
            vc[a] = px_clampi(px_f2i(floorf(ro[a] + rd[a] * (tb + 1e-3f))), lo[a], lo[a] + 7);
          }
          uint32_t gid = g & 0xffu;
          if ((g & PX_UNIFORM_BIT) != 0u && !((s_glass[gid >> 5] >> (gid & 31u)) & 1u)) {
            hit = true;
            done = true;
            ht = tb;
            hface = nb;
            hval = g & 0xffffu;
            hkey = (uint32_t)vc[0] + (uint32_t)w.dims[0] * ((uint32_t)vc[1] + (uint32_t)w.dims[1] * (uint32_t)vc[2]);
          } else {
            for (int a = 0; a < 3; a++) {
              float pos = rd[a] >= 0.0f ? 1.0f : 0.0f;
              tMaxV[a] = ((float)vc[a] + pos - ro[a]) * inv[a];
            }
            tv = tb;
            nv = nb;
            iters &= 0x3FFFu;  // voxel counter restarts
            level = kLvlVoxel;
          }
        } else {
          inGlass = false;
          adv = kLvlBrick;
        }
      }
    } else {
      if (((iters >> 14) & 0x1Fu) >= 25u) {
        adv = kLvlBrick;
      } else {
        iters += 1u << 14;
        steps++;
        uint32_t v = px_brick_voxel(w.pool, g, vc[0] & 7, vc[1] & 7, vc[2] & 7);
        if (v != 0u) {
          uint32_t id = v & 0xffu;
          bool stop = true;
          if ((s_glass[id >> 5] >> (id & 31u)) & 1u) {
            if (!inGlass || id != glassId) glass_surface(mats, trans, id, nv, rd, SHADOW);
            inGlass = true;
            glassId = id;
            stop = fmaxf(trans[0], fmaxf(trans[1], trans[2])) < 0.02f;
          }
          if (stop) {
            hit = true;
            done = true;
            ht = tv;
            hface = nv;
            hval = v;
            hkey = (uint32_t)vc[0] + (uint32_t)w.dims[0] * ((uint32_t)vc[1] + (uint32_t)w.dims[1] * (uint32_t)vc[2]);
          }
        } else {
          inGlass = false;
        }
        if (!done) adv = kLvlVoxel;
      }
    }
    // ---- DDA steps, cascading voxel -> brick -> super ----
    if (adv == kLvlVoxel) {
      if (tMaxV[0] < tMaxV[1] && tMaxV[0] < tMaxV[2]) {
        tv = tMaxV[0]; vc[0] += stp[0]; tMaxV[0] += fabsf(inv[0]); nv = step_face(0, stp[0]);
      } else if (tMaxV[1] < tMaxV[2]) {
        tv = tMaxV[1]; vc[1] += stp[1]; tMaxV[1] += fabsf(inv[1]); nv = step_face(1, stp[1]);
      } else {
        tv = tMaxV[2]; vc[2] += stp[2]; tMaxV[2] += fabsf(inv[2]); nv = step_face(2, stp[2]);
      }
      int32_t lo0 = bc[0] * 8, lo1 = bc[1] * 8, lo2 = bc[2] * 8;
      if (tv > tExit) done = true;
      else if (vc[0] < lo0 || vc[1] < lo1 || vc[2] < lo2 || vc[0] > lo0 + 7 || vc[1] > lo1 + 7 || vc[2] > lo2 + 7)
        adv = kLvlBrick;
      else
        level = kLvlVoxel;
    }
    if (adv == kLvlBrick) {
      if (tMaxB[0] < tMaxB[1] && tMaxB[0] < tMaxB[2]) {
        tb = tMaxB[0]; bc[0] += stp[0]; tMaxB[0] += fabsf(inv[0]) * 8.0f; nb = step_face(0, stp[0]);
      } else if (tMaxB[1] < tMaxB[2]) {
        tb = tMaxB[1]; bc[1] += stp[1]; tMaxB[1] += fabsf(inv[1]) * 8.0f; nb = step_face(1, stp[1]);
      } else {
        tb = tMaxB[2]; bc[2] += stp[2]; tMaxB[2] += fabsf(inv[2]) * 8.0f; nb = step_face(2, stp[2]);
      }
      int32_t sLo0 = sc[0] * 4, sLo1 = sc[1] * 4, sLo2 = sc[2] * 4;
      int32_t sHi0 = min(sLo0 + 3, w.bd[0] - 1), sHi1 = min(sLo1 + 3, w.bd[1] - 1), sHi2 = min(sLo2 + 3, w.bd[2] - 1);
      if (tb > tExit) done = true;
      else if (bc[0] < sLo0 || bc[1] < sLo1 || bc[2] < sLo2 || bc[0] > sHi0 || bc[1] > sHi1 || bc[2] > sHi2)
        adv = kLvlSuper;
      else
        level = kLvlBrick;
    }
    if (adv == kLvlSuper) {
      if (tMaxS[0] < tMaxS[1] && tMaxS[0] < tMaxS[2]) {
        t = tMaxS[0]; sc[0] += stp[0]; tMaxS[0] += fabsf(inv[0]) * 32.0f; normal = step_face(0, stp[0]);
      } else if (tMaxS[1] < tMaxS[2]) {
        t = tMaxS[1]; sc[1] += stp[1]; tMaxS[1] += fabsf(inv[1]) * 32.0f; normal = step_face(1, stp[1]);
      } else {
        t = tMaxS[2]; sc[2] += stp[2]; tMaxS[2] += fabsf(inv[2]) * 32.0f; normal = step_face(2, stp[2]);
      }
      if (t > tExit || sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || sc[0] >= w.sd[0] || sc[1] >= w.sd[1] || sc[2] >= w.sd[2])
        done = true;
      else
        level = kLvlSuper;
    }
    if (done) {
      if (hit) store_hit4(out + slot, hkey, px_canon0(ht), steps, hface, hval);
      else store_hit4(out + slot, PX_NONE, 0.0f, steps, 0xFFu, 0u);
      active = false;
    }
  }
}

// Shadow rays of shade() (raymarch.wgsl:323-356) for every primary hit, in
// slot order, compacted into a queue (one atomic per wavefront and slot).
__global__ __launch_bounds__(256) void k_shadow_gen(PxWorldView w, PxShadowDesc sd, const px_ray* __restrict__ rays,
                                                    uint32_t n, const px_hit* __restrict__ prim,
                                                    px_hit* __restrict__ sh_out, px_ray* __restrict__ queue,
                                                    uint32_t* __restrict__ counter) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < n;
  float ro[3] = {0, 0, 0}, rd[3] = {0, 0, 1};
  uint32_t face = 0xFFu, mat = 0u;
  float ht = 0.0f;
  if (live) {
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    float4 r0 = rp[0], r1 = rp[1];
    ro[0] = r0.x; ro[1] = r0.y; ro[2] = r0.z;
    rd[0] = r0.w; rd[1] = r1.x; rd[2] = r1.y;
    uint4 h = *reinterpret_cast<const uint4*>(prim + i);
    ht = __uint_as_float(h.y);
    face = h.w & 0xffu;
    mat = (h.w >> 8) & 0xffu;
    uint4 nc = make_uint4(PX_NONE, 0u, 0u, 0xFEu);
    for (uint32_t s = 0; s < sd.slots; s++) *reinterpret_cast<uint4*>(sh_out + (size_t)i * sd.slots + s) = nc;
  }
  bool shade = live && face != 0xFFu && !(w.mats[mat * 20u + 3u] > 0.0f);
  float N[3] = {0.0f, 0.0f, 0.0f};
  if (face < 6u) N[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  float P[3], O[3];
  for (int a = 0; a < 3; a++) {
    P[a] = ro[a] + rd[a] * ht;
    O[a] = P[a] + N[a] * 0.5f;
  }
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
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
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

// Per-super-cell 64-bit brick occupancy (bit = lx + 4 (ly + 4 lz)).
__global__ void k_super_bm(const uint32_t* __restrict__ grid, int32_t bx, int32_t by, int32_t bz, int32_t sx,
                           int32_t sy, int32_t sz, uint64_t* __restrict__ out) {
  uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t ns = (uint32_t)(sx * sy * sz);
  if (si >= ns) return;
  int32_t s0 = (int32_t)(si % (uint32_t)sx), s1 = (int32_t)((si / (uint32_t)sx) % (uint32_t)sy);
  int32_t s2 = (int32_t)(si / ((uint32_t)sx * (uint32_t)sy));
  uint64_t m = 0ull;
  for (int32_t lz = 0; lz < 4; lz++)
    for (int32_t ly = 0; ly < 4; ly++)
      for (int32_t lx = 0; lx < 4; lx++) {
        int32_t x = s0 * 4 + lx, y = s1 * 4 + ly, z = s2 * 4 + lz;
        if (x >= bx || y >= by || z >= bz) continue;
        if (grid[(uint32_t)(x + bx * (y + by * z))] != 0u) m |= 1ull << (lx + 4 * (ly + 4 * lz));
      }
  out[si] = m;
}

__global__ void k_zero_counter(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0u;
}

}  // namespace

int px_launch_super_bm(px_ctx* c) {
  uint32_t ns = c->nsuper;
  hipLaunchKernelGGL(k_super_bm, dim3((ns + 255u) / 256u), dim3(256), 0, c->stream, c->d_grid, (int32_t)c->bd[0],
                     (int32_t)c->bd[1], (int32_t)c->bd[2], (int32_t)c->sd[0], (int32_t)c->sd[1], (int32_t)c->sd[2],
                     c->d_super_bm);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

static size_t march_lds(const px_ctx* c) { return (8u + (c->nsuper + 31u) / 32u) * 4u; }

bool px_march_fits(const px_ctx* c) { return march_lds(c) <= 64u * 1024u; }

// rays[0..n) -> out; kind PRIMARY or SHADOW
int px_launch_march(px_ctx* c, bool shadow, const px_ray* d_rays, uint32_t n, px_hit* d_out) {
  if (n == 0) return PX_OK;
  PxWorldView w = px_view(c);
  TuneParams tp{c->tune.chunk, c->tune.refill_min};
  size_t lds = march_lds(c);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, c->d_counters + 1);
  dim3 grid(c->tune.blocks), block(c->tune.block);
  PxProfScope prof(c, PX_PROF_TRACE);
  if (shadow)
    hipLaunchKernelGGL((k_march_persistent<0, true>), grid, block, lds, c->stream, w, d_rays, n, nullptr, d_out,
                       c->d_counters + 1, tp);
  else
    hipLaunchKernelGGL((k_march_persistent<0, false>), grid, block, lds, c->stream, w, d_rays, n, nullptr, d_out,
                       c->d_counters + 1, tp);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

// The frame: primary march, shadow-ray generation into the queue, shadow march.
int px_launch_frame_persistent(px_ctx* c, const px_ray* d_primary, uint32_t n, px_hit* d_prim, px_hit* d_shadow) {
  PxWorldView w = px_view(c);
  TuneParams tp{c->tune.chunk, c->tune.refill_min};
  size_t lds = march_lds(c);
  dim3 grid(c->tune.blocks), block(c->tune.block);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, c->d_counters + 0);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, c->d_counters + 1);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, c->d_counters + 2);
  {
    PxProfScope prof(c, PX_PROF_PRIMARY);
    hipLaunchKernelGGL((k_march_persistent<0, false>), grid, block, lds, c->stream, w, d_primary, n, nullptr, d_prim,
                       c->d_counters + 1, tp);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  {
    PxProfScope prof(c, PX_PROF_SHADOW_GEN);
    hipLaunchKernelGGL(k_shadow_gen, dim3((n + 255u) / 256u), dim3(256), 0, c->stream, w, c->shadow, d_primary, n,
                       d_prim, d_shadow, c->d_shadow_rays, c->d_counters + 0);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  {
    PxProfScope prof(c, PX_PROF_SHADOW);
    hipLaunchKernelGGL((k_march_persistent<1, true>), grid, block, lds, c->stream, w, c->d_shadow_rays, 0u,
                       c->d_counters + 0, d_shadow, c->d_counters + 2, tp);
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}
