// px_particles.hip - debris particles for gfx950.
//
//   spawn()        src/render/shaders/sim.wgsl:166-170
//   pushWrite      src/render/shaders/sim.wgsl:186-192
//   frameBegin     src/render/shaders/sim.wgsl:363-369
//   pushAlive      src/render/shaders/sim.wgsl:1082-1085
//   particleSim    src/render/shaders/sim.wgsl:1088-1173
//
// The reference integrates every ring slot in one dispatch and appends the
// alive slots and the deposit writes to global lists with one atomicAdd per
// particle, so the list order is the GPU's schedule and, past MAX_WRITES, the
// winning requests are whichever came first.  Here a step is three launches
// with no global atomics, over tiles of kTile slots (one workgroup each):
//
//   k_particle_sim    one lane per slot: the physics, the new record, a
//                     status byte (alive, write request) and per-tile counts
//                     of both (wave ballots);
//   k_particle_scan   one workgroup: exclusive offsets of the tile counts,
//                     added to the list lengths already in args / ctrl;
//   k_particle_emit   one lane per slot: alive and write entries at their
//                     slot-order rank; a write request past max_writes gets
//                     the failed-write life, the others the deposit life.
//
// So both lists are in slot order and the lowest slots win the writes: the
// serial schedule, reproduced bit for bit by the oracle.  The FP32 arithmetic
// keeps the WGSL's operation order (-ffp-contract=off); length() is
// sqrt((x*x + y*y) + z*z).
#include "px_internal.h"
#include "px_march.h"

#pragma clang fp contract(off)

namespace {

constexpr uint32_t kBlock = 1024;          // 16 wavefronts
constexpr uint32_t kSub = 1;                // sub-tiles of kBlock slots per workgroup
constexpr uint32_t kTile = kBlock * kSub;  // slots per workgroup
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kA_DRAWP = 36;  // layout.ts:41-45: 12 dispatch triples, then A_DRAWP
constexpr uint8_t kAlive = 1;
constexpr uint8_t kWrite = 2;

PX_DEV uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

PX_DEV bool solid_at(const PxWorldView& w, int32_t x, int32_t y, int32_t z) {
  return px::read_voxel(w, x, y, z) != 0u;
}

// Rank of this lane among the block's lanes with `flag`, and the block total.
// Block-cooperative.
PX_DEV uint32_t block_rank(bool flag, uint32_t* s_wave, uint32_t* total) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(flag);
  if (lane == 0u) s_wave[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0u, sum = 0u;
  for (uint32_t k = 0; k < kBlock / 64u; k++) {
    const uint32_t c = s_wave[k];
    before += k < wave ? c : 0u;
    sum += c;
  }
  *total = sum;
  __syncthreads();  // s_wave may be rewritten by the next call
  return before + lanes_below(m);
}

// Inclusive prefix sum over the wavefront.
PX_DEV uint32_t wave_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// particleSim (sim.wgsl:1088-1173) for slot i; returns its status byte.
PX_DEV uint8_t sim_slot(const PxWorldView& w, float4* __restrict__ P, uint32_t i, float dt, float grav,
                        float rest_speed) {
  uint8_t st = 0;
  // both halves at once: one memory round trip for live slots
  const float4 a = P[2u * i];       // pos.xyz, value
  const float4 b = P[2u * i + 1u];  // vel.xyz, life
  float life = b.w;
  if (!(life <= 0.0f)) {
    life = life - dt;
    if (life <= 0.0f) {
      P[2u * i + 1u].w = 0.0f;
    } else {
      uint32_t value = __float_as_uint(a.w);
      if ((value & PX_P_SETTLED) != 0u) {
        P[2u * i + 1u].w = life;
        st = kAlive;
      } else {
        const bool dust = (value & PX_P_DUST) != 0u;
        float pos[3] = {a.x, a.y, a.z};
        float vel[3] = {b.x, b.y, b.z};
        const float g = dust ? grav * 0.15f : grav;
        vel[1] = vel[1] - g * dt;
        const float drag = 1.0f - (dust ? 2.5f : 0.2f) * dt;
        for (int k = 0; k < 3; k++) vel[k] = vel[k] * drag;
        int32_t cell[3];
        for (int k = 0; k < 3; k++) cell[k] = px_f2i(floorf(pos[k]));
        bool write_req = false;
        if (solid_at(w, cell[0], cell[1], cell[2])) {
          // spawned inside something (or the world changed under it): pop upwards
          pos[1] = pos[1] + 1.0f;
          st = kAlive;
        } else {
          float mv[3];
          for (int k = 0; k < 3; k++) mv[k] = vel[k] * dt;
          const float dist = sqrtf(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
          if (dist > 1e-6f) {
            float dir[3], t = 0.0f;
            uint32_t v = 0u, steps = 0u;
            const int axis = px::segment_walk(w, cell, pos, mv, dist, dir, &t, &v, &steps);
            if (axis < 0) {
              for (int k = 0; k < 3; k++) pos[k] = pos[k] + mv[k];
            } else {
              const float s = fmaxf(t - 1e-3f, 0.0f);
              for (int k = 0; k < 3; k++) pos[k] = pos[k] + dir[k] * s;
              const float fr = axis == 1 ? 0.55f : 0.75f;
              for (int k = 0; k < 3; k++) vel[k] = k == axis ? -vel[k] * 0.25f : vel[k] * fr;
            }
          }
          // rest check: slow and supported from below
          int32_t c[3];
          for (int k = 0; k < 3; k++) c[k] = px_f2i(floorf(pos[k]));
          const float speed = sqrtf(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
          if (speed < rest_speed && solid_at(w, c[0], c[1] - 1, c[2])) {
            for (int k = 0; k < 3; k++) {
              pos[k] = (float)c[k] + 0.5f;
              vel[k] = 0.0f;
            }
            value |= PX_P_SETTLED;
            const bool in_world = c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < w.dims[0] && c[1] < w.dims[1] &&
                                  c[2] < w.dims[2];
            if ((value & PX_P_DEPOSIT) != 0u && in_world) {
              // life resolved by k_particle_emit once the request's rank is known
              write_req = true;
            } else {
              life = fminf(life, dust ? 0.8f : 4.0f);
            }
          }
          if (pos[1] < -200.0f) life = 0.0f;
          st = (write_req ? kWrite : 0u) | (life > 0.0f ? kAlive : 0u);
        }
        P[2u * i] = make_float4(pos[0], pos[1], pos[2], __uint_as_float(value));
        P[2u * i + 1u] = make_float4(vel[0], vel[1], vel[2], life);
      }
    }
  }
  return st;
}

__global__ __launch_bounds__(kBlock) void k_particle_sim(PxWorldView w, float4* __restrict__ P, float dt, float grav,
                                                         float rest_speed, uint8_t* __restrict__ status,
                                                         uint32_t* __restrict__ bcount) {
  __shared__ uint32_t s_a[kBlock / 64], s_w[kBlock / 64];
  uint32_t ca = 0u, cw = 0u;
  for (uint32_t s = 0; s < kSub; s++) {
    const uint32_t i = blockIdx.x * kTile + s * kBlock + threadIdx.x;
    const uint8_t st = sim_slot(w, P, i, dt, grav, rest_speed);
    status[i] = st;
    uint32_t ta, tw;
    block_rank((st & kAlive) != 0, s_a, &ta);
    block_rank((st & kWrite) != 0, s_w, &tw);
    ca += ta;
    cw += tw;
  }
  if (threadIdx.x == 0) {
    bcount[2u * blockIdx.x] = ca;
    bcount[2u * blockIdx.x + 1u] = cw;
  }
}

// Exclusive offsets of the per-tile counts (in place: bcount[2b], [2b+1]
// become the first alive slot and the first write index of tile b) and the
// new list lengths.  One workgroup, each thread owning `per` consecutive tiles.
__global__ __launch_bounds__(kScanThreads) void k_particle_scan(uint32_t* __restrict__ bcount, uint32_t nb,
                                                                uint32_t* __restrict__ ctrl,
                                                                uint32_t* __restrict__ args) {
  __shared__ uint32_t s_a[kScanThreads / 64], s_w[kScanThreads / 64];
  const uint32_t base_a = args[kA_DRAWP + 1u];
  const uint32_t base_w = ctrl[PX_C_WRITES];
  const uint32_t per = (nb + kScanThreads - 1u) / kScanThreads;
  const uint32_t b0 = threadIdx.x * per;
  uint32_t sa = 0u, sw = 0u;
  for (uint32_t k = 0; k < per; k++) {
    const uint32_t b = b0 + k;
    if (b < nb) {
      sa += bcount[2u * b];
      sw += bcount[2u * b + 1u];
    }
  }
  const uint32_t ia = wave_scan(sa), iw = wave_scan(sw);
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 63u) {
    s_a[wave] = ia;
    s_w[wave] = iw;
  }
  __syncthreads();  // also orders every thread's read of the bases before the update below
  uint32_t pa = 0u, pw = 0u, ta = 0u, tw = 0u;
  for (uint32_t k = 0; k < kScanThreads / 64u; k++) {
    pa += k < wave ? s_a[k] : 0u;
    pw += k < wave ? s_w[k] : 0u;
    ta += s_a[k];
    tw += s_w[k];
  }
  uint32_t oa = base_a + pa + ia - sa, ow = base_w + pw + iw - sw;
  for (uint32_t k = 0; k < per; k++) {
    const uint32_t b = b0 + k;
    if (b < nb) {
      const uint32_t ca = bcount[2u * b], cw = bcount[2u * b + 1u];
      bcount[2u * b] = oa;
      bcount[2u * b + 1u] = ow;
      oa += ca;
      ow += cw;
    }
  }
  if (threadIdx.x == 0) {
    args[kA_DRAWP + 1u] = base_a + ta;
    ctrl[PX_C_WRITES] = base_w + tw;  // requests, granted or not (pushWrite counts every call)
  }
}

__global__ __launch_bounds__(kBlock) void k_particle_emit(PxWorldView w, float4* __restrict__ P,
                                                          const uint8_t* __restrict__ status,
                                                          const uint32_t* __restrict__ boff,
                                                          uint32_t* __restrict__ alive, uint32_t alive_cap,
                                                          uint32_t* __restrict__ wlist, uint32_t max_writes) {
  __shared__ uint32_t s_a[kBlock / 64], s_w[kBlock / 64];
  uint32_t oa = boff[2u * blockIdx.x], ow = boff[2u * blockIdx.x + 1u];
  for (uint32_t s = 0; s < kSub; s++) {
    const uint32_t i = blockIdx.x * kTile + s * kBlock + threadIdx.x;
    const uint8_t st = status[i];
    uint32_t ta, tw;
    const uint32_t ra = block_rank((st & kAlive) != 0, s_a, &ta);
    const uint32_t rw = block_rank((st & kWrite) != 0, s_w, &tw);
    if (st & kAlive) {
      const uint32_t slot = oa + ra;
      if (slot < alive_cap) alive[slot] = i;
    }
    if (st & kWrite) {
      const uint32_t k = ow + rw;
      const float4 a = P[2u * i];
      const float life = P[2u * i + 1u].w;
      const uint32_t value = __float_as_uint(a.w);
      float nl;
      if (k < max_writes) {
        // settled at the cell centre: floor(pos) is the cell
        const int32_t cx = px_f2i(floorf(a.x)), cy = px_f2i(floorf(a.y)), cz = px_f2i(floorf(a.z));
        wlist[2u * k] = px::key_of(w, cx, cy, cz);
        wlist[2u * k + 1u] = (value & 0xffffu) | 0x80u;  // rubbleOf (sim.wgsl:65)
        nl = fminf(life, 0.5f);
      } else {
        nl = fminf(life, (value & PX_P_DUST) != 0u ? 0.8f : 4.0f);
      }
      P[2u * i + 1u].w = nl;
    }
    oa += ta;
    ow += tw;
  }
}

__global__ void k_frame_begin(uint32_t* __restrict__ ctrl, uint32_t* __restrict__ args) {
  if (threadIdx.x == 0) {
    ctrl[PX_C_SPAWNED] = 0u;
    ctrl[PX_C_DIRTY] = 0u;
    ctrl[PX_C_WRITES] = 0u;
    ctrl[PX_C_ALLOCS] = 0u;
    args[kA_DRAWP] = 36u;
    args[kA_DRAWP + 1u] = 0u;
    args[kA_DRAWP + 2u] = 0u;
    args[kA_DRAWP + 3u] = 0u;
  }
}

// spawn() for records src[0..n): every lane reads the same counters (nothing
// writes them in this launch); record j is stored when it is within the
// budget and no later accepted record maps to the same ring slot.
__global__ __launch_bounds__(kBlock) void k_spawn(float4* __restrict__ P, uint32_t mask,
                                                  const float4* __restrict__ src, uint32_t n,
                                                  const uint32_t* __restrict__ ctrl, uint32_t budget) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t spawned = ctrl[PX_C_SPAWNED], cursor = ctrl[PX_C_PCURSOR];
  const uint32_t room = spawned >= budget ? 0u : budget - spawned;
  const uint32_t accepted = min(n, room);
  if (j >= accepted || accepted - j > mask + 1u) return;
  const uint32_t slot = (cursor + j) & mask;
  P[2u * slot] = src[2u * j];
  P[2u * slot + 1u] = src[2u * j + 1u];
}

__global__ void k_spawn_commit(uint32_t* __restrict__ ctrl, uint32_t n, uint32_t budget) {
  if (threadIdx.x == 0) {
    const uint32_t spawned = ctrl[PX_C_SPAWNED];
    const uint32_t room = spawned >= budget ? 0u : budget - spawned;
    ctrl[PX_C_SPAWNED] = spawned + n;
    ctrl[PX_C_PCURSOR] = ctrl[PX_C_PCURSOR] + min(n, room);
  }
}

int need_particles(px_ctx* c, const char* what) {
  if (!c->d_particles) return px_fail(c, PX_ERR_STATE, what, "px_particles_init has not been called");
  return PX_OK;
}

int launch_spawn(px_ctx* c, const float4* d_src, uint32_t n) {
  if (n == 0) return PX_OK;
  PxProfScope prof(c, PX_PROF_PARTICLES);
  const uint32_t budget = c->lim.spawn_budget;
  hipLaunchKernelGGL(k_spawn, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->d_particles,
                     c->max_particles - 1u, d_src, n, c->d_ctrl, budget);
  hipLaunchKernelGGL(k_spawn_commit, dim3(1), dim3(64), 0, c->stream, c->d_ctrl, n, budget);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

}  // namespace

void px_particles_free(px_ctx* c) {
  void* bufs[] = {c->d_particles, c->d_alive, c->d_pstatus, c->d_pblock, c->d_spawn_stage};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  c->d_particles = nullptr;
  c->d_alive = c->d_pblock = nullptr;
  c->d_pstatus = nullptr;
  c->d_spawn_stage = nullptr;
  c->spawn_cap = 0;
  c->max_particles = 0;
}

int px_launch_frame_begin(px_ctx* c) {
  hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, c->stream, c->d_ctrl, c->d_args);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_launch_particle_step(px_ctx* c) {
  if (int rc = need_particles(c, "px_particle_step")) return rc;
  const float dt = c->h_uniforms[24], grav = c->h_uniforms[25], vs = c->h_uniforms[3];
  const float rest = 0.6f / vs;
  const uint32_t nb = c->max_particles / kTile;
  PxProfScope prof(c, PX_PROF_PARTICLES);
  PxWorldView w = px_view(c);
  {
    PxProfScope prof_sim(c, PX_PROF_PARTICLE_SIM);
    hipLaunchKernelGGL(k_particle_sim, dim3(nb), dim3(kBlock), 0, c->stream, w, c->d_particles, dt, grav, rest,
                       c->d_pstatus, c->d_pblock);
  }
  hipLaunchKernelGGL(k_particle_scan, dim3(1), dim3(kScanThreads), 0, c->stream, c->d_pblock, nb, c->d_ctrl,
                     c->d_args);
  hipLaunchKernelGGL(k_particle_emit, dim3(nb), dim3(kBlock), 0, c->stream, w, c->d_particles, c->d_pstatus,
                     c->d_pblock, c->d_alive, c->max_particles, c->d_write_list, c->lim.max_writes);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

extern "C" {

int px_particles_init(px_ctx* c, uint32_t max_particles) {
  if (!c) return PX_ERR_ARG;
  if (max_particles == 0) max_particles = 1u << 18;
  if ((max_particles & (max_particles - 1u)) != 0u || max_particles < (1u << 12) || max_particles > (1u << 21))
    return px_fail(c, PX_ERR_ARG, "px_particles_init", "max_particles must be a power of two in [2^12, 2^21]");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  px_particles_free(c);
  if (int rc = px_writes_alloc(c)) return rc;
  const size_t np = max_particles, nb = np / kTile;
  int rc = PX_OK;
#define TRY(x)                                                     \
  do {                                                             \
    hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) {                                        \
      rc = px_fail(c, PX_ERR_HIP, #x, hipGetErrorString(e_));      \
      px_particles_free(c);                                        \
      return rc;                                                   \
    }                                                              \
  } while (0)
  TRY(hipMalloc(&c->d_particles, np * sizeof(px_particle)));
  TRY(hipMalloc(&c->d_alive, np * 4));
  TRY(hipMalloc(&c->d_pstatus, np));
  TRY(hipMalloc(&c->d_pblock, nb * 8));
  TRY(hipMemsetAsync(c->d_particles, 0, np * sizeof(px_particle), c->stream));
  TRY(hipMemsetAsync(c->d_alive, 0, np * 4, c->stream));
  TRY(hipMemsetAsync(c->d_ctrl + PX_C_PCURSOR, 0, 4, c->stream));
  TRY(hipStreamSynchronize(c->stream));
#undef TRY
  c->max_particles = max_particles;
  return PX_OK;
}

int px_frame_begin(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_frame_begin(c);
}

int px_spawn_device(px_ctx* c, const px_particle* d_src, uint32_t n) {
  if (!c || (n && !d_src)) return PX_ERR_ARG;
  if (int rc = need_particles(c, "px_spawn_device")) return rc;
  if (reinterpret_cast<uintptr_t>(d_src) & 15u)
    return px_fail(c, PX_ERR_ARG, "px_spawn_device", "d_src must be 16-byte aligned");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return launch_spawn(c, reinterpret_cast<const float4*>(d_src), n);
}

int px_spawn(px_ctx* c, const px_particle* src, uint32_t n) {
  if (!c || (n && !src)) return PX_ERR_ARG;
  if (int rc = need_particles(c, "px_spawn")) return rc;
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (n > c->spawn_cap) {
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (c->d_spawn_stage) (void)hipFree(c->d_spawn_stage);
    c->d_spawn_stage = nullptr;
    c->spawn_cap = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_spawn_stage, (size_t)n * sizeof(px_particle)));
    c->spawn_cap = n;
  }
  PX_HIP_CHECK(c, hipMemcpyAsync(c->d_spawn_stage, src, (size_t)n * sizeof(px_particle), hipMemcpyHostToDevice,
                                 c->stream));
  int rc = launch_spawn(c, c->d_spawn_stage, n);
  if (rc) return rc;
  // the caller may reuse `src` on return: wait for the pageable copy
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_particle_step(px_ctx* c) {
  if (!c) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  return px_launch_particle_step(c);
}

int px_read_particles(px_ctx* c, uint32_t first, uint32_t count, px_particle* out) {
  if (!c || (count && !out)) return PX_ERR_ARG;
  if (int rc = need_particles(c, "px_read_particles")) return rc;
  if ((uint64_t)first + count > c->max_particles)
    return px_fail(c, PX_ERR_ARG, "px_read_particles", "range beyond the ring");
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (count)
    PX_HIP_CHECK(c, hipMemcpyAsync(out, c->d_particles + 2u * first, (size_t)count * sizeof(px_particle),
                                   hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return PX_OK;
}

int px_read_alive(px_ctx* c, uint32_t* out, uint32_t cap, uint32_t* n) {
  if (!c || !n || (cap && !out)) return PX_ERR_ARG;
  if (int rc = need_particles(c, "px_read_alive")) return rc;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  uint32_t cnt = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&cnt, c->d_args + kA_DRAWP + 1u, 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  *n = cnt;
  const uint32_t m = min(min(cnt, cap), c->max_particles);
  if (m) {
    PX_HIP_CHECK(c, hipMemcpyAsync(out, c->d_alive, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return PX_OK;
}

int px_read_writes(px_ctx* c, uint32_t* pairs, uint32_t cap, uint32_t* n) {
  if (!c || !n || (cap && !pairs)) return PX_ERR_ARG;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  if (int rc = px_writes_alloc(c)) return rc;
  uint32_t cnt = 0;
  PX_HIP_CHECK(c, hipMemcpyAsync(&cnt, c->d_ctrl + PX_C_WRITES, 4, hipMemcpyDeviceToHost, c->stream));
  PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  cnt = min(cnt, c->lim.max_writes);
  *n = cnt;
  const uint32_t m = min(cnt, cap);
  if (m) {
    PX_HIP_CHECK(c, hipMemcpyAsync(pairs, c->d_write_list, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return PX_OK;
}

int px_particle_buffers(px_ctx* c, void** particles, void** alive, void** writes) {
  if (!c) return PX_ERR_ARG;
  if (int rc = need_particles(c, "px_particle_buffers")) return rc;
  if (particles) *particles = c->d_particles;
  if (alive) *alive = c->d_alive;
  if (writes) *writes = c->d_write_list;
  return PX_OK;
}

}  // extern "C"
