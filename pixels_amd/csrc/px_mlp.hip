// px_mlp.hip - learned ray-cost predictor for gfx950: a 16-64-64-1 MLP
// (ReLU) that predicts log2 of a ray's DDA step count, trained on the GPU
// with Adam, used to bin rays into cost-homogeneous wavefronts before
// traversal.  No counterpart in the reference: its traversal is the plain
// DDA, which here stays both the fallback and the judge (binning only
// reorders rays; every ray is traced by the same kernel, so outputs are
// bit-identical).
//
// Numerics: weights and activations are BF16 (round to nearest even), every
// product accumulates in FP32 on MFMA (v_mfma_f32_16x16x16_bf16), biases,
// losses, gradients and the Adam state are FP32.
//
// Layout on CDNA4: the network runs feature-major, D = W^T X^T, one wavefront
// per 64 samples (four 16-sample column groups).  The 16x16 output tile of
// v_mfma_f32_16x16x16_bf16 holds rows 4g+j (g = lane / 16, j = 0..3) of
// column lane % 16, which is exactly the B-operand layout of the next
// product, so activations chain from layer to layer in registers.  The
// weight gradients contract over samples; for those the activations are
// staged feature-major in LDS and re-read as MFMA fragments.
#include <cstring>
#include <vector>

#include "px_internal.h"
#include "px_march.h"

#pragma clang fp contract(off)

namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 16, kH = 64;
// master parameter vector (FP32), canonical order:
//   W1[f][o] (16x64, input-major), b1[64], W2[i][o] (64x64), b2[64], w3[64], b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH, kParams = kOffB3 + 1;
// per-wavefront partial record: gradients in the parameter order, then the loss sum
constexpr int kPartFloats = kParams + 1;

PX_DEV uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
PX_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

PX_DEV f32x4 mfma(s16x4 a, s16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Fragment of a 16x16 tile from row-major bf16 storage S[r][k] (leading dim
// ld): lane reads S[r0 + lane % 16][k0 + 4 (lane / 16) .. +3].  This is the A
// operand of an [r][k] matrix and the B operand of a [k][c] matrix stored as
// S[c][k].
// (S, ld and k0 are multiples of 4 elements, so each lane reads 8 aligned bytes.)
PX_DEV s16x4 frag(const uint16_t* S, int ld, int r0, int k0) {
  const uint32_t l = __lane_id();
  return *reinterpret_cast<const s16x4*>(S + (r0 + (int)(l & 15u)) * ld + k0 + 4 * (int)(l >> 4));
}

PX_DEV s16x4 pack4(float a, float b, float c, float d) {
  s16x4 v;
  v.x = (short)f2bf(a);
  v.y = (short)f2bf(b);
  v.z = (short)f2bf(c);
  v.w = (short)f2bf(d);
  return v;
}

// bf16 copies of the weights in the layouts the fragments read
struct alignas(16) MlpLds {
  uint16_t w1t[kH * kIn];  // W1T[o][f]
  uint16_t w2t[kH * kH];   // W2T[o][i]
  uint16_t w2[kH * kH];    // W2[i][o]
  uint16_t w3p[16 * kH];   // row 0 = w3, rows 1..15 zero (M padded to 16)
  float b1[kH], b2[kH], b3;
};

PX_DEV void load_weights(MlpLds& L, const uint16_t* __restrict__ wbf, const float* __restrict__ par) {
  // wbf: bf16 of the parameter vector (same order as par)
  for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) {
    const int f = i / kH, o = i % kH;
    L.w1t[o * kIn + f] = wbf[kOffW1 + i];
  }
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
    const int in = i / kH, o = i % kH;
    L.w2t[o * kH + in] = wbf[kOffW2 + i];
    L.w2[i] = wbf[kOffW2 + i];
  }
  for (int i = threadIdx.x; i < 16 * kH; i += blockDim.x) L.w3p[i] = i < kH ? wbf[kOffW3 + i] : (uint16_t)0;
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    L.b1[i] = par[kOffB1 + i];
    L.b2[i] = par[kOffB2 + i];
  }
  if (threadIdx.x == 0) L.b3 = par[kOffB3];
  __syncthreads();
}

// Forward pass of 64 samples [s0, s0 + 64) on one wavefront.  feat: bf16
// [n][16].  Samples >= n read as zeros.  Outputs the activations (bf16, in
// the output-tile layout: h1[t][c] = features 16t + 4g + j of sample 16c +
// lane % 16) and y for lanes 0..15 of each column group (y[c]).
struct Fwd {
  s16x4 x[4];       // B fragments of X^T, per column group
  s16x4 h1[4][4];   // [t][c]
  s16x4 h2[4][4];
  float y[4];       // valid in lanes 0..15
};

PX_DEV void forward(const MlpLds& L, const uint16_t* __restrict__ feat, uint32_t n, uint32_t s0, Fwd& F) {
  const uint32_t l = __lane_id();
  const int g = (int)(l >> 4);
  for (int c = 0; c < 4; c++) {
    const uint32_t s = s0 + 16u * c + (l & 15u);
    if (s < n) {
      F.x[c] = *reinterpret_cast<const s16x4*>(feat + (size_t)s * kIn + 4 * g);
    } else {
      F.x[c] = s16x4{0, 0, 0, 0};
    }
  }
  // layer 1: D1[t][c] = W1T[t] . X^T[c] + b1, ReLU
  for (int t = 0; t < 4; t++) {
    const s16x4 a = frag(L.w1t, kIn, 16 * t, 0);
    float b[4];
    for (int j = 0; j < 4; j++) b[j] = L.b1[16 * t + 4 * g + j];
    for (int c = 0; c < 4; c++) {
      f32x4 d = mfma(a, F.x[c], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      F.h1[t][c] = pack4(fmaxf(d.x + b[0], 0.0f), fmaxf(d.y + b[1], 0.0f), fmaxf(d.z + b[2], 0.0f),
                         fmaxf(d.w + b[3], 0.0f));
    }
  }
  // layer 2: D2[t][c] = sum_kt W2T[t][kt] . H1[kt][c] + b2, ReLU
  for (int t = 0; t < 4; t++) {
    s16x4 a[4];
    for (int kt = 0; kt < 4; kt++) a[kt] = frag(L.w2t, kH, 16 * t, 16 * kt);
    float b[4];
    for (int j = 0; j < 4; j++) b[j] = L.b2[16 * t + 4 * g + j];
    for (int c = 0; c < 4; c++) {
      f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      for (int kt = 0; kt < 4; kt++) d = mfma(a[kt], F.h1[kt][c], d);
      F.h2[t][c] = pack4(fmaxf(d.x + b[0], 0.0f), fmaxf(d.y + b[1], 0.0f), fmaxf(d.z + b[2], 0.0f),
                         fmaxf(d.w + b[3], 0.0f));
    }
  }
  // layer 3: y[c] = w3 . H2[c] + b3 (row 0 of a tile padded to 16 rows)
  s16x4 a3[4];
  for (int kt = 0; kt < 4; kt++) a3[kt] = frag(L.w3p, kH, 0, 16 * kt);
  for (int c = 0; c < 4; c++) {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kt = 0; kt < 4; kt++) d = mfma(a3[kt], F.h2[kt][c], d);
    F.y[c] = d.x + L.b3;  // row 0 lives in lanes 0..15, element 0
  }
}

// 4 wavefronts per block share the LDS weights; each takes 64 samples a turn.
__global__ __launch_bounds__(256) void k_mlp_predict(const uint16_t* __restrict__ wbf, const float* __restrict__ par,
                                                     const uint16_t* __restrict__ feat, uint32_t n,
                                                     float* __restrict__ pred) {
  __shared__ MlpLds L;
  load_weights(L, wbf, par);
  const uint32_t l = __lane_id();
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (uint32_t s0 = (blockIdx.x * waves + wave) * 64u; s0 < n; s0 += gridDim.x * waves * 64u) {
    Fwd F;
    forward(L, feat, n, s0, F);
    if (l < 16u)
      for (int c = 0; c < 4; c++) {
        const uint32_t s = s0 + 16u * c + l;
        if (s < n) pred[s] = F.y[c];
      }
  }
}

// Forward + backward of 64 samples per block (one wavefront) for the loss
// (1/n_batch) sum (y - target)^2; writes this block's gradient partials and
// loss sum to part[blockIdx.x * kPartFloats ...].
__global__ __launch_bounds__(64) void k_mlp_grad(const uint16_t* __restrict__ wbf, const float* __restrict__ par,
                                                 const uint16_t* __restrict__ feat, const float* __restrict__ target,
                                                 uint32_t n, float inv_batch, float* __restrict__ part) {
  __shared__ MlpLds L;
  // fragments read 8 bytes per lane: every array is 16-byte aligned
  __shared__ alignas(16) uint16_t sA[kH * 64];    // feature-major [f][s]: H1, then dH1
  __shared__ alignas(16) uint16_t sB[kH * 64];    // feature-major [f][s]: H2, then dH2
  __shared__ alignas(16) uint16_t sX[kIn * 64];   // X^T feature-major [f][s]
  __shared__ alignas(16) uint16_t sD[16 * 64];    // row 0 = dy (bf16), rows 1..15 zero
  __shared__ alignas(16) uint16_t sOne[16 * 64];  // row 0 = 1.0: the B fragment of a column of ones
  load_weights(L, wbf, par);
  const uint32_t l = __lane_id();
  const int g = (int)(l >> 4), col = (int)(l & 15u);
  const uint32_t s0 = blockIdx.x * 64u;
  Fwd F;
  forward(L, feat, n, s0, F);
  float* out = part + (size_t)blockIdx.x * kPartFloats;

  // dy = 2 (y - target) / n_batch, zero for padding samples
  float loss = 0.0f;
  for (int c = 0; c < 4; c++) {
    const uint32_t s = s0 + 16u * c + (uint32_t)col;
    float dy = 0.0f;
    if (l < 16u && s < n) {
      const float e = F.y[c] - target[s];
      loss += e * e;
      dy = 2.0f * e * inv_batch;
    }
    // broadcast dy of sample 16c + col to every lane of that column
    const float dyb = __shfl(dy, col);
    if (l < 16u) sD[16 * c + col] = f2bf(dy);
    F.y[c] = dyb;  // reuse: dy per column, every lane
  }
  for (int i = (int)l; i < 15 * 64; i += 64) sD[64 + i] = 0;
  for (int i = (int)l; i < 16 * 64; i += 64) sOne[i] = i < 64 ? (uint16_t)0x3F80u : (uint16_t)0;  // 1.0 bf16
  // stage X^T, H1, H2 feature-major
  for (int c = 0; c < 4; c++) {
    const int s = 16 * c + col;
    for (int j = 0; j < 4; j++) sX[(4 * g + j) * 64 + s] = (uint16_t)(j == 0 ? F.x[c].x : j == 1 ? F.x[c].y : j == 2 ? F.x[c].z : F.x[c].w);
    for (int t = 0; t < 4; t++) {
      const s16x4 a = F.h1[t][c], b = F.h2[t][c];
      sA[(16 * t + 4 * g + 0) * 64 + s] = (uint16_t)a.x;
      sA[(16 * t + 4 * g + 1) * 64 + s] = (uint16_t)a.y;
      sA[(16 * t + 4 * g + 2) * 64 + s] = (uint16_t)a.z;
      sA[(16 * t + 4 * g + 3) * 64 + s] = (uint16_t)a.w;
      sB[(16 * t + 4 * g + 0) * 64 + s] = (uint16_t)b.x;
      sB[(16 * t + 4 * g + 1) * 64 + s] = (uint16_t)b.y;
      sB[(16 * t + 4 * g + 2) * 64 + s] = (uint16_t)b.z;
      sB[(16 * t + 4 * g + 3) * 64 + s] = (uint16_t)b.w;
    }
  }
  __syncthreads();

  // dW3[k] = sum_s dy[s] H2[k][s]  (M = 1 padded, K = samples, N = k)
  for (int t = 0; t < 4; t++) {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 4; kc++) d = mfma(frag(sD, 64, 0, 16 * kc), frag(sB, 64, 16 * t, 16 * kc), d);
    if (l < 16u) out[kOffW3 + 16 * t + col] = d.x;
  }
  // db3 = sum_s dy[s]
  {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 4; kc++) d = mfma(frag(sD, 64, 0, 16 * kc), frag(sOne, 64, 0, 16 * kc), d);
    if (l == 0u) out[kOffB3] = d.x;
  }
  // dH2 = w3[k] dy[s] where H2 > 0 (output-tile layout), then feature-major in sB
  s16x4 dh2[4][4];
  for (int t = 0; t < 4; t++) {
    float w[4];
    for (int j = 0; j < 4; j++) w[j] = bf2f(L.w3p[16 * t + 4 * g + j]);
    for (int c = 0; c < 4; c++) {
      const s16x4 h = F.h2[t][c];
      const float dy = F.y[c];
      dh2[t][c] = pack4(h.x > 0 ? w[0] * dy : 0.0f, h.y > 0 ? w[1] * dy : 0.0f, h.z > 0 ? w[2] * dy : 0.0f,
                        h.w > 0 ? w[3] * dy : 0.0f);
    }
  }
  __syncthreads();
  for (int c = 0; c < 4; c++) {
    const int s = 16 * c + col;
    for (int t = 0; t < 4; t++) {
      const s16x4 a = dh2[t][c];
      sB[(16 * t + 4 * g + 0) * 64 + s] = (uint16_t)a.x;
      sB[(16 * t + 4 * g + 1) * 64 + s] = (uint16_t)a.y;
      sB[(16 * t + 4 * g + 2) * 64 + s] = (uint16_t)a.z;
      sB[(16 * t + 4 * g + 3) * 64 + s] = (uint16_t)a.w;
    }
  }
  __syncthreads();
  // dW2[i][o] = sum_s H1[i][s] dH2[o][s]: tile (M = o, K = s, N = i) gives dW2T[o][i]
  for (int to = 0; to < 4; to++)
    for (int ti = 0; ti < 4; ti++) {
      f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      for (int kc = 0; kc < 4; kc++) d = mfma(frag(sB, 64, 16 * to, 16 * kc), frag(sA, 64, 16 * ti, 16 * kc), d);
      // element (o = 16 to + 4g + j, i = 16 ti + col) -> W2[i][o]
      for (int j = 0; j < 4; j++) out[kOffW2 + (16 * ti + col) * kH + 16 * to + 4 * g + j] = d[j];
    }
  // db2[o] = sum_s dH2[o][s]
  for (int to = 0; to < 4; to++) {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 4; kc++) d = mfma(frag(sB, 64, 16 * to, 16 * kc), frag(sOne, 64, 0, 16 * kc), d);
    if (col == 0)
      for (int j = 0; j < 4; j++) out[kOffB2 + 16 * to + 4 * g + j] = d[j];
  }
  // dH1 = W2 . dH2 where H1 > 0 (M = i, K = o, N = samples; B = dh2 in registers)
  s16x4 dh1[4][4];
  for (int ti = 0; ti < 4; ti++) {
    s16x4 a[4];
    for (int ko = 0; ko < 4; ko++) a[ko] = frag(L.w2, kH, 16 * ti, 16 * ko);
    for (int c = 0; c < 4; c++) {
      f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      for (int ko = 0; ko < 4; ko++) d = mfma(a[ko], dh2[ko][c], d);
      const s16x4 h = F.h1[ti][c];
      dh1[ti][c] = pack4(h.x > 0 ? d.x : 0.0f, h.y > 0 ? d.y : 0.0f, h.z > 0 ? d.z : 0.0f, h.w > 0 ? d.w : 0.0f);
    }
  }
  __syncthreads();
  for (int c = 0; c < 4; c++) {
    const int s = 16 * c + col;
    for (int t = 0; t < 4; t++) {
      const s16x4 a = dh1[t][c];
      sA[(16 * t + 4 * g + 0) * 64 + s] = (uint16_t)a.x;
      sA[(16 * t + 4 * g + 1) * 64 + s] = (uint16_t)a.y;
      sA[(16 * t + 4 * g + 2) * 64 + s] = (uint16_t)a.z;
      sA[(16 * t + 4 * g + 3) * 64 + s] = (uint16_t)a.w;
    }
  }
  __syncthreads();
  // dW1[f][o] = sum_s X[s][f] dH1[o][s]: tile (M = o, K = s, N = f)
  for (int to = 0; to < 4; to++) {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 4; kc++) d = mfma(frag(sA, 64, 16 * to, 16 * kc), frag(sX, 64, 0, 16 * kc), d);
    for (int j = 0; j < 4; j++) out[kOffW1 + col * kH + 16 * to + 4 * g + j] = d[j];
  }
  // db1[o] = sum_s dH1[o][s]
  for (int to = 0; to < 4; to++) {
    f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < 4; kc++) d = mfma(frag(sA, 64, 16 * to, 16 * kc), frag(sOne, 64, 0, 16 * kc), d);
    if (col == 0)
      for (int j = 0; j < 4; j++) out[kOffB1 + 16 * to + 4 * g + j] = d[j];
  }
  // loss sum of the block
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  if (l == 0u) out[kParams] = loss;
}

// Sums the partials of nb records in a fixed order (deterministic), one
// parameter (or the loss) per thread; blockIdx.y takes records
// [y * per, (y + 1) * per) and writes record y of `out`.
__global__ void k_mlp_reduce(const float* __restrict__ part, uint32_t nb, uint32_t per, float* __restrict__ out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kPartFloats) return;
  const uint32_t b0 = blockIdx.y * per, b1 = min(b0 + per, nb);
  float s = 0.0f;
  for (uint32_t b = b0; b < b1; b++) s += part[(size_t)b * kPartFloats + p];
  out[(size_t)blockIdx.y * kPartFloats + p] = s;
}

// Adam (Kingma & Ba 2015, bias-corrected), then the bf16 copy of every weight.
__global__ void k_mlp_adam(float* __restrict__ par, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, uint16_t* __restrict__ wbf, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kParams) return;
  const float gr = grad[p];
  const float mm = b1 * m[p] + (1.0f - b1) * gr;
  const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / c1, vh = vv / c2;
  const float w = par[p] - lr * mh / (sqrtf(vh) + eps);
  par[p] = w;
  wbf[p] = f2bf(w);
}

__global__ void k_mlp_cast(const float* __restrict__ par, uint16_t* __restrict__ wbf) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < (uint32_t)kParams) wbf[p] = f2bf(par[p]);
}

// Features of a ray (16, in [-1, 1]), bf16: origin / dims, unit direction,
// super-cell occupancy at 8 points evenly spread over the ray's clipped
// extent (capped at `horizon` voxels), log2 extent, and the kind.
template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_mlp_features(PxWorldView w, const px_ray* __restrict__ rays, uint32_t n,
                                                      uint32_t kind, float horizon, uint16_t* __restrict__ feat) {
  extern __shared__ uint32_t s_mem[];
  const px::MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* rp = reinterpret_cast<const float4*>(rays + i);
  const float4 r0 = rp[0], r1 = rp[1];
  const float ro[3] = {r0.x, r0.y, r0.z};
  float rd[3] = {r0.w, r1.x, r1.y};
  const float maxT = r1.z;
  float f[16];
  const float len = sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
  float tmn = 0.0f, tmx = maxT;
  for (int a = 0; a < 3; a++) {
    f[a] = fminf(fmaxf(ro[a] / w.fdims[a] * 2.0f - 1.0f, -1.0f), 1.0f);
    f[3 + a] = len > 0.0f ? rd[a] / len : 0.0f;
    const float d = fabsf(rd[a]) < 1e-6f ? (rd[a] < 0.0f ? -1e-6f : 1e-6f) : rd[a];
    const float inv = 1.0f / d;
    const float t0 = (0.0f - ro[a]) * inv, t1 = (w.fdims[a] - ro[a]) * inv;
    tmn = fmaxf(tmn, fminf(t0, t1));
    tmx = fminf(tmx, fmaxf(t0, t1));
  }
  const float t1 = fminf(tmx, tmn + horizon);
  const float ext = fmaxf(t1 - tmn, 0.0f);
  for (int k = 0; k < 8; k++) {
    float occ = -1.0f;
    if (ext > 0.0f) {
      const float t = tmn + ext * ((float)k + 0.5f) * 0.125f;
      int32_t sc[3];
      bool in = true;
      for (int a = 0; a < 3; a++) {
        sc[a] = (int32_t)floorf((ro[a] + rd[a] * t) / 32.0f);
        in = in && sc[a] >= 0 && sc[a] < w.sd[a];
      }
      if (in && px::bit_of(T.sup, (uint32_t)(sc[0] + w.sd[0] * (sc[1] + w.sd[1] * sc[2])))) occ = 1.0f;
    }
    f[6 + k] = occ;
  }
  f[14] = fminf(log2f(1.0f + ext) / log2f(1.0f + horizon), 1.0f) * 2.0f - 1.0f;
  f[15] = kind == PX_RAY_PRIMARY ? -1.0f : (kind == PX_RAY_SHADOW ? 1.0f : 0.0f);
  uint4 o0, o1;
  o0.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
  o0.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
  o0.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
  o0.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
  o1.x = (uint32_t)f2bf(f[8]) | ((uint32_t)f2bf(f[9]) << 16);
  o1.y = (uint32_t)f2bf(f[10]) | ((uint32_t)f2bf(f[11]) << 16);
  o1.z = (uint32_t)f2bf(f[12]) | ((uint32_t)f2bf(f[13]) << 16);
  o1.w = (uint32_t)f2bf(f[14]) | ((uint32_t)f2bf(f[15]) << 16);
  uint4* op = reinterpret_cast<uint4*>(feat + (size_t)i * kIn);
  op[0] = o0;
  op[1] = o1;
}

// Binning: within each window of `win` rays, a stable counting sort of the
// rays by cost bucket (predicted log2 steps quantised into `nb` buckets over
// [lo, hi)).  One block of 256 threads per window; perm[k] = source index.
PX_DEV uint32_t bucket_of(float p, float lo, float scale, uint32_t nb) {
  int b = (int)floorf((p - lo) * scale);
  return (uint32_t)min(max(b, 0), (int)nb - 1);
}

__global__ __launch_bounds__(256) void k_bin(const float* __restrict__ pred, uint32_t n, uint32_t win, uint32_t nb,
                                             float lo, float hi, uint32_t* __restrict__ perm) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t sb[256];
  const uint32_t w0 = blockIdx.x * win;
  if (w0 >= n) return;
  const uint32_t w1 = min(w0 + win, n);
  const uint32_t tid = threadIdx.x;
  const float scale = (float)nb / (hi - lo);
  run[tid] = 0u;
  __syncthreads();
  for (uint32_t i = w0 + tid; i < w1; i += 256u) atomicAdd(&run[bucket_of(pred[i], lo, scale, nb)], 1u);
  __syncthreads();
  if (tid == 0u) {
    uint32_t acc = 0u;
    for (uint32_t b = 0; b < nb; b++) {
      const uint32_t c = run[b];
      run[b] = acc;
      acc += c;
    }
  }
  __syncthreads();
  // chunks of 256 in index order; inside a chunk, rank = earlier rays of the
  // same bucket (stable)
  for (uint32_t c0 = w0; c0 < w1; c0 += 256u) {
    const uint32_t i = c0 + tid;
    const uint32_t b = i < w1 ? bucket_of(pred[i], lo, scale, nb) : 0xFFFFu;
    sb[tid] = b;
    __syncthreads();
    uint32_t rank = 0u;
    for (uint32_t k = 0; k < tid; k++) rank += sb[k] == b ? 1u : 0u;
    if (i < w1) perm[w0 + run[b] + rank] = i;
    __syncthreads();
    if (i < w1) atomicAdd(&run[b], 1u);
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
struct px_mlp {
  px_ctx* c = nullptr;
  float* d_par = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_grad = nullptr;
  uint16_t* d_wbf = nullptr;
  float* d_part = nullptr;
  uint32_t part_cap = 0;  // blocks
  uint32_t step = 0;
};

static int mlp_fail(px_mlp* m, int code, const char* what, const char* detail) {
  return px_fail(m ? m->c : nullptr, code, what, detail);
}

#define MLP_HIP(m, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return mlp_fail((m), PX_ERR_HIP, #expr, hipGetErrorString(e_)); \
  } while (0)

extern "C" {

uint32_t px_mlp_param_count(void) { return (uint32_t)kParams; }

int px_mlp_create(px_ctx* c, uint64_t seed, px_mlp** out) {
  if (!c || !out) return PX_ERR_ARG;
  *out = nullptr;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  px_mlp* m = new (std::nothrow) px_mlp();
  if (!m) return px_fail(c, PX_ERR_ARG, "px_mlp_create", "out of host memory");
  m->c = c;
  auto bad = [&](const char* what) {
    (void)hipFree(m->d_par);
    (void)hipFree(m->d_m);
    (void)hipFree(m->d_v);
    (void)hipFree(m->d_grad);
    (void)hipFree(m->d_wbf);
    delete m;
    return px_fail(c, PX_ERR_HIP, "px_mlp_create", what);
  };
  if (hipMalloc(&m->d_par, kParams * 4) != hipSuccess || hipMalloc(&m->d_m, kParams * 4) != hipSuccess ||
      hipMalloc(&m->d_v, kParams * 4) != hipSuccess || hipMalloc(&m->d_grad, kPartFloats * 4) != hipSuccess ||
      hipMalloc(&m->d_wbf, kParams * 2) != hipSuccess)
    return bad("hipMalloc");
  // He-uniform initialisation from a splitmix64 stream, biases zero
  std::vector<float> p(kParams, 0.0f);
  uint64_t s = seed ? seed : 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int count, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; i++) p[off + i] = (float)((next() * 2.0 - 1.0) * a);
  };
  fill(kOffW1, kIn * kH, kIn);
  fill(kOffW2, kH * kH, kH);
  fill(kOffW3, kH, kH);
  std::vector<float> z(kParams, 0.0f);
  if (hipMemcpy(m->d_par, p.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_m, z.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_v, z.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess)
    return bad("upload");
  hipLaunchKernelGGL(k_mlp_cast, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, m->d_par, m->d_wbf);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bad("cast");
  *out = m;
  return PX_OK;
}

int px_mlp_destroy(px_mlp* m) {
  if (!m) return PX_ERR_ARG;
  (void)hipSetDevice(m->c->device);
  (void)hipStreamSynchronize(m->c->stream);
  (void)hipFree(m->d_par);
  (void)hipFree(m->d_m);
  (void)hipFree(m->d_v);
  (void)hipFree(m->d_grad);
  (void)hipFree(m->d_wbf);
  (void)hipFree(m->d_part);
  delete m;
  return PX_OK;
}

int px_mlp_get_params(px_mlp* m, float* h_params) {
  if (!m || !h_params) return PX_ERR_ARG;
  MLP_HIP(m, hipStreamSynchronize(m->c->stream));
  MLP_HIP(m, hipMemcpy(h_params, m->d_par, kParams * 4, hipMemcpyDeviceToHost));
  return PX_OK;
}

int px_mlp_set_params(px_mlp* m, const float* h_params) {
  if (!m || !h_params) return PX_ERR_ARG;
  MLP_HIP(m, hipStreamSynchronize(m->c->stream));
  MLP_HIP(m, hipMemcpy(m->d_par, h_params, kParams * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_cast, dim3((kParams + 255) / 256), dim3(256), 0, m->c->stream, m->d_par, m->d_wbf);
  MLP_HIP(m, hipGetLastError());
  m->step = 0;
  std::vector<float> z(kParams, 0.0f);
  MLP_HIP(m, hipMemcpy(m->d_m, z.data(), kParams * 4, hipMemcpyHostToDevice));
  MLP_HIP(m, hipMemcpy(m->d_v, z.data(), kParams * 4, hipMemcpyHostToDevice));
  return PX_OK;
}

int px_mlp_features(px_ctx* c, uint32_t kind, const px_ray* d_rays, uint32_t n, uint16_t* d_feat) {
  if (!c || (n && (!d_rays || !d_feat))) return PX_ERR_ARG;
  if (kind > 3u) return px_fail(c, PX_ERR_ARG, "px_mlp_features", "unknown ray kind");
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  int rc = px_refresh_masks(c);
  if (rc) return rc;
  PxWorldView w = px_view(c);
  const size_t full = (8u + (c->nsuper + 31u) / 32u) * 4u;
  const float horizon = (float)std::max(c->dims[0], std::max(c->dims[1], c->dims[2]));
  PxProfScope prof(c, PX_PROF_MLP);
  if (full <= 64u * 1024u)
    hipLaunchKernelGGL(k_mlp_features<true>, dim3((n + 255u) / 256u), dim3(256), full, c->stream, w, d_rays, n, kind,
                       horizon, d_feat);
  else
    hipLaunchKernelGGL(k_mlp_features<false>, dim3((n + 255u) / 256u), dim3(256), 32, c->stream, w, d_rays, n, kind,
                       horizon, d_feat);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_mlp_predict(px_mlp* m, const uint16_t* d_feat, uint32_t n, float* d_pred) {
  if (!m || (n && (!d_feat || !d_pred))) return PX_ERR_ARG;
  if (n == 0) return PX_OK;
  px_ctx* c = m->c;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  uint32_t blocks = (n + 255u) / 256u;
  const uint32_t cap = c->n_cu * 2u;
  if (blocks > cap) blocks = cap;
  PxProfScope prof(c, PX_PROF_MLP);
  hipLaunchKernelGGL(k_mlp_predict, dim3(blocks), dim3(256), 0, c->stream, m->d_wbf, m->d_par, d_feat, n, d_pred);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

static int mlp_grad(px_mlp* m, const uint16_t* d_feat, const float* d_target, uint32_t n) {
  px_ctx* c = m->c;
  const uint32_t nb = (n + 63u) / 64u;
  if (nb > m->part_cap) {
    (void)hipFree(m->d_part);
    m->d_part = nullptr;
    m->part_cap = 0;
    // per-block records, then the first-pass sums (one per 32 records)
    MLP_HIP(m, hipMalloc(&m->d_part, ((size_t)nb + (nb + 31u) / 32u) * kPartFloats * 4));
    m->part_cap = nb;
  }
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(64), 0, c->stream, m->d_wbf, m->d_par, d_feat, d_target, n,
                     1.0f / (float)n, m->d_part);
  MLP_HIP(m, hipGetLastError());
  // two fixed-order passes: nb records -> splits records -> 1
  const uint32_t per = 32u, splits = (nb + per - 1u) / per;
  float* second = m->d_part + (size_t)nb * kPartFloats;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kPartFloats + 255) / 256, splits), dim3(256), 0, c->stream, m->d_part, nb,
                     per, second);
  MLP_HIP(m, hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kPartFloats + 255) / 256, 1), dim3(256), 0, c->stream, second, splits,
                     splits, m->d_grad);
  MLP_HIP(m, hipGetLastError());
  return PX_OK;
}

int px_mlp_grad(px_mlp* m, const uint16_t* d_feat, const float* d_target, uint32_t n, float* h_grad, float* h_loss) {
  if (!m || !n || !d_feat || !d_target || !h_grad) return PX_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->c->device));
  int rc = mlp_grad(m, d_feat, d_target, n);
  if (rc) return rc;
  std::vector<float> g(kPartFloats);
  MLP_HIP(m, hipMemcpyAsync(g.data(), m->d_grad, kPartFloats * 4, hipMemcpyDeviceToHost, m->c->stream));
  MLP_HIP(m, hipStreamSynchronize(m->c->stream));
  memcpy(h_grad, g.data(), kParams * 4);
  if (h_loss) *h_loss = g[kParams] / (float)n;
  return PX_OK;
}

int px_mlp_train_step(px_mlp* m, const uint16_t* d_feat, const float* d_target, uint32_t n, float lr, float* h_loss) {
  if (!m || !n || !d_feat || !d_target) return PX_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->c->device));
  px_ctx* c = m->c;
  PxProfScope prof(c, PX_PROF_MLP);
  int rc = mlp_grad(m, d_feat, d_target, n);
  if (rc) return rc;
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - (float)std::pow((double)b1, (double)m->step);
  const float c2 = 1.0f - (float)std::pow((double)b2, (double)m->step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, m->d_par, m->d_grad, m->d_m,
                     m->d_v, m->d_wbf, lr, b1, b2, 1e-8f, c1, c2);
  MLP_HIP(m, hipGetLastError());
  if (h_loss) {
    float l = 0.0f;
    MLP_HIP(m, hipMemcpyAsync(&l, m->d_grad + kParams, 4, hipMemcpyDeviceToHost, c->stream));
    MLP_HIP(m, hipStreamSynchronize(c->stream));
    *h_loss = l / (float)n;
  }
  return PX_OK;
}

int px_bin_rays(px_ctx* c, const float* d_pred, uint32_t n, uint32_t window, uint32_t buckets, float lo, float hi,
                uint32_t* d_perm) {
  if (!c || (n && (!d_pred || !d_perm))) return PX_ERR_ARG;
  if (buckets == 0u || buckets > 256u) return px_fail(c, PX_ERR_ARG, "px_bin_rays", "buckets: 1..256");
  if (window == 0u) return px_fail(c, PX_ERR_ARG, "px_bin_rays", "window must be at least 1");
  if (!(hi > lo)) return px_fail(c, PX_ERR_ARG, "px_bin_rays", "need hi > lo");
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  PxProfScope prof(c, PX_PROF_MLP);
  hipLaunchKernelGGL(k_bin, dim3((n + window - 1u) / window), dim3(256), 0, c->stream, d_pred, n, window, buckets,
                     lo, hi, d_perm);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

}  // extern "C"
