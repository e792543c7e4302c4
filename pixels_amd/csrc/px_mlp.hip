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

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
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
// v_mfma_f32_16x16x32_bf16 (gfx950): K = 32 per product.  A and B map their
// (lane / 16, element) pairs to k the same way, so any fixed order of the 32
// k indices, used for both operands, gives the same contraction.
PX_DEV f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
PX_DEV s16x8 cat8(s16x4 a, s16x4 b) { return s16x8{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}; }

// bf16(max(x, 0)) of four FP32 values: v_cvt_pk_bf16_f32 rounds to nearest
// even (as f2bf for finite values), then a signed 16-bit max with +0 maps
// every negative bf16 (sign bit set) to +0.  Rounding keeps the sign, so this
// equals rounding max(x, 0).
PX_DEV s16x4 relu_bf16(f32x4 d) {
  const s16x2 z = {0, 0};
  const s16x2 lo = __builtin_elementwise_max(__builtin_bit_cast(s16x2, __builtin_convertvector(f32x2{d.x, d.y}, bf16x2)), z);
  const s16x2 hi = __builtin_elementwise_max(__builtin_bit_cast(s16x2, __builtin_convertvector(f32x2{d.z, d.w}, bf16x2)), z);
  return s16x4{lo.x, lo.y, hi.x, hi.y};
}

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

// Row strides (bf16 elements) of the LDS weight copies.  A fragment read
// takes 16 rows at once; rows 128 bytes apart would all start in the same
// LDS bank, so the rows are padded: 68 elements (136 bytes) and 20 (40 bytes)
// spread 16 consecutive rows over distinct banks.  Multiples of 4, as frag()
// needs.
constexpr int kLdIn = 20, kLdH = 68;

// bf16 copies of the weights in the layouts the fragments read
struct alignas(16) MlpLds {
  uint16_t w1t[kH * kLdIn];  // W1T[o][f]
  uint16_t w2t[kH * kLdH];   // W2T[o][i]
  uint16_t w2[kH * kLdH];    // W2[i][o]
  uint16_t w3p[16 * kLdH];   // row 0 = w3, rows 1..15 zero (M padded to 16)
  float b1[kH], b2[kH], b3;
};

PX_DEV void load_weights(MlpLds& L, const uint16_t* __restrict__ wbf, const float* __restrict__ par) {
  // wbf: bf16 of the parameter vector (same order as par)
  for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) {
    const int f = i / kH, o = i % kH;
    L.w1t[o * kLdIn + f] = wbf[kOffW1 + i];
  }
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
    const int in = i / kH, o = i % kH;
    L.w2t[o * kLdH + in] = wbf[kOffW2 + i];
    L.w2[in * kLdH + o] = wbf[kOffW2 + i];
  }
  for (int i = threadIdx.x; i < 16 * kLdH; i += blockDim.x) L.w3p[i] = i < kH ? wbf[kOffW3 + i] : (uint16_t)0;
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    L.b1[i] = par[kOffB1 + i];
    L.b2[i] = par[kOffB2 + i];
  }
  if (threadIdx.x == 0) L.b3 = par[kOffB3];
  __syncthreads();
}

// The forward pass reads only these (no W2 copy for the backward product).
struct alignas(16) MlpFwdLds {
  uint16_t w1t[kH * kLdIn];
  uint16_t w2t[kH * kLdH];
  uint16_t w3p[16 * kLdH];
  float b1[kH], b2[kH], b3;
};

PX_DEV void load_fwd_weights(MlpFwdLds& L, const uint16_t* __restrict__ wbf, const float* __restrict__ par) {
  for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) L.w1t[(i % kH) * kLdIn + i / kH] = wbf[kOffW1 + i];
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) L.w2t[(i % kH) * kLdH + i / kH] = wbf[kOffW2 + i];
  for (int i = threadIdx.x; i < 16 * kLdH; i += blockDim.x) L.w3p[i] = i < kH ? wbf[kOffW3 + i] : (uint16_t)0;
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    L.b1[i] = par[kOffB1 + i];
    L.b2[i] = par[kOffB2 + i];
  }
  if (threadIdx.x == 0) L.b3 = par[kOffB3];
  __syncthreads();
}

// Forward pass of 64 samples on one wavefront (four 16-sample column groups).
// F.x holds the B fragments of X^T: F.x[c] = features 4g..4g+3 of sample
// 16c + lane % 16.  Outputs the activations (bf16, in the output-tile layout:
// h1(t, c) = features 16t + 4g + j of sample 16c + lane % 16) and y for
// lanes 0..15 of each column group (y[c]).  Each bias is the accumulator's
// initial value.  Layers 2 and 3 contract K = 64 as two K = 32 products
// whose k order is (tile 2q rows, tile 2q + 1 rows) for both operands.
struct Fwd {
  s16x4 x[4];       // B fragments of X^T, per column group
  // activations as the K = 32 operands of the next layer: hq[q][c] holds
  // tile 2q in its low half and tile 2q + 1 in its high half
  s16x8 h1q[2][4];
  s16x8 h2q[2][4];
  float y[4];       // valid in lanes 0..15
  PX_DEV s16x4 h1(int t, int c) const { return (t & 1) ? h1q[t >> 1][c].hi : h1q[t >> 1][c].lo; }
  PX_DEV s16x4 h2(int t, int c) const { return (t & 1) ? h2q[t >> 1][c].hi : h2q[t >> 1][c].lo; }
};

template <class W>
PX_DEV void forward_x(const W& L, Fwd& F) {
  const int g = (int)(__lane_id() >> 4);
  // layer 1: D1[t][c] = W1T[t] . X^T[c] + b1, ReLU
  for (int t = 0; t < 4; t++) {
    const s16x4 a = frag(L.w1t, kLdIn, 16 * t, 0);
    const f32x4 b = {L.b1[16 * t + 4 * g], L.b1[16 * t + 4 * g + 1], L.b1[16 * t + 4 * g + 2], L.b1[16 * t + 4 * g + 3]};
    for (int c = 0; c < 4; c++) {
      const s16x4 h = relu_bf16(mfma(a, F.x[c], b));
      if (t & 1)
        F.h1q[t >> 1][c].hi = h;
      else
        F.h1q[t >> 1][c].lo = h;
    }
  }
  // layer 2: D2[t][c] = sum_q W2T[t][q] . H1[q][c] + b2, ReLU
  for (int t = 0; t < 4; t++) {
    s16x8 a[2];
    for (int q = 0; q < 2; q++)
      a[q] = cat8(frag(L.w2t, kLdH, 16 * t, 32 * q), frag(L.w2t, kLdH, 16 * t, 32 * q + 16));
    const f32x4 b = {L.b2[16 * t + 4 * g], L.b2[16 * t + 4 * g + 1], L.b2[16 * t + 4 * g + 2], L.b2[16 * t + 4 * g + 3]};
    for (int c = 0; c < 4; c++) {
      f32x4 d = mfma32(a[0], F.h1q[0][c], b);
      d = mfma32(a[1], F.h1q[1][c], d);
      const s16x4 h = relu_bf16(d);
      if (t & 1)
        F.h2q[t >> 1][c].hi = h;
      else
        F.h2q[t >> 1][c].lo = h;
    }
  }
  // layer 3: y[c] = w3 . H2[c] + b3 (row 0 of a tile padded to 16 rows)
  s16x8 a3[2];
  for (int q = 0; q < 2; q++) a3[q] = cat8(frag(L.w3p, kLdH, 0, 32 * q), frag(L.w3p, kLdH, 0, 32 * q + 16));
  const f32x4 b3 = {L.b3, 0.0f, 0.0f, 0.0f};
  for (int c = 0; c < 4; c++) {
    f32x4 d = mfma32(a3[0], F.h2q[0][c], b3);
    d = mfma32(a3[1], F.h2q[1][c], d);
    F.y[c] = d.x;  // row 0 lives in lanes 0..15, element 0
  }
}

// feat: bf16 [n][16]; samples >= n read as zeros.
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
  forward_x(L, F);
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

// Forward + backward for the loss (1/n_batch) sum (y - target)^2.  Blocks of
// kGradWaves wavefronts, one per SIMD; wavefront w of block b takes the
// 64-sample tiles v, v + W, v + 2W .. (v = b kGradWaves + w, W = all
// wavefronts) in that order and keeps its gradient in the MFMA accumulators
// (each tile's products are accumulated onto the running sums as the C
// operand).  The wavefronts of a block add their sums in LDS in wavefront
// order and the block writes one record (gradient + loss sum) to
// part[blockIdx.x]; k_mlp_sum adds the records in a fixed order.  Every sum
// runs in an order fixed by the indices, so the result does not depend on
// scheduling.
constexpr uint32_t kGradWaves = 4;

PX_DEV void grad_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kGradWaves) void k_mlp_grad(const uint16_t* __restrict__ wbf, const float* __restrict__ par,
                                                 const uint16_t* __restrict__ feat, const float* __restrict__ target,
                                                 uint32_t n, float inv_batch, float* __restrict__ part) {
  __shared__ MlpLds L;
  // per wavefront; fragments read 8 bytes per lane: every array is 16-byte aligned
  __shared__ alignas(16) uint16_t s_a[kGradWaves][kH * 64];    // feature-major [f][s]: H1, then dH1
  __shared__ alignas(16) uint16_t s_b[kGradWaves][kH * 64];    // feature-major [f][s]: H2, then dH2
  __shared__ alignas(16) uint16_t s_x[kGradWaves][kIn * 64];   // X^T feature-major [f][s]
  __shared__ alignas(16) uint16_t s_d[kGradWaves][16 * 64];    // row 0 = dy (bf16), rows 1..15 zero
  __shared__ alignas(16) uint16_t s_one[kGradWaves][16 * 64];  // row 0 = 1.0: B fragment of a column of ones
  __shared__ float s_red[kPartFloats];                          // the block's record
  load_weights(L, wbf, par);
  const uint32_t l = __lane_id(), wave = threadIdx.x >> 6;
  uint16_t* sA = s_a[wave];
  uint16_t* sB = s_b[wave];
  uint16_t* sX = s_x[wave];
  uint16_t* sD = s_d[wave];
  uint16_t* sOne = s_one[wave];
  const int g = (int)(l >> 4), col = (int)(l & 15u);
  for (int i = (int)l; i < 15 * 64; i += 64) sD[64 + i] = 0;
  for (int i = (int)l; i < 16 * 64; i += 64) sOne[i] = i < 64 ? (uint16_t)0x3F80u : (uint16_t)0;  // 1.0 bf16
  const f32x4 z = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  f32x4 aW3[4], aB3 = z, aW2[4][4], aB2[4], aW1[4], aB1[4];
  for (int a = 0; a < 4; a++) {
    aW3[a] = aB2[a] = aW1[a] = aB1[a] = z;
    for (int b = 0; b < 4; b++) aW2[a][b] = z;
  }
  float loss = 0.0f;
  const uint32_t tiles = (n + 63u) / 64u;
  for (uint32_t tile = blockIdx.x * kGradWaves + wave; tile < tiles; tile += gridDim.x * kGradWaves) {
    const uint32_t s0 = tile * 64u;
    Fwd F;
    forward(L, feat, n, s0, F);
    // dy = 2 (y - target) / n_batch, zero for padding samples
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
    // stage X^T, H1, H2 feature-major
    for (int c = 0; c < 4; c++) {
      const int s = 16 * c + col;
      for (int j = 0; j < 4; j++)
        sX[(4 * g + j) * 64 + s] = (uint16_t)(j == 0 ? F.x[c].x : j == 1 ? F.x[c].y : j == 2 ? F.x[c].z : F.x[c].w);
      for (int t = 0; t < 4; t++) {
        const s16x4 a = F.h1(t, c), b = F.h2(t, c);
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
    grad_wave_sync();
    // dW3[k] += sum_s dy[s] H2[k][s] (M = 1 padded, K = samples, N = k); db3 += sum_s dy[s]
    for (int t = 0; t < 4; t++)
      for (int kc = 0; kc < 4; kc++) aW3[t] = mfma(frag(sD, 64, 0, 16 * kc), frag(sB, 64, 16 * t, 16 * kc), aW3[t]);
    for (int kc = 0; kc < 4; kc++) aB3 = mfma(frag(sD, 64, 0, 16 * kc), frag(sOne, 64, 0, 16 * kc), aB3);
    // dH2 = w3[k] dy[s] where H2 > 0 (output-tile layout), then feature-major in sB
    s16x4 dh2[4][4];
    for (int t = 0; t < 4; t++) {
      float w[4];
      for (int j = 0; j < 4; j++) w[j] = bf2f(L.w3p[16 * t + 4 * g + j]);
      for (int c = 0; c < 4; c++) {
        const s16x4 h = F.h2(t, c);
        const float dy = F.y[c];
        dh2[t][c] = pack4(h.x > 0 ? w[0] * dy : 0.0f, h.y > 0 ? w[1] * dy : 0.0f, h.z > 0 ? w[2] * dy : 0.0f,
                          h.w > 0 ? w[3] * dy : 0.0f);
      }
    }
    grad_wave_sync();
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
    grad_wave_sync();
    // dW2[i][o] += sum_s H1[i][s] dH2[o][s]: tile (M = o, K = s, N = i) gives dW2T[o][i];
    // db2[o] += sum_s dH2[o][s]
    for (int to = 0; to < 4; to++) {
      for (int ti = 0; ti < 4; ti++)
        for (int kc = 0; kc < 4; kc++)
          aW2[to][ti] = mfma(frag(sB, 64, 16 * to, 16 * kc), frag(sA, 64, 16 * ti, 16 * kc), aW2[to][ti]);
      for (int kc = 0; kc < 4; kc++)
        aB2[to] = mfma(frag(sB, 64, 16 * to, 16 * kc), frag(sOne, 64, 0, 16 * kc), aB2[to]);
    }
    // dH1 = W2 . dH2 where H1 > 0 (M = i, K = o, N = samples; B = dh2 in registers)
    s16x4 dh1[4][4];
    for (int ti = 0; ti < 4; ti++) {
      s16x4 a[4];
      for (int ko = 0; ko < 4; ko++) a[ko] = frag(L.w2, kLdH, 16 * ti, 16 * ko);
      for (int c = 0; c < 4; c++) {
        f32x4 d = z;
        for (int ko = 0; ko < 4; ko++) d = mfma(a[ko], dh2[ko][c], d);
        const s16x4 h = F.h1(ti, c);
        dh1[ti][c] = pack4(h.x > 0 ? d.x : 0.0f, h.y > 0 ? d.y : 0.0f, h.z > 0 ? d.z : 0.0f, h.w > 0 ? d.w : 0.0f);
      }
    }
    grad_wave_sync();
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
    grad_wave_sync();
    // dW1[f][o] += sum_s X[s][f] dH1[o][s]: tile (M = o, K = s, N = f); db1[o] += sum_s dH1[o][s]
    for (int to = 0; to < 4; to++) {
      for (int kc = 0; kc < 4; kc++)
        aW1[to] = mfma(frag(sA, 64, 16 * to, 16 * kc), frag(sX, 64, 0, 16 * kc), aW1[to]);
      for (int kc = 0; kc < 4; kc++)
        aB1[to] = mfma(frag(sA, 64, 16 * to, 16 * kc), frag(sOne, 64, 0, 16 * kc), aB1[to]);
    }
    grad_wave_sync();  // the next tile overwrites sA, sB, sX, sD
  }
  // this block's record: the wavefronts' sums added in wavefront order
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  __syncthreads();
  for (uint32_t w = 0; w < kGradWaves; w++) {
    if (wave == w) {
      float* out = s_red;
      const bool first = w == 0u;
#define PX_ACC(idx, v) (out[idx] = first ? (v) : out[idx] + (v))
      if (l < 16u)
        for (int t = 0; t < 4; t++) PX_ACC(kOffW3 + 16 * t + col, aW3[t].x);
      if (l == 0u) PX_ACC(kOffB3, aB3.x);
      for (int to = 0; to < 4; to++) {
        // element (o = 16 to + 4g + j, i = 16 ti + col) of dW2T -> W2[i][o]
        for (int ti = 0; ti < 4; ti++)
          for (int j = 0; j < 4; j++) PX_ACC(kOffW2 + (16 * ti + col) * kH + 16 * to + 4 * g + j, aW2[to][ti][j]);
        if (col == 0)
          for (int j = 0; j < 4; j++) PX_ACC(kOffB2 + 16 * to + 4 * g + j, aB2[to][j]);
        for (int j = 0; j < 4; j++) PX_ACC(kOffW1 + col * kH + 16 * to + 4 * g + j, aW1[to][j]);
        if (col == 0)
          for (int j = 0; j < 4; j++) PX_ACC(kOffB1 + 16 * to + 4 * g + j, aB1[to][j]);
      }
      if (l == 0u) PX_ACC(kParams, loss);
#undef PX_ACC
    }
    __syncthreads();
  }
  float* rec = part + (size_t)blockIdx.x * kPartFloats;
  for (uint32_t p = threadIdx.x; p < (uint32_t)kPartFloats; p += blockDim.x) rec[p] = s_red[p];
}

// Sum of the nb block records of k_mlp_grad: a block takes 64 parameters (or
// the loss) per lane, its wavefront w the records w, w + 16, w + 32 ..; the
// 16 wavefront sums are then added in wavefront order.  A fixed order.
constexpr uint32_t kSumWaves = 16;
__global__ __launch_bounds__(64 * kSumWaves) void k_mlp_sum(const float* __restrict__ part, uint32_t nb,
                                                            float* __restrict__ out) {
  __shared__ float s_part[kSumWaves][64];
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t p = blockIdx.x * 64u + lane;
  float s = 0.0f;
  if (p < (uint32_t)kPartFloats)
    for (uint32_t b = wave; b < nb; b += kSumWaves) s += part[(size_t)b * kPartFloats + p];
  s_part[wave][lane] = s;
  __syncthreads();
  if (wave == 0u && p < (uint32_t)kPartFloats) {
    float t = 0.0f;
    for (uint32_t w = 0; w < kSumWaves; w++) t += s_part[w][lane];
    out[p] = t;
  }
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

// Per-launch constants of the features (host-computed once per launch).
struct FeatK {
  float sx[3];      // 2 / dims: origin -> [-1, 1]
  float dims[3];
  float horizon;    // cap of the sampled extent, voxels
  float inv_log_h;  // 1 / log2(1 + horizon)
  float kind;       // -1 primary, 1 shadow, 0 otherwise
};

PX_DEV uint32_t pack_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

// Features of a ray (16, in [-1, 1]), bf16.  None depends on where the
// camera stands except the origin's height, so a network trained on some
// cameras carries over to others:
//   f0  log2(1 + entry t) over the world (0 when the origin is inside),
//   f1  origin height y / dims.y,
//   f2  log2(1 + clipped extent) (to the exit or max t, capped at `horizon`),
//   f3..f5  unit direction,
//   f6..f10  brick occupancy (grid != 0) at 0.5, 4, 8, 16, 32 voxels past the
//            entry, f11..f14 super-cell occupancy at 64, 128, 256, 512
//            voxels: +1 occupied, -1 empty or past the extent,
//   f15 the kind.
// The near samples tell how soon the ray meets geometry, the far ones how
// cluttered its path stays - what the step count of march() follows.
// r0, r1 are the two halves of the 32-byte ray record.  Returned packed two
// per word, feature 2k in the low half of word k.  The arithmetic is the
// hardware's (v_rcp_f32, v_rsq_f32, v_log_f32, v_cvt_pk_bf16_f32) and
// branch-free: features are inputs to the network, not outputs, and the same
// function makes the training features and the ones prediction uses.
PX_DEV void ray_features(const PxWorldView& w, const px::MarchTables& T, float4 r0, float4 r1, const FeatK& K,
                         uint4& o0, uint4& o1) {
  const float ro[3] = {r0.x, r0.y, r0.z};
  const float rd[3] = {r0.w, r1.x, r1.y};
  const float maxT = r1.z;
  float f[16];
  const float len2 = rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2];
  const float rl = len2 > 0.0f ? __builtin_amdgcn_rsqf(len2) : 0.0f;
  float tmn = 0.0f, tmx = maxT;
  for (int a = 0; a < 3; a++) {
    f[3 + a] = __builtin_amdgcn_fmed3f(rd[a] * rl, -1.0f, 1.0f);
    const float d = fabsf(rd[a]) < 1e-6f ? (rd[a] < 0.0f ? -1e-6f : 1e-6f) : rd[a];
    const float inv = __builtin_amdgcn_rcpf(d);
    const float t0 = -ro[a] * inv, t1 = (K.dims[a] - ro[a]) * inv;
    tmn = fmaxf(tmn, fminf(t0, t1));
    tmx = fminf(tmx, fmaxf(t0, t1));
  }
  const float len = len2 * rl;  // |rd|
  const float ext = fmaxf(fminf(tmx, tmn + K.horizon) - tmn, 0.0f) * len;  // voxels
  const float ent = tmn * len;
  f[0] = fminf(__builtin_amdgcn_logf(1.0f + ent) * K.inv_log_h, 1.0f) * 2.0f - 1.0f;
  f[1] = -1.0f;
  f[2] = fminf(__builtin_amdgcn_logf(1.0f + ext) * K.inv_log_h, 1.0f) * 2.0f - 1.0f;
  // sample k at `dist[k]` voxels past the entry: point = ro + rd (tmn + dist rl)
  const float dist[10] = {0.5f, 4.0f, 8.0f, 16.0f, 32.0f, 64.0f, 128.0f, 256.0f, 512.0f, 2.0f};
  for (int k = 0; k < 10; k++) {
    const float t = __builtin_fmaf(dist[k], rl, tmn);
    const bool along = ext > 0.0f && dist[k] <= ext;
    const bool brick = k < 5 || k == 9;
    const float sh = brick ? 0.125f : 0.03125f;  // brick or super-cell units
    uint32_t c[3];
    bool in = along;
    for (int a = 0; a < 3; a++) {
      c[a] = (uint32_t)(int32_t)floorf(__builtin_fmaf(rd[a], t, ro[a]) * sh);
      in = in && c[a] < (uint32_t)(brick ? w.bd[a] : w.sd[a]);
    }
    bool occ = false;
    if (brick) {
      const uint32_t idx = in ? c[0] + __umul24((uint32_t)w.bd[0], c[1] + __umul24((uint32_t)w.bd[1], c[2])) : 0u;
      occ = in && ((w.brick_mask[idx >> 5] >> (idx & 31u)) & 1u);
    } else {
      const uint32_t idx = in ? c[0] + __umul24((uint32_t)w.sd[0], c[1] + __umul24((uint32_t)w.sd[1], c[2])) : 0u;
      occ = in && ((T.sup[idx >> 5] >> (idx & 31u)) & 1u);
    }
    f[k == 9 ? 1 : 6 + k] = occ ? 1.0f : -1.0f;
  }
  f[15] = K.kind;
  o0 = uint4{pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3]), pack_bf16(f[4], f[5]), pack_bf16(f[6], f[7])};
  o1 = uint4{pack_bf16(f[8], f[9]), pack_bf16(f[10], f[11]), pack_bf16(f[12], f[13]), pack_bf16(f[14], f[15])};
}

template <bool SUP_LDS>
__global__ __launch_bounds__(256) void k_mlp_features(PxWorldView w, const px_ray* __restrict__ rays, uint32_t n,
                                                      FeatK K, uint16_t* __restrict__ feat) {
  extern __shared__ uint32_t s_mem[];
  const px::MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* rp = reinterpret_cast<const float4*>(rays + i);
  uint4 o0, o1;
  ray_features(w, T, rp[0], rp[1], K, o0, o1);
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

// ---------------------------------------------------------------------------
// Prediction and a stable global counting sort by cost key, fused.
//
// The rays are cut into tiles of 256, 64 per wavefront; block g owns tiles
// [g R, (g + 1) R).  Each wavefront computes its 64 rays' features, stages
// them in LDS as MFMA fragments, runs the network, and turns the prediction
// into a bucket (bucket_of, as the windowed sort); the sort key is the bucket,
// or (face of the direction, bucket) with PX_SORT_FACES.  A lane's rank among
// the wavefront's rays of its key comes from one ballot per key bit; the
// block keeps running key counts over its tiles in LDS, so each ray gets its
// rank inside (block range, key): keys[i] = key << 24 | rank.  The block's
// key totals go to part[key * G + g] (key-major).  k_sort_offsets scans each
// key's row of part (one block per key), and k_sort_scatter adds the totals
// of the earlier keys, which gives each (key, block) run's first sorted
// position, and writes perm[position + rank] = i.  The order is by key, then
// by ray index: the stable sort of the whole launch.  (A
// last-block-done scan inside the first kernel needs a device-scope release
// per block; on gfx950 each is an L2 write-back, and that form measured
// 0.20 ms per 2M rays against 0.14 ms for separate launches.)
constexpr uint32_t kSortMaxKeys = 256;  // 8 key bits; the block size of the sort kernels

PX_DEV void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Major axis and sign of a ray direction (0..5 = +x, -x, +y, -y, +z, -z; the
// first of equal magnitudes wins).  r0.w, r1.x, r1.y are the direction.
PX_DEV uint32_t ray_face(float4 r0, float4 r1) {
  uint32_t a = 0u;
  float m = fabsf(r0.w), d = r0.w;
  if (fabsf(r1.x) > m) {
    a = 1u;
    m = fabsf(r1.x);
    d = r1.x;
  }
  if (fabsf(r1.y) > m) {
    a = 2u;
    d = r1.y;
  }
  return a * 2u + (d < 0.0f ? 1u : 0u);
}

template <bool SUP_LDS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_mlp_sort_keys(
    PxWorldView w, const px_ray* __restrict__ rays, uint32_t n, FeatK K, const uint16_t* __restrict__ wbf,
    const float* __restrict__ par, uint32_t nb, uint32_t faces, float lo, float hi, uint32_t tiles_per_block,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ part) {
  extern __shared__ uint32_t s_mem[];
  __shared__ MlpFwdLds L;
  // staged features, 24 elements (48 bytes) per sample: 16-byte aligned
  // stores, and fragment reads of 16 samples spread over the banks
  constexpr int kLdX = 24;
  __shared__ alignas(16) uint16_t s_x[4][64 * kLdX];
  __shared__ uint32_t s_b[4][64];
  __shared__ uint32_t s_wc[2][4][kSortMaxKeys];  // per-wavefront key counts, by tile parity
  __shared__ uint32_t s_run[kSortMaxKeys];
  const uint32_t tid = threadIdx.x, l = __lane_id(), wave = tid >> 6;
  const uint32_t nkeys = faces ? 6u * nb : nb;
  s_run[tid] = 0u;  // blockDim == kSortMaxKeys
  const px::MarchTables T = px::load_march_tables<SUP_LDS>(w, s_mem);
  load_fwd_weights(L, wbf, par);
  const float scale = (float)nb / (hi - lo);
  const uint32_t ntiles = (n + 255u) / 256u;
  const uint32_t t0 = blockIdx.x * tiles_per_block, t1 = min(t0 + tiles_per_block, ntiles);
  // the next tile's ray record is loaded while this tile's network runs
  float4 n0 = float4{0.0f, 0.0f, 0.0f, 0.0f}, n1 = n0;
  if (t0 < t1) {
    const uint32_t i = t0 * 256u + tid;
    if (i < n) {
      const float4* rp = reinterpret_cast<const float4*>(rays + i);
      n0 = rp[0];
      n1 = rp[1];
    }
  }
  for (uint32_t tile = t0; tile < t1; tile++) {
    const uint32_t i = tile * 256u + tid;
    const bool valid = i < n;
    const float4 r0 = n0, r1 = n1;
    if (tile + 1u < t1 && i + 256u < n) {
      const float4* rp = reinterpret_cast<const float4*>(rays + i + 256u);
      n0 = rp[0];
      n1 = rp[1];
    }
    uint4 q0 = uint4{0u, 0u, 0u, 0u}, q1 = q0;
    if (valid) ray_features(w, T, r0, r1, K, q0, q1);
    uint4* xp = reinterpret_cast<uint4*>(&s_x[wave][l * kLdX]);
    xp[0] = q0;
    xp[1] = q1;
    wave_lds_sync();
    Fwd F;
    for (int c = 0; c < 4; c++) F.x[c] = frag(s_x[wave], kLdX, 16 * c, 0);
    forward_x(L, F);
    if (l < 16u)
      for (int c = 0; c < 4; c++) s_b[wave][16 * c + l] = bucket_of(F.y[c], lo, scale, nb);
    wave_lds_sync();
    // sort key: the cost bucket, after the direction's face when asked
    const uint32_t b = s_b[wave][l];
    const uint32_t key = faces ? ray_face(r0, r1) * nb + b : b;
    // lanes with this lane's key, from one ballot per key bit
    uint64_t mo = __ballot(valid);
    for (uint32_t k = 0; k < 8u; k++) {
      const uint64_t v = __ballot((key >> k) & 1u);
      mo &= ((key >> k) & 1u) ? v : ~v;
    }
    const uint32_t rank = (uint32_t)__popcll(mo & ((1ull << l) - 1ull));
    // the wavefront's count of each key, written by the key's first lane
    uint32_t(*wc)[kSortMaxKeys] = s_wc[tile & 1u];
    for (uint32_t j = l; j < nkeys; j += 64u) wc[wave][j] = 0u;
    wave_lds_sync();
    if (valid && rank == 0u) wc[wave][key] = (uint32_t)__popcll(mo);
    __syncthreads();
    if (valid) {
      uint32_t r = s_run[key] + rank;
      for (uint32_t v = 0; v < wave; v++) r += wc[v][key];
      keys[i] = (key << 24) | r;
    }
    __syncthreads();
    if (tid < nkeys) s_run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
  }
  __syncthreads();
  if (tid < nkeys) part[tid * gridDim.x + blockIdx.x] = s_run[tid];
}

// Per bucket b (one block each): the exclusive prefix of part[b][0..G) over
// the blocks, into off[b][..], and the bucket's total into tot[b].  Each
// thread owns consecutive 16-byte groups (G is a multiple of 4).
__global__ __launch_bounds__(256) void k_sort_offsets(const uint32_t* __restrict__ part, uint32_t g4,
                                                      uint32_t* __restrict__ off, uint32_t* __restrict__ tot) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t tid = threadIdx.x, l = __lane_id(), wv = tid >> 6;
  const uint32_t per = (g4 + 255u) / 256u;
  const uint32_t q0 = min(tid * per, g4), q1 = min(q0 + per, g4);
  const uint4* p4 = reinterpret_cast<const uint4*>(part) + (size_t)blockIdx.x * g4;
  uint4* o4 = reinterpret_cast<uint4*>(off) + (size_t)blockIdx.x * g4;
  uint32_t sum = 0;
  for (uint32_t q = q0; q < q1; q++) {
    const uint4 v = p4[q];
    sum += v.x + v.y + v.z + v.w;
  }
  // exclusive prefix of the thread sums over the block
  uint32_t x = sum;
  for (uint32_t o = 1; o < 64u; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  if (l == 63u) s_tmp[wv] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (uint32_t k = 0; k < wv; k++) run += s_tmp[k];
  if (tid == 255u) tot[blockIdx.x] = run + sum;
  for (uint32_t q = q0; q < q1; q++) {
    const uint4 v = p4[q];
    uint4 o;
    o.x = run;
    run += v.x;
    o.y = run;
    run += v.y;
    o.z = run;
    run += v.z;
    o.w = run;
    run += v.w;
    o4[q] = o;
  }
}

// perm[base[key] + off[key][g] - off[key][g0] + rank] = i for the rays of
// the window of blocks [g0, g1) (all blocks when the sort has no windows):
// base[key] = the window's first ray + the window's counts of the keys
// before it, each count off[k][g1] - off[k][g0] (tot[k] stands for
// off[k][nblk]).  Windows are whole tiles of 256 rays, so a block of this
// kernel lies in one.
__global__ __launch_bounds__(256) void k_sort_scatter(const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ tot, uint32_t nkeys, uint32_t n,
                                                      uint32_t nblk, uint32_t rays_per_block, uint32_t window_log2,
                                                      uint32_t* __restrict__ perm) {
  __shared__ uint32_t s_base[kSortMaxKeys];
  __shared__ uint32_t s_w0[kSortMaxKeys];
  __shared__ uint32_t s_tmp[4];
  const uint32_t i0 = blockIdx.x * 256u;
  uint32_t first = 0u, g0 = 0u, g1 = nblk;
  if (window_log2) {
    const uint32_t win = i0 >> window_log2, bpw = (1u << window_log2) / rays_per_block;
    first = win << window_log2;
    g0 = win * bpw;
    g1 = min(g0 + bpw, nblk);
  }
  {
    const uint32_t t = threadIdx.x, l = __lane_id(), wv = t >> 6;
    uint32_t v = 0u, p0 = 0u;
    if (t < nkeys) {
      p0 = g0 < nblk ? off[t * nblk + g0] : tot[t];
      v = (g1 < nblk ? off[t * nblk + g1] : tot[t]) - p0;
    }
    s_w0[t] = p0;
    uint32_t x = v;
    for (uint32_t o = 1; o < 64u; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (l >= o) x += y;
    }
    if (l == 63u) s_tmp[wv] = x;
    __syncthreads();
    uint32_t base = first;
    for (uint32_t k = 0; k < wv; k++) base += s_tmp[k];
    s_base[t] = base + x - v;
  }
  __syncthreads();
  const uint32_t i = i0 + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i], key = k >> 24;
  const uint32_t dst = s_base[key] + (off[key * nblk + i / rays_per_block] - s_w0[key]) + (k & 0xFFFFFFu);
  if (dst < n) perm[dst] = i;  // always true for consistent counts
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
  // global sort: part and off (keys * G each), key totals (256), keys (n)
  uint32_t* d_sort = nullptr;
  size_t sort_cap = 0;  // words
};

static int mlp_fail(px_mlp* m, int code, const char* what, const char* detail) {
  return px_fail(m ? m->c : nullptr, code, what, detail);
}

#define MLP_HIP(m, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return mlp_fail((m), PX_ERR_HIP, #expr, hipGetErrorString(e_)); \
  } while (0)

static FeatK feat_consts(const px_ctx* c, uint32_t kind) {
  FeatK K;
  for (int a = 0; a < 3; a++) {
    K.dims[a] = (float)c->dims[a];
    K.sx[a] = 2.0f / K.dims[a];
  }
  K.horizon = (float)std::max(c->dims[0], std::max(c->dims[1], c->dims[2]));
  K.inv_log_h = (float)(1.0 / std::log2(1.0 + (double)K.horizon));
  K.kind = kind == PX_RAY_PRIMARY ? -1.0f : (kind == PX_RAY_SHADOW ? 1.0f : 0.0f);
  return K;
}

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
  (void)hipFree(m->d_sort);
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
  const FeatK K = feat_consts(c, kind);
  PxProfScope prof(c, PX_PROF_MLP);
  if (full <= 64u * 1024u)
    hipLaunchKernelGGL(k_mlp_features<true>, dim3((n + 255u) / 256u), dim3(256), full, c->stream, w, d_rays, n, K,
                       d_feat);
  else
    hipLaunchKernelGGL(k_mlp_features<false>, dim3((n + 255u) / 256u), dim3(256), 32, c->stream, w, d_rays, n, K,
                       d_feat);
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
  const uint32_t tiles = (n + 63u) / 64u;
  // one block of kGradWaves wavefronts per CU (one wavefront per SIMD)
  const uint32_t G = std::min<uint32_t>((tiles + kGradWaves - 1u) / kGradWaves, c->n_cu);
  if (G > m->part_cap) {
    (void)hipFree(m->d_part);
    m->d_part = nullptr;
    m->part_cap = 0;
    MLP_HIP(m, hipMalloc(&m->d_part, (size_t)G * kPartFloats * 4));
    m->part_cap = G;
  }
  hipLaunchKernelGGL(k_mlp_grad, dim3(G), dim3(64 * kGradWaves), 0, c->stream, m->d_wbf, m->d_par, d_feat, d_target,
                     n, 1.0f / (float)n, m->d_part);
  MLP_HIP(m, hipGetLastError());
  hipLaunchKernelGGL(k_mlp_sum, dim3((kPartFloats + 63) / 64), dim3(64 * kSumWaves), 0, c->stream, m->d_part, G,
                     m->d_grad);
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

int px_mlp_sort_rays(px_mlp* m, uint32_t kind, const px_ray* d_rays, uint32_t n, uint32_t buckets, float lo,
                     float hi, uint32_t flags, uint32_t* d_perm) {
  if (!m || (n && (!d_rays || !d_perm))) return PX_ERR_ARG;
  px_ctx* c = m->c;
  if (kind > 3u) return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "unknown ray kind");
  if (flags & ~(PX_SORT_FACES | PX_SORT_WINDOW(31u))) return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "unknown flags");
  const uint32_t faces = (flags & PX_SORT_FACES) ? 1u : 0u;
  const uint32_t wlog = (flags >> 8) & 31u;
  if (wlog != 0u && (wlog < 8u || wlog > 30u))
    return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "window: 2^8 .. 2^30 rays");
  if (buckets == 0u || (uint64_t)buckets * (faces ? 6u : 1u) > kSortMaxKeys)
    return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "buckets: 1..256, or 1..42 with PX_SORT_FACES");
  if (!(hi > lo)) return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "need hi > lo");
  if (n == 0) return PX_OK;
  PX_HIP_CHECK(c, hipSetDevice(c->device));
  int rc = px_refresh_masks(c);
  if (rc) return rc;
  const uint32_t nkeys = buckets * (faces ? 6u : 1u);
  // block ranges: about four blocks per CU, a whole number of 256-ray tiles
  // each; the block count is padded to a multiple of 4 (16-byte scan groups)
  const uint32_t ntiles = (n + 255u) / 256u;
  const uint32_t target = std::max(4u, c->n_cu * 4u);
  uint32_t per = (ntiles + target - 1u) / target;
  if (wlog) {
    // a power of two of tiles that divides the window's, so no block
    // straddles two windows: the one with the fewest turns per block over
    // the rounds of resident blocks (about `target` at a time), counting a
    // block's start (network weights into LDS) as 4 tile turns
    const uint32_t tw = 1u << (wlog - 8u);
    uint32_t best = 1u;
    uint64_t best_cost = ~0ull;
    for (uint32_t p = 1u; p <= per && p <= tw; p *= 2u) {
      const uint64_t blocks = (ntiles + p - 1u) / p;
      const uint64_t cost = ((blocks + target - 1u) / target) * (p + 4u);
      if (cost <= best_cost) {
        best_cost = cost;
        best = p;
      }
    }
    per = best;
  }
  const uint32_t nblk = ((ntiles + per - 1u) / per + 3u) & ~3u;
  if ((size_t)per * 256u >= (size_t)1 << 24) return px_fail(c, PX_ERR_ARG, "px_mlp_sort_rays", "too many rays");
  const size_t need = 2u * (size_t)nkeys * nblk + kSortMaxKeys + n;
  if (need > m->sort_cap) {
    MLP_HIP(m, hipStreamSynchronize(c->stream));
    (void)hipFree(m->d_sort);
    m->d_sort = nullptr;
    m->sort_cap = 0;
    MLP_HIP(m, hipMalloc(&m->d_sort, need * 4));
    m->sort_cap = need;
  }
  uint32_t* part = m->d_sort;
  uint32_t* off = part + (size_t)nkeys * nblk;
  uint32_t* tot = off + (size_t)nkeys * nblk;
  uint32_t* keys = tot + kSortMaxKeys;
  PxWorldView w = px_view(c);
  const size_t full = (8u + (c->nsuper + 31u) / 32u) * 4u;
  const FeatK K = feat_consts(c, kind);
  PxProfScope prof(c, PX_PROF_MLP);
  if (full <= 24u * 1024u)
    hipLaunchKernelGGL(k_mlp_sort_keys<true>, dim3(nblk), dim3(256), full, c->stream, w, d_rays, n, K, m->d_wbf,
                       m->d_par, buckets, faces, lo, hi, per, keys, part);
  else
    hipLaunchKernelGGL(k_mlp_sort_keys<false>, dim3(nblk), dim3(256), 32, c->stream, w, d_rays, n, K, m->d_wbf,
                       m->d_par, buckets, faces, lo, hi, per, keys, part);
  PX_HIP_CHECK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sort_offsets, dim3(nkeys), dim3(256), 0, c->stream, part, nblk / 4u, off, tot);
  PX_HIP_CHECK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sort_scatter, dim3((n + 255u) / 256u), dim3(256), 0, c->stream, keys, off, tot, nkeys, n,
                     nblk, per * 256u, wlog, d_perm);
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}

int px_mlp_trace_sorted(px_mlp* m, uint32_t kind, const px_ray* d_rays, uint32_t n, uint32_t buckets, float lo,
                        float hi, uint32_t flags, uint32_t* d_perm, px_hit* d_out) {
  if (!m) return PX_ERR_ARG;
  int rc = px_mlp_sort_rays(m, kind, d_rays, n, buckets, lo, hi, flags, d_perm);
  if (rc) return rc;
  return px_trace_perm_device(m->c, kind, d_rays, d_perm, n, d_out);
}

}  // extern "C"
