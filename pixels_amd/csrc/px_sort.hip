// px_sort.hip - stable LSD radix sort of 64-bit keys on gfx950, 8 bits per
// pass over bits [begin_bit, end_bit):
//   k_radix_hist     per tile of 1024 keys, the count of each digit, stored
//                    digit-major (hist[digit * tiles + tile]);
//   k_radix_rows     one workgroup per digit: the exclusive scan of the
//                    digit's row of the table and the row's total;
//   k_radix_scatter  each tile again, in key order: the digit's first output
//                    position is the totals of the smaller digits (an LDS
//                    scan of the 256 row totals) plus the row's offset; a
//                    key's rank among the earlier keys of its tile with the
//                    same digit comes from eight wavefront ballots (lanes of
//                    equal digit) plus the counts of the earlier wavefronts
//                    and chunks, so equal digits keep their input order.
// Used to group the voxel write list by world column (keys column << 32 |
// list index, sorted from bit 32 so that the list order stays inside a
// column) and the scene builder's (cell, primitive) pairs.  Only kernel
// launches once the temporary buffer has its size: no memset, no host wait,
// so a caller may capture the sort in a graph.
#include <algorithm>

#include "px_internal.h"

namespace {

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortChunks = 4;
constexpr uint32_t kSortTile = kSortThreads * kSortChunks;  // keys per tile

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t* __restrict__ keys, uint32_t n,
                                                              uint32_t shift, uint32_t* __restrict__ hist,
                                                              uint32_t tiles) {
  __shared__ uint32_t s_h[256];
  const uint32_t t = threadIdx.x;
  s_h[t] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t c = 0; c < kSortChunks; c++) {
    const uint32_t i = base + c * kSortThreads + t;
    if (i < n) atomicAdd(&s_h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[t * tiles + blockIdx.x] = s_h[t];
}

// Row d of the digit-major table (hist[d * tiles ..]): exclusive prefix sum
// in place, the row's total into tot[d].  One workgroup per digit; each
// thread sums a contiguous run, the run sums are scanned in LDS.
__global__ __launch_bounds__(kSortThreads) void k_radix_rows(uint32_t* __restrict__ hist, uint32_t tiles,
                                                              uint32_t* __restrict__ tot) {
  __shared__ uint32_t s[kSortThreads];
  const uint32_t t = threadIdx.x;
  uint32_t* row = hist + (size_t)blockIdx.x * tiles;
  const uint32_t per = (tiles + kSortThreads - 1u) / kSortThreads;
  const uint32_t lo = min(t * per, tiles), hi = min(lo + per, tiles);
  uint32_t sum = 0u;
  for (uint32_t i = lo; i < hi; i++) sum += row[i];
  s[t] = sum;
  __syncthreads();
  for (uint32_t o = 1; o < kSortThreads; o <<= 1) {
    const uint32_t v = t >= o ? s[t - o] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  uint32_t run = s[t] - sum;
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
  if (t == kSortThreads - 1u) tot[blockIdx.x] = s[t];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* __restrict__ in,
                                                                 uint64_t* __restrict__ out, uint32_t n,
                                                                 uint32_t shift, const uint32_t* __restrict__ off,
                                                                 uint32_t tiles, const uint32_t* __restrict__ tot) {
  constexpr uint32_t kWaves = kSortThreads / 64u;
  __shared__ uint32_t s_base[256];        // next output position of each digit for this tile
  __shared__ uint32_t s_wcnt[kWaves][256];  // keys of each digit per wavefront, current chunk
  const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
  const uint64_t lt = (1ull << lane) - 1ull;
  // first position of digit t: the keys of the smaller digits, then this
  // tile's offset inside the digit's row
  s_base[t] = tot[t];
  __syncthreads();
  for (uint32_t o = 1; o < 256u; o <<= 1) {
    const uint32_t v = t >= o ? s_base[t - o] : 0u;
    __syncthreads();
    s_base[t] += v;
    __syncthreads();
  }
  s_base[t] = s_base[t] - tot[t] + off[t * tiles + blockIdx.x];
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t c = 0; c < kSortChunks; c++) {
    for (uint32_t w = 0; w < kWaves; w++) s_wcnt[w][t] = 0u;
    const uint32_t i = base + c * kSortThreads + t;
    const bool valid = i < n;
    const uint64_t k = valid ? in[i] : 0ull;
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    uint64_t same = __ballot(valid);
    for (int b = 0; b < 8; b++) {
      const bool bit = ((d >> b) & 1u) != 0u;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lt);
    __syncthreads();  // s_wcnt cleared, s_base of the previous chunk advanced
    if (valid && rank == (uint32_t)__popcll(same) - 1u) s_wcnt[wave][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t before = 0u;
      for (uint32_t w = 0; w < wave; w++) before += s_wcnt[w][d];
      out[s_base[d] + before + rank] = k;
    }
    __syncthreads();
    uint32_t add = 0u;
    for (uint32_t w = 0; w < kWaves; w++) add += s_wcnt[w][t];
    s_base[t] += add;
  }
}

__global__ void k_copy_u64(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[i];
}

}  // namespace

// Sorts n keys from d_in into d_out by bits [begin_bit, end_bit), stably, on
// the context's stream; d_in is left as it was.  The temporary storage (n
// keys + the digit table) is the context's, grown on demand: the first call
// for a size allocates and waits, later calls only launch kernels.
int px_sort_u64(px_ctx* c, const uint64_t* d_in, uint64_t* d_out, uint32_t n, uint32_t begin_bit,
                uint32_t end_bit) {
  if (n == 0) return PX_OK;
  const uint32_t tiles = (n + kSortTile - 1u) / kSortTile;
  const uint32_t passes = end_bit > begin_bit ? (end_bit - begin_bit + 7u) / 8u : 0u;
  const size_t key_bytes = ((size_t)n * 8 + 255) & ~(size_t)255;
  const size_t bytes = key_bytes + (size_t)256 * tiles * 4 + 256 * 4;
  if (bytes > c->sort_tmp_bytes) {
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (c->d_sort_tmp) (void)hipFree(c->d_sort_tmp);
    c->d_sort_tmp = nullptr;
    c->sort_tmp_bytes = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_sort_tmp, bytes));
    c->sort_tmp_bytes = bytes;
  }
  uint64_t* tmp = reinterpret_cast<uint64_t*>(c->d_sort_tmp);
  uint32_t* hist = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->d_sort_tmp) + key_bytes);
  uint32_t* tot = hist + (size_t)256 * tiles;
  hipStream_t s = c->stream;
  if (passes == 0) {
    hipLaunchKernelGGL(k_copy_u64, dim3(std::min<uint32_t>((n + 255u) / 256u, 4096u)), dim3(256), 0, s, d_in, d_out,
                       n);
    PX_HIP_CHECK(c, hipGetLastError());
    return PX_OK;
  }
  const uint64_t* src = d_in;
  for (uint32_t p = 0; p < passes; p++) {
    // the last pass lands in d_out
    uint64_t* dst = ((passes - 1u - p) & 1u) == 0u ? d_out : tmp;
    const uint32_t shift = begin_bit + 8u * p;
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, s, src, n, shift, hist, tiles);
    hipLaunchKernelGGL(k_radix_rows, dim3(256), dim3(kSortThreads), 0, s, hist, tiles, tot);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, src, dst, n, shift,
                       (const uint32_t*)hist, tiles, (const uint32_t*)tot);
    src = dst;
  }
  PX_HIP_CHECK(c, hipGetLastError());
  return PX_OK;
}
