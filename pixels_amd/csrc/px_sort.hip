// px_sort.hip - device radix sort of 64-bit keys (rocPRIM), used to group
// the voxel write list by world column in list order.  Kept in its own
// translation unit so the rocPRIM templates compile once.
#include <rocprim/device/device_radix_sort.hpp>

#include "px_internal.h"

// Sorts n keys from d_in into d_out over bits [0, end_bit), on the context's
// stream.  The temporary storage is the context's, grown on demand (the
// first call for a size allocates; later calls do not).
int px_sort_u64(px_ctx* c, const uint64_t* d_in, uint64_t* d_out, uint32_t n, uint32_t end_bit) {
  if (n == 0) return PX_OK;
  size_t bytes = 0;
  PX_HIP_CHECK(c, rocprim::radix_sort_keys(nullptr, bytes, d_in, d_out, n, 0u, end_bit, c->stream));
  if (bytes > c->sort_tmp_bytes) {
    PX_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (c->d_sort_tmp) (void)hipFree(c->d_sort_tmp);
    c->d_sort_tmp = nullptr;
    c->sort_tmp_bytes = 0;
    PX_HIP_CHECK(c, hipMalloc(&c->d_sort_tmp, bytes));
    c->sort_tmp_bytes = bytes;
  }
  size_t have = c->sort_tmp_bytes;
  PX_HIP_CHECK(c, rocprim::radix_sort_keys(c->d_sort_tmp, have, d_in, d_out, n, 0u, end_bit, c->stream));
  return PX_OK;
}
