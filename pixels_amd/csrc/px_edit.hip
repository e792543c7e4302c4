// px_edit.hip - edit chain on the heap brick pool (placeholder: filled in by
// the carve / flood / maint / island kernels).
#include "px_internal.h"

int px_edit_alloc(px_ctx* c) {
  PX_HIP_CHECK(c, hipMalloc(&c->d_impacts, 64 * 16 * sizeof(float)));
  return PX_OK;
}

void px_edit_free(px_ctx* c) {
  if (c->d_impacts) hipFree(c->d_impacts);
  c->d_impacts = nullptr;
}

int px_launch_island_pass(px_ctx* c) { return px_fail(c, PX_ERR_STATE, "px_island_pass", "not available yet"); }

int px_read_island_labels(px_ctx* c, uint32_t*, uint32_t*, uint32_t, uint32_t* n) {
  *n = 0;
  return px_fail(c, PX_ERR_STATE, "px_island_labels", "not available yet");
}
