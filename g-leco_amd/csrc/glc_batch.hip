// glc_batch.hip -- batched form of libglc (many columns, one launch).
// Placeholder entry points: they report GLC_ERR_UNSUPPORTED until the batched
// kernels land.
#include <hip/hip_runtime.h>

#include "glc.h"
#include "glc_internal.hpp"

using glc::fail;

extern "C" {
int glc_batch_compress(int, const int32_t*, const void* const*, const int64_t*, int32_t, void*, glc_batch** out) {
    if (out) *out = nullptr;
    return fail(GLC_ERR_UNSUPPORTED, "batched compress not available in this build");
}
int glc_batch_decompress(const glc_batch*, void* const*, void*) {
    return fail(GLC_ERR_UNSUPPORTED, "batched decompress not available in this build");
}
int glc_batch_decompress_graph(glc_batch*, void* const*, void*) {
    return fail(GLC_ERR_UNSUPPORTED, "batched decompress not available in this build");
}
int glc_batch_set_schedule(glc_batch*, int32_t, int32_t) { return fail(GLC_ERR_UNSUPPORTED, "no batch"); }
int glc_batch_get_column(const glc_batch*, int, const glc_column**) { return fail(GLC_ERR_UNSUPPORTED, "no batch"); }
int glc_batch_num_columns(const glc_batch*) { return 0; }
int glc_batch_free(glc_batch*) { return GLC_OK; }
}
