// The stable pair sort behind the traversal order (sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hdns {

// Sort `count` (key, value) pairs by the low `key_bits` bits of the key, stably, in double buffers
// (a holds the input; the result lands in a or b, *result_in_b says which).  With temp == nullptr
// only *temp_bytes is set, to the scratch the sort of `count` pairs needs.
hipError_t tile_sort_pairs(void *temp, size_t *temp_bytes, uint32_t *keys_a, uint32_t *keys_b, int32_t *vals_a,
                           int32_t *vals_b, int64_t count, int key_bits, hipStream_t stream, bool *result_in_b);

}  // namespace hdns
