// The traversal order's sort: a stable LSD radix sort of (tile key, tracer index) pairs, from
// rocPRIM.  Being stable, it orders the tracers of one tile by their index, so the order is a
// function of the positions alone: a rebuilt order equals a freshly built one
// (test/particles.jl:418-423), which an atomic counting-sort scatter cannot promise.
#include <rocprim/device/device_radix_sort.hpp>

#include "sort.hpp"

namespace hdns {

// rocPRIM takes its merge-sort path up to 2^20 pairs by default (a block sort and ten merge passes
// for 2^20 pairs); a 15-bit key at 256^3 needs only two onesweep radix passes.  Both paths are
// stable, so the order does not depend on the path taken.
using TileSortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                  rocprim::default_config, 1u << 16>;

hipError_t tile_sort_pairs(void *temp, size_t *temp_bytes, uint32_t *keys_a, uint32_t *keys_b, int32_t *vals_a,
                           int32_t *vals_b, int64_t count, int key_bits, hipStream_t stream, bool *result_in_b) {
  rocprim::double_buffer<uint32_t> keys(keys_a, keys_b);
  rocprim::double_buffer<int32_t> vals(vals_a, vals_b);
  size_t bytes = temp ? *temp_bytes : 0;
  const hipError_t e = rocprim::radix_sort_pairs<TileSortConfig>(temp, bytes, keys, vals, (size_t)count, 0u,
                                                                     (unsigned)key_bits, stream);
  if (!temp) {
    *temp_bytes = bytes;
    return e;
  }
  if (result_in_b) *result_in_b = vals.current() == vals_b;
  return e;
}

}  // namespace hdns
