// Buffer-resource access to device arrays: a uniform base in scalar registers plus a per-lane
// 32-bit byte offset and a uniform 32-bit scalar offset.  Compared with 64-bit flat addresses this
// holds one vector register per access pattern instead of two, and the uniform part of an index
// (a component, a pass's stride times q) rides in the scalar offset.  Accesses beyond the
// resource's byte range read zeros and drop writes.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace hdns {

constexpr int kBufferWord3 = 0x00020000;   // raw 32-bit buffer, CDNA

template <typename T>
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  // a resource over `elems` complex values from p (the range must stay below 2^31 bytes)
  __device__ __forceinline__ Buf(const cx<T> *p, size_t elems)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<cx<T> *>(p), (short)0,
                                            (int)(elems * sizeof(cx<T>) < 0x7fffffffu ? elems * sizeof(cx<T>)
                                                                                       : 0x7fffffffu),
                                            kBufferWord3)) {}
  // element voff (per lane) + soff (uniform)
  __device__ __forceinline__ cx<T> load(uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(T) == 4) {
      return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, voff * 8u, soff * 8u, 0));
    } else {
      return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, voff * 16u, soff * 16u, 0));
    }
  }
  __device__ __forceinline__ void store(cx<T> v, uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(T) == 4) {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, voff * 8u, soff * 8u, 0);
    } else {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0u, 0u, 0));
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, voff * 16u, soff * 16u, 0);
    }
  }
};

}  // namespace hdns
