// Block placement across the XCDs.  The blocks of a launch go to the 8 XCDs in turn (block b, in
// x-fastest linear order, on XCD b mod 8), and each XCD has its own L2.  A kernel whose
// neighbouring chunks of work share cache lines hands the blocks of one XCD consecutive chunks,
// so the sharing chunks run back to back on one XCD.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hdns {

// The chunk of work block b of nb takes so that the blocks of one XCD take consecutive chunks:
// a bijection of [0, nb).
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t nb) {
  constexpr int64_t X = 8;
  const int64_t per = (nb + X - 1) / X, rem = nb % X, r = b % X;
  const int64_t start = rem == 0 || r <= rem ? r * per : r * per - (r - rem);
  return start + b / X;
}

}  // namespace hdns
