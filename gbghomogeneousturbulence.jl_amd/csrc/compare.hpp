// The distance between the states of two contexts on one grid, on the device: the FP32-bulk
// error report against an FP64 run of the same flow, without moving either state to the host.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace hdns {

constexpr int DIST_BLOCKS = 1024;

// Per-block sums of w |a - b|^2 and w |b|^2 over the retained cube of `ncomp` compact components
// (the Parseval weight w is 1 on the kx = 0 and kx = N/2 planes and 2 inside).  The
// element types and padded row widths of the two layouts may differ.  partials: [ncomp][2][blocks].
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_state_distance(const cx<TA> *__restrict__ a, size_t comp_a, int kxp_a,
                                                        const cx<TB> *__restrict__ b, size_t comp_b, int kxp_b,
                                                        int k, int m, int n, int ncomp, double *__restrict__ partials) {
  __shared__ double red[4][8][2];
  const size_t modes = (size_t)k * k * (m + 1);
  for (int c = 0; c < ncomp; ++c) {
    double dd = 0.0, bb = 0.0;
    for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < modes; l += (size_t)gridDim.x * blockDim.x) {
      const int kx = (int)(l % (m + 1));
      const size_t row = l / (m + 1);
      const cx<TA> va = a[c * comp_a + row * kxp_a + kx];
      const cx<TB> vb = b[c * comp_b + row * kxp_b + kx];
      const double w = (kx == 0 || 2 * kx == n) ? 1.0 : 2.0;
      const double dr = (double)va.re - (double)vb.re, di = (double)va.im - (double)vb.im;
      dd += w * (dr * dr + di * di);
      bb += w * ((double)vb.re * (double)vb.re + (double)vb.im * (double)vb.im);
    }
    for (int off = 32; off > 0; off >>= 1) {
      dd += __shfl_down(dd, off, 64);
      bb += __shfl_down(bb, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[c][threadIdx.x >> 6][0] = dd;
      red[c][threadIdx.x >> 6][1] = bb;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * ncomp) {
    const int c = threadIdx.x >> 1, which = threadIdx.x & 1;
    double s = 0.0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) s += red[c][wv][which];
    partials[((size_t)c * 2 + which) * gridDim.x + blockIdx.x] = s;
  }
}

}  // namespace hdns
