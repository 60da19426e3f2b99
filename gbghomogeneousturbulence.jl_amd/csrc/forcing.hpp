// The post-step forcing projection VarianceRescaling (src/Forcing.jl:575-596): after the last
// stage, the band's modes are rescaled by one factor per box that holds the band's kinetic energy
// (over band) or the whole field's (over total) at the target.  Three small launches on the state
// after the step's projections, before the closing deposit reads it:
//   k_band_sums      per-block sums of w |u|^2 over the whole state and over the band (FP64);
//   k_rescale_factor one block per box: the compensated totals, the empty-band floor and the factor;
//   k_rescale_apply  u <- u + (s - 1) u on the band's modes, inside the band's bounding cube.
// The rescale multiplies whole modes by one real number, so it commutes with the Leray and the
// reality projections that the reference applies after it.
#pragma once
#include <hip/hip_runtime.h>

#include "spectral.hpp"

namespace hdns {

constexpr int RESCALE_BLOCKS = 256;

struct RescaleArgs {
  double target;     // the held value: the band's or the whole field's kinetic energy
  int over_total;    // 1: hold the whole field's value
  double inv_norm;   // 1/N^6
  double eps_t;      // eps of the run's element type (the empty-band floor)
};

// partials [box][RESCALE_BLOCKS][2]: sum w |u|^2 over every retained mode, and over the band
template <typename T>
__global__ __launch_bounds__(256) void k_band_sums(const cx<T> *__restrict__ uh, Geo g, BandSpec band,
                                                   double *__restrict__ partials) {
  __shared__ double red[4][2];
  const int box = blockIdx.z;
  const size_t comp = comp_of(g);
  uh += (size_t)box * 3 * comp;
  const size_t modes = (size_t)g.k * g.k * (g.m + 1);
  double whole = 0.0, inband = 0.0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < modes; l += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(l % (g.m + 1));
    const size_t row = l / (g.m + 1);
    const int ky_idx = (int)(row % g.k), kz_idx = (int)(row / g.k);
    const size_t at = row * g.kxp + kx;
    double e = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const cx<T> v = uh[c * comp + at];
      e += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
    }
    const double w = kx == 0 ? 1.0 : 2.0;
    const int my = signed_of(ky_idx, g), mz = signed_of(kz_idx, g);
    whole += w * e;
    if (in_band(kx * kx + my * my + mz * mz, band)) inband += w * e;
  }
  for (int off = 32; off > 0; off >>= 1) {
    whole += __shfl_down(whole, off, 64);
    inband += __shfl_down(inband, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6][0] = whole;
    red[threadIdx.x >> 6][1] = inband;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int wv = 0; wv < 4; ++wv) s += red[wv][threadIdx.x];
    partials[((size_t)box * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = s;
  }
}

// One block per box: scalars[8 box + 6] = s - 1, with s = sqrt(target_band / k_band), target_band
// the target (over band) or max(0, target - k + k_band) (over total); 0 for an empty band, which
// is left alone (src/Forcing.jl:110-111, :575-596).
__global__ __launch_bounds__(256) void k_rescale_factor(const double *__restrict__ partials, double *__restrict__ scalars,
                                                        RescaleArgs ra) {
  __shared__ double red[256][2], cor[256][2];
  const int box = blockIdx.x, tid = threadIdx.x;
  partials += (size_t)box * RESCALE_BLOCKS * 2;
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  for (int b = tid; b < RESCALE_BLOCKS; b += 256) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double x = partials[(size_t)b * 2 + i], t = s[i] + x;
      c[i] += fabs(s[i]) >= fabs(x) ? (s[i] - t) + x : (x - t) + s[i];
      s[i] = t;
    }
  }
  red[tid][0] = s[0];
  red[tid][1] = s[1];
  cor[tid][0] = c[0];
  cor[tid][1] = c[1];
  __syncthreads();
  if (tid == 0) {
    double tot[2];
    for (int i = 0; i < 2; ++i) {
      double S = 0.0, C = 0.0;
      for (int t = 0; t < 256; ++t) {
        const double x = red[t][i], u = S + x;
        C += fabs(S) >= fabs(x) ? (S - u) + x : (x - u) + S;
        S = u;
        C += cor[t][i];
      }
      tot[i] = (S + C) * ra.inv_norm;
    }
    const double whole_raw = tot[0], band_raw = tot[1];
    double f = 0.0;
    if (isfinite(band_raw) && band_raw > 10.0 * ra.eps_t * whole_raw) {
      const double band = 0.5 * band_raw, whole = 0.5 * whole_raw;
      const double target_band = ra.over_total ? fmax(0.0, ra.target - whole + band) : ra.target;
      f = sqrt(target_band / band) - 1.0;
    }
    scalars[8 * box + 6] = f;
  }
}

// u <- u + (s - 1) u on the band's modes; the grid covers the cube |m| <= r around the origin
// (r = the band's highest shell, or M for a band with no upper edge) for every box.
template <typename T>
__global__ __launch_bounds__(256) void k_rescale_apply(cx<T> *__restrict__ uh, Geo g, BandSpec band, int r,
                                                       const double *__restrict__ scalars) {
  const int box = blockIdx.z;
  const size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int side = 2 * r + 1;
  if (l >= (size_t)side * side * (r + 1)) return;
  const int kx = (int)(l % (r + 1));
  const int my = (int)((l / (r + 1)) % side) - r, mz = (int)(l / ((size_t)(r + 1) * side)) - r;
  if (!in_band(kx * kx + my * my + mz * mz, band)) return;
  const T f = T(scalars[8 * box + 6]);
  const int ky_idx = my < 0 ? my + g.k : my, kz_idx = mz < 0 ? mz + g.k : mz;
  const size_t comp = comp_of(g);
  cx<T> *u = uh + (size_t)box * 3 * comp + ((size_t)kz_idx * g.k + ky_idx) * g.kxp + kx;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cx<T> v = u[c * comp];
    u[c * comp] = {v.re + f * v.re, v.im + f * v.im};
  }
}

}  // namespace hdns
