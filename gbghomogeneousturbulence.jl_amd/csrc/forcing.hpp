// Band factors measured on the state of one field, on the device (no host readback):
//   * VarianceRescaling's post-step rescale (src/Forcing.jl:575-596): after the last stage the band's
//     modes are rescaled by one factor per box that holds the band's value (over band) or the whole
//     field's (over total) at the target;
//   * the amplitude of a ConstantPower or VarianceController scalar forcing at a stage's scalar
//     (src/Forcing.jl:1041-1117), where the velocity's comes from the deposit's reductions.
// The variance factor is 1/2 for the velocity (kinetic energy) and 1 for the scalar (src/Forcing.jl:456).
// Three small launches:
//   k_band_sums      per-block sums of w |f|^2 over the whole field and over the band (FP64);
//   k_band_factor    one block per box: the compensated totals, the empty-band floor and the factor;
//   k_rescale_apply  f <- f + (s - 1) f on the band's modes, inside the band's bounding cube.
// The rescale multiplies whole modes by one real number, so it commutes with the Leray and the
// reality projections that the reference applies after it.
#pragma once
#include <hip/hip_runtime.h>

#include "spectral.hpp"

namespace hdns {

constexpr int RESCALE_BLOCKS = 256;

struct BandArgs {
  int mode;           // 0: rescale factor s - 1; 1: ConstantPower amplitude; 2: VarianceController amplitude
  double factor;      // 1/2 (velocity) or 1 (scalar)
  int slot;           // scalars[8 box + slot] receives the result
  double target;      // the held value (rescale, controller)
  int over_total;     // rescale: hold the whole field's value
  double rate;        // ConstantPower: the injection rate
  double gain, turnover_time;   // controller
  const double *diffusivity;    // controller: per box, the field's diffusivity
  int dirichlet_slot;           // controller: scalars[8 box + this] = sum w |k|^2 |f|^2 / N^6
  double inv_norm;    // 1/N^6
  double eps_t;       // eps of the run's element type (the empty-band floor)
};

// partials [box][RESCALE_BLOCKS][2]: sum w |f|^2 over every retained mode, and over the band, of
// the ncomp compact components of f (box b's start at b * ncomp * comp)
template <typename T>
__global__ __launch_bounds__(256) void k_band_sums(const cx<T> *__restrict__ f, int ncomp, Geo g, BandSpec band,
                                                   double *__restrict__ partials) {
  __shared__ double red[4][2];
  const int box = blockIdx.z;
  const size_t comp = comp_of(g);
  f += (size_t)box * ncomp * comp;
  const size_t modes = (size_t)g.k * g.k * (g.m + 1);
  double whole = 0.0, inband = 0.0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < modes; l += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(l % (g.m + 1));
    const size_t row = l / (g.m + 1);
    const int ky_idx = (int)(row % g.k), kz_idx = (int)(row / g.k);
    const size_t at = row * g.kxp + kx;
    double e = 0.0;
    for (int c = 0; c < ncomp; ++c) {
      const cx<T> v = f[c * comp + at];
      e += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
    }
    const double w = parseval_w(kx, g.n);
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

// One block per box.  With W = factor * whole and B = factor * band (an empty band, the relative
// floor of src/Forcing.jl:110-111, gives 0):
//   mode 0: s - 1, s = sqrt(target_band / B), target_band = target or max(0, target - W + B);
//   mode 1: A = rate / (2 B);
//   mode 2: A = (2 factor D <f, |k|^2 f> - G (W - target) / t_l) / (2 B).
__global__ __launch_bounds__(256) void k_band_factor(const double *__restrict__ partials, double *__restrict__ scalars,
                                                     BandArgs ba) {
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
      tot[i] = (S + C) * ba.inv_norm;
    }
    const double whole_raw = tot[0], band_raw = tot[1];
    const double W = ba.factor * whole_raw, B = ba.factor * band_raw;
    double out = 0.0;
    if (isfinite(band_raw) && band_raw > 10.0 * ba.eps_t * whole_raw) {
      if (ba.mode == 0) {
        const double target_band = ba.over_total ? fmax(0.0, ba.target - W + B) : ba.target;
        out = sqrt(target_band / B) - 1.0;
      } else if (ba.mode == 1) {
        out = ba.rate / (2.0 * B);
      } else {
        const double dissipation = 2.0 * ba.factor * ba.diffusivity[box] * scalars[8 * box + ba.dirichlet_slot];
        out = (dissipation - ba.gain * (W - ba.target) / ba.turnover_time) / (2.0 * B);
      }
    }
    scalars[8 * box + ba.slot] = out;
  }
}

// f <- f + (s - 1) f on the band's modes of the ncomp components; the grid covers the cube
// |m| <= r around the origin (r = the band's highest shell, or M for a band with no upper edge)
// for every box.
template <typename T>
__global__ __launch_bounds__(256) void k_rescale_apply(cx<T> *__restrict__ f, int ncomp, Geo g, BandSpec band, int r,
                                                       const double *__restrict__ scalars, int slot) {
  const int box = blockIdx.z;
  const size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int side = 2 * r + 1;
  if (l >= (size_t)side * side * (r + 1)) return;
  const int kx = (int)(l % (r + 1));
  const int my = (int)((l / (r + 1)) % side) - r, mz = (int)(l / ((size_t)(r + 1) * side)) - r;
  if (!in_band(kx * kx + my * my + mz * mz, band)) return;
  const T s1 = T(scalars[8 * box + slot]);
  const int ky_idx = my < 0 ? my + g.k : my, kz_idx = mz < 0 ? mz + g.k : mz;
  const size_t comp = comp_of(g);
  cx<T> *u = f + (size_t)box * ncomp * comp + ((size_t)kz_idx * g.k + ky_idx) * g.kxp + kx;
  for (int c = 0; c < ncomp; ++c) {
    const cx<T> v = u[c * comp];
    u[c * comp] = {v.re + s1 * v.re, v.im + s1 * v.im};
  }
}

}  // namespace hdns

namespace hdns {

// StochasticOU's additive force f on the band (src/Forcing.jl:1404-1405), held over the step and
// added at every stage.  The stage update is linear in its source, so the force is added to what
// the update just wrote, with the stage's own integrating factor E, evaluated as the update does:
//   2N schemes:         q += E (b (dt f)),  and h += E (dt f) except at the last stage;
//   Shu-Osher schemes:  q += E^pb (cb (dt f)).
// f is transverse (velocity) and a real field's spectrum, so the state projections that follow
// the last stage leave it as it is.  src: [box][ncomp][2r+1][2r+1][r+1], FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_band_source(cx<T> *__restrict__ q, cx<T> *__restrict__ h, int ncomp, Geo g,
                                                     BandSpec band, int r, const cx<double> *__restrict__ src,
                                                     StageArgs sa, int scalar) {
  const int box = blockIdx.z;
  const size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int side = 2 * r + 1;
  const size_t cube = (size_t)side * side * (r + 1);
  if (l >= cube) return;
  const int mx = (int)(l % (r + 1));
  const int my = (int)((l / (r + 1)) % side) - r, mz = (int)(l / ((size_t)(r + 1) * side)) - r;
  if (!in_band(mx * mx + my * my + mz * mz, band)) return;
  const double kxd = g.k0 * mx, kyd = g.k0 * my, kzd = g.k0 * mz;
  const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
  const double xe = -(scalar ? sa.kappa[box] : sa.visc[box]) * stage_sub(sa) * k2d;
  const T dt = T(stage_dt(sa));
  const int ky_idx = my < 0 ? my + g.k : my, kz_idx = mz < 0 ? mz + g.k : mz;
  const size_t comp = comp_of(g), at = ((size_t)kz_idx * g.k + ky_idx) * g.kxp + mx;
  q += (size_t)box * ncomp * comp + at;
  h += (size_t)box * ncomp * comp + at;
  src += (size_t)box * ncomp * cube + l;
  for (int c = 0; c < ncomp; ++c) {
    const cx<double> fd = src[c * cube];
    const cx<T> df = scale(cx<T>{T(fd.re), T(fd.im)}, dt);
    if (sa.ssp) {
      q[c * comp] = q[c * comp] + scale(scale(df, T(sa.cb)), T(exp(sa.pb * xe)));
    } else {
      const T E = T(exp(xe));
      q[c * comp] = q[c * comp] + scale(scale(df, T(sa.b)), E);
      if (!sa.last) h[c * comp] = h[c * comp] + scale(df, E);
    }
  }
}

}  // namespace hdns
