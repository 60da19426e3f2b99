// Off-step observables on the device: the velocity-derivative skewness
// S = C * 3 <det S> / <S:S>^(3/2) (src/Observables.jl:465-477).  The strain S = sym grad u is
// formed spectrally in the z inverse pass's loader, brought to physical space by the inverse y pass
// and an x pass that forms det S at every point and sums it in FP64; <S:S> = (1/2) <|w|^2> comes
// from the deposit's reductions.
#pragma once
#include "pencil.hpp"

namespace hdns {

// the six independent strain components: xx, yy, zz, xy, xz, yz
__host__ __device__ constexpr int strain_i(int s) { return s < 3 ? s : (s == 5 ? 1 : 0); }
__host__ __device__ constexpr int strain_j(int s) { return s < 3 ? s : (s == 3 ? 1 : 2); }

// z pencils Q [6][z][ky][kx] of S_hat_ij = i (k_i u_j + k_j u_i) / 2, scaled (grid: tiles, K, 6).
// Tiles whose line misses the ball are skipped, as the inverse y pass never reads them.
template <typename T, int N, int TX, int NT>
__global__ __launch_bounds__(NT) void k_zstrain(const cx<T> *__restrict__ uh, cx<T> *__restrict__ q, Geo g,
                                                const cx<double> *__restrict__ tw, T scale_by) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, s = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  if (!in_disk(line_start<T>(kx0), my, g)) return;
  stage_twiddles<T, N, NT>(twl, tw);
  const int ci = strain_i(s), cj = strain_j(s);
  const size_t comp = comp_of(g), pencil = (size_t)N * g.k * g.kxp, zstride = (size_t)g.k * g.kxp;
  const cx<T> *ui = uh + ci * comp + (size_t)ky_idx * g.kxp + kx0;
  const cx<T> *uj = uh + cj * comp + (size_t)ky_idx * g.kxp + kx0;
  cx<T> *dst = q + s * pencil + (size_t)ky_idx * g.kxp + kx0;
  fft_io<T, N, TX, NT, +1, true, false>(
      buf, twl,
      [&](int, int, int n, int col) {
        const int kz_idx = idx_of_bin(n, g);
        if (kz_idx < 0) return czero<T>();
        const int mx = kx0 + col, mz = signed_of(kz_idx, g);
        if (mx > g.m || !keeps(mx * mx + my * my + mz * mz, g)) return czero<T>();
        const int mv[3] = {mx, my, mz};
        const T ki = T(g.k0 * mv[ci]), kj = T(g.k0 * mv[cj]);
        const cx<T> a = uj[(size_t)kz_idx * zstride + col], b = ui[(size_t)kz_idx * zstride + col];
        const cx<T> t = {ki * a.re + kj * b.re, ki * a.im + kj * b.im};   // k_i u_j + k_j u_i
        return cx<T>{T(-0.5) * scale_by * t.im, T(0.5) * scale_by * t.re};   // (i / 2) t
      },
      [&](int, int, int n, int col, cx<T> v) { dst[(size_t)n * zstride + col] = v; });
}

// Per pair of y rows: the inverse x transforms of the six strain components (R [6][z][y][kx]), the
// determinant of the symmetric strain at every point of both rows, summed in FP64; one partial
// per block.  The packing of two rows as one complex transform is the forward x pass's.
template <typename T, int N, int XT>
__global__ __launch_bounds__(XT) void k_xdet(const cx<T> *__restrict__ rin, Geo g, const cx<double> *__restrict__ tw,
                                             double *__restrict__ partials) {
  constexpr int LPR = xpass_lanes<T, N, XT>();
  constexpr bool WAVE = LPR == 64;
  constexpr int XR = XT / LPR;
  constexpr int E = (N + LPR - 1) / LPR;
  constexpr int M = N / 3;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[XT / 64];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + XR * N;
  stage_twiddles<T, N, XT>(twl, tw);
  __syncthreads();
  const int rows_per_plane = N / (2 * XR);
  const int z = blockIdx.x / rows_per_plane;
  const int w = threadIdx.x / LPR, lane = threadIdx.x % LPR;
  const int ya = (blockIdx.x % rows_per_plane) * (2 * XR) + 2 * w;
  const size_t plane = (size_t)N * g.kxp;
  cx<T> *wb = buf + w * N;

  cx<T> f[6][E];   // S_xx, S_yy, S_zz, S_xy, S_xz, S_yz at x = lane + LPR q: (row a, row b)
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const cx<T> *src = rin + (size_t)c * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
    for (int e = lane; e < N; e += LPR) wb[e] = czero<T>();
    pass_sync<WAVE>();
    for (int kx = lane; kx <= M; kx += LPR) {
      cx<T> a = src[kx], b = src[g.kxp + kx];
      if (kx == 0) {   // the self-conjugate coefficient: only its real part is seen
        a.im = T(0);
        b.im = T(0);
      }
      wb[kx] = {a.re - b.im, a.im + b.re};
      if (kx > 0) wb[N - kx] = {a.re + b.im, b.re - a.im};
    }
    pass_sync<WAVE>();
    fft_columns<T, N, 1, LPR, +1, false, WAVE>(wb, twl);
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (lane + LPR * q < N) f[c][q] = wb[lane + LPR * q];
    pass_sync<WAVE>();
  }
  auto det = [](T a, T b, T c, T d, T e, T f_) { return a * (b * c - f_ * f_) - d * (d * c - e * f_) + e * (d * f_ - b * e); };
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    if (lane + LPR * q >= N) continue;
    acc += (double)det(f[0][q].re, f[1][q].re, f[2][q].re, f[3][q].re, f[4][q].re, f[5][q].re);
    acc += (double)det(f[0][q].im, f[1][q].im, f[2][q].im, f[3][q].im, f[4][q].im, f[5][q].im);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < XT / 64; ++i) s += red[i];
    partials[blockIdx.x] = s;
  }
}

}  // namespace hdns

namespace hdns {

// kinetic_energy_and_dissipation_below_scale (src/Observables.jl:195-264): for every side
// l = n L / N, the retained sums sum w W_l^2 |u|^2 and sum w |k|^2 W_l^2 |u|^2 of the box-mean
// window W_l(k) = prod_a sinc(k_a l / 2).  t[(n - 1) (M + 1) + |m|] = sinc^2(pi |m| n / N) is the
// per-axis factor; each block takes BS_SIDES sides, so one read of a mode serves all of them.
// partials: [side][blocks][2], FP64.
constexpr int BS_SIDES = 8;
constexpr int BS_BLOCKS = 128;

template <typename T>
__global__ __launch_bounds__(256) void k_below_scale(const cx<T> *__restrict__ uh, Geo g, const double *__restrict__ t,
                                                     int nsides, double *__restrict__ partials) {
  __shared__ double red[4][BS_SIDES][2];
  const size_t comp = comp_of(g);
  const int s0 = blockIdx.y * BS_SIDES;
  const size_t modes = (size_t)g.k * g.k * (g.m + 1);
  double acc[BS_SIDES][2];
#pragma unroll
  for (int i = 0; i < BS_SIDES; ++i) acc[i][0] = acc[i][1] = 0.0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < modes; l += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(l % (g.m + 1));
    const size_t row = l / (g.m + 1);
    const int my = signed_of((int)(row % g.k), g), mz = signed_of((int)(row / g.k), g);
    const size_t at = row * g.kxp + kx;
    double e = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const cx<T> v = uh[c * comp + at];
      e += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
    }
    if (e == 0.0) continue;
    const double w = parseval_w(kx, g.n);
    const double kxd = g.k0 * kx, kyd = g.k0 * my, kzd = g.k0 * mz;
    const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
    const int ay = my < 0 ? -my : my, az = mz < 0 ? -mz : mz;
#pragma unroll
    for (int i = 0; i < BS_SIDES; ++i) {
      if (s0 + i >= nsides) break;
      const double *ts = t + (size_t)(s0 + i) * (g.m + 1);
      const double w2 = ts[kx] * ts[ay] * ts[az];
      acc[i][0] += w * w2 * e;
      acc[i][1] += w * k2d * w2 * e;
    }
  }
#pragma unroll
  for (int i = 0; i < BS_SIDES; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      double v = acc[i][j];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i][j] = v;
    }
  __syncthreads();
  if (threadIdx.x < 2 * BS_SIDES) {
    const int i = threadIdx.x >> 1, j = threadIdx.x & 1;
    if (s0 + i < nsides) {
      double v = 0.0;
      for (int wv = 0; wv < 4; ++wv) v += red[wv][i][j];
      partials[((size_t)(s0 + i) * gridDim.x + blockIdx.x) * 2 + j] = v;
    }
  }
}

}  // namespace hdns
