// FP32 z passes on column pairs.  An element of these transforms is two adjacent complex FP32
// values of a tile row -- columns 2i and 2i+1, 16 bytes, the size of one FP64 complex value -- and
// every butterfly acts on both columns at once: the same passes, twiddles and LDS layout as the
// FP64 transforms, with each real operation done twice in FP32.  A thread then moves 16 bytes per
// access (as an FP64 transform does) and the tile's latency chain carries twice the bytes of the
// one-column-per-element FP32 transform.
#pragma once
#include "spectral.hpp"

namespace hdns {

// Two FP32 lanes of one real quantity (column 2i in x, 2i+1 in y); each operation is one FP32
// operation per lane, in the order written.
struct f2 {
  float x, y;
  f2() = default;
  __host__ __device__ constexpr explicit f2(double v) : x(float(v)), y(float(v)) {}
  __host__ __device__ constexpr f2(float a, float b) : x(a), y(b) {}
};
__device__ __forceinline__ f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ f2 operator-(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ f2 operator*(f2 a, f2 b) { return {a.x * b.x, a.y * b.y}; }
__device__ __forceinline__ f2 operator-(f2 a) { return {-a.x, -a.y}; }

// the pair at p (two adjacent complex values, 16-byte aligned): HBM holds re0 im0 re1 im1
__device__ __forceinline__ cx<f2> load_pair(const cx<float> *p) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  return {f2(v.x, v.z), f2(v.y, v.w)};
}
__device__ __forceinline__ void store_pair(cx<float> *p, cx<f2> v) {
  *reinterpret_cast<float4 *>(p) = make_float4(v.re.x, v.im.x, v.re.y, v.im.y);
}
__device__ __forceinline__ cx<float> half_of(const cx<f2> &v, int h) {
  return h ? cx<float>{v.re.y, v.im.y} : cx<float>{v.re.x, v.im.x};
}

// The pair z passes' tile: TXP pairs (2 TXP columns; 16 columns, 128-byte rows, up to N = 512,
// 8 columns at N = 1024), on as many threads as the first pass has butterflies, in whole
// wavefronts, at most 512.
template <int N>
struct PairShape {
  static constexpr int TXP = N >= 1024 ? 4 : 8;
  static constexpr int B1 = (N / first_radix<N>()) * TXP;   // first-pass butterflies of a tile
  static constexpr int NT = B1 >= 512 ? 512 : (B1 + 63) / 64 * 64;
  static constexpr size_t LDS = ((size_t)N * TXP + N) * sizeof(cx<f2>);   // tile + twiddles
};

// The inverse z pass of the state (k_zi) on column pairs: u, w = i k x u and phi, with the 1/N^3
// of the inverse folded in, into Q, and the per-block FP64 partial sums in the same order of
// modes per thread as k_zi takes them (each thread owns both columns of its pair).  Modes outside
// the ball read as the zeros the state holds there.  2/3 rule only.
template <int N, bool SCALAR>
__global__ __launch_bounds__(PairShape<N>::NT) void k_zi_pairs(const cx<float> *__restrict__ uh,
                                                               const cx<float> *__restrict__ ph,
                                                               cx<float> *__restrict__ q, Geo g,
                                                               const cx<double> *__restrict__ tw, BandSpec band,
                                                               double *__restrict__ partials) {
  using S = PairShape<N>;
  constexpr int TXP = S::TXP, NT = S::NT, TX = 2 * TXP;
  constexpr int K = band_k(N, DEALIAS_TWO_THIRDS);
  constexpr int R1 = first_radix<N>();
  using S1 = PassShape<N, R1, TXP, NT>;
  constexpr int PER = S1::PER;
  constexpr int NCIN = SCALAR ? 7 : 6;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<f2> *buf = reinterpret_cast<cx<f2> *>(smem_raw);
  cx<f2> *twl = buf + N * TXP;
  __shared__ double red[NT / 64][NRED];
  const int tid = threadIdx.x;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, box = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  const size_t part = ((size_t)box * gridDim.y * gridDim.x + (size_t)blockIdx.y * gridDim.x + blockIdx.x) * NRED;
  if (!in_disk(line_start<float>(kx0), my, g)) {   // the tile's line misses the ball: never read
    if (tid < NRED) partials[part + tid] = 0.0;
    return;
  }
  const size_t comp = (size_t)K * K * g.kxp;
  const size_t pencil = (size_t)N * K * g.kxp;
  const size_t zstride = (size_t)K * g.kxp;
  uh += box * 3 * comp;
  if constexpr (SCALAR) ph += box * comp;
  q += box * g.box_a;
  stage_twiddles<f2, N, NT>(twl, tw);
  const float inv = float(1.0 / ((double)N * N * N));

  cx<f2> u[3][PER][R1];
  cx<f2> p[PER][R1];
  double acc[NRED] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int pp = 0; pp < PER; ++pp) {
    const int bb = tid + pp * NT;
    const bool live = S1::TOT % NT == 0 || bb < S1::TOT;
    const int cp = bb % TXP, j = bb / TXP;
    const int mx0 = kx0 + 2 * cp;
#pragma unroll
    for (int qq = 0; qq < R1; ++qq) {
      const int kz_idx = live ? idx_of_bin(j + qq * S1::NB, g) : -1;
      if (kz_idx < 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c][pp][qq] = czero<f2>();
        p[pp][qq] = czero<f2>();
        continue;
      }
      const int mz = signed_of(kz_idx, g);
      const size_t at = (size_t)kz_idx * zstride + (size_t)ky_idx * g.kxp + mx0;
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c][pp][qq] = load_pair(uh + c * comp + at);
      p[pp][qq] = SCALAR ? load_pair(ph + at) : czero<f2>();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int mx = mx0 + h;
        if (!keeps(mx * mx + my * my + mz * mz, g)) continue;   // zero outside the ball
        const double kxd = g.k0 * mx, kyd = g.k0 * my, kzd = g.k0 * mz;
        const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
        const double w = parseval_w(mx, N);
        double e2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const cx<float> v = half_of(u[c][pp][qq], h);
          e2 += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
        }
        acc[0] += w * e2;
        if (in_band(mx * mx + my * my + mz * mz, band)) acc[1] += w * e2;
        acc[2] += w * k2d * e2;
        if constexpr (SCALAR) {
          const cx<float> v = half_of(p[pp][qq], h);
          const double v2 = (double)v.re * (double)v.re + (double)v.im * (double)v.im;
          acc[3] += w * v2;
          acc[4] += w * k2d * v2;
        }
      }
    }
  }

#pragma unroll
  for (int c = 0; c < NCIN; ++c) {
    cx<float> *qc = q + (size_t)c * pencil + (size_t)ky_idx * g.kxp + kx0;
    fft_io<f2, N, TXP, NT, +1, true, false>(
        buf, twl,
        [&](int pp, int qq, int n, int cp) {
          cx<f2> v;
          if (c < 3) {
            v = u[c][pp][qq];
          } else if (c < 6) {  // i k x u
            const int mx0 = kx0 + 2 * cp, mz = signed_of(idx_of_bin(n, g), g);
            const f2 kx(float(g.k0 * mx0), float(g.k0 * (mx0 + 1)));
            const f2 ky(g.k0 * my), kz(g.k0 * mz);
            cx<f2> t;
            if (c == 3) t = scale(u[2][pp][qq], ky) - scale(u[1][pp][qq], kz);
            else if (c == 4) t = scale(u[0][pp][qq], kz) - scale(u[2][pp][qq], kx);
            else t = scale(u[1][pp][qq], kx) - scale(u[0][pp][qq], ky);
            v = {-t.im, t.re};
          } else {
            v = p[pp][qq];
          }
          return scale(v, f2(inv, inv));
        },
        [&](int, int, int n, int cp, cx<f2> v) { store_pair(qc + (size_t)n * zstride + 2 * cp, v); });
  }

  // deterministic block reduction: wave shuffles, then the waves in order
#pragma unroll
  for (int i = 0; i < NRED; ++i) {
    double v = acc[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((tid & 63) == 0) red[tid >> 6][i] = v;
  }
  __syncthreads();
  if (tid < NRED) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; ++w) s += red[w][tid];
    partials[part + tid] = s;
  }
}

}  // namespace hdns
