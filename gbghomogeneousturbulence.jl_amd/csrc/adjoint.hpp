// Kernels of the adjoint step: the transpose of one IF-RK step's derivative, stage by stage.
//
// Per stage (reverse order), with lam_q the adjoint of the stage's output state and lam_h that
// of its output register (zero after the last stage):
//   k_adj_pre    lam_h1 = b E lam_q + E lam_h;  mu = dt lam_h1;
//                nu = M P Z mu_u and xi = i k M mu_phi (the spectra of the physical adjoints m, s);
//                partial sums of <mu_u, 1_band u_i>_w for the forcing's rank-one term
//   transforms   m, s to physical (plain inverse z pass, inverse y pass, x pass) and, in the same
//                x pass, the ten products w x m, m x u, phi s, s . u with the stage's recorded
//                physical u, w, phi; forward x, forward y, plain forward z (scaled by 1/N^3)
//   k_adj_post   g_u = F(w x m) + i k x F(m x u) + F(phi s) + A 1_b mu_u
//                      - (A / e_b) <mu_u, 1_b u_i>_w/N^6 1_b u_i - G mu_phi z
//                g_phi = F(s . u);   lam_q <- E lam_q + g;   lam_h <- a lam_h1
// (the derivation, with the transposes of each operation, is stated with the CPU restatement the
// tests hold these kernels to).  The state arrays hold [u_x, u_y, u_z, phi] as four compact
// components; every kernel writes zeros outside the 2/3 ball.
#pragma once
#include "spectral.hpp"

namespace hdns {

struct AdjStage {
  double a, b;          // the stage's coefficients, rounded to T
  double dt;            // dt rounded to T
  double sub;           // d_i dt
  double visc, kappa;   // nu, nu/Sc
  double mean_gradient;
  int last;             // the stage has no output register
  int scalar;           // phi is carried
  BandSpec band;
  const double *scal;   // the stage's recorded sums: [0] A, [2] <u, 1_b u>_w / N^6
  const double *proj;   // [0] <mu_u, 1_b u_i>_w / N^6, from k_adj_pre's partial sums
};

// the mode of compact element l: (kz_idx, ky_idx, mx), its signed indices and Parseval weight
struct Mode {
  int mx, my, mz, m2;
  double k2;
  bool in;
};

__device__ __forceinline__ Mode mode_of(size_t l, const Geo &g) {
  Mode m;
  m.mx = (int)(l % g.kxp);
  const int ky_idx = (int)((l / g.kxp) % g.k), kz_idx = (int)(l / ((size_t)g.kxp * g.k));
  m.my = signed_of(ky_idx, g);
  m.mz = signed_of(kz_idx, g);
  m.m2 = m.mx * m.mx + m.my * m.my + m.mz * m.mz;
  const double kx = g.k0 * m.mx, ky = g.k0 * m.my, kz = g.k0 * m.mz;
  m.k2 = (kx * kx + ky * ky) + kz * kz;
  m.in = m.mx <= g.m && keeps(m.m2, g);
  return m;
}

template <typename T>
__device__ __forceinline__ void leray3(cx<T> *v, const Mode &m, const Geo &g) {
  if (m.m2 == 0) return;   // P is the identity at k = 0
  const T kx = T(g.k0 * m.mx), ky = T(g.k0 * m.my), kz = T(g.k0 * m.mz), k2 = T(m.k2);
  const cx<T> s = {(kx * v[0].re + ky * v[1].re + kz * v[2].re) / k2,
                   (kx * v[0].im + ky * v[1].im + kz * v[2].im) / k2};
  v[0] = v[0] - scale(s, kx);
  v[1] = v[1] - scale(s, ky);
  v[2] = v[2] - scale(s, kz);
}

constexpr int ADJ_BLOCKS = 1024;   // grid of the pointwise adjoint kernels (fixed: partial sums)

// lam_h1 and the spectra of the physical adjoints (t6 = [nu_x, nu_y, nu_z, xi_x, xi_y, xi_z])
template <typename T>
__global__ __launch_bounds__(256) void k_adj_pre(const cx<T> *__restrict__ lq, cx<T> *__restrict__ lh,
                                                 const cx<T> *__restrict__ qi, cx<T> *__restrict__ t6, Geo g,
                                                 AdjStage as, double *__restrict__ partials) {
  const size_t comp = comp_of(g);
  __shared__ double red[4];
  double acc = 0.0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < comp; l += (size_t)gridDim.x * blockDim.x) {
    const Mode m = mode_of(l, g);
    const int nc = as.scalar ? 4 : 3;
    cx<T> h1[4], nu[3];
    const T Eu = T(exp(-as.visc * as.sub * m.k2)), Ep = T(exp(-as.kappa * as.sub * m.k2));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= nc) {
        h1[c] = czero<T>();
        continue;
      }
      const T E = c < 3 ? Eu : Ep;
      cx<T> v = scale(lq[c * comp + l], T(as.b) * E);
      if (!as.last) v = v + scale(lh[c * comp + l], E);
      h1[c] = m.in ? v : czero<T>();
      lh[c * comp + l] = h1[c];
    }
    const T dt = T(as.dt);
    // nu = M P Z mu_u
#pragma unroll
    for (int c = 0; c < 3; ++c) nu[c] = (m.in && m.m2 > 0) ? scale(h1[c], dt) : czero<T>();
    leray3(nu, m, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) t6[c * comp + l] = nu[c];
    // xi = i k M mu_phi
    const cx<T> mp = scale(h1[3], dt);
    const T kx = T(g.k0 * m.mx), ky = T(g.k0 * m.my), kz = T(g.k0 * m.mz);
    t6[3 * comp + l] = m.in ? cx<T>{-kx * mp.im, kx * mp.re} : czero<T>();
    t6[4 * comp + l] = m.in ? cx<T>{-ky * mp.im, ky * mp.re} : czero<T>();
    t6[5 * comp + l] = m.in ? cx<T>{-kz * mp.im, kz * mp.re} : czero<T>();
    // <mu_u, 1_band u_i>_w
    if (m.in && in_band(m.m2, as.band)) {
      const double w = parseval_w(m.mx, g.n);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cx<T> u = qi[c * comp + l];
        const cx<T> mu = scale(h1[c], dt);
        acc += w * ((double)mu.re * (double)u.re + (double)mu.im * (double)u.im);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// the sum of the partials (one block, Neumaier), times a scale
static __global__ void k_adj_sum(const double *__restrict__ partials, int n, double scale_by, double *__restrict__ out) {
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double x = partials[i], t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  ss[threadIdx.x] = s;
  cc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int t = 0; t < 256; ++t) {
      const double x = ss[t], u = S + x;
      C += fabs(S) >= fabs(x) ? (S - u) + x : (x - u) + S;
      S = u;
      C += cc[t];
    }
    out[0] = (S + C) * scale_by;
  }
}

// lam_q <- E lam_q + g;  lam_h <- a lam_h1  (t10 = [F(w x m) 3, F(m x u) 3, F(phi s) 3, F(s . u)],
// already divided by N^3)
template <typename T>
__global__ __launch_bounds__(256) void k_adj_post(cx<T> *__restrict__ lq, cx<T> *__restrict__ lh,
                                                  const cx<T> *__restrict__ qi, const cx<T> *__restrict__ t10, Geo g,
                                                  AdjStage as) {
  const size_t comp = comp_of(g);
  const double A = as.band.on ? as.scal[0] : 0.0;
  const double eb = 0.5 * as.scal[2];
  const double rank1 = (A != 0.0 && eb != 0.0) ? (A / eb) * as.proj[0] : 0.0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < comp; l += (size_t)gridDim.x * blockDim.x) {
    const Mode m = mode_of(l, g);
    const int nc = as.scalar ? 4 : 3;
    if (!m.in) {
      for (int c = 0; c < nc; ++c) {
        lq[c * comp + l] = czero<T>();
        lh[c * comp + l] = czero<T>();
      }
      continue;
    }
    const T Eu = T(exp(-as.visc * as.sub * m.k2)), Ep = T(exp(-as.kappa * as.sub * m.k2));
    const T dt = T(as.dt), a = T(as.a);
    const T kx = T(g.k0 * m.mx), ky = T(g.k0 * m.my), kz = T(g.k0 * m.mz);
    cx<T> gu[3], b3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gu[c] = t10[c * comp + l] + t10[(6 + c) * comp + l];   // F(w x m) + F(phi s)
      b3[c] = t10[(3 + c) * comp + l];
    }
    // + i k x F(m x u)
    const cx<T> cx_ = {ky * b3[2].re - kz * b3[1].re, ky * b3[2].im - kz * b3[1].im};
    const cx<T> cy_ = {kz * b3[0].re - kx * b3[2].re, kz * b3[0].im - kx * b3[2].im};
    const cx<T> cz_ = {kx * b3[1].re - ky * b3[0].re, kx * b3[1].im - ky * b3[0].im};
    gu[0] = gu[0] + cx<T>{-cx_.im, cx_.re};
    gu[1] = gu[1] + cx<T>{-cy_.im, cy_.re};
    gu[2] = gu[2] + cx<T>{-cz_.im, cz_.re};
    if (in_band(m.m2, as.band)) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cx<T> mu = scale(lh[c * comp + l], dt);
        gu[c] = gu[c] + scale(mu, T(A)) - scale(qi[c * comp + l], T(rank1));
      }
    }
    if (as.scalar) gu[2] = gu[2] - scale(scale(lh[3 * comp + l], dt), T(as.mean_gradient));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lq[c * comp + l] = scale(lq[c * comp + l], Eu) + gu[c];
      lh[c * comp + l] = scale(lh[c * comp + l], a);
    }
    if (as.scalar) {
      lq[3 * comp + l] = scale(lq[3 * comp + l], Ep) + t10[9 * comp + l];
      lh[3 * comp + l] = scale(lh[3 * comp + l], a);
    }
  }
}

// the adjoint of J = k + <phi^2>: lam_u = u/N^6, lam_phi = 2 phi/N^6; lam_h = 0
template <typename T>
__global__ __launch_bounds__(256) void k_adj_objective(cx<T> *__restrict__ lq, cx<T> *__restrict__ lh,
                                                       const cx<T> *__restrict__ uh, const cx<T> *__restrict__ ph,
                                                       Geo g, int scalar) {
  const size_t comp = comp_of(g);
  const T inv6 = T(1.0 / pow((double)g.n, 6.0));
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < comp; l += (size_t)gridDim.x * blockDim.x) {
    const Mode m = mode_of(l, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lq[c * comp + l] = m.in ? scale(uh[c * comp + l], inv6) : czero<T>();
      lh[c * comp + l] = czero<T>();
    }
    if (scalar) {
      lq[3 * comp + l] = m.in ? scale(ph[l], T(2) * inv6) : czero<T>();
      lh[3 * comp + l] = czero<T>();
    }
  }
}

// the Leray projection of lam_u in place (the transpose of the state projection), masked
template <typename T>
__global__ __launch_bounds__(256) void k_adj_leray(cx<T> *__restrict__ lq, Geo g) {
  const size_t comp = comp_of(g);
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < comp; l += (size_t)gridDim.x * blockDim.x) {
    const Mode m = mode_of(l, g);
    cx<T> v[3] = {lq[l], lq[comp + l], lq[2 * comp + l]};
    leray3(v, m, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) lq[c * comp + l] = m.in ? v[c] : czero<T>();
  }
}

// Plain inverse z transform: compact components -> z pencils Q [c][z][ky][kx] (grid: tiles, K, nc).
// Tiles whose line misses the ball are skipped, as the inverse y pass never reads them.
template <typename T, int N, int TX, int NT>
__global__ __launch_bounds__(NT) void k_zinv_plain(const cx<T> *__restrict__ src, cx<T> *__restrict__ q, Geo g,
                                                   const cx<double> *__restrict__ tw, T scale_by) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, c = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  if (!in_disk(line_start<T>(kx0), my, g)) return;
  stage_twiddles<T, N, NT>(twl, tw);
  const size_t comp = comp_of(g), pencil = (size_t)N * g.k * g.kxp, zstride = (size_t)g.k * g.kxp;
  const cx<T> *s = src + c * comp + (size_t)ky_idx * g.kxp + kx0;
  cx<T> *dst = q + c * pencil + (size_t)ky_idx * g.kxp + kx0;
  fft_io<T, N, TX, NT, +1, true, false>(
      buf, twl,
      [&](int, int, int n, int col) {
        const int kz_idx = idx_of_bin(n, g);
        if (kz_idx < 0) return czero<T>();
        const int mx = kx0 + col, mz = signed_of(kz_idx, g);
        if (mx > g.m || !keeps(mx * mx + my * my + mz * mz, g)) return czero<T>();
        return scale(s[(size_t)kz_idx * zstride + col], scale_by);
      },
      [&](int, int, int n, int col, cx<T> v) { dst[(size_t)n * zstride + col] = v; });
}

// Plain forward z transform: z pencils PF [c][z][ky][kx] -> compact components, scaled, zero
// outside the ball (grid: tiles, K, nc).  A tile outside the disk (its PF lines unwritten) writes
// zeros.
template <typename T, int N, int TX, int NT>
__global__ __launch_bounds__(NT) void k_zfwd_plain(const cx<T> *__restrict__ pf, cx<T> *__restrict__ dst, Geo g,
                                                   const cx<double> *__restrict__ tw, T scale_by) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, c = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  const size_t comp = comp_of(g), pencil = (size_t)N * g.k * g.kxp, zstride = (size_t)g.k * g.kxp;
  cx<T> *d = dst + c * comp + (size_t)ky_idx * g.kxp + kx0;
  if (!in_disk(kx0, my, g)) {
    for (int e = threadIdx.x; e < g.k * TX; e += NT) d[(size_t)(e / TX) * zstride + e % TX] = czero<T>();
    return;
  }
  stage_twiddles<T, N, NT>(twl, tw);
  const cx<T> *s = pf + c * pencil + (size_t)ky_idx * g.kxp + kx0;
  fft_io<T, N, TX, NT, -1, true, false>(
      buf, twl, [&](int, int, int n, int col) { return s[(size_t)n * zstride + col]; },
      [&](int, int, int n, int col, cx<T> v) {
        const int kz_idx = idx_of_bin(n, g);
        if (kz_idx < 0) return;
        const int mx = kx0 + col, mz = signed_of(kz_idx, g);
        const bool in = mx <= g.m && keeps(mx * mx + my * my + mz * mz, g);
        d[(size_t)kz_idx * zstride + col] = in ? scale(v, scale_by) : czero<T>();
      });
}

// The adjoint x pass: per pair of y rows, the inverse x transforms of m (R components 0..2) and
// s (3..5), the recorded physical u, w, phi of the stage, the ten products and their forward x
// transforms into PX (components: w x m, m x u, phi s, s . u).  The layout and the two-rows-as-
// one-complex packing are those of the forward x pass.
template <typename T, int N, int XT>
__global__ __launch_bounds__(XT) void k_xadj(const cx<T> *__restrict__ rin, const T *__restrict__ phys,
                                             const T *__restrict__ physw, cx<T> *__restrict__ pxo, Geo g,
                                             const cx<double> *__restrict__ tw, int scalar) {
  constexpr int LPR = xpass_lanes<T, N, XT>();
  constexpr bool WAVE = LPR == 64;
  constexpr int XR = XT / LPR;
  constexpr int E = (N + LPR - 1) / LPR;
  constexpr int M = N / 3;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + XR * N;
  stage_twiddles<T, N, XT>(twl, tw);
  __syncthreads();
  const int rows_per_plane = N / (2 * XR);
  const int z = blockIdx.x / rows_per_plane;
  const int w = threadIdx.x / LPR, lane = threadIdx.x % LPR;
  const int ya = (blockIdx.x % rows_per_plane) * (2 * XR) + 2 * w;
  const size_t plane = (size_t)N * g.kxp;
  const size_t vol = (size_t)N * N * N;
  const size_t rowa = ((size_t)z * N + ya) * N, rowb = rowa + N;
  cx<T> *wb = buf + w * N;

  cx<T> f[6][E];   // m (0..2) and s (3..5) at x = lane + LPR q: (row a, row b)
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const cx<T> *src = rin + (size_t)c * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
    for (int e = lane; e < N; e += LPR) wb[e] = czero<T>();
    pass_sync<WAVE>();
    for (int kx = lane; kx <= M; kx += LPR) {
      cx<T> a = src[kx], b = src[g.kxp + kx];
      if (kx == 0) {
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

#pragma unroll
  for (int d = 0; d < 10; ++d) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int x = lane + LPR * q;
      if (x >= N) continue;
      auto ld = [&](const T *p, int c) -> cx<T> { return {p[c * vol + rowa + x], p[c * vol + rowb + x]}; };
      cx<T> v = czero<T>();
      if (d < 3) {         // w x m
        const int c1 = (d + 1) % 3, c2 = (d + 2) % 3;
        const cx<T> w1 = ld(physw, c1), w2 = ld(physw, c2);
        v = {w1.re * f[c2][q].re - w2.re * f[c1][q].re, w1.im * f[c2][q].im - w2.im * f[c1][q].im};
      } else if (d < 6) {  // m x u
        const int e = d - 3, c1 = (e + 1) % 3, c2 = (e + 2) % 3;
        const cx<T> u1 = ld(phys, c1), u2 = ld(phys, c2);
        v = {f[c1][q].re * u2.re - f[c2][q].re * u1.re, f[c1][q].im * u2.im - f[c2][q].im * u1.im};
      } else if (scalar) {
        if (d < 9) {       // phi s
          const cx<T> ph = ld(phys, 3);
          v = {ph.re * f[d - 3][q].re, ph.im * f[d - 3][q].im};
        } else {           // s . u
          const cx<T> u0 = ld(phys, 0), u1 = ld(phys, 1), u2 = ld(phys, 2);
          v = {f[3][q].re * u0.re + f[4][q].re * u1.re + f[5][q].re * u2.re,
               f[3][q].im * u0.im + f[4][q].im * u1.im + f[5][q].im * u2.im};
        }
      }
      wb[x] = v;
    }
    pass_sync<WAVE>();
    fft_columns<T, N, 1, LPR, -1, false, WAVE>(wb, twl);
    cx<T> *dst = pxo + (size_t)d * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
    for (int kx = lane; kx < g.kxp; kx += LPR) {
      cx<T> A = czero<T>(), B = czero<T>();
      if (kx <= M) {
        const cx<T> C = wb[kx];
        const cx<T> Cm = wb[kx == 0 ? 0 : N - kx];
        A = {T(0.5) * (C.re + Cm.re), T(0.5) * (C.im - Cm.im)};
        B = {T(0.5) * (C.im + Cm.im), T(0.5) * (Cm.re - C.re)};
      }
      dst[kx] = A;
      dst[g.kxp + kx] = B;
    }
    pass_sync<WAVE>();
  }
}

}  // namespace hdns
