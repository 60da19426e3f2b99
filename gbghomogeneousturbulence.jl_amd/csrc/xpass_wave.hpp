// The x pass with one wavefront per row pair (or per row) and the data in registers: FP32 at
// N = 256 and 512 (pairs) and 1024 (rows), FP64 at N = 256 (pairs) and 512 (rows).
//
// A wavefront transforms H complex points, H = 512 (FP32 from N = 512 on) or 256, S = H / 64 per
// lane: three radix-8 or four radix-4 Stockham passes in which lane j is butterfly j of every
// pass: it reads elements j + 64 q (q = 0..S-1) and, after the last pass, holds the outputs
// j + 64 q.  So the inverse transforms leave lane j with the physical values at x = j + 64 q of
// every component, the products are formed lane-locally, and they feed the forward transforms'
// first pass directly.  Only the exchanges between passes go through LDS (a wave-private, padded
// buffer).  The per-lane twiddles of the last pass are held in registers for the whole kernel;
// the earlier passes' come from a 64-entry table of W_64 in LDS.  A component costs S complex
// values per lane: 16 registers (8 for FP32 at N = 256).
//
// With H = N a wavefront takes a row pair: the Hermitian pack of two real rows (row a in the real
// part, row b in the imaginary part) is read into the first pass: element n of the packed
// spectrum is built from (a, b) at kx = n (n <= M) or kx = N - n (n >= N - M), and is zero in
// between.  The forward transforms' outputs are unpacked with the conjugate partner C[N - kx],
// which lives in lane (64 - j) mod 64 and is fetched with a cross-lane permute.
//
// With N = 2 H a wavefront takes one real row, as the H-point complex sequence
// z[n] = x[2n] + i x[2n+1].  With W = exp(-2 pi i / N):
//   inverse:  Z[m] = (X[m] + conj X[H-m]) + i W^-m (X[m] - conj X[H-m]),  z = IFFT_H(Z);
//   forward:  Z = FFT_H(z),  X[k] = (Z[k] + conj Z[H-k]) / 2 + W^k (Z[k] - conj Z[H-k]) / (2 i).
// Lane j then holds the physical pairs x[2n], x[2n+1] at n = j + 64 q, and the products stay
// lane-local exactly as for a row pair.
//
// The components are streamed so that at most u, two partial Lamb components and the component
// in flight are live: u, then phi and the three fluxes u phi, then w_x, w_y, w_z, each adding
// its terms of u x w, each Lamb component transformed as soon as it is complete.
#pragma once
#include "pencil.hpp"

namespace hdns {

// The wave-private exchange buffer's padding: one pad slot per 8 (FP32) or 4 (FP64) values
// spreads a pass's strided accesses over the LDS banks.
template <typename T>
constexpr int xpad_shift() { return sizeof(T) == 4 ? 3 : 2; }
template <typename T>
__device__ __forceinline__ int xpad(int i) { return i + (i >> xpad_shift<T>()); }

// The shape of the register-resident x pass for (T, N).
template <typename T, int N>
struct XWave {
  // the wavefront's complex transform length: 512 FP32 points from N = 512 on, else 256
  static constexpr int H = sizeof(T) == 4 && N >= 512 ? 512 : 256;
  static constexpr int S = H / 64;                       // values per lane per component
  static constexpr bool SUPPORTED = N == H || N == 2 * H;
  static constexpr bool HALF = N == 2 * H;               // one real row per wavefront (else a pair)
  static constexpr int RW = HALF ? 1 : 2;                // rows per wavefront
  static constexpr int M = N / 3;
  static constexpr int TXW = sizeof(T) == 4 ? 16 : 8;    // the tile width that pads every row
  static constexpr int KXP = (M + 1 + TXW - 1) / TXW * TXW;
  static constexpr int QF = (KXP + 63) / 64;             // slots holding stored outputs kx < KXP
  static constexpr int NP = H + (H >> xpad_shift<T>());  // the padded exchange buffer
};

template <int SIGN, typename T>
__device__ __forceinline__ cx<T> twiddled(cx<T> v, cx<T> w) {
  return v * cx<T>{w.re, SIGN < 0 ? w.im : -w.im};
}

// x: element j + 64 q of the input in, element j + 64 q of the transform out (lane j).
// t64[m] = W_64^m, t3[q - 1] = W_512^{q j}.
template <int SIGN, typename T>
__device__ __forceinline__ void wave_fft512(cx<T> (&x)[8], cx<T> *wb, const cx<T> *t64, const cx<T> (&t3)[7],
                                            int j) {
  dft8<SIGN>(x);                                     // NS = 1: outputs 8 j + q
#pragma unroll
  for (int q = 0; q < 8; ++q) wb[xpad<T>(8 * j + q)] = x[q];
  pass_sync<true>();
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = wb[xpad<T>(j + 64 * q)];
  pass_sync<true>();
#pragma unroll
  for (int q = 1; q < 8; ++q) x[q] = twiddled<SIGN>(x[q], t64[q * (j & 7)]);
  dft8<SIGN>(x);                                     // NS = 8: outputs 64 (j/8) + j mod 8 + 8 q
  const int base = (j >> 3) * 64 + (j & 7);
#pragma unroll
  for (int q = 0; q < 8; ++q) wb[xpad<T>(base + 8 * q)] = x[q];
  pass_sync<true>();
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = wb[xpad<T>(j + 64 * q)];
  pass_sync<true>();                                 // the buffer may be rewritten at once
#pragma unroll
  for (int q = 1; q < 8; ++q) x[q] = twiddled<SIGN>(x[q], t3[q - 1]);
  dft8<SIGN>(x);                                     // NS = 64: outputs j + 64 q
}

// The 256-point transform in four radix-4 passes: t64[m] = W_64^m (W_16^m = t64[4 m]),
// t3[q - 1] = W_256^{q j}.
template <int SIGN, typename T>
__device__ __forceinline__ void wave_fft256(cx<T> (&x)[4], cx<T> *wb, const cx<T> *t64, const cx<T> (&t3)[3],
                                            int j) {
  auto exchange = [&](int base, int stride) {        // outputs base + stride q -> inputs j + 64 q
#pragma unroll
    for (int q = 0; q < 4; ++q) wb[xpad<T>(base + stride * q)] = x[q];
    pass_sync<true>();
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = wb[xpad<T>(j + 64 * q)];
    pass_sync<true>();
  };
  dft4<SIGN>(x);                                     // NS = 1: outputs 4 j + q
  exchange(4 * j, 1);
#pragma unroll
  for (int q = 1; q < 4; ++q) x[q] = twiddled<SIGN>(x[q], t64[4 * q * (j & 3)]);
  dft4<SIGN>(x);                                     // NS = 4: outputs 16 (j/4) + j mod 4 + 4 q
  exchange((j >> 2) * 16 + (j & 3), 4);
#pragma unroll
  for (int q = 1; q < 4; ++q) x[q] = twiddled<SIGN>(x[q], t64[q * (j & 15)]);
  dft4<SIGN>(x);                                     // NS = 16: outputs 64 (j/16) + j mod 16 + 16 q
  exchange((j >> 4) * 64 + (j & 15), 16);
#pragma unroll
  for (int q = 1; q < 4; ++q) x[q] = twiddled<SIGN>(x[q], t3[q - 1]);
  dft4<SIGN>(x);                                     // NS = 64: outputs j + 64 q
}

template <int SIGN, typename T, int S>
__device__ __forceinline__ void wave_fft(cx<T> (&x)[S], cx<T> *wb, const cx<T> *t64, const cx<T> (&t3)[S - 1],
                                         int j) {
  if constexpr (S == 8) wave_fft512<SIGN>(x, wb, t64, t3, j);
  else wave_fft256<SIGN>(x, wb, t64, t3, j);
}

// the value v of lane `lane` (all lanes active)
__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(lane << 2, (int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(lane << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// How the coefficients of each component reach the wavefront:
//   XFETCH_JIT: loaded into registers just before the component's transform (the wavefront waits
//               for HBM, the other waves of the SIMD fill the gap);
//   XFETCH_LDS: the wavefront's rows of the next component (RW * kxp coefficients, contiguous in
//               R) are copied HBM -> LDS by direct-to-LDS loads while the current component
//               transforms; the copy holds no registers, and the first pass reads the staged rows.
enum { XFETCH_JIT = 0, XFETCH_LDS = 1 };

// IL: the deposit's variant for runs with tracers, which stores u and phi interleaved
// ((u_x, u_y, u_z, phi) per point); the stage passes and runs without tracers use IL = false.
template <typename T, int N, int NCIN, int XT, int FETCH, bool IL = false>
__global__ __launch_bounds__(XT, 3) void k_xpass_wave(const cx<T> *__restrict__ rin, cx<T> *__restrict__ pxo, Geo g,
                                                      const cx<double> *__restrict__ tw, XArgs xa) {
  using X = XWave<T, N>;
  static_assert(X::SUPPORTED, "a row pair of H points or a row of 2 H points per wavefront");
  static_assert(NCIN == 7 || NCIN == 6, "u, w and optionally phi");
  constexpr bool HALF = X::HALF;
  constexpr int H = X::H, S = X::S, RW = X::RW, M = X::M, NP = X::NP;
  constexpr int XR = XT / 64;          // wavefronts per block
  // The exchange buffers are a static array and the stages the dynamic one, so the compiler
  // tells their accesses apart: waiting for a stage copy never holds up an exchange.
  __shared__ __align__(16) cx<T> xbuf[XR * NP];
  extern __shared__ __align__(16) unsigned char smem_raw[];   // XFETCH_LDS: XR stages of RW * kxp
  __shared__ double red[XR];
  __shared__ cx<T> t64[64];
  const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
  cx<T> *wb = xbuf + w * NP;
  // the padded row of R (kx 0..M and zeros to the tile width): the launch checks g.kxp == KXP
  constexpr int SROW = RW * X::KXP;                 // coefficients of one component's rows
  cx<T> *stage = reinterpret_cast<cx<T> *>(smem_raw) + w * SROW;

  const int rows_per_plane = N / (RW * XR);
  const int z = blockIdx.x / rows_per_plane, box = blockIdx.y;
  const int ya = (blockIdx.x % rows_per_plane) * (RW * XR) + RW * w;
  rin += box * g.box_b;
  pxo += box * g.box_a;
  const size_t plane = (size_t)N * g.kxp;           // one z plane of one component in R / PX
  const size_t vol = (size_t)N * N * N;
  const cx<T> *rbase = rin + (size_t)z * plane + (size_t)ya * g.kxp;

  // the rows of component c -> this wavefront's stage, 64 16-byte pieces per load instruction
  // (lane l's piece lands at 16 l bytes past the instruction's stage address)
  auto stage_in = [&](int c) {
    const char *src = reinterpret_cast<const char *>(rbase + (size_t)c * N * plane);
    char *dst = reinterpret_cast<char *>(stage);
    constexpr int PIECES = SROW * (int)sizeof(cx<T>) / 16;
#pragma unroll
    for (int i = 0; i < PIECES; i += 64)
      if (PIECES % 64 == 0 || i + j < PIECES)
        __builtin_amdgcn_global_load_lds(src + 16 * (i + j), dst + 16 * i, 16, 0, 0);
  };
  if constexpr (FETCH == XFETCH_LDS) stage_in(0);

  if (threadIdx.x < 64) t64[j] = {T(tw[(N / 64) * j].re), T(tw[(N / 64) * j].im)};   // W_64^j
  __syncthreads();

  cx<T> t3[S - 1];                                  // W_H^{q j}
#pragma unroll
  for (int q = 1; q < S; ++q) t3[q - 1] = {T(tw[(N / H) * q * j].re), T(tw[(N / H) * q * j].im)};
  cx<T> wq[HALF ? S : 1];                           // W_N^{j + 64 q} (HALF)
  if constexpr (HALF) {
#pragma unroll
    for (int q = 0; q < S; ++q) wq[q] = {T(tw[j + 64 * q].re), T(tw[j + 64 * q].im)};
  }

  // The coefficients of a component that the first pass of its inverse transform reads, from
  // `src` (its rows in R, or their staged copy; row b follows row a at kxp): slot q holds (a, b)
  // at kx = j + 64 q (when <= M) or at kx = N - j - 64 q (when >= N - M), and zeros in between.
  // (HALF: slot q holds X[m] and X[H - m] of the row at m = j + 64 q, zero beyond M; X[H] is
  // outside the ball)
  cx<T> ra[S], rb[S];
  auto coefficients = [&](const cx<T> *src) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int n = j + 64 * q;
      ra[q] = czero<T>();
      rb[q] = czero<T>();
      if constexpr (HALF) {
        if (n <= M) ra[q] = src[n];
        if (H - n <= M) rb[q] = src[H - n];
      } else if (n <= M) {
        ra[q] = src[n];
        rb[q] = src[g.kxp + n];
      } else if (n >= N - M) {
        ra[q] = src[N - n];
        rb[q] = src[g.kxp + N - n];
      }
    }
  };
  // pack the coefficients of component c as the spectrum of (row a) + i (row b), start the copy of
  // component `next` (if any) into the stage, and transform to physical space: x[q] = (row a,
  // row b) at x = j + 64 q
  auto inverse = [&](cx<T>(&x)[S], int c, int next) {
    if constexpr (FETCH == XFETCH_LDS) {
      // the copy has landed: direct-to-LDS loads count in vmcnt until their data are in LDS, and
      // the compiler does not always order LDS reads after them by itself
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      coefficients(stage);
    } else {
      coefficients(rbase + (size_t)c * N * plane);
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int n = j + 64 * q;
      cx<T> a = ra[q], b = rb[q];
      if constexpr (HALF) {
        if (n == 0) a.im = T(0);   // the self-conjugate coefficient: only its real part is seen
        const cx<T> sm = {a.re + b.re, a.im - b.im};          // X[m] + conj X[H-m]
        const cx<T> df = {a.re - b.re, a.im + b.im};          // X[m] - conj X[H-m]
        const cx<T> p = twiddled<+1>(df, wq[q]);              // W^-m (..)
        x[q] = {sm.re - p.im, sm.im + p.re};
        continue;
      }
      if (n <= M) {
        if (n == 0) {  // the self-conjugate coefficient: only its real part is seen
          a.im = T(0);
          b.im = T(0);
        }
        x[q] = {a.re - b.im, a.im + b.re};
      } else {  // the mirrored coefficient, or zero
        x[q] = {a.re + b.im, b.re - a.im};
      }
    }
    if constexpr (FETCH == XFETCH_LDS) {
      if (next >= 0) {
        // the stage's reads have returned (x holds them) before the copy overwrites it
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        stage_in(next);
      }
    }
    wave_fft<+1>(x, wb, t64, t3, j);
  };

  // product d to the spectrum, kx <= M kept, zeros up to kxp
  auto forward = [&](int d, cx<T>(&x)[S]) {
    wave_fft<-1>(x, wb, t64, t3, j);
    cx<T> *dst = pxo + (size_t)d * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
    const int partner = (64 - j) & 63;
    if constexpr (HALF) {
      // X[k] for k = j + 64 q (q < QF): Z[H - k] is slot S - 1 - q of lane 64 - j, or slot
      // (S - q) mod S of lane 0
#pragma unroll
      for (int q = 0; q < X::QF; ++q) {
        cx<T> zp = {lane_value(x[S - 1 - q].re, partner), lane_value(x[S - 1 - q].im, partner)};
        if (j == 0) zp = x[(S - q) & (S - 1)];
        const int kx = j + 64 * q;
        if (kx <= M) {
          const cx<T> Z = x[q];
          const cx<T> e = {T(0.5) * (Z.re + zp.re), T(0.5) * (Z.im - zp.im)};     // (Z + conj Zp) / 2
          const cx<T> o = {T(0.5) * (Z.im + zp.im), T(0.5) * (zp.re - Z.re)};     // (Z - conj Zp) / 2i
          const cx<T> t = twiddled<-1>(o, wq[q]);
          dst[kx] = {e.re + t.re, e.im + t.im};
        } else if (kx < g.kxp) {
          dst[kx] = czero<T>();
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < X::QF; ++q) {
      // C[N - kx] for kx = j + 64 q: slot S - 1 - q of lane 64 - j, or slot (S - q) mod S of lane 0
      cx<T> cm = {lane_value(x[S - 1 - q].re, partner), lane_value(x[S - 1 - q].im, partner)};
      if (j == 0) cm = x[(S - q) & (S - 1)];
      const int kx = j + 64 * q;
      if (kx <= M) {
        const cx<T> C = x[q];
        dst[kx] = {T(0.5) * (C.re + cm.re), T(0.5) * (C.im - cm.im)};
        dst[g.kxp + kx] = {T(0.5) * (C.im + cm.im), T(0.5) * (cm.re - C.re)};
      } else if (kx < g.kxp) {
        dst[kx] = czero<T>();
        dst[g.kxp + kx] = czero<T>();
      }
    }
  };

  const size_t rowa = ((size_t)z * N + ya) * N, rowb = rowa + N;
  T *phys = reinterpret_cast<T *>(xa.phys) + box * xa.box_phys;
  T *physw = reinterpret_cast<T *>(xa.physw) + box * xa.box_physw;
  auto store_phys = [&](T *base, const cx<T>(&x)[S]) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if constexpr (HALF) {   // (x[2n], x[2n+1]) at n = j + 64 q
        *reinterpret_cast<cx<T> *>(base + rowa + 2 * (j + 64 * q)) = x[q];
      } else {
        base[rowa + j + 64 * q] = x[q].re;
        base[rowb + j + 64 * q] = x[q].im;
      }
    }
  };

  // the interleaved copy for the tracers: (u_x, u_y, u_z, phi) at the points of slot q of the
  // row(s), the values passed one by one so that they stay in registers
  auto store_points = [&](int q, cx<T> u0, cx<T> u1, cx<T> u2, cx<T> ph) {
    if constexpr (HALF) {   // (x[2n], x[2n+1]) at n = j + 64 q
      const size_t at = rowa + 2 * (j + 64 * q);
      store_point4(phys + 4 * at, u0.re, u1.re, u2.re, ph.re);
      store_point4(phys + 4 * (at + 1), u0.im, u1.im, u2.im, ph.im);
    } else {
      store_point4(phys + 4 * (rowa + j + 64 * q), u0.re, u1.re, u2.re, ph.re);
      store_point4(phys + 4 * (rowb + j + 64 * q), u0.im, u1.im, u2.im, ph.im);
    }
  };

  constexpr int W0 = 3, PHI = 6;   // components of R: u (0..2), w (3..5), phi (6)
  cx<T> u[3][S];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    inverse(u[c], c, c < 2 ? c + 1 : (NCIN == 7 ? PHI : W0));
    if constexpr (!IL) {
      if (xa.store_u) store_phys(phys + c * vol, u[c]);
    }
  }
  if constexpr (IL && NCIN == 6) {
    if (xa.store_u) {
#pragma unroll
      for (int q = 0; q < S; ++q) store_points(q, u[0][q], u[1][q], u[2][q], czero<T>());
    }
  }

  double mx = 0.0;
  if (xa.want_max) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const T sa = u[0][q].re * u[0][q].re + u[1][q].re * u[1][q].re + u[2][q].re * u[2][q].re;
      const T sb = u[0][q].im * u[0][q].im + u[1][q].im * u[1][q].im + u[2][q].im * u[2][q].im;
      const double da = (double)sa, db = (double)sb;
      if (da > mx || da != da) mx = da;
      if (db > mx || db != db) mx = db;
    }
  }

  if constexpr (NCIN == 7) {  // the fluxes u phi (multiply!(u, phi, u), src/Generator.jl:727)
    cx<T> ph[S];
    inverse(ph, PHI, W0);
    if constexpr (IL) {
      if (xa.store_u || xa.store_phi) {
#pragma unroll
        for (int q = 0; q < S; ++q) store_points(q, u[0][q], u[1][q], u[2][q], ph[q]);
      }
    } else {
      if (xa.store_phi) store_phys(phys + 3 * vol, ph);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cx<T> p[S];
#pragma unroll
      for (int q = 0; q < S; ++q) p[q] = {u[d][q].re * ph[q].re, u[d][q].im * ph[q].im};
      forward(3 + d, p);
    }
  }

  // the negated Lamb vector L = u x w (cross!(w, u, w), src/Generator.jl:529):
  //   L_0 = u_1 w_2 - u_2 w_1,  L_1 = u_2 w_0 - u_0 w_2,  L_2 = u_0 w_1 - u_1 w_0
  cx<T> l0[S], l1[S];
  {
    cx<T> w0[S];
    inverse(w0, W0, W0 + 1);
    if (xa.store_w) store_phys(physw, w0);
    cx<T> l2[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      l1[q] = {u[2][q].re * w0[q].re, u[2][q].im * w0[q].im};
      l2[q] = {-(u[1][q].re * w0[q].re), -(u[1][q].im * w0[q].im)};
    }
    cx<T> w1[S];
    inverse(w1, W0 + 1, W0 + 2);
    if (xa.store_w) store_phys(physw + vol, w1);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      l0[q] = {-(u[2][q].re * w1[q].re), -(u[2][q].im * w1[q].im)};
      l2[q] = {u[0][q].re * w1[q].re + l2[q].re, u[0][q].im * w1[q].im + l2[q].im};
    }
    forward(2, l2);
  }
  {
    cx<T> w2[S];
    inverse(w2, W0 + 2, -1);
    if (xa.store_w) store_phys(physw + 2 * vol, w2);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      l0[q] = {u[1][q].re * w2[q].re + l0[q].re, u[1][q].im * w2[q].im + l0[q].im};
      l1[q] = {l1[q].re - u[0][q].re * w2[q].re, l1[q].im - u[0][q].im * w2[q].im};
    }
    forward(0, l0);
    forward(1, l1);
  }

  if (xa.want_max) {
    // wave then block max, NaN-propagating, then one atomic per block
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(mx, off, 64);
      if (o > mx || o != o) mx = o;
    }
    if (j == 0) red[w] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = red[0];
      for (int i = 1; i < XR; ++i)
        if (red[i] > b || red[i] != red[i]) b = red[i];
      atomicMax(xa.max_bits + box, (unsigned long long)__double_as_longlong(b));
    }
  }
}

}  // namespace hdns
