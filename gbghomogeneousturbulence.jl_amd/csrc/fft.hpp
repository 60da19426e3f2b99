// Device building blocks of the pencil transforms: a complex type, the small DFT butterflies,
// and a block-cooperative Stockham (autosort) FFT over a batch of columns held in LDS.
//
// The batch layout in LDS is either
//   interleaved  elem(n, col) = n*TX + col   (the column tile exactly as it sits in HBM, so the
//                                            load and the store are straight copies), or
//   contiguous   elem(n, col) = col*N + n    (one column after another; used for x rows).
// Each pass with radix R over sub-transform length NS maps butterfly j (0 <= j < N/R) to
//   inputs  j + q*N/R,  twiddle W_{NS*R}^{q*(j mod NS)},  outputs (j/NS)*NS*R + (j mod NS) + q*NS,
// the Stockham self-sorting form, so the result lands in natural order with no bit reversal.
// Twiddles come from an FP64 table of W_N^m = exp(-2 pi i m/N) (each entry the rounding of an
// extended-precision value), staged once per block into LDS rounded to the working type, so a
// pass reads them at LDS latency.  SIGN = -1 is the forward (unnormalised) transform, +1 the inverse.
#pragma once
#include <hip/hip_runtime.h>

namespace hdns {

template <typename T>
struct cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }

// a * (SIGN * i)
template <int SIGN, typename T>
__device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  if constexpr (SIGN > 0) return {-a.im, a.re};
  else return {a.im, -a.re};
}

// a * (c + SIGN*i*s)
template <int SIGN, typename T>
__device__ __forceinline__ cx<T> mul_cs(cx<T> a, T c, T s) {
  if constexpr (SIGN > 0) return {a.re * c - a.im * s, a.im * c + a.re * s};
  else return {a.re * c + a.im * s, a.im * c - a.re * s};
}

template <typename T> struct fftc {
  static constexpr T r2 = T(0.70710678118654752440084436210484903928);   // sqrt(2)/2
  static constexpr T c8 = T(0.92387953251128675612818318939678828682);   // cos(pi/8)
  static constexpr T s8 = T(0.38268343236508977172845998403039886676);   // sin(pi/8)
  static constexpr T h3 = T(0.86602540378443864676372317075293618347);   // sqrt(3)/2
};

template <int SIGN, typename T>
__device__ __forceinline__ void dft2(cx<T> *v) {
  cx<T> a = v[0];
  v[0] = a + v[1];
  v[1] = a - v[1];
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft3(cx<T> *v) {
  cx<T> t = v[1] + v[2];
  cx<T> d = v[1] - v[2];
  cx<T> m = {v[0].re - T(0.5) * t.re, v[0].im - T(0.5) * t.im};
  cx<T> r = mul_si<SIGN>(scale(d, fftc<T>::h3));
  v[0] = v[0] + t;
  v[1] = m + r;
  v[2] = m - r;
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft4(cx<T> *v) {
  cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  cx<T> t2 = v[1] + v[3], t3 = mul_si<SIGN>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft8(cx<T> *v) {
  cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN>(e);
  dft4<SIGN>(o);
  const T r = fftc<T>::r2;
  o[1] = mul_cs<SIGN>(o[1], r, r);
  o[2] = mul_si<SIGN>(o[2]);
  o[3] = mul_cs<SIGN>(o[3], -r, r);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = e[k] + o[k];
    v[k + 4] = e[k] - o[k];
  }
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft16(cx<T> *v) {
  cx<T> e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8<SIGN>(e);
  dft8<SIGN>(o);
  const T r = fftc<T>::r2, c = fftc<T>::c8, s = fftc<T>::s8;
  o[1] = mul_cs<SIGN>(o[1], c, s);
  o[2] = mul_cs<SIGN>(o[2], r, r);
  o[3] = mul_cs<SIGN>(o[3], s, c);
  o[4] = mul_si<SIGN>(o[4]);
  o[5] = mul_cs<SIGN>(o[5], -s, c);
  o[6] = mul_cs<SIGN>(o[6], -r, r);
  o[7] = mul_cs<SIGN>(o[7], -c, s);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = e[k] + o[k];
    v[k + 8] = e[k] - o[k];
  }
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cx<T> *v) {
  if constexpr (R == 2) dft2<SIGN>(v);
  else if constexpr (R == 3) dft3<SIGN>(v);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else if constexpr (R == 8) dft8<SIGN>(v);
  else if constexpr (R == 16) dft16<SIGN>(v);
  else static_assert(R == 2, "unsupported radix");
}

template <int N, int TX, bool IL>
__device__ __forceinline__ int lds_index(int n, int col) {
  if constexpr (IL) return n * TX + col;
  else return col * N + n;
}

// The barrier between the read and write phases of a pass: the whole block, or (WAVE) only the
// wavefront that owns a private LDS region -- LDS requests of one wave complete in order, so a
// compiler fence around the wave barrier is all a wave-private exchange needs.
template <bool WAVE>
__device__ __forceinline__ void pass_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham pass over TX columns of length N held in LDS, by NT cooperating threads: the
// threads t with t % NT equal to the same index form the group (NT = 64 lanes of one wavefront
// when WAVE; otherwise every group of the block passes the same block-wide barriers).
template <typename T, int N, int R, int NS, int TX, int NT, int SIGN, bool IL, bool WAVE>
__device__ __forceinline__ void stockham_pass(cx<T> *buf, const cx<T> *tw) {
  static_assert(!WAVE || NT == 64, "a wave-private transform runs on the 64 lanes of one wavefront");
  constexpr int NB = N / R;
  constexpr int TOT = NB * TX;
  constexpr int PER = (TOT + NT - 1) / NT;
  const int tid = (int)(threadIdx.x % NT);   // NT threads of the block (or of a wave) own the transform
  cx<T> v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int col = IL ? b % TX : b / NB;
      const int j = IL ? b / TX : b % NB;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = buf[lds_index<N, TX, IL>(j + q * NB, col)];
    }
  }
  pass_sync<WAVE>();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int col = IL ? b % TX : b / NB;
      const int j = IL ? b / TX : b % NB;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const cx<T> w = tw[q * k * (N / (NS * R))];
          const cx<T> wt = {w.re, SIGN < 0 ? w.im : -w.im};
          v[p][q] = v[p][q] * wt;
        }
      }
      dft<R, SIGN>(v[p]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[lds_index<N, TX, IL>(base + q * NS, col)] = v[p][q];
    }
  }
  pass_sync<WAVE>();
}

template <typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE, int NS, int R, int... REST>
__device__ __forceinline__ void run_passes(cx<T> *buf, const cx<T> *tw) {
  stockham_pass<T, N, R, NS, TX, NT, SIGN, IL, WAVE>(buf, tw);
  if constexpr (sizeof...(REST) > 0) run_passes<T, N, TX, NT, SIGN, IL, WAVE, NS * R, REST...>(buf, tw);
}

// Stage the FP64 table into LDS in the working type (the caller synchronises before use).
template <typename T, int N, int NT>
__device__ __forceinline__ void stage_twiddles(cx<T> *twl, const cx<double> *__restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += NT) twl[i] = {T(tw[i].re), T(tw[i].im)};
}

// The radix plan of each supported length.
template <typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE = false>
__device__ __forceinline__ void fft_columns(cx<T> *buf, const cx<T> *tw) {
  if constexpr (N == 16) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 4, 4>(buf, tw);
  else if constexpr (N == 32) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 4>(buf, tw);
  else if constexpr (N == 64) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 8>(buf, tw);
  else if constexpr (N == 128) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 4, 4>(buf, tw);
  else if constexpr (N == 256) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 8, 4>(buf, tw);
  else if constexpr (N == 512) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 8, 8>(buf, tw);
  else if constexpr (N == 1024) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 8, 4, 4>(buf, tw);
  else if constexpr (N == 48) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 4, 4, 3>(buf, tw);
  else if constexpr (N == 96) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 4, 3>(buf, tw);
  else if constexpr (N == 192) run_passes<T, N, TX, NT, SIGN, IL, WAVE, 1, 8, 8, 3>(buf, tw);
  else static_assert(N == 16, "unsupported transform length");
}

// ---------------------------------------------------------------------------------------------
// Transforms whose first pass reads its inputs through a loader and whose last pass hands its
// outputs to a storer, so the data enter and leave through registers: HBM (or the caller's
// registers) -> pass 1 -> LDS -> middle passes -> LDS -> last pass -> registers.  Butterfly b of a
// pass is (col, j) as above; the loader is called as load(p, q, n, col) for input element n of the
// p-th butterfly of this thread, the storer as store(p, q, n, col, value) for output element n.
// The callbacks are inlined, so arrays they index with (p, q) stay in registers.

template <int N, int R, int TX, int NT>
struct PassShape {
  static constexpr int NB = N / R;
  static constexpr int TOT = NB * TX;
  static constexpr int PER = (TOT + NT - 1) / NT;
};

// First pass (NS = 1: no twiddles).  SYNC_READ: the loader reads the buffer this pass writes.
template <typename T, int N, int R, int TX, int NT, int SIGN, bool IL, bool WAVE, bool SYNC_READ, class Load>
__device__ __forceinline__ void first_pass(cx<T> *buf, Load &&load) {
  using S = PassShape<N, R, TX, NT>;
  const int tid = (int)(threadIdx.x % NT);
  cx<T> v[S::PER][R];
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * NT;
    if (S::TOT % NT == 0 || b < S::TOT) {
      const int col = IL ? b % TX : b / S::NB;
      const int j = IL ? b / TX : b % S::NB;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = load(p, q, j + q * S::NB, col);
    }
  }
  if constexpr (SYNC_READ) pass_sync<WAVE>();
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * NT;
    if (S::TOT % NT == 0 || b < S::TOT) {
      const int col = IL ? b % TX : b / S::NB;
      const int j = IL ? b / TX : b % S::NB;
      dft<R, SIGN>(v[p]);
#pragma unroll
      for (int q = 0; q < R; ++q) buf[lds_index<N, TX, IL>(j * R + q, col)] = v[p][q];
    }
  }
  pass_sync<WAVE>();
}

// Last pass: inputs from LDS, outputs to the storer; ends with a barrier so the buffer may be
// rewritten at once.
template <typename T, int N, int R, int NS, int TX, int NT, int SIGN, bool IL, bool WAVE, class Store>
__device__ __forceinline__ void last_pass(const cx<T> *buf, const cx<T> *tw, Store &&store) {
  using S = PassShape<N, R, TX, NT>;
  const int tid = (int)(threadIdx.x % NT);
#pragma unroll
  for (int p = 0; p < S::PER; ++p) {
    const int b = tid + p * NT;
    if (S::TOT % NT == 0 || b < S::TOT) {
      const int col = IL ? b % TX : b / S::NB;
      const int j = IL ? b / TX : b % S::NB;
      const int k = j % NS;
      cx<T> v[R];
#pragma unroll
      for (int q = 0; q < R; ++q) v[q] = buf[lds_index<N, TX, IL>(j + q * S::NB, col)];
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const cx<T> w = tw[q * k * (N / (NS * R))];
        v[q] = v[q] * cx<T>{w.re, SIGN < 0 ? w.im : -w.im};
      }
      dft<R, SIGN>(v);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) store(p, q, base + q * NS, col, v[q]);
    }
  }
  pass_sync<WAVE>();
}

template <int... RS>
struct Radices {};

template <typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE, int NS, int R, int... REST, class Store>
__device__ __forceinline__ void middle_and_last(Radices<R, REST...>, cx<T> *buf, const cx<T> *tw, Store &&store) {
  if constexpr (sizeof...(REST) == 0) {
    last_pass<T, N, R, NS, TX, NT, SIGN, IL, WAVE>(buf, tw, store);
  } else {
    stockham_pass<T, N, R, NS, TX, NT, SIGN, IL, WAVE>(buf, tw);
    middle_and_last<T, N, TX, NT, SIGN, IL, WAVE, NS * R>(Radices<REST...>{}, buf, tw, store);
  }
}

template <typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE, bool SYNC_READ, int R1, int... REST,
          class Load, class Store>
__device__ __forceinline__ void fft_io_plan(Radices<R1, REST...>, cx<T> *buf, const cx<T> *tw, Load &&load,
                                            Store &&store) {
  first_pass<T, N, R1, TX, NT, SIGN, IL, WAVE, SYNC_READ>(buf, load);
  middle_and_last<T, N, TX, NT, SIGN, IL, WAVE, R1>(Radices<REST...>{}, buf, tw, store);
}

// The radix plan of each length in the register-in/register-out form (first_radix / last_radix
// in spectral.hpp name its first and last pass).
template <int N> struct IoPlan;
template <> struct IoPlan<16> { using type = Radices<4, 4>; };
template <> struct IoPlan<32> { using type = Radices<8, 4>; };
template <> struct IoPlan<64> { using type = Radices<8, 8>; };
template <> struct IoPlan<128> { using type = Radices<8, 4, 4>; };
template <> struct IoPlan<256> { using type = Radices<8, 8, 4>; };
template <> struct IoPlan<512> { using type = Radices<8, 8, 8>; };
template <> struct IoPlan<1024> { using type = Radices<8, 8, 4, 4>; };
template <> struct IoPlan<48> { using type = Radices<4, 4, 3>; };
template <> struct IoPlan<96> { using type = Radices<8, 4, 3>; };
template <> struct IoPlan<192> { using type = Radices<8, 8, 3>; };

template <typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE, bool SYNC_READ = false, class Load,
          class Store>
__device__ __forceinline__ void fft_io(cx<T> *buf, const cx<T> *tw, Load &&load, Store &&store) {
  fft_io_plan<T, N, TX, NT, SIGN, IL, WAVE, SYNC_READ>(typename IoPlan<N>::type{}, buf, tw, load, store);
}

// fft_io with hook() run between the first pass and the rest: by then every thread has passed the
// first pass's closing barrier, so whatever the loader read may be overwritten.
template <typename T, int N, int TX, int NT, int SIGN, bool IL, int R1, int... REST, class Load, class Store,
          class Hook>
__device__ __forceinline__ void fft_io_hook_plan(Radices<R1, REST...>, cx<T> *buf, const cx<T> *tw, Load &&load,
                                                 Store &&store, Hook &&hook) {
  first_pass<T, N, R1, TX, NT, SIGN, IL, false, false>(buf, load);
  hook();
  middle_and_last<T, N, TX, NT, SIGN, IL, false, R1>(Radices<REST...>{}, buf, tw, store);
}
template <typename T, int N, int TX, int NT, int SIGN, bool IL, class Load, class Store, class Hook>
__device__ __forceinline__ void fft_io_hook(cx<T> *buf, const cx<T> *tw, Load &&load, Store &&store, Hook &&hook) {
  fft_io_hook_plan<T, N, TX, NT, SIGN, IL>(typename IoPlan<N>::type{}, buf, tw, load, store, hook);
}

// The second radix plan of each length, led by a radix-16 pass (one pass fewer at 256 and 1024);
// the autotuner picks between it and fft_io's plan for the y passes.
template <int N>
struct AltPlan;
template <> struct AltPlan<16> { using type = Radices<2, 8>; };
template <> struct AltPlan<32> { using type = Radices<16, 2>; };
template <> struct AltPlan<64> { using type = Radices<16, 4>; };
template <> struct AltPlan<128> { using type = Radices<16, 8>; };
template <> struct AltPlan<256> { using type = Radices<16, 16>; };
template <> struct AltPlan<512> { using type = Radices<16, 8, 4>; };
template <> struct AltPlan<1024> { using type = Radices<16, 8, 8>; };
template <> struct AltPlan<48> { using type = Radices<16, 3>; };
template <> struct AltPlan<96> { using type = Radices<16, 2, 3>; };
template <> struct AltPlan<192> { using type = Radices<16, 4, 3>; };

// fft_io under radix plan PLAN (0: fft_io's, 1: AltPlan's)
template <int PLAN, typename T, int N, int TX, int NT, int SIGN, bool IL, bool WAVE, bool SYNC_READ = false,
          class Load, class Store>
__device__ __forceinline__ void fft_io_planned(cx<T> *buf, const cx<T> *tw, Load &&load, Store &&store) {
  if constexpr (PLAN == 0) fft_io<T, N, TX, NT, SIGN, IL, WAVE, SYNC_READ>(buf, tw, load, store);
  else fft_io_plan<T, N, TX, NT, SIGN, IL, WAVE, SYNC_READ>(typename AltPlan<N>::type{}, buf, tw, load, store);
}

}  // namespace hdns
