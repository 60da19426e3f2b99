// The pencil kernels of the DNS stage: the two y passes, the x pass that forms the physical
// products, and the two z passes that carry every spectral pointwise operation of the stage.
//
// Retained modes only.  Every field the step keeps lies inside the cube |mx|, |my|, |mz| <= M of
// its dealiasing strategy: M = floor(N/3) under the 2/3 rule, the largest m with 9 m^2 <= 2 N^2
// under PhaseShift (the ball of radius sqrt(2) N/3), and N/2 under HouLi, which keeps every mode.
// Spectral data is stored compactly as
//   [c][kz_idx][ky_idx][kx]   kx in [0, KXP) (padded past M with zeros),
//                             ky_idx, kz_idx in [0, K), K = min(2M + 1, N), m = idx <= M ? idx : idx - K
//                             (under HouLi K = N and the Nyquist index N/2 carries m = +N/2),
// and the transforms are pruned: a z pass only ever touches the K retained ky columns, a y pass
// only the kx < KXP columns, and the x pass zero-fills M < kx < KXP.  The half-transformed
// pencils are
//   PX [c][z][y][kx]   after the forward x pass      (6 product components)
//   PF [c][z][ky][kx]  after the forward y pass
//   Q  [c][z][ky][kx]  after the inverse z pass      (7 components: u, w, phi)
//   R  [c][z][y][kx]   after the inverse y pass
// PX/Q share one buffer and PF/R another: each is dead before its twin is written.
#pragma once
#include "bufio.hpp"
#include "fft.hpp"

namespace hdns {

// The dealiasing strategies (src/SimulationDomain.jl:44-195): the sharp 2/3 ball, Hou-Li's smooth
// filter (every mode kept), and the phase-shifted evaluation with the ball of radius sqrt(2) N/3.
enum { DEALIAS_TWO_THIRDS = 0, DEALIAS_HOU_LI = 1, DEALIAS_PHASE_SHIFT = 2 };

// Largest retained |m| per axis, and the retained signed wavenumbers per y/z axis.
__host__ __device__ constexpr int band_m(int n, int da) {
  if (da == DEALIAS_HOU_LI) return n / 2;
  if (da == DEALIAS_PHASE_SHIFT) {
    int m = 0;
    while (9 * (m + 1) * (m + 1) <= 2 * n * n) ++m;
    return m;
  }
  return n / 3;
}
__host__ __device__ constexpr int band_k(int n, int da) { return 2 * band_m(n, da) + 1 > n ? n : 2 * band_m(n, da) + 1; }
// The ball a strategy keeps, as the bound of 9 |m|^2 (no bound under HouLi).
__host__ __device__ constexpr int band_ball9(int n, int da) {
  return da == DEALIAS_HOU_LI ? 0x7fffffff : (da == DEALIAS_PHASE_SHIFT ? 2 * n * n : n * n);
}

struct Geo {
  int n;      // grid points per axis
  int m;      // largest retained |m| per axis (band_m)
  int kxp;    // padded kx extent of every compact array
  int k;      // retained signed wavenumbers on y and z (band_k)
  double k0;  // 2 pi / L
  int ball9;  // a mode is kept iff 9 |m|^2 <= ball9 (band_ball9)
  // A batch of boxes of one size lies back to back in every array; a kernel finds its box in the
  // grid's last dimension.  Per box: buffer A and B hold box_a / box_b complex values, the state
  // [3][comp] (u, h_u) or [comp] (phi, h_phi).
  size_t box_a, box_b;
  int ncf, nci;   // product and inverse components (the y passes' grid z is components x boxes)
};

// compact elements of one component
__host__ __device__ __forceinline__ size_t comp_of(const Geo &g) { return (size_t)g.k * g.k * g.kxp; }

// FFT bin of a compact signed-wavenumber index.
__device__ __forceinline__ int bin_of(int idx, const Geo &g) { return idx <= g.m ? idx : g.n - g.k + idx; }
// Signed wavenumber index of a compact index.
__device__ __forceinline__ int signed_of(int idx, const Geo &g) { return idx <= g.m ? idx : idx - g.k; }
// Compact index of an FFT bin, or -1 when the bin is not retained.
__device__ __forceinline__ int idx_of_bin(int b, const Geo &g) {
  if (b <= g.m) return b;
  if (b >= g.n - g.m) return b - g.n + g.k;
  return -1;
}

template <typename T>
__device__ __forceinline__ cx<T> czero() { return {T(0), T(0)}; }

// The (kx, ky) column of a z pencil meets the retained ball 9|m|^2 <= ball9 only when
// 9 (mx^2 + my^2) <= ball9.  Every retained state mode outside the ball is zero and stays zero
// (the state is adopted truncated, and the mask, the Leray projection, the forcing and the
// integrating factor all keep zeros zero), so the passes skip the (kx, ky) columns outside this
// disk, at the granularity of whole 128-byte lines (partly written lines cost more than they
// save): a ky row of PF or Q is written and read only in the lines whose first column lies in
// the disk.  In a written Q line the columns outside the disk hold exact zeros (the transform of
// zero modes), so the inverse y pass reads the whole line.
__host__ __device__ __forceinline__ bool in_disk(int mx, int my, const Geo &g) {
  return 9 * (mx * mx + my * my) <= g.ball9;
}
// The retained ball: a mode the strategy keeps (9|m|^2 <= N^2 under the 2/3 rule, the exact-integer
// closed ball of src/SimulationDomain.jl:342-376).
__host__ __device__ __forceinline__ bool keeps(int m2, const Geo &g) { return 9 * m2 <= g.ball9; }
// Parseval weight of a half-spectrum mode: 1 on the self-conjugate kx = 0 and kx = N/2 planes, 2 inside.
__host__ __device__ __forceinline__ double parseval_w(int mx, int n) { return (mx == 0 || 2 * mx == n) ? 1.0 : 2.0; }

// First column of the 128-byte line that holds column mx.
template <typename T>
__host__ __device__ __forceinline__ int line_start(int mx) { return mx & ~(int)(128 / sizeof(cx<T>) - 1); }

// ---------------------------------------------------------------------------------------------
// y passes: a block transforms TX adjacent kx columns of one z plane of one component.

template <typename T, int N, int TX, int NT, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_yfwd(const cx<T> *__restrict__ px, cx<T> *__restrict__ pf, Geo g,
                                             const cx<double> *__restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  stage_twiddles<T, N, NT>(twl, tw);
  const int kx0 = blockIdx.x * TX, z = blockIdx.y, c = blockIdx.z % g.ncf, box = blockIdx.z / g.ncf;
  const cx<T> *src = px + box * g.box_a + ((size_t)c * N + z) * (size_t)N * g.kxp + kx0;
  cx<T> *dst = pf + box * g.box_b + ((size_t)c * N + z) * (size_t)g.k * g.kxp + kx0;
  static_assert(TX * sizeof(cx<T>) <= 128, "a y tile is at most one 128-byte line of a row");
  fft_io_planned<PLAN, T, N, TX, NT, -1, true, false>(
      buf, twl, [&](int, int, int n, int col) { return src[(size_t)n * g.kxp + col]; },
      [&](int, int, int n, int col, cx<T> v) {
        const int idx = idx_of_bin(n, g);
        if (idx >= 0 && in_disk(kx0, signed_of(idx, g), g)) dst[(size_t)idx * g.kxp + col] = v;
      });
}

template <typename T, int N, int TX, int NT, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_yinv(const cx<T> *__restrict__ q, cx<T> *__restrict__ r, Geo g,
                                             const cx<double> *__restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  stage_twiddles<T, N, NT>(twl, tw);
  const int kx0 = blockIdx.x * TX, z = blockIdx.y, c = blockIdx.z % g.nci, box = blockIdx.z / g.nci;
  const cx<T> *src = q + box * g.box_a + ((size_t)c * N + z) * (size_t)g.k * g.kxp + kx0;
  cx<T> *dst = r + box * g.box_b + ((size_t)c * N + z) * (size_t)N * g.kxp + kx0;
  fft_io_planned<PLAN, T, N, TX, NT, +1, true, false>(
      buf, twl,
      [&](int, int, int n, int col) {
        const int idx = idx_of_bin(n, g);
        return idx >= 0 && in_disk(kx0, signed_of(idx, g), g) ? src[(size_t)idx * g.kxp + col] : czero<T>();
      },
      [&](int, int, int n, int col, cx<T> v) { dst[(size_t)n * g.kxp + col] = v; });
}

// ---------------------------------------------------------------------------------------------
// x pass: LPR lanes own one pair of adjacent y rows of one z plane (XT threads per block).  Two
// real rows travel as one complex transform (row a in the real part, row b in the imaginary
// part), so each component costs one complex FFT per row pair.  At FP32 and N <= 512 a single
// wavefront owns a row pair and synchronises with wave barriers only; wider transforms spread over
// 2 or 4 wavefronts so a lane holds at most 8 FP32 (4 FP64) values per component.  The kernel
// inverts the 7 (or 6) components to physical space, forms u x w and u phi there, transforms the
// 6 (or 3) products forward and keeps kx <= M.  On request it also stores the physical u and phi
// (the deposit) and the max of |u|^2.

struct XArgs {
  int store_u, store_phi, store_w;
  int want_max;
  void *phys;             // per box [4][N][N][N] real: u_x, u_y, u_z, phi; interleaved: [N][N][N][4]
  void *physw;            // per box [3][N][N][N] real: w
  unsigned long long *max_bits;  // per box: max |u|^2 as the bits of a non-negative double
  size_t box_phys, box_physw;    // reals per box in phys / physw
  int interleave;         // store u and phi as (u_x, u_y, u_z, phi) per point (phi 0 without a scalar)
};

// One point of the interleaved physical fields: four reals in 16-byte stores.
typedef float point4f __attribute__((ext_vector_type(4)));
typedef double point2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_point4(float *p, float a, float b, float c, float d) {
  *reinterpret_cast<point4f *>(p) = point4f{a, b, c, d};
}
__device__ __forceinline__ void store_point4(double *p, double a, double b, double c, double d) {
  reinterpret_cast<point2d *>(p)[0] = point2d{a, b};
  reinterpret_cast<point2d *>(p)[1] = point2d{c, d};
}

// lanes per row pair: a lane holds N/LPR values of each component.  BATCH: below 64 lanes a row
// pair takes the power of two at or above N/e lanes, and a wavefront holds 64/LPR row pairs and
// transforms them as one batch (else a row pair takes a whole wavefront).
template <typename T, int N, int XT, bool BATCH = false>
constexpr int xpass_lanes() {
  constexpr int e = sizeof(T) == 4 ? 8 : 4;
  constexpr int l = N / e;
  if constexpr (BATCH && l < 64) {
    int p = 1;
    while (p < l) p *= 2;
    return p;
  } else {
    return l < 64 ? 64 : (l > XT ? XT : l);
  }
}

// The transform of this thread's row pair, whose buffer wb is N values.  LPR < 64: the wavefront's
// 64/LPR row pairs have consecutive buffers and go through the passes together, one butterfly per
// lane where the old layout left all but N/R lanes idle; LPR = 64: one row pair per wavefront;
// LPR > 64: the row pair spans wavefronts and the passes synchronise the block.
template <typename T, int N, int LPR, int SIGN>
__device__ __forceinline__ void xrow_fft(cx<T> *wb, const cx<T> *twl) {
  if constexpr (LPR < 64) {
    constexpr int G = 64 / LPR;
    cx<T> *first = wb - (int)((threadIdx.x % 64) / LPR) * N;
    fft_columns<T, N, G, 64, SIGN, false, true>(first, twl);
  } else {
    fft_columns<T, N, 1, LPR, SIGN, false, LPR == 64>(wb, twl);
  }
}

// IL: the deposit's variant for runs with tracers, which stores u and phi interleaved.
template <typename T, int N, int NCIN, int XT, int DA = DEALIAS_TWO_THIRDS, bool IL = false>
__global__ __launch_bounds__(XT) void k_xpass(const cx<T> *__restrict__ rin,
                                                              cx<T> *__restrict__ pxo, Geo g,
                                                              const cx<double> *__restrict__ tw, XArgs xa) {
  constexpr int LPR = xpass_lanes<T, N, XT, true>();
  constexpr bool WAVE = LPR <= 64;                 // a wavefront holds whole row pairs
  constexpr int NT = XT;
  constexpr int XR = NT / LPR;                     // row pairs per block
  static_assert((N * N / 2) % XR == 0, "the blocks tile the N^2/2 row pairs of a box");
  constexpr int E = (N + LPR - 1) / LPR;           // physical elements per lane per row pair
  constexpr int M = band_m(N, DA);
  constexpr int PF = (M + 1 + LPR - 1) / LPR;      // retained coefficients per lane per row
  static_assert(NCIN == 7 || NCIN == 6, "u, w and optionally phi");
  constexpr int NCOUT = NCIN == 7 ? 6 : 3;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + XR * N;
  __shared__ double red[NT / 64];
  stage_twiddles<T, N, NT>(twl, tw);
  __syncthreads();

  const int w = threadIdx.x / LPR, lane = threadIdx.x % LPR, box = blockIdx.y;
  const int rp = blockIdx.x * XR + w;              // row pair (z, ya, ya + 1), y fastest
  const int z = rp / (N / 2), ya = 2 * (rp % (N / 2));
  rin += box * g.box_b;
  pxo += box * g.box_a;
  const size_t plane = (size_t)N * g.kxp;           // one z plane of one component in R / PX
  const size_t vol = (size_t)N * N * N;
  cx<T> *wb = buf + w * N;                         // this row pair's transform buffer

  cx<T> pa[PF], pb[PF];
  auto fetch = [&](int c) {
    const cx<T> *src = rin + (size_t)c * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int kx = lane + LPR * i;
      if (kx <= M) {
        pa[i] = src[kx];
        pb[i] = src[g.kxp + kx];
      }
    }
  };

  cx<T> f[NCIN][E];
  fetch(0);
#pragma unroll
  for (int c = 0; c < NCIN; ++c) {
    for (int e = lane; e < N - 2 * M - 1; e += LPR) wb[M + 1 + e] = czero<T>();
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int kx = lane + LPR * i;
      if (kx <= M) {
        cx<T> a = pa[i], b = pb[i];
        if (kx == 0 || 2 * kx == N) {  // the self-conjugate coefficients: only their real parts are seen
          a.im = T(0);
          b.im = T(0);
        }
        wb[kx] = {a.re - b.im, a.im + b.re};
        if (kx > 0 && 2 * kx != N) wb[N - kx] = {a.re + b.im, b.re - a.im};
      }
    }
    if (c + 1 < NCIN) fetch(c + 1);
    pass_sync<WAVE>();
    xrow_fft<T, N, LPR, +1>(wb, twl);
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (lane + LPR * q < N) f[c][q] = wb[lane + LPR * q];
    pass_sync<WAVE>();
  }

  // physical space: (re, im) = (row a, row b) of every component
  double mx = 0.0;
  if (xa.want_max) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (lane + LPR * q >= N) continue;
      const T sa = f[0][q].re * f[0][q].re + f[1][q].re * f[1][q].re + f[2][q].re * f[2][q].re;
      const T sb = f[0][q].im * f[0][q].im + f[1][q].im * f[1][q].im + f[2][q].im * f[2][q].im;
      const double da = (double)sa, db = (double)sb;
      if (da > mx || da != da) mx = da;
      if (db > mx || db != db) mx = db;
    }
  }
  if (xa.store_u || xa.store_phi || xa.store_w) {
    const size_t rowa = ((size_t)z * N + ya) * N, rowb = rowa + N;
    T *ph = reinterpret_cast<T *>(xa.phys) + box * xa.box_phys;
    T *pw = reinterpret_cast<T *>(xa.physw) + box * xa.box_physw;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int x = lane + LPR * q;
      if (x >= N) continue;
      if constexpr (IL) {
        if (xa.store_u || xa.store_phi) {
          const T pa = NCIN == 7 ? f[NCIN - 1][q].re : T(0), pb = NCIN == 7 ? f[NCIN - 1][q].im : T(0);
          store_point4(ph + 4 * (rowa + x), f[0][q].re, f[1][q].re, f[2][q].re, pa);
          store_point4(ph + 4 * (rowb + x), f[0][q].im, f[1][q].im, f[2][q].im, pb);
        }
      } else if (xa.store_u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          ph[c * vol + rowa + x] = f[c][q].re;
          ph[c * vol + rowb + x] = f[c][q].im;
        }
      }
      if constexpr (NCIN == 7) {
        if (xa.store_phi && !IL) {
          ph[3 * vol + rowa + x] = f[NCIN - 1][q].re;
          ph[3 * vol + rowb + x] = f[NCIN - 1][q].im;
        }
      }
      if (xa.store_w) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pw[c * vol + rowa + x] = f[3 + c][q].re;
          pw[c * vol + rowb + x] = f[3 + c][q].im;
        }
      }
    }
  }

#pragma unroll
  for (int d = 0; d < NCOUT; ++d) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (lane + LPR * q >= N) continue;
      cx<T> v;
      if (d < 3) {  // the negated Lamb vector u x w (cross!(w, u, w), src/Generator.jl:529)
        const int c1 = (d + 1) % 3, c2 = (d + 2) % 3;
        v.re = f[c1][q].re * f[3 + c2][q].re - f[c2][q].re * f[3 + c1][q].re;
        v.im = f[c1][q].im * f[3 + c2][q].im - f[c2][q].im * f[3 + c1][q].im;
      } else {      // the flux u phi (multiply!(u, phi, u), src/Generator.jl:727)
        v.re = f[d - 3][q].re * f[NCIN - 1][q].re;
        v.im = f[d - 3][q].im * f[NCIN - 1][q].im;
      }
      wb[lane + LPR * q] = v;
    }
    pass_sync<WAVE>();
    xrow_fft<T, N, LPR, -1>(wb, twl);
    cx<T> *dst = pxo + (size_t)d * N * plane + (size_t)z * plane + (size_t)ya * g.kxp;
    for (int kx = lane; kx < g.kxp; kx += LPR) {   // zeros past M keep every line whole
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

  if (xa.want_max) {
    // wave then block max, NaN-propagating, then one atomic per block
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(mx, off, 64);
      if (o > mx || o != o) mx = o;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = red[0];
      for (int i = 1; i < NT / 64; ++i)
        if (red[i] > b || red[i] != red[i]) b = red[i];
      atomicMax(xa.max_bits + box, (unsigned long long)__double_as_longlong(b));
    }
  }
}

}  // namespace hdns
