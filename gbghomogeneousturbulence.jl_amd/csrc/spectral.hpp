// The z passes, which carry every spectral pointwise operation of a stage, and the small
// kernels around them (reality projection, the FP64 reductions, spectra, frame tables).
//
// A z-pass block owns the TX adjacent kx columns of one retained ky, over all z.  Its threads
// own retained modes (kz_idx, col) in a fixed round-robin; the column transform runs in LDS in
// the interleaved layout [z][TX], which is exactly the tile as it sits in HBM.
#pragma once
#include "pencil.hpp"

namespace hdns {


struct BandSpec {
  int on;      // a band forcing present (every forcing but NoForcing)
  int lo4;     // 4|m|^2 >= lo4 = (2*lowest - 1)^2
  int hi4;     // 4|m|^2 <  hi4 = (2*highest + 1)^2, or -1 for no upper edge
};

__device__ __forceinline__ bool in_band(int m2, const BandSpec &b) {
  const int f = 4 * m2;
  return b.on && f >= b.lo4 && (b.hi4 < 0 || f < b.hi4);
}

struct StageArgs {
  int stage;          // 0-based stage index
  int last;           // 1 at the final stage
  double a, b;        // low-storage coefficients, already rounded to the element type
  double dt;          // the step, rounded to the element type for the accumulator
  double sub;         // d_i * dt: the exponent's substep (the exponent per |k|^2 is -D * sub)
  const double *visc;   // per box: nu
  const double *kappa;  // per box: nu / Sc
  double mean_gradient;
  BandSpec band;
  const double *scalars;  // per box [8]: [0] = the ConstantPower amplitude at this stage's state
  double lin_rate;        // Linear forcing: the rate A folded into the band's exponent, else 0
  // the scalar forcing, read by the scalar kernels built with SF: the source A_phi 1_band phi_hat
  // (A_phi = scalars[8 box + 7], 0 but for ConstantPower and VarianceController) and a Linear rate
  // folded into the scalar's integrating factor on its band
  BandSpec sband;
  double slin_rate;
  // Shu-Osher form (IFSSPRK3 src/Propagator.jl:599-691, IFHeun :456-517), read by the kernels built with SSP:
  // q <- cb E^pb (q + dt N) + E^pa (ca h), with E the factor over the half step sub = dt / 2 and
  // h the step's initial state, which stage 0 saves
  int ssp;
  int pa, pb;
  double ca, cb;          // the Shu-Osher weights, rounded to the element type
  // HouLi (src/SimulationDomain.jl:365-376): the filter exp(-s (|k|/k_Ny)^p) of each |m|^2,
  // 0 .. 3 (N/2)^2, applied to the nonlinear terms (src/Generator.jl:502-519, :668-704)
  const double *filter;
  // PhaseShift (src/Generator.jl:502-519, :675-704): the raw transforms of the products formed on
  // the grid displaced by half a cell, per box [ncf][comp] compact (the Lamb vector, then the flux)
  const void *shifted;
  // when set, dt and sub are read on the device from dyn[2 stage] and dyn[2 stage + 1] instead:
  // a captured step then serves every dt (the host writes them before each replay)
  const double *dyn;
};

__device__ __forceinline__ double stage_dt(const StageArgs &sa) { return sa.dyn ? sa.dyn[2 * sa.stage] : sa.dt; }
__device__ __forceinline__ double stage_sub(const StageArgs &sa) { return sa.dyn ? sa.dyn[2 * sa.stage + 1] : sa.sub; }

// PhaseShift's translation of a mode by half a cell along every axis: e^{i s pi (mx + my + mz)/N},
// s = +1 samples the field on the displaced grid, s = -1 shifts a displaced term back
// (src/Generator.jl:166-175).  FP64, rounded to T.
template <typename T>
__device__ __forceinline__ cx<T> half_cell_phase(int msum, int n, int sign) {
  double sn, cs;
  sincospi((double)(sign * msum) / (double)n, &sn, &cs);
  return {T(cs), T(sn)};
}

template <typename T>
__device__ __forceinline__ cx<T> cmul(const cx<T> &a, const cx<T> &b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Radix of the first and the last pass of a length-N transform (fft_io's plans).
template <int N>
constexpr int first_radix() { return (N == 16 || N == 48) ? 4 : 8; }
template <int N>
constexpr int last_radix() { return N % 3 == 0 ? 3 : ((N == 64 || N == 512) ? 8 : 4); }

// The flux tile of PF component c for the staged scalar half: N rows of TX values, copied HBM ->
// LDS 16 bytes per lane by direct-to-LDS loads (no registers held), into `stage` as [n][TX].  Lane
// l of a wavefront's instruction lands at 16 l bytes past the instruction's (uniform) LDS address.
template <typename T, int N, int TX, int NT>
__device__ __forceinline__ void stage_tile(cx<T> *stage, const cx<T> *src, size_t zstride) {
  constexpr int PR = TX * (int)sizeof(cx<T>) / 16;   // 16-byte pieces per tile row
  constexpr int PIECES = N * PR;
  static_assert(TX * sizeof(cx<T>) % 16 == 0, "a tile row holds whole 16-byte pieces");
  const int lane = threadIdx.x & 63, w0 = threadIdx.x & ~63;
  const char *s = reinterpret_cast<const char *>(src);
  char *d = reinterpret_cast<char *>(stage);
#pragma unroll
  for (int i0 = w0; i0 < PIECES; i0 += NT) {
    const int i = i0 + lane, n = i / PR, r = i % PR;
    if (PIECES % NT == 0 || i < PIECES)
      __builtin_amdgcn_global_load_lds(s + (size_t)n * zstride * sizeof(cx<T>) + 16 * r, d + 16 * i0, 16, 0, 0);
  }
}

// The scalar half of a stage (src/Generator.jl:648-730, src/Propagator.jl:366-391): the three
// fluxes' forward z transforms, N_phi = -mask i k . F - G u_z at the stage velocity, and the
// phi / h_phi update.  It runs before the velocity update, which overwrites u_z.
// STAGE: each flux tile reaches the first pass through `stage` in LDS, the next one streaming in
// while the current one transforms (the caller has started component 0's copy).
template <typename T, int N, int TX, int NT, bool SSP, bool SF, int DA = DEALIAS_TWO_THIRDS, bool STAGE = false>
__device__ __forceinline__ void zf_scalar_part(cx<T> *buf, const cx<T> *twl, const cx<T> *__restrict__ pf,
                                               const cx<T> *__restrict__ uh, cx<T> *__restrict__ ph,
                                               cx<T> *__restrict__ hp, const Geo &g, const StageArgs &sa, int kx0,
                                               int ky_idx, int my, int box, cx<T> *stage = nullptr) {
  constexpr int K = band_k(N, DA);
  constexpr int RL = last_radix<N>();
  using SL = PassShape<N, RL, TX, NT>;
  constexpr int PER = SL::PER;
  constexpr int QS = RL % 2 == 0 ? 2 : RL;   // modes whose state loads are in flight together
  static_assert(RL % QS == 0, "the load groups tile the last pass's outputs");
  const int tid = threadIdx.x;
  const size_t comp = (size_t)K * K * g.kxp;
  const size_t pencil = (size_t)N * K * g.kxp;
  const size_t zstride = (size_t)K * g.kxp;
  cx<T> G[PER][RL];
#pragma unroll
  for (int p = 0; p < PER; ++p)
#pragma unroll
    for (int q = 0; q < RL; ++q) G[p][q] = czero<T>();
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __builtin_amdgcn_sched_barrier(0);   // one component's loads in flight at a time
    const cx<T> *src = pf + (size_t)(3 + c) * pencil + (size_t)ky_idx * g.kxp + kx0;
    auto accumulate = [&](int p, int q, int n, int col, cx<T> v) {
      const int mc = c == 0 ? kx0 + col : (c == 1 ? my : signed_of(idx_of_bin(n, g), g));
      const T kc = T(g.k0 * mc);
      G[p][q] = {G[p][q].re + kc * v.re, G[p][q].im + kc * v.im};
    };
    if constexpr (STAGE) {
      // the copy has landed: direct-to-LDS loads count in vmcnt until their data are in LDS, and
      // the other threads' pieces are seen after the barrier
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      fft_io_hook<T, N, TX, NT, -1, true>(
          buf, twl, [&](int, int, int n, int col) { return stage[n * TX + col]; }, accumulate,
          [&]() {   // every thread has read the stage: the next flux streams in during this transform
            if (c < 2)
              stage_tile<T, N, TX, NT>(stage, pf + (size_t)(4 + c) * pencil + (size_t)ky_idx * g.kxp + kx0, zstride);
          });
    } else {
      fft_io<T, N, TX, NT, -1, true, false>(
          buf, twl, [&](int, int, int n, int col) { return src[(size_t)n * zstride + col]; }, accumulate);
    }
  }
  const T Gm = T(sa.mean_gradient);
  const double sub = stage_sub(sa);
  const T dt = T(stage_dt(sa)), a = T(sa.a), b = T(sa.b);
  const double coef_p = -sa.kappa[box] * sub;
  const T Aphi = SF ? T(sa.scalars[8 * box + 7]) : T(0);
  __builtin_amdgcn_sched_barrier(0);   // the state loads stay behind the transforms
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int bb = tid + p * NT;
    if (!(SL::TOT % NT == 0 || bb < SL::TOT)) continue;
    const int col = bb % TX, j = bb / TX;
    const int mx = kx0 + col;
#pragma unroll
    for (int q0 = 0; q0 < RL; q0 += QS) {
      // the group's loads first, unconditionally (a mode outside the ball, which is zero and
      // stays zero, reads mode 0 and is not stored), so their latencies overlap
      int at[QS];
      cx<T> uz[QS], pv[QS], hv[QS];
#pragma unroll
      for (int i = 0; i < QS; ++i) {
        const int kz_idx = idx_of_bin(j + (q0 + i) * SL::NB, g);
        const int mz = kz_idx < 0 ? 0 : signed_of(kz_idx, g);
        at[i] = kz_idx >= 0 && keeps(mx * mx + my * my + mz * mz, g)
                    ? kz_idx * (int)zstride + ky_idx * g.kxp + mx
                    : -1;
        const int ld = at[i] < 0 ? 0 : at[i];
        uz[i] = uh[2 * comp + ld];   // the stage velocity: the velocity update comes after
        pv[i] = ph[ld];
      }
      if (sa.stage > 0) {
#pragma unroll
        for (int i = 0; i < QS; ++i) hv[i] = hp[at[i] < 0 ? 0 : at[i]];
      } else {
#pragma unroll
        for (int i = 0; i < QS; ++i) hv[i] = czero<T>();
      }
#pragma unroll
      for (int i = 0; i < QS; ++i) {
        if (at[i] < 0) continue;
        const int q = q0 + i;
        const int mz = signed_of(idx_of_bin(j + q * SL::NB, g), g);
        const double kxd = g.k0 * mx, kyd = g.k0 * my, kzd = g.k0 * mz;
        const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
        cx<T> kF = G[p][q];   // k . F
        if constexpr (DA == DEALIAS_PHASE_SHIFT) {   // the two-grid average, the displaced term shifted back
          const cx<T> *sh = reinterpret_cast<const cx<T> *>(sa.shifted) + (size_t)box * g.ncf * comp + at[i];
          const cx<T> back = half_cell_phase<T>(mx + my + mz, N, -1);
          const cx<T> ks = cmul(scale(sh[3 * comp], T(g.k0 * mx)) + scale(sh[4 * comp], T(g.k0 * my)) +
                                    scale(sh[5 * comp], T(g.k0 * mz)),
                                back);
          kF = scale(kF + ks, T(0.5));
        }
        if constexpr (DA == DEALIAS_HOU_LI) kF = scale(kF, T(sa.filter[mx * mx + my * my + mz * mz]));
        cx<T> nl = cx<T>{kF.im, -kF.re} - scale(uz[i], Gm);   // -i k.F - G u_z
        double xe = coef_p * k2d;   // the scalar's integrating-factor exponent over the substep
        if constexpr (SF) {
          if (in_band(mx * mx + my * my + mz * mz, sa.sband)) {   // src/Forcing.jl:1104-1117, :600-642
            nl = nl + scale(pv[i], Aphi);
            xe += sa.slin_rate * sub;
          }
        }
        if constexpr (SSP) {
          const cx<T> t = scale(pv[i] + scale(nl, dt), T(exp(sa.pb * xe)));
          if (sa.stage == 0) {
            ph[at[i]] = t;
            hp[at[i]] = pv[i];
          } else {
            ph[at[i]] = scale(t, T(sa.cb)) + scale(scale(hv[i], T(sa.ca)), T(exp(sa.pa * xe)));
          }
          continue;
        }
        const T E = T(exp(xe));
        cx<T> h = scale(nl, dt);
        if (sa.stage > 0) h = scale(hv[i], a) + h;
        ph[at[i]] = scale(pv[i] + scale(h, b), E);
        if (!sa.last) hp[at[i]] = scale(h, E);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Every spectral operation of one stage (src/Propagator.jl:338-392), fused:
//   N_u  = P mask FFT_z(PF_u), mean mode zeroed, + A 1_band u        (src/Generator.jl:484-534,
//                                                                      src/Forcing.jl:1104-1117)
//   N_phi = -mask i k . FFT_z(PF_flux) - G u_z                        (src/Generator.jl:648-730)
//   h <- a h + dt N;  q <- E (q + b h);  h <- E h  (not at the last)  (src/Propagator.jl:366-391)
//   or, for IFSSPRK3 and IFHeun, q <- cb E^pb (q + dt N) + E^pa (ca h), h = u_n
//                                                                      (src/Propagator.jl:456-517, :599-669)
//   and at the last stage the state Leray projection                  (src/Propagator.jl:401, :677).
// E = exp(coef |k|^2) is evaluated in FP64 per mode and rounded to T.  A thread owns the modes
// its last transform pass produces: kz bins n = j + q N/R of column col, for its butterflies
// (col, j); the product tiles stream from HBM straight into the first pass.  FL: the first two
// transformed velocity products wait in LDS (retained bins only, [kz_idx][col]) instead of
// registers, so the stage update holds only the third and more state loads can be in flight;
// each thread reads back only what it wrote itself.
template <int N, int TX, typename T, int DA = DEALIAS_TWO_THIRDS>
constexpr size_t zf_lds_bytes(bool fl) {
  return ((size_t)N * TX + N + (fl ? 2 * (size_t)band_k(N, DA) * TX : 0)) * sizeof(cx<T>);
}

template <typename T, int N, int TX, int NT, bool SCALAR, bool FL = false, bool SSP = false, int DA = DEALIAS_TWO_THIRDS>
__global__ __launch_bounds__(NT) void k_zf(const cx<T> *__restrict__ pf, cx<T> *__restrict__ uh,
                                           cx<T> *__restrict__ hu, cx<T> *__restrict__ ph,
                                           cx<T> *__restrict__ hp, Geo g, const cx<double> *__restrict__ tw,
                                           StageArgs sa) {
  constexpr int K = band_k(N, DA);
  constexpr int RL = last_radix<N>();
  using SL = PassShape<N, RL, TX, NT>;
  constexpr int PER = SL::PER;
  constexpr int QG = RL % 2 != 0 ? RL : (PER > 1 ? 2 : (RL < 4 ? RL : 4));   // state loads in flight together
  static_assert(RL % QG == 0, "the load groups tile the last pass's outputs");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  const int tid = threadIdx.x;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, box = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  if (!in_disk(kx0, my, g)) return;   // no column of this tile meets the ball (its PF lines are unwritten)
  stage_twiddles<T, N, NT>(twl, tw);
  const T ky = T(g.k0 * my);
  const size_t comp = (size_t)K * K * g.kxp;           // one compact component
  const size_t pencil = (size_t)N * K * g.kxp;         // one PF component
  const size_t zstride = (size_t)K * g.kxp;
  pf += box * g.box_b;
  uh += box * 3 * comp;
  hu += box * 3 * comp;
  if constexpr (SCALAR) {
    ph += box * comp;
    hp += box * comp;
  }

  if constexpr (SCALAR)
    zf_scalar_part<T, N, TX, NT, SSP, false, DA>(buf, twl, pf, uh, ph, hp, g, sa, kx0, ky_idx, my, box);

  // velocity products, forward z transform, straight into the owned modes.  The compiler
  // memory barrier keeps the scalar part's u_z loads from living on into the velocity part.
  __asm__ volatile("" ::: "memory");
  cx<T> F[FL ? 1 : 3][PER][RL];
  cx<T> *fl = twl + N;   // FL: [2][K][TX]
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __builtin_amdgcn_sched_barrier(0);   // one component's loads in flight at a time
    const cx<T> *src = pf + (size_t)c * pencil + (size_t)ky_idx * g.kxp + kx0;
    fft_io<T, N, TX, NT, -1, true, false>(
        buf, twl,
        [&](int, int, int n, int col) { return src[(size_t)n * zstride + col]; },
        [&](int p, int q, int n, int col, cx<T> v) {
          if constexpr (FL) {
            if (c < 2) {
              const int kz_idx = idx_of_bin(n, g);
              if (kz_idx >= 0) fl[(c * K + kz_idx) * TX + col] = v;
            } else {
              F[0][p][q] = v;
            }
          } else {
            F[c][p][q] = v;
          }
        });
  }
  // the transformed product c of mode (p, q) of this thread, kz_idx its compact index
  auto Fv = [&](int c, int p, int q, int kz_idx, int col) -> cx<T> {
    if constexpr (FL) return c < 2 ? fl[(c * K + kz_idx) * TX + col] : F[0][p][q];
    else return F[c][p][q];
  };

  const T A = sa.band.on ? T(sa.scalars[8 * box]) : T(0);
  const double sub = stage_sub(sa);
  const T a = T(sa.a), b = T(sa.b), dt = T(stage_dt(sa));
  const double coef_u = -sa.visc[box] * sub;
  __builtin_amdgcn_sched_barrier(0);   // the state loads stay behind the transforms
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int bb = tid + p * NT;
    if (!(SL::TOT % NT == 0 || bb < SL::TOT)) continue;
    const int col = bb % TX, j = bb / TX;
    const int mx = kx0 + col;
#pragma unroll
    for (int q0 = 0; q0 < RL; q0 += QG) {
      // the group's loads first, unconditionally (a mode outside the ball, which is zero and
      // stays zero, reads mode 0 and is not stored), so their latencies overlap
      int at[QG];
      cx<T> u[QG][3], hv[QG][3];
#pragma unroll
      for (int i = 0; i < QG; ++i) {
        const int kz_idx = idx_of_bin(j + (q0 + i) * SL::NB, g);
        const int mz = kz_idx < 0 ? 0 : signed_of(kz_idx, g);
        at[i] = kz_idx >= 0 && keeps(mx * mx + my * my + mz * mz, g)
                    ? kz_idx * (int)zstride + ky_idx * g.kxp + mx
                    : -1;
        const int ld = at[i] < 0 ? 0 : at[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[i][c] = uh[c * comp + ld];
      }
      if (sa.stage > 0) {
#pragma unroll
        for (int i = 0; i < QG; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) hv[i][c] = hu[c * comp + (at[i] < 0 ? 0 : at[i])];
      } else {
#pragma unroll
        for (int i = 0; i < QG; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) hv[i][c] = czero<T>();
      }
#pragma unroll
      for (int i = 0; i < QG; ++i) {
        if (at[i] < 0) continue;
        const int q = q0 + i;
        const int kz_idx = idx_of_bin(j + q * SL::NB, g);
        const int mz = signed_of(kz_idx, g);
        const int m2 = mx * mx + my * my + mz * mz;
        const T kx = T(g.k0 * mx), kz = T(g.k0 * mz);
        const double kxd = g.k0 * mx, kyd = g.k0 * my, kzd = g.k0 * mz;
        const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
        cx<T> nl[3];
        if (m2 > 0) {
          const T k2 = T(k2d);
          cx<T> f0 = Fv(0, p, q, kz_idx, col), f1 = Fv(1, p, q, kz_idx, col), f2 = Fv(2, p, q, kz_idx, col);
          if constexpr (DA == DEALIAS_PHASE_SHIFT) {   // the two-grid average, the displaced term shifted back
            const cx<T> *sh = reinterpret_cast<const cx<T> *>(sa.shifted) + (size_t)box * g.ncf * comp + at[i];
            const cx<T> back = half_cell_phase<T>(mx + my + mz, N, -1);
            f0 = scale(f0 + cmul(sh[0], back), T(0.5));
            f1 = scale(f1 + cmul(sh[comp], back), T(0.5));
            f2 = scale(f2 + cmul(sh[2 * comp], back), T(0.5));
          }
          const cx<T> kf = {kx * f0.re + ky * f1.re + kz * f2.re, kx * f0.im + ky * f1.im + kz * f2.im};
          const cx<T> s = {kf.re / k2, kf.im / k2};
          nl[0] = f0 - scale(s, kx);
          nl[1] = f1 - scale(s, ky);
          nl[2] = f2 - scale(s, kz);
          if constexpr (DA == DEALIAS_HOU_LI) {
            const T F = T(sa.filter[m2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) nl[c] = scale(nl[c], F);
          }
        } else {
          nl[0] = nl[1] = nl[2] = czero<T>();
        }
        const bool banded = in_band(m2, sa.band);
        if (banded) {
#pragma unroll
          for (int c = 0; c < 3; ++c) nl[c] = nl[c] + scale(u[i][c], A);
        }
        // Linear forcing rides the integrating factor: exp((A - nu |k|^2) d_i dt) in the band
        // (src/Generator.jl:565-581, src/Forcing.jl:600-640)
        const double xe = coef_u * k2d + (banded ? sa.lin_rate * sub : 0.0);
        cx<T> un[3];
        if constexpr (SSP) {
          const T Eb = T(exp(sa.pb * xe)), Ea = sa.stage > 0 ? T(exp(sa.pa * xe)) : T(0);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const cx<T> t = scale(u[i][c] + scale(nl[c], dt), Eb);
            if (sa.stage == 0) {
              un[c] = t;
              hu[c * comp + at[i]] = u[i][c];
            } else {
              un[c] = scale(t, T(sa.cb)) + scale(scale(hv[i][c], T(sa.ca)), Ea);
            }
          }
        } else {
          const T E = T(exp(xe));
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            cx<T> h = scale(nl[c], dt);
            if (sa.stage > 0) h = scale(hv[i][c], a) + h;
            un[c] = scale(u[i][c] + scale(h, b), E);
            if (!sa.last) hu[c * comp + at[i]] = scale(h, E);
          }
        }
        if (sa.last && m2 > 0) {  // helmholtz_projection!(u, u): P without the mask
          const T k2 = T(k2d);
          const cx<T> kf = {kx * un[0].re + ky * un[1].re + kz * un[2].re,
                            kx * un[0].im + ky * un[1].im + kz * un[2].im};
          const cx<T> s = {kf.re / k2, kf.im / k2};
          un[0] = un[0] - scale(s, kx);
          un[1] = un[1] - scale(s, ky);
          un[2] = un[2] - scale(s, kz);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) uh[c * comp + at[i]] = un[c];
      }
      // the next group's loads stay behind this group's stores: hoisting them all costs occupancy
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The scalar half on its own (the velocity half then runs as k_zf without SCALAR): two smaller
// kernels hold fewer registers than the fused one.
// STAGE: the flux tiles stream through a third LDS tile (zs_lds_bytes), the first one's copy
// started before the twiddles are staged.
template <int N, int TX, typename T>
constexpr size_t zs_lds_bytes(bool stage) {
  return ((size_t)N * TX + N + (stage ? (size_t)N * TX : 0)) * sizeof(cx<T>);
}

template <typename T, int N, int TX, int NT, bool SSP = false, bool SF = false, int DA = DEALIAS_TWO_THIRDS,
          bool STAGE = false>
__global__ __launch_bounds__(NT) void k_zs(const cx<T> *__restrict__ pf, const cx<T> *__restrict__ uh,
                                           cx<T> *__restrict__ ph, cx<T> *__restrict__ hp, Geo g,
                                           const cx<double> *__restrict__ tw, StageArgs sa) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  cx<T> *stage = twl + N;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, box = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  if (!in_disk(kx0, my, g)) return;   // no column of this tile meets the ball (its PF lines are unwritten)
  const cx<T> *pfb = pf + box * g.box_b;
  if constexpr (STAGE)
    stage_tile<T, N, TX, NT>(stage, pfb + (size_t)3 * N * g.k * g.kxp + (size_t)ky_idx * g.kxp + kx0,
                             (size_t)g.k * g.kxp);
  stage_twiddles<T, N, NT>(twl, tw);
  const size_t comp = comp_of(g);
  zf_scalar_part<T, N, TX, NT, SSP, SF, DA, STAGE>(buf, twl, pfb, uh + box * 3 * comp, ph + box * comp,
                                                   hp + box * comp, g, sa, kx0, ky_idx, my, box, stage);
}

// The inverse z pass of the state: u, w = i k x u and phi (with the 1/N^3 of the inverse
// transform folded in) into Q, plus the per-block FP64 partial sums every later consumer needs:
// [0] sum w|u|^2, [1] band sum w|u|^2, [2] sum w |k|^2 |u|^2, [3] sum w|phi|^2, [4] sum w |k|^2 |phi|^2
// (w the Parseval weight of the half spectrum).  This is to_physical! + vorticity_physical!'s
// spectral half (src/Propagator.jl:349-363, src/Observables.jl:329-333) and the state side of
// weighted_inner (src/Forcing.jl:1109, src/Observables.jl:150-173).  A thread owns the modes its
// first transform pass reads (kz bins n = j + q N/R of column col) and streams the transformed
// columns from the last pass straight to HBM.
constexpr int NRED = 5;

// SHIFT (PhaseShift's displaced evaluation, src/Generator.jl:505-513): the state is first
// translated by half a cell along every axis, so Q holds u, w and phi sampled on the displaced grid.
// WIDE: the last pass's outputs go to a second LDS tile, and the finished tile is written to Q 16
// bytes per lane (two FP32 values of a row), half the store instructions of the 8-byte stores
// straight from the last pass.
template <int N, int TX, typename T>
constexpr size_t zi_lds_bytes(bool wide) {
  return ((size_t)N * TX + N + (wide ? (size_t)N * TX : 0)) * sizeof(cx<T>);
}

template <typename T, int N, int TX, int NT, bool SCALAR, int DA = DEALIAS_TWO_THIRDS, bool SHIFT = false,
          bool WIDE = false>
__global__ __launch_bounds__(NT) void k_zi(const cx<T> *__restrict__ uh, const cx<T> *__restrict__ ph,
                                           cx<T> *__restrict__ q, Geo g, const cx<double> *__restrict__ tw,
                                           BandSpec band, double *__restrict__ partials) {
  constexpr int M = band_m(N, DA), K = band_k(N, DA);
  constexpr int R1 = first_radix<N>();
  using S1 = PassShape<N, R1, TX, NT>;
  constexpr int PER = S1::PER;
  constexpr int NCIN = SCALAR ? 7 : 6;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *twl = buf + N * TX;
  __shared__ double red[NT / 64][NRED];
  const int tid = threadIdx.x;
  const int kx0 = blockIdx.x * TX, ky_idx = blockIdx.y, box = blockIdx.z;
  const int my = signed_of(ky_idx, g);
  // the box's partials follow those of the boxes before it
  const size_t part = ((size_t)box * gridDim.y * gridDim.x + (size_t)blockIdx.y * gridDim.x + blockIdx.x) * NRED;
  if (!in_disk(line_start<T>(kx0), my, g)) {   // the tile's line misses the ball: never read
    if (tid < NRED) partials[part + tid] = 0.0;
    return;
  }
  uh += box * 3 * comp_of(g);
  if constexpr (SCALAR) ph += box * comp_of(g);
  q += box * g.box_a;
  stage_twiddles<T, N, NT>(twl, tw);
  const size_t comp = (size_t)K * K * g.kxp;
  const size_t pencil = (size_t)N * K * g.kxp;
  const size_t zstride = (size_t)K * g.kxp;
  const T inv = T(1.0 / ((double)N * N * N));

  cx<T> u[3][PER][R1];
  cx<T> p[PER][R1];
  double acc[NRED] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int pp = 0; pp < PER; ++pp) {
    const int bb = tid + pp * NT;
    const bool live = S1::TOT % NT == 0 || bb < S1::TOT;
    const int col = bb % TX, j = bb / TX;
    const int mx = kx0 + col;
#pragma unroll
    for (int qq = 0; qq < R1; ++qq) {
      const int kz_idx = live ? idx_of_bin(j + qq * S1::NB, g) : -1;
      const int mz = kz_idx < 0 ? 0 : signed_of(kz_idx, g);
      if (kz_idx < 0 || !keeps(mx * mx + my * my + mz * mz, g)) {   // zero outside the ball
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c][pp][qq] = czero<T>();
        p[pp][qq] = czero<T>();
        continue;
      }
      const size_t at = (size_t)kz_idx * zstride + (size_t)ky_idx * g.kxp + mx;
      const double kxd = g.k0 * mx, kyd = g.k0 * my, kzd = g.k0 * mz;
      const double k2d = (kxd * kxd + kyd * kyd) + kzd * kzd;
      const double w = parseval_w(mx, N);
      double e2 = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        u[c][pp][qq] = uh[c * comp + at];
        e2 += (double)u[c][pp][qq].re * (double)u[c][pp][qq].re + (double)u[c][pp][qq].im * (double)u[c][pp][qq].im;
      }
      if constexpr (SHIFT) {
        const cx<T> sh = half_cell_phase<T>(mx + my + mz, N, +1);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c][pp][qq] = cmul(u[c][pp][qq], sh);
      }
      acc[0] += w * e2;
      if (in_band(mx * mx + my * my + mz * mz, band)) acc[1] += w * e2;
      acc[2] += w * k2d * e2;
      if constexpr (SCALAR) {
        p[pp][qq] = ph[at];
        if constexpr (SHIFT) p[pp][qq] = cmul(p[pp][qq], half_cell_phase<T>(mx + my + mz, N, +1));
        const double v2 = (double)p[pp][qq].re * (double)p[pp][qq].re + (double)p[pp][qq].im * (double)p[pp][qq].im;
        acc[3] += w * v2;
        acc[4] += w * k2d * v2;
      } else {
        p[pp][qq] = czero<T>();
      }
    }
  }

  // buffer stores need every offset of a Q component below 2^31 bytes (kxp is at most M+1
  // rounded up to 16); beyond that (FP64 at 1024^3) the stores take 64-bit addresses
  constexpr bool BUFQ = (size_t)N * K * ((M + 16) / 16 * 16) * sizeof(cx<T>) < 0x7fff0000u;
  cx<T> *obuf = twl + N;   // WIDE: the transformed tile [n][TX]
#pragma unroll
  for (int c = 0; c < NCIN; ++c) {
    cx<T> *qc = q + (size_t)c * pencil + (size_t)ky_idx * g.kxp + kx0;
    const Buf<T> dst(qc, pencil);
    fft_io<T, N, TX, NT, +1, true, false>(
        buf, twl,
        [&](int pp, int qq, int n, int col) {
          cx<T> v;
          if (c < 3) {
            v = u[c][pp][qq];
          } else if (c < 6) {  // i k x u
            const int mx = kx0 + col, mz = signed_of(idx_of_bin(n, g), g);
            const T kx = T(g.k0 * mx), ky = T(g.k0 * my), kz = T(g.k0 * mz);
            cx<T> t;
            if (c == 3) t = scale(u[2][pp][qq], ky) - scale(u[1][pp][qq], kz);
            else if (c == 4) t = scale(u[0][pp][qq], kz) - scale(u[2][pp][qq], kx);
            else t = scale(u[1][pp][qq], kx) - scale(u[0][pp][qq], ky);
            v = {-t.im, t.re};
          } else {
            v = p[pp][qq];
          }
          return scale(v, inv);
        },
        [&](int, int, int n, int col, cx<T> v) {
          if constexpr (WIDE) obuf[n * TX + col] = v;
          else if constexpr (BUFQ) dst.store(v, (uint32_t)(n * (int)zstride + col), 0u);
          else qc[(size_t)n * zstride + col] = v;
        });
    if constexpr (WIDE) {
      // the last pass ended with a barrier: the tile is complete.  Piece i of a row is its 16
      // bytes at column V i; the next component's last pass writes obuf only after the barriers
      // of its first and middle passes, by which time these reads have returned.
      constexpr int V = 16 / (int)sizeof(cx<T>), PR = TX / V, PIECES = N * PR;
      static_assert(TX % V == 0, "a tile row holds whole 16-byte pieces");
      using P16 = decltype(__builtin_amdgcn_raw_buffer_load_b128(dst.r, 0u, 0u, 0));
#pragma unroll
      for (int i = threadIdx.x; i < PIECES; i += NT) {
        const int n = i / PR, col = (i % PR) * V;
        const P16 v = *reinterpret_cast<const P16 *>(obuf + n * TX + col);
        if constexpr (BUFQ)
          __builtin_amdgcn_raw_buffer_store_b128(v, dst.r, (uint32_t)(n * (int)zstride + col) * (uint32_t)sizeof(cx<T>),
                                                 0u, 0);
        else
          *reinterpret_cast<P16 *>(qc + (size_t)n * zstride + col) = v;
      }
    }
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

// reality_projection! on the kx = 0 plane of u_hat and phi_hat, and on the kx = N/2 plane when the
// strategy keeps it (src/Propagator.jl:411-412): every conjugate pair becomes (f(k) + conj f(-k))/2;
// a self-paired mode keeps its real part.  Grid: (K^2 / 256, boxes, planes).
template <typename T>
__global__ void k_reality(cx<T> *__restrict__ uh, cx<T> *__restrict__ ph, Geo g, int ncomp) {
  const int K = g.k;
  const int l = blockIdx.x * blockDim.x + threadIdx.x, box = blockIdx.y;
  if (l >= K * K) return;
  const int plane = blockIdx.z == 0 ? 0 : g.n / 2;   // the kx column of the plane
  uh += box * 3 * comp_of(g);
  if (ncomp > 3) ph += box * comp_of(g);
  const int kz_idx = l / K, ky_idx = l % K;
  const int pz = (K - kz_idx) % K, py = (K - ky_idx) % K;
  const int lp = pz * K + py;
  if (lp < l) return;
  const size_t comp = (size_t)K * K * g.kxp;
  const size_t a0 = (size_t)kz_idx * K * g.kxp + (size_t)ky_idx * g.kxp + plane;
  const size_t a1 = (size_t)pz * K * g.kxp + (size_t)py * g.kxp + plane;
  for (int c = 0; c < ncomp; ++c) {
    cx<T> *f = c < 3 ? uh + c * comp : ph;
    const cx<T> x = f[a0], y = f[a1];
    const cx<T> nx = {T(0.5) * (x.re + y.re), T(0.5) * (x.im - y.im)};
    f[a0] = nx;
    if (lp != l) f[a1] = conj(nx);
  }
}

// Neumaier-compensated FP64 sum of the per-block partials, then the ConstantPower amplitude of
// the state they measured (src/Forcing.jl:1091-1117, the empty-band clamp at :110-111).
// scalars: [0] amplitude A, [1] <u,u>, [2] <u,1_band u>, [3] <u,|k|^2 u>, [4] <phi,phi>,
//          [5] <phi,|k|^2 phi>   (normalised inner products)
struct FinalizeArgs {
  int nblocks;
  double inv_norm;       // 1/N^6
  int forcing;           // the amplitude A: 0 none, 1 ConstantPower, 2 VarianceController
  double injection_rate;
  double eps_t;          // eps of the run's element type
  double target, gain, turnover_time;   // VarianceController k_inf, G, t_l
  const double *visc;    // per box: nu
};

static __global__ void k_finalize(const double *__restrict__ partials, double *__restrict__ scalars, FinalizeArgs fa) {
  __shared__ double ss[256][NRED], cc[256][NRED];
  const int tid = threadIdx.x, box = blockIdx.x;   // one block per box
  partials += (size_t)box * fa.nblocks * NRED;
  scalars += 8 * box;
  double s[NRED] = {0, 0, 0, 0, 0}, c[NRED] = {0, 0, 0, 0, 0};
  for (int b = tid; b < fa.nblocks; b += 256) {
#pragma unroll
    for (int i = 0; i < NRED; ++i) {
      const double x = partials[(size_t)b * NRED + i];
      const double t = s[i] + x;
      c[i] += fabs(s[i]) >= fabs(x) ? (s[i] - t) + x : (x - t) + s[i];
      s[i] = t;
    }
  }
#pragma unroll
  for (int i = 0; i < NRED; ++i) {
    ss[tid][i] = s[i];
    cc[tid][i] = c[i];
  }
  __syncthreads();
  if (tid < NRED) {
    double S = 0.0, C = 0.0;
    for (int t = 0; t < 256; ++t) {
      const double x = ss[t][tid];
      const double u = S + x;
      C += fabs(S) >= fabs(x) ? (S - u) + x : (x - u) + S;
      S = u;
      C += cc[t][tid];
    }
    scalars[1 + tid] = (S + C) * fa.inv_norm;
  }
  __syncthreads();
  if (tid == 0) {
    // ConstantPower: A = eps_in / (2 k_band) (src/Forcing.jl:1091-1117); VarianceController:
    // A = (eps - G (k - k_inf) / t_l) / (2 k_band) with eps = nu <u, |k|^2 u> (src/Forcing.jl:1041-1065);
    // both 0 for an empty band (src/Forcing.jl:110-111, :1093-1103)
    double A = 0.0;
    if (fa.forcing) {
      const double band = 0.5 * scalars[2], whole = 0.5 * scalars[1];
      const bool empty = !(isfinite(band) && band > 10.0 * fa.eps_t * whole);
      double num = fa.injection_rate;
      if (fa.forcing == 2) num = fa.visc[box] * scalars[3] - fa.gain * (whole - fa.target) / fa.turnover_time;
      A = empty ? 0.0 : num / (2.0 * band);
    }
    scalars[0] = A;
  }
}

// Shell-binned spectrum of the compact state (src/Observables.jl:502-514): shell n >= 0 of a mode
// is the integer with n - 1/2 <= |m| < n + 1/2, i.e. (2n - 1)^2 <= 4|m|^2 < (2n + 1)^2.  Each
// block sums one shell over a fixed share of its (my, mz) rows: along a row the shell's modes are
// one run of kx, found from integer bounds.  A thread adds its rows in order and the block adds
// its threads in a fixed tree, so every partial, and the spectrum, is the same to the last bit on
// every call.  Parseval weights: 1 on the kx = 0 and kx = N/2 planes, 2 inside.
constexpr int SPEC_CHUNKS = 16;    // blocks per shell
constexpr int SPEC_THREADS = 256;

template <typename T>
__global__ void __launch_bounds__(SPEC_THREADS) k_spectrum_shells(const cx<T> *__restrict__ f, int ncomp, Geo g,
                                                                   double *__restrict__ partials) {
  const int n = blockIdx.x, s = blockIdx.y, t = threadIdx.x;
  // the stored signed indices run from -(K - 1 - M) to M (under HouLi the Nyquist index is +N/2 only)
  const int hi_m = min(n, g.m), lo_m = -min(n, g.k - 1 - g.m), side = hi_m - lo_m + 1;
  const long rows = (long)side * side;
  const int hi2 = (2 * n + 1) * (2 * n + 1), lo2 = n > 0 ? (2 * n - 1) * (2 * n - 1) : 0;
  const size_t comp = comp_of(g);
  double acc = 0.0;
  for (long j = s + (long)SPEC_CHUNKS * t; j < rows; j += (long)SPEC_CHUNKS * SPEC_THREADS) {
    const int my = (int)(j % side) + lo_m, mz = (int)(j / side) + lo_m;
    const int r4 = 4 * (my * my + mz * mz);
    if (r4 >= hi2) continue;
    int lo = 0;
    if (r4 < lo2) {   // the least mx with 4 (mx^2 + r^2) >= (2n - 1)^2
      lo = (int)sqrt((double)(lo2 - r4) * 0.25);
      while (4 * lo * lo + r4 < lo2) ++lo;
      while (lo > 0 && 4 * (lo - 1) * (lo - 1) + r4 >= lo2) --lo;
    }
    int hi = (int)sqrt((double)(hi2 - r4) * 0.25);   // the greatest mx with 4 (mx^2 + r^2) < (2n + 1)^2
    while (hi >= 0 && 4 * hi * hi + r4 >= hi2) --hi;
    while (4 * (hi + 1) * (hi + 1) + r4 < hi2) ++hi;
    hi = min(hi, g.m);
    const int ky_idx = my >= 0 ? my : my + g.k, kz_idx = mz >= 0 ? mz : mz + g.k;
    const size_t row = ((size_t)kz_idx * g.k + ky_idx) * g.kxp;
    for (int mx = lo; mx <= hi; ++mx) {
      double e2 = 0.0;
      for (int c = 0; c < ncomp; ++c) {
        const cx<T> v = f[c * comp + row + mx];
        e2 += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
      }
      acc += (mx == 0 || 2 * mx == g.n) ? e2 : 2.0 * e2;
    }
  }
  __shared__ double red[SPEC_THREADS];
  red[t] = acc;
  __syncthreads();
  for (int off = SPEC_THREADS / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) partials[(size_t)n * SPEC_CHUNKS + s] = red[0];
}

// bins[n] = the shell's partials added in chunk order.
static __global__ void k_spectrum_sum(const double *__restrict__ partials, int nbins, double *__restrict__ bins) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nbins) return;
  double s = 0.0;
  for (int c = 0; c < SPEC_CHUNKS; ++c) s += partials[(size_t)n * SPEC_CHUNKS + c];
  bins[n] = s;
}

// The frame tables as the device computes them: signed indices of the three axes and the mask
// over the reference's full half-spectrum layout [N][N][N/2+1].
static __global__ void k_frame_tables(int n, int ball9, int *mx, int *my, int *mz, unsigned char *mask) {
  const size_t hx = n / 2 + 1;
  const size_t total = (size_t)n * n * hx;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < total; l += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(l % hx), iy = (int)((l / hx) % n), iz = (int)(l / (hx * n));
    const int sx = ix, sy = iy <= n / 2 ? iy : iy - n, sz = iz <= n / 2 ? iz : iz - n;
    if (iy == 0 && iz == 0) mx[ix] = sx;
    if (ix == 0 && iz == 0) my[iy] = sy;
    if (ix == 0 && iy == 0) mz[iz] = sz;
    mask[l] = 9 * (sx * sx + sy * sy + sz * sz) <= ball9 ? 1 : 0;
  }
}

static __global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static __global__ void k_fill_f64(double *p, double v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace hdns
