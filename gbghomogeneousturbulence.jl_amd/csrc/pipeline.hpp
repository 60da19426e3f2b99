// Launch shapes of the stage kernels for one (element type, grid size) pair.
//
// Tile width TX: 16 complex FP32 or 8 complex FP64 = 128 contiguous bytes per z or y row of a
// tile.  The column buffer of a y or z block is N*TX complex values (64 KiB at N = 512).
#pragma once
#include <algorithm>
#include <type_traits>

#include "adjoint.hpp"
#include "context.hpp"
#include "observe.hpp"
#include "pairs.hpp"
#include "sort.hpp"
#include "tracers.hpp"
#include "xpass_wave.hpp"

namespace hdns {

template <typename T>
constexpr int tile_width_of() { return sizeof(T) == 4 ? 16 : 8; }

// Launch schedules per stage kernel family (two; three for the z forward, four for the z inverse),
// chosen per context by hdns_autotune:
//   y passes: 256 or 512 threads per block (512 or 1024 from N = 512 on);
//   z forward: (0) the scalar half and the velocity half as two kernels on 64*TX threads, one
//              last-pass butterfly per thread at N = 512, the first two transformed velocity
//              products staged in LDS where they fit (N <= 512); or (1) one fused kernel on 512;
//              or (2) schedule 0 with the scalar half's flux tiles streamed through LDS;
//   z inverse: a TX-wide tile on 64*TX threads (one first-pass butterfly per thread at N = 512)
//              or on 32*TX threads (two), or (2) schedule 0's tile storing through LDS 16 bytes
//              per lane, or (3) FP32 on column pairs (pairs.hpp);
//   x pass: 256 or 512 threads per block, or for the register-resident pass (FP32 at N = 256,
//           512 and 1024, FP64 at N = 256 and 512) the next component staged in LDS by
//           direct-to-LDS loads, or each component loaded into registers just before its
//           transform;
//   y radix plan: fft_io's plan or AltPlan's, led by a radix-16 pass.
// At N >= 1024 the z schedules use narrower tiles (below).
constexpr int kYThreads[2] = {256, 512};
constexpr int kXThreads[2] = {256, 512};
template <typename T>
struct ZSched {
  static constexpr int zf_nt[2] = {64 * tile_width_of<T>(), 512};
  static constexpr int zi_tx[2] = {tile_width_of<T>(), tile_width_of<T>()};
  static constexpr int zi_nt[2] = {64 * tile_width_of<T>(), 32 * tile_width_of<T>()};
};

// DA: the dealiasing strategy (DEALIAS_TWO_THIRDS, _HOU_LI, _PHASE_SHIFT), which fixes the retained
// cube (M, K) and the nonlinear terms' dealiasing in the z forward pass.  The adjoint and the
// strain-determinant diagnostics are built for the 2/3 rule only.
template <typename T, int N, int DA = DEALIAS_TWO_THIRDS>
struct PipelineImpl final : Pipeline {
  static constexpr int TX = tile_width_of<T>();
  static constexpr int M = band_m(N, DA), K = band_k(N, DA);
  static constexpr bool TWO_THIRDS = DA == DEALIAS_TWO_THIRDS;
  // the FP32 z passes on column pairs (pairs.hpp): 2/3 rule, where a tile row holds whole pairs
  static constexpr bool PAIRS = std::is_same<T, float>::value && TWO_THIRDS && N >= 16;
  static constexpr size_t LDS_COL = ((size_t)N * TX + N) * sizeof(cx<T>);   // tile + twiddles
  // row pairs per block and LDS of the x kernels on XT threads (BATCH: the forward x pass, whose
  // wavefronts batch several row pairs at small N)
  template <int XT, bool BATCH = false>
  static constexpr int xr() { return XT / xpass_lanes<T, N, XT, BATCH>(); }
  template <int XT, bool BATCH = false>
  static constexpr size_t lds_x() { return ((size_t)xr<XT, BATCH>() * N + N) * sizeof(cx<T>); }
  // the forward x pass's threads per block under schedule V: no more row pairs than a box has
  template <int V>
  static constexpr int xthreads() {
    constexpr int cap = xpass_lanes<T, N, kXThreads[V], true>() * (N * N / 2);
    return kXThreads[V] < cap ? kXThreads[V] : cap;
  }
  // the register-resident x pass, one wavefront per row pair or per row: FP32 at N = 256, 512 and
  // 1024 (a 256- or 512-point complex transform per wavefront), FP64 at N = 256 and 512 (256)
  static constexpr bool XWAVE = TWO_THIRDS && XWave<T, N>::SUPPORTED;
  static constexpr int XW_ROWS = XWave<T, N>::RW;   // rows per wavefront
  // the register-resident x pass's stages (XFETCH_LDS): one component's rows per wavefront
  static constexpr int KXP = (M + 1 + TX - 1) / TX * TX;   // the padded row of every compact array
  template <int XT>
  static constexpr size_t lds_xw() { return (size_t)(XT / 64) * XW_ROWS * KXP * sizeof(cx<T>); }

  static void set_lds(const void *fn, size_t bytes) {
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  }

  // The z passes take the schedules of ZSched, except at N >= 1024: there a full-width tile would
  // give each thread two first-pass butterflies (which spill), so the tile is only as wide as one
  // butterfly per thread allows on 1024 threads (FP32) or on 512 (FP64, whose values need the
  // 256 registers a thread keeps at two waves per SIMD), in every z schedule.
  static constexpr bool ZWIDE = N >= 1024;
  // y passes: from N = 512 on the two schedules are 512 and 1024 threads (a column tile of 64 or
  // 128 KiB leaves room for two blocks per CU, or one)
  // Up to N = 128, schedule 0 of the y and z passes gives a block no more threads than the first
  // pass of its tile has butterflies (whole wavefronts), where the wider blocks would leave most
  // lanes idle through it; schedule 1 keeps the wider block.
  static constexpr int fit0(int nt, int tx) {
    const int b = (N / first_radix<N>() * tx + 63) / 64 * 64;
    return N <= 128 && b < nt ? b : nt;
  }
  static constexpr int YT[2] = {fit0(N >= 512 ? 512 : kYThreads[0], TX), N >= 512 ? 1024 : kYThreads[1]};
  static constexpr int ZNT_WIDE = sizeof(T) == 8 ? 512 : 1024;
  static constexpr int ZTX = ZWIDE ? std::min(TX, ZNT_WIDE / (N / first_radix<N>())) : TX;
  static constexpr int ZF_NT[2] = {ZWIDE ? ZNT_WIDE : fit0(ZSched<T>::zf_nt[0], ZTX),
                                   ZWIDE ? ZNT_WIDE : ZSched<T>::zf_nt[1]};
  static constexpr int ZI_TX[2] = {ZWIDE ? ZTX : ZSched<T>::zi_tx[0], ZWIDE ? ZTX : ZSched<T>::zi_tx[1]};
  static constexpr int ZI_NT[2] = {ZWIDE ? ZNT_WIDE : fit0(ZSched<T>::zi_nt[0], ZSched<T>::zi_tx[0]),
                                   ZWIDE ? ZNT_WIDE : ZSched<T>::zi_nt[1]};
  static constexpr size_t LDS_Z = ((size_t)N * ZTX + N) * sizeof(cx<T>);   // z tile + twiddles

  // the z forward schedule 0 stages two velocity products in LDS when they fit beside the tile
  static constexpr bool ZFL = zf_lds_bytes<N, ZTX, T, DA>(true) <= 160 * 1024;
  static constexpr size_t ZFL_BYTES = zf_lds_bytes<N, ZTX, T, DA>(ZFL);

  template <int V>
  static void set_all() {
    set_lds((const void *)k_yfwd<T, N, TX, YT[V], 0>, LDS_COL);
    set_lds((const void *)k_yinv<T, N, TX, YT[V], 0>, LDS_COL);
    set_lds((const void *)k_yfwd<T, N, TX, YT[V], 1>, LDS_COL);
    set_lds((const void *)k_yinv<T, N, TX, YT[V], 1>, LDS_COL);
    if constexpr (V == 0) {
      constexpr int ZN = ZF_NT[0];
      set_lds((const void *)k_zs<T, N, ZTX, ZN, false, false, DA>, LDS_Z);
      set_lds((const void *)k_zf<T, N, ZTX, ZN, false, ZFL, false, DA>, ZFL_BYTES);
      set_lds((const void *)k_zs<T, N, ZTX, ZN, true, false, DA>, LDS_Z);
      set_lds((const void *)k_zs<T, N, ZTX, ZN, false, true, DA>, LDS_Z);
      set_lds((const void *)k_zs<T, N, ZTX, ZN, true, true, DA>, LDS_Z);
      set_lds((const void *)k_zf<T, N, ZTX, ZN, false, ZFL, true, DA>, ZFL_BYTES);
      // the z forward's third schedule: schedule 0 with the flux tiles streamed through LDS
      set_lds((const void *)k_zs<T, N, ZTX, ZN, false, false, DA, true>, zs_lds_bytes<N, ZTX, T>(true));
    } else {
      constexpr int ZN = ZF_NT[1];
      set_lds((const void *)k_zf<T, N, ZTX, ZN, true, false, false, DA>, LDS_Z);
      set_lds((const void *)k_zf<T, N, ZTX, ZN, false, false, false, DA>, LDS_Z);
    }
    constexpr int ZT = ZI_TX[V], ZN = ZI_NT[V];
    constexpr size_t ZL = ((size_t)N * ZT + N) * sizeof(cx<T>);
    set_lds((const void *)k_zi<T, N, ZT, ZN, true, DA>, ZL);
    set_lds((const void *)k_zi<T, N, ZT, ZN, false, DA>, ZL);
    if constexpr (V == 0) {   // the z inverse's third schedule: schedule 0's tile with 16-byte stores
      set_lds((const void *)k_zi<T, N, ZT, ZN, true, DA, false, true>, zi_lds_bytes<N, ZT, T>(true));
      set_lds((const void *)k_zi<T, N, ZT, ZN, false, DA, false, true>, zi_lds_bytes<N, ZT, T>(true));
      if constexpr (PAIRS) {   // the fourth: FP32 on column pairs
        set_lds((const void *)k_zi_pairs<N, true>, PairShape<N>::LDS);
        set_lds((const void *)k_zi_pairs<N, false>, PairShape<N>::LDS);
      }
    }
    if constexpr (DA == DEALIAS_PHASE_SHIFT) {
      set_lds((const void *)k_zi<T, N, ZT, ZN, true, DA, true>, ZL);
      set_lds((const void *)k_zi<T, N, ZT, ZN, false, DA, true>, ZL);
    }
    if constexpr (XWAVE) {
      constexpr int F = V == 0 ? XFETCH_LDS : XFETCH_JIT;
      set_lds((const void *)k_xpass_wave<T, N, 7, 256, F>, F == XFETCH_LDS ? lds_xw<256>() : 0);
      set_lds((const void *)k_xpass_wave<T, N, 6, 256, F>, F == XFETCH_LDS ? lds_xw<256>() : 0);
      set_lds((const void *)k_xpass_wave<T, N, 7, 256, F, true>, F == XFETCH_LDS ? lds_xw<256>() : 0);
      set_lds((const void *)k_xpass_wave<T, N, 6, 256, F, true>, F == XFETCH_LDS ? lds_xw<256>() : 0);
    } else {
      constexpr int XT = xthreads<V>();
      set_lds((const void *)k_xpass<T, N, 7, XT, DA>, lds_x<XT, true>());
      set_lds((const void *)k_xpass<T, N, 6, XT, DA>, lds_x<XT, true>());
      set_lds((const void *)k_xpass<T, N, 7, XT, DA, true>, lds_x<XT, true>());
      set_lds((const void *)k_xpass<T, N, 6, XT, DA, true>, lds_x<XT, true>());
    }
  }

  static constexpr int ZNT_PLAIN = 256;   // threads of the plain z transforms of the adjoint
  static constexpr int XT_ADJ = 256;

  PipelineImpl() {
    set_all<0>();
    set_all<1>();
    set_lds((const void *)k_zfwd_plain<T, N, TX, ZNT_PLAIN>, LDS_COL);
    if constexpr (TWO_THIRDS) {
      set_lds((const void *)k_zinv_plain<T, N, TX, ZNT_PLAIN>, LDS_COL);
      set_lds((const void *)k_xadj<T, N, XT_ADJ>, lds_x<XT_ADJ>());
      set_lds((const void *)k_zstrain<T, N, TX, ZNT_PLAIN>, LDS_COL);
      set_lds((const void *)k_xdet<T, N, XT_ADJ>, lds_x<XT_ADJ>());
    }
  }

  // --- PhaseShift's displaced evaluation (src/Generator.jl:502-519, :675-704) ------------------
  // The state translated by half a cell on every axis, recovered on the displaced grid, the
  // products formed there and transformed forward into c.shifted ([box][ncf][comp], raw and
  // unshifted: the z forward pass shifts them back and averages them with the primary products).
  // It overwrites the pencil buffers, so it runs before the primary inverse pass.
  void shifted(Ctx &c) override {
    if constexpr (DA == DEALIAS_PHASE_SHIFT) {
      constexpr int ZT = ZI_TX[0], ZN = ZI_NT[0];
      constexpr size_t ZL = ((size_t)N * ZT + N) * sizeof(cx<T>);
      const dim3 grid(c.geo.kxp / ZT, K, c.nbox);
      if (c.cfg.has_scalar)
        hipLaunchKernelGGL((k_zi<T, N, ZT, ZN, true, DA, true>), grid, dim3(ZN), ZL, c.stream, (const cx<T> *)c.uh,
                           (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band, c.partials);
      else
        hipLaunchKernelGGL((k_zi<T, N, ZT, ZN, false, DA, true>), grid, dim3(ZN), ZL, c.stream, (const cx<T> *)c.uh,
                           (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band, c.partials);
      yinv(c);
      XArgs none{};   // no physical field is kept and no max taken on the displaced grid
      xpass(c, none);
      yfwd(c);
      const size_t comp = c.comp;
      for (int b = 0; b < c.nbox; ++b)
        hipLaunchKernelGGL((k_zfwd_plain<T, N, TX, ZNT_PLAIN>), dim3(tiles(c), K, c.geo.ncf), dim3(ZNT_PLAIN), LDS_COL,
                           c.stream, (const cx<T> *)c.bufb + b * c.geo.box_b, (cx<T> *)c.shifted + (size_t)b * c.geo.ncf * comp,
                           c.geo, (const cx<double> *)c.tw, T(1));
    }
  }

  // --- the derivative skewness's strain determinant -------------------------------------------
  int strain_det_blocks(const Ctx &) const override { return N * N / (2 * xr<XT_ADJ>()); }
  int strain_det(Ctx &c, double *partials) override {
    if constexpr (!TWO_THIRDS) {
      return 0;
    } else {
    const cx<double> *tw = (const cx<double> *)c.tw;
    cx<T> *pa = (cx<T> *)c.bufa + c.box * c.geo.box_a, *pb = (cx<T> *)c.bufb + c.box * c.geo.box_b;
    hipLaunchKernelGGL((k_zstrain<T, N, TX, ZNT_PLAIN>), dim3(tiles(c), K, 6), dim3(ZNT_PLAIN), LDS_COL, c.stream,
                       (const cx<T> *)c.uh + c.box * 3 * c.comp, pa, c.geo, tw, T(1.0 / ((double)N * N * N)));
    Geo g6 = c.geo;
    g6.nci = 6;
    hipLaunchKernelGGL((k_yinv<T, N, TX, 256>), dim3(tiles(c), N, 6), dim3(256), LDS_COL, c.stream, (const cx<T> *)pa,
                       pb, g6, tw);
    const int blocks = strain_det_blocks(c);
    hipLaunchKernelGGL((k_xdet<T, N, XT_ADJ>), dim3(blocks), dim3(XT_ADJ), lds_x<XT_ADJ>(), c.stream, (const cx<T> *)pb,
                       c.geo, tw, partials);
    return blocks;
    }
  }

  // --- the adjoint step -------------------------------------------------------------------
  size_t adj_pencil_a(const Ctx &c) const override {
    return std::max((size_t)6 * N * K, (size_t)10 * N * N) * c.geo.kxp;   // Q (6) or PX (10)
  }
  size_t adj_pencil_b(const Ctx &c) const override {
    return std::max((size_t)6 * N * N, (size_t)10 * N * K) * c.geo.kxp;   // R (6) or PF (10)
  }
  void adj_objective(Ctx &c) override {
    if constexpr (TWO_THIRDS) {
    hipLaunchKernelGGL((k_adj_objective<T>), dim3(ADJ_BLOCKS), dim3(256), 0, c.stream, (cx<T> *)c.adj.lq,
                       (cx<T> *)c.adj.lh, (const cx<T> *)c.uh, (const cx<T> *)c.ph, c.geo, c.cfg.has_scalar);
    }
  }
  void adj_begin(Ctx &c) override {   // the closing reality and Leray projections, transposed
    if constexpr (TWO_THIRDS) {
    cx<T> *lq = (cx<T> *)c.adj.lq;
    hipLaunchKernelGGL((k_reality<T>), dim3((K * K + 255) / 256, 1), dim3(256), 0, c.stream, lq, lq + 3 * c.comp, c.geo,
                       c.cfg.has_scalar ? 4 : 3);
    hipLaunchKernelGGL((k_adj_leray<T>), dim3(ADJ_BLOCKS), dim3(256), 0, c.stream, lq, c.geo);
    }
  }
  void adj_stage(Ctx &c, const AdjStage &as, int i) override {
    if constexpr (TWO_THIRDS) {
    AdjBufs &a = c.adj;
    const cx<double> *tw = (const cx<double> *)c.tw;
    hipLaunchKernelGGL((k_adj_pre<T>), dim3(ADJ_BLOCKS), dim3(256), 0, c.stream, (const cx<T> *)a.lq, (cx<T> *)a.lh,
                       (const cx<T> *)a.rq[i], (cx<T> *)a.t6, c.geo, as, a.partials);
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, c.stream, (const double *)a.partials, ADJ_BLOCKS,
                       1.0 / pow((double)N, 6.0), a.proj);
    hipLaunchKernelGGL((k_zinv_plain<T, N, TX, ZNT_PLAIN>), dim3(tiles(c), K, 6), dim3(ZNT_PLAIN), LDS_COL, c.stream,
                       (const cx<T> *)a.t6, (cx<T> *)a.pa, c.geo, tw, T(1));
    Geo g6 = c.geo, g10 = c.geo;
    g6.nci = 6;
    g10.ncf = 10;
    hipLaunchKernelGGL((k_yinv<T, N, TX, 256>), dim3(tiles(c), N, 6), dim3(256), LDS_COL, c.stream,
                       (const cx<T> *)a.pa, (cx<T> *)a.pb, g6, tw);
    hipLaunchKernelGGL((k_xadj<T, N, XT_ADJ>), dim3(N * N / (2 * xr<XT_ADJ>())), dim3(XT_ADJ), lds_x<XT_ADJ>(),
                       c.stream, (const cx<T> *)a.pb, (const T *)a.rphys[i], (const T *)a.rphysw[i], (cx<T> *)a.pa,
                       c.geo, tw, c.cfg.has_scalar);
    hipLaunchKernelGGL((k_yfwd<T, N, TX, 256>), dim3(tiles(c), N, 10), dim3(256), LDS_COL, c.stream,
                       (const cx<T> *)a.pa, (cx<T> *)a.pb, g10, tw);
    hipLaunchKernelGGL((k_zfwd_plain<T, N, TX, ZNT_PLAIN>), dim3(tiles(c), K, 10), dim3(ZNT_PLAIN), LDS_COL, c.stream,
                       (const cx<T> *)a.pb, (cx<T> *)a.t10, c.geo, tw, T(1.0 / ((double)N * N * N)));
    hipLaunchKernelGGL((k_adj_post<T>), dim3(ADJ_BLOCKS), dim3(256), 0, c.stream, (cx<T> *)a.lq, (cx<T> *)a.lh,
                       (const cx<T> *)a.rq[i], (const cx<T> *)a.t10, c.geo, as);
    }
  }

  int tiles(const Ctx &c) const { return c.geo.kxp / TX; }
  int tile_width() const override { return TX; }
  int z_blocks(const Ctx &c) const override {
    if constexpr (PAIRS)
      if (c.sched[2] == 3) return c.geo.kxp / (2 * PairShape<N>::TXP) * K;
    return c.geo.kxp / ZI_TX[c.sched[2] == 1 ? 1 : 0] * K;
  }

  template <int NTY, int PLAN>
  void yfwd_v(Ctx &c) {
    hipLaunchKernelGGL((k_yfwd<T, N, TX, NTY, PLAN>), dim3(tiles(c), N, c.geo.ncf * c.nbox), dim3(NTY), LDS_COL,
                       c.stream, (const cx<T> *)c.bufa, (cx<T> *)c.bufb, c.geo, (const cx<double> *)c.tw);
  }
  void yfwd(Ctx &c) override {
    if (c.sched[4]) c.sched[0] ? yfwd_v<YT[1], 1>(c) : yfwd_v<YT[0], 1>(c);
    else c.sched[0] ? yfwd_v<YT[1], 0>(c) : yfwd_v<YT[0], 0>(c);
  }

  // SSP: the Shu-Osher stage update (IFSSPRK3, IFHeun); SF: a scalar forcing.  Both are built for
  // schedule 0 only, which those runs always take.
  template <int V, bool SSP = false, bool SF = false, bool STAGE = false>
  void zf_v(Ctx &c, const StageArgs &sa) {
    constexpr int ZN = ZF_NT[V];
    const dim3 grid(c.geo.kxp / ZTX, K, c.nbox);
    const cx<T> *pf = (const cx<T> *)c.bufb;
    cx<T> *uh = (cx<T> *)c.uh, *hu = (cx<T> *)c.hu, *ph = (cx<T> *)c.ph, *hp = (cx<T> *)c.hp;
    const cx<double> *tw = (const cx<double> *)c.tw;
    if constexpr (V == 0) {
      if (c.cfg.has_scalar)   // the scalar half reads the stage u_z before the velocity half updates it
        hipLaunchKernelGGL((k_zs<T, N, ZTX, ZN, SSP, SF, DA, STAGE>), grid, dim3(ZN), (zs_lds_bytes<N, ZTX, T>(STAGE)),
                           c.stream, pf, (const cx<T> *)uh, ph, hp, c.geo, tw, sa);
      hipLaunchKernelGGL((k_zf<T, N, ZTX, ZN, false, ZFL, SSP, DA>), grid, dim3(ZN), ZFL_BYTES, c.stream, pf, uh, hu, ph,
                         hp, c.geo, tw, sa);
    } else {
      if (c.cfg.has_scalar)
        hipLaunchKernelGGL((k_zf<T, N, ZTX, ZN, true, false, false, DA>), grid, dim3(ZN), LDS_Z, c.stream, pf, uh, hu,
                           ph, hp, c.geo, tw, sa);
      else
        hipLaunchKernelGGL((k_zf<T, N, ZTX, ZN, false, false, false, DA>), grid, dim3(ZN), LDS_Z, c.stream, pf, uh, hu,
                           ph, hp, c.geo, tw, sa);
    }
  }
  void zf(Ctx &c, const StageArgs &sa) override {
    // the scalar forcings acting inside the stage update (VarianceRescaling acts after the step,
    // StochasticOU's force is added after the update)
    const bool sf = c.cfg.has_scalar && c.cfg.scalar_forcing != HDNS_NO_FORCING &&
                    c.cfg.scalar_forcing != HDNS_VARIANCE_RESCALING && c.cfg.scalar_forcing != HDNS_STOCHASTIC_OU;
    if (sa.ssp) sf ? zf_v<0, true, true>(c, sa) : zf_v<0, true>(c, sa);
    else if (sf) zf_v<0, false, true>(c, sa);
    else if (c.sched[1] == 2) zf_v<0, false, false, true>(c, sa);
    else c.sched[1] ? zf_v<1>(c, sa) : zf_v<0>(c, sa);
  }

  void reality(Ctx &c) override {
    const int n = K * K;
    const int planes = 2 * M == N ? 2 : 1;   // kx = 0, and kx = N/2 when the strategy keeps it
    hipLaunchKernelGGL((k_reality<T>), dim3((n + 255) / 256, c.nbox, planes), dim3(256), 0, c.stream, (cx<T> *)c.uh,
                       (cx<T> *)c.ph, c.geo, c.cfg.has_scalar ? 4 : 3);
  }

  // z inverse schedule V (0, 1) on its tile; WIDE: schedule 0's tile with 16-byte stores (schedule 2)
  template <int V, bool WIDE = false>
  void zi_v(Ctx &c) {
    constexpr int ZT = ZI_TX[V], ZN = ZI_NT[V];
    constexpr size_t ZL = zi_lds_bytes<N, ZT, T>(WIDE);
    const dim3 grid(c.geo.kxp / ZT, K, c.nbox);
    if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_zi<T, N, ZT, ZN, true, DA, false, WIDE>), grid, dim3(ZN), ZL, c.stream, (const cx<T> *)c.uh,
                         (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band, c.partials);
    else
      hipLaunchKernelGGL((k_zi<T, N, ZT, ZN, false, DA, false, WIDE>), grid, dim3(ZN), ZL, c.stream,
                         (const cx<T> *)c.uh, (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw,
                         c.band, c.partials);
  }
  void zi_pairs(Ctx &c) {
    if constexpr (PAIRS) {
      using S = PairShape<N>;
      const dim3 grid(c.geo.kxp / (2 * S::TXP), K, c.nbox);
      if (c.cfg.has_scalar)
        hipLaunchKernelGGL((k_zi_pairs<N, true>), grid, dim3(S::NT), S::LDS, c.stream, (const cx<float> *)c.uh,
                           (const cx<float> *)c.ph, (cx<float> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band,
                           c.partials);
      else
        hipLaunchKernelGGL((k_zi_pairs<N, false>), grid, dim3(S::NT), S::LDS, c.stream, (const cx<float> *)c.uh,
                           (const cx<float> *)c.ph, (cx<float> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band,
                           c.partials);
    } else {
      zi_v<0>(c);   // no pair kernels for this type or strategy: schedule 0
    }
  }
  void zi(Ctx &c) override {
    if (c.sched[2] == 3) zi_pairs(c);
    else if (c.sched[2] == 2) zi_v<0, true>(c);
    else c.sched[2] ? zi_v<1>(c) : zi_v<0>(c);
  }

  template <int NTY, int PLAN>
  void yinv_v(Ctx &c) {
    hipLaunchKernelGGL((k_yinv<T, N, TX, NTY, PLAN>), dim3(tiles(c), N, c.geo.nci * c.nbox), dim3(NTY), LDS_COL,
                       c.stream, (const cx<T> *)c.bufa, (cx<T> *)c.bufb, c.geo, (const cx<double> *)c.tw);
  }
  void yinv(Ctx &c) override {
    if (c.sched[4]) c.sched[0] ? yinv_v<YT[1], 1>(c) : yinv_v<YT[0], 1>(c);
    else c.sched[0] ? yinv_v<YT[1], 0>(c) : yinv_v<YT[0], 0>(c);
  }

  // schedule 0 stages the next component in LDS, schedule 1 loads each component just before its
  // transform (256 threads)
  template <int F>
  void xpass_wave_v(Ctx &c, const XArgs &xa) {
    constexpr int XT = 256;
    const int blocks = N * N / (XW_ROWS * (XT / 64));
    static_assert(!XWAVE || KXP == XWave<T, N>::KXP, "the staged rows are the kernel's padded rows");
    if (F == XFETCH_LDS && c.geo.kxp != KXP) return xpass_wave_v<XFETCH_JIT>(c, xa);   // rows of another padding
    const size_t lds = F == XFETCH_LDS ? lds_xw<XT>() : 0;
    const bool il = xa.interleave != 0;
    if (c.cfg.has_scalar && il)
      hipLaunchKernelGGL((k_xpass_wave<T, N, 7, XT, F, true>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_xpass_wave<T, N, 7, XT, F>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else if (il)
      hipLaunchKernelGGL((k_xpass_wave<T, N, 6, XT, F, true>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else
      hipLaunchKernelGGL((k_xpass_wave<T, N, 6, XT, F>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
  }

  template <int V>
  void xpass_v(Ctx &c, const XArgs &xa) {
    constexpr int XT = xthreads<V>();
    const int blocks = N * N / (2 * xr<XT, true>());
    constexpr size_t lds = lds_x<XT, true>();
    const bool il = xa.interleave != 0;
    if (c.cfg.has_scalar && il)
      hipLaunchKernelGGL((k_xpass<T, N, 7, XT, DA, true>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_xpass<T, N, 7, XT, DA>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else if (il)
      hipLaunchKernelGGL((k_xpass<T, N, 6, XT, DA, true>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else
      hipLaunchKernelGGL((k_xpass<T, N, 6, XT, DA>), dim3(blocks, c.nbox), dim3(XT), lds, c.stream,
                         (const cx<T> *)c.bufb, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
  }
  void xpass(Ctx &c, const XArgs &xa) override {
    if constexpr (XWAVE) c.sched[3] ? xpass_wave_v<XFETCH_JIT>(c, xa) : xpass_wave_v<XFETCH_LDS>(c, xa);
    else c.sched[3] ? xpass_v<1>(c, xa) : xpass_v<0>(c, xa);
  }

  void spectrum(Ctx &c, int which, double *partials, double *bins) override {   // the selected box
    const size_t comp = c.comp;
    const cx<T> *f = which == HDNS_FIELD_SCALAR ? (const cx<T> *)c.ph + c.box * comp
                                                : (const cx<T> *)c.uh + c.box * 3 * comp;
    const int nc = which == HDNS_FIELD_SCALAR ? 1 : 3;
    hipLaunchKernelGGL((k_spectrum_shells<T>), dim3(c.nbins, SPEC_CHUNKS), dim3(SPEC_THREADS), 0, c.stream, f, nc,
                       c.geo, partials);
    hipLaunchKernelGGL(k_spectrum_sum, dim3((c.nbins + 255) / 256), dim3(256), 0, c.stream, (const double *)partials,
                       c.nbins, bins);
  }
  TracerArgs targs(const Ctx &c) const {
    return TracerArgs{c.m, N, c.cfg.side_length, c.cfg.tracer_capacity, c.box_phys, c.ntiles,
                      (int64_t)c.trail_cap * 3 * c.cfg.tracer_capacity};
  }
  int tblocks(const Ctx &c) const { return (int)((c.m + 255) / 256); }
  // The gathering kernels' launch (k_advect, k_sample, k_advect_sample).  Capped: 128-thread
  // blocks, at most four per CU (eight waves) -- a 40 KiB LDS reservation the kernels leave
  // unused -- so fewer 4^3 stencils are in flight per XCD and more of their field lines stay in L2
  // between the tracers that share them.  Uncapped: 256-thread blocks at the occupancy the
  // registers allow.  Each tracer's arithmetic is the same under both.  Measured push + sample
  // (2^20 tracers at 256^3; 2^22 at 512^3), capped against uncapped: FP64 0.96 against 1.22 ms,
  // FP32 equal at 256^3 and 3.35 against 3.59 ms at 512^3; two blocks per CU were slower everywhere.
  static constexpr size_t GATHER_CAP_LDS = 40 * 1024;
  int gthreads(const Ctx &c) const { return tracer_gather_capped(c) ? 128 : 256; }
  int gblocks(const Ctx &c) const {   // TRACER_LANES lanes per tracer
    return (int)((TRACER_LANES * c.m + gthreads(c) - 1) / gthreads(c));
  }
  size_t glds(const Ctx &c) const { return tracer_gather_capped(c) ? GATHER_CAP_LDS : 0; }
  void advect(Ctx &c, const AdamsArgs &aa) override {
    if (c.m == 0) return;
    hipLaunchKernelGGL((k_advect<T>), dim3(gblocks(c), c.nbox), dim3(gthreads(c)), glds(c), c.stream, (T *)c.pos,
                       (T *)c.upos,
                       (const T *)c.vel, (const T *)c.phys, nullptr, targs(c), aa);
  }
  void trail_combine(Ctx &c, const TrailCombine &tc, void *out) override {   // the selected box
    if (c.m == 0) return;
    const TracerArgs ta = targs(c);
    hipLaunchKernelGGL((k_trail_combine<T>), dim3(tblocks(c)), dim3(256), 0, c.stream,
                       (const T *)c.vel + (size_t)c.box * ta.vstride, ta, tc, (T *)out);
  }
  void sample(Ctx &c) override {
    if (c.m == 0) return;
    T *slot = (T *)c.vel + (size_t)c.trail_newest * 3 * c.cfg.tracer_capacity;
    hipLaunchKernelGGL((k_sample<T>), dim3(gblocks(c), c.nbox), dim3(gthreads(c)), glds(c), c.stream,
                       (const T *)c.pos, slot,
                       c.cfg.has_scalar ? (T *)c.phis : nullptr, (const T *)c.phys, nullptr, targs(c));
  }
  void advect_sample(Ctx &c, const AdamsArgs &aa) override {
    if (c.m == 0) return;
    T *slot = (T *)c.vel + (size_t)c.trail_newest * 3 * c.cfg.tracer_capacity;
    hipLaunchKernelGGL((k_advect_sample<T>), dim3(gblocks(c), c.nbox), dim3(gthreads(c)), glds(c), c.stream,
                       (T *)c.pos, (T *)c.upos,
                       (const T *)c.vel, slot, c.cfg.has_scalar ? (T *)c.phis : nullptr, (const T *)c.phys,
                       nullptr, targs(c), aa);
  }
  void cell_keys(Ctx &c, int64_t *keys) override {   // the selected box
    if (c.m == 0) return;
    hipLaunchKernelGGL((k_cell_keys<T>), dim3(tblocks(c)), dim3(256), 0, c.stream,
                       (const T *)c.pos + (size_t)c.box * 3 * c.cfg.tracer_capacity, keys, targs(c));
  }
  hipError_t tile_sort(Ctx &c) override {
    if (c.m == 0) return hipSuccess;
    const size_t per = (size_t)c.nbox * c.cfg.tracer_capacity;
    uint32_t *ka = c.skeys, *kb = c.skeys + per;
    int32_t *va = c.svals, *vb = c.svals + per;
    hipLaunchKernelGGL((k_tile_keys<T>), dim3(tblocks(c), c.nbox), dim3(256), 0, c.stream, (const T *)c.pos,
                       (const int32_t *)c.ids, ka, va, targs(c));
    bool in_b = false;
    size_t bytes = c.sort_temp_bytes;
    const hipError_t e = tile_sort_pairs(c.sort_temp, &bytes, ka, kb, va, vb, (int64_t)c.nbox * c.m, c.sort_key_bits,
                                         c.stream, &in_b);
    if (e != hipSuccess) return e;
    // every record into the new order (and the order itself); the gathers then visit the slots in order
    hipLaunchKernelGGL((k_slot_permute<T>), dim3(tblocks(c), c.nbox), dim3(256), 0, c.stream,
                       (const int32_t *)(in_b ? vb : va), (const int32_t *)c.ids, c.ids2, c.order, (const T *)c.pos,
                       (T *)c.pos2, (const T *)c.upos, (T *)c.upos2, (const T *)c.vel, (T *)c.vel2, c.trail_cap,
                       (const T *)c.phis, (T *)c.phis2, targs(c));
    std::swap(c.pos, c.pos2);
    std::swap(c.upos, c.upos2);
    std::swap(c.vel, c.vel2);
    std::swap(c.phis, c.phis2);
    std::swap(c.ids, c.ids2);
    return hipSuccess;
  }
};

template <typename T>
std::unique_ptr<Pipeline> make_pipeline_t(int n) {
  switch (n) {
    case 16: return std::make_unique<PipelineImpl<T, 16>>();
    case 32: return std::make_unique<PipelineImpl<T, 32>>();
    case 48: return std::make_unique<PipelineImpl<T, 48>>();
    case 96: return std::make_unique<PipelineImpl<T, 96>>();
    case 192: return std::make_unique<PipelineImpl<T, 192>>();
    case 64: return std::make_unique<PipelineImpl<T, 64>>();
    case 128: return std::make_unique<PipelineImpl<T, 128>>();
    case 256: return std::make_unique<PipelineImpl<T, 256>>();
    case 512: return std::make_unique<PipelineImpl<T, 512>>();
    case 1024: return std::make_unique<PipelineImpl<T, 1024>>();
    default: return nullptr;
  }
}

// HouLi and PhaseShift pipelines: the grids they are built for (WIDE_GRIDS in the context code).
template <typename T, int DA>
std::unique_ptr<Pipeline> make_pipeline_wide_t(int n) {
  switch (n) {
    case 16: return std::make_unique<PipelineImpl<T, 16, DA>>();
    case 32: return std::make_unique<PipelineImpl<T, 32, DA>>();
    case 48: return std::make_unique<PipelineImpl<T, 48, DA>>();
    case 64: return std::make_unique<PipelineImpl<T, 64, DA>>();
    case 128: return std::make_unique<PipelineImpl<T, 128, DA>>();
    case 256: return std::make_unique<PipelineImpl<T, 256, DA>>();
    default: return nullptr;
  }
}

}  // namespace hdns
