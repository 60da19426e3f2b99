// Launch shapes of the stage kernels for one (element type, grid size) pair.
//
// Tile width TX: 16 complex FP32 or 8 complex FP64 = 128 contiguous bytes per z or y row of a
// tile.  The column buffer of a y or z block is N*TX complex values (64 KiB at N = 512).
#pragma once
#include "context.hpp"
#include "tracers.hpp"

namespace hdns {

template <typename T>
constexpr int tile_width_of() { return sizeof(T) == 4 ? 16 : 8; }

template <typename T, int N>
struct PipelineImpl final : Pipeline {
  static constexpr int TX = tile_width_of<T>();
  static constexpr int M = N / 3, K = 2 * M + 1;
  static constexpr int NTY = 512;
  static constexpr int NTZ = N >= 512 ? 1024 : 512;
  static constexpr size_t LDS_COL = ((size_t)N * TX + N) * sizeof(cx<T>);   // tile + twiddles
  static constexpr int XR = XPASS_THREADS / xpass_lanes<T, N>();
  static constexpr size_t LDS_X = ((size_t)XR * N + N) * sizeof(cx<T>);

  static void set_lds(const void *fn, size_t bytes) {
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  }

  PipelineImpl() {
    set_lds((const void *)k_yfwd<T, N, TX, NTY>, LDS_COL);
    set_lds((const void *)k_yinv<T, N, TX, NTY>, LDS_COL);
    set_lds((const void *)k_zf<T, N, TX, NTZ, true>, LDS_COL);
    set_lds((const void *)k_zf<T, N, TX, NTZ, false>, LDS_COL);
    set_lds((const void *)k_zi<T, N, TX, NTZ, true>, LDS_COL);
    set_lds((const void *)k_zi<T, N, TX, NTZ, false>, LDS_COL);
    set_lds((const void *)k_xpass<T, N, 7>, LDS_X);
    set_lds((const void *)k_xpass<T, N, 6>, LDS_X);
  }

  int tiles(const Ctx &c) const { return c.geo.kxp / TX; }
  int tile_width() const override { return TX; }

  void yfwd(Ctx &c) override {
    const int nc = c.cfg.has_scalar ? 6 : 3;
    hipLaunchKernelGGL((k_yfwd<T, N, TX, NTY>), dim3(tiles(c), N, nc), dim3(NTY), LDS_COL, c.stream,
                       (const cx<T> *)c.bufa, (cx<T> *)c.bufb, c.geo, (const cx<double> *)c.tw);
  }
  void zf(Ctx &c, const StageArgs &sa) override {
    const dim3 grid(tiles(c), K);
    if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_zf<T, N, TX, NTZ, true>), grid, dim3(NTZ), LDS_COL, c.stream, (const cx<T> *)c.bufb,
                         (cx<T> *)c.uh, (cx<T> *)c.hu, (cx<T> *)c.ph, (cx<T> *)c.hp, c.geo,
                         (const cx<double> *)c.tw, sa);
    else
      hipLaunchKernelGGL((k_zf<T, N, TX, NTZ, false>), grid, dim3(NTZ), LDS_COL, c.stream, (const cx<T> *)c.bufb,
                         (cx<T> *)c.uh, (cx<T> *)c.hu, (cx<T> *)c.ph, (cx<T> *)c.hp, c.geo,
                         (const cx<double> *)c.tw, sa);
  }
  void reality(Ctx &c) override {
    const int n = K * K;
    hipLaunchKernelGGL((k_reality<T>), dim3((n + 255) / 256), dim3(256), 0, c.stream, (cx<T> *)c.uh, (cx<T> *)c.ph,
                       c.geo, c.cfg.has_scalar ? 4 : 3);
  }
  void zi(Ctx &c) override {
    const dim3 grid(tiles(c), K);
    if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_zi<T, N, TX, NTZ, true>), grid, dim3(NTZ), LDS_COL, c.stream, (const cx<T> *)c.uh,
                         (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band, c.partials);
    else
      hipLaunchKernelGGL((k_zi<T, N, TX, NTZ, false>), grid, dim3(NTZ), LDS_COL, c.stream, (const cx<T> *)c.uh,
                         (const cx<T> *)c.ph, (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, c.band, c.partials);
  }
  void yinv(Ctx &c) override {
    const int nc = c.cfg.has_scalar ? 7 : 6;
    hipLaunchKernelGGL((k_yinv<T, N, TX, NTY>), dim3(tiles(c), N, nc), dim3(NTY), LDS_COL, c.stream,
                       (const cx<T> *)c.bufa, (cx<T> *)c.bufb, c.geo, (const cx<double> *)c.tw);
  }
  void xpass(Ctx &c, const XArgs &xa) override {
    const int blocks = N * N / (2 * XR);
    if (c.cfg.has_scalar)
      hipLaunchKernelGGL((k_xpass<T, N, 7>), dim3(blocks), dim3(XPASS_THREADS), LDS_X, c.stream, (const cx<T> *)c.bufb,
                         (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
    else
      hipLaunchKernelGGL((k_xpass<T, N, 6>), dim3(blocks), dim3(XPASS_THREADS), LDS_X, c.stream, (const cx<T> *)c.bufb,
                         (cx<T> *)c.bufa, c.geo, (const cx<double> *)c.tw, xa);
  }
  void spectrum(Ctx &c, int which, double *bins) override {
    const cx<T> *f = which == HDNS_FIELD_SCALAR ? (const cx<T> *)c.ph : (const cx<T> *)c.uh;
    const int nc = which == HDNS_FIELD_SCALAR ? 1 : 3;
    hipLaunchKernelGGL((k_spectrum<T>), dim3(1024), dim3(256), 0, c.stream, f, nc, c.geo, bins);
  }
  TracerArgs targs(const Ctx &c) const { return TracerArgs{c.m, N, c.cfg.side_length}; }
  int tblocks(const Ctx &c) const { return (int)((c.m + 255) / 256); }
  void advect(Ctx &c, double b1, double beta1, double c2) override {
    if (c.m == 0) return;
    hipLaunchKernelGGL((k_advect<T>), dim3(tblocks(c)), dim3(256), 0, c.stream, (T *)c.pos, (const T *)c.vel,
                       (const T *)c.phys, c.order_valid ? c.order : nullptr, targs(c), T(b1), T(beta1), T(c2));
  }
  void sample(Ctx &c) override {
    if (c.m == 0) return;
    const size_t vol = (size_t)N * N * N;
    const T *phi = c.cfg.has_scalar ? (const T *)c.phys + 3 * vol : nullptr;
    hipLaunchKernelGGL((k_sample<T>), dim3(tblocks(c)), dim3(256), 0, c.stream, (const T *)c.pos, (T *)c.vel,
                       c.cfg.has_scalar ? (T *)c.phis : nullptr, (const T *)c.phys, phi,
                       c.order_valid ? c.order : nullptr, targs(c));
  }
  void cell_keys(Ctx &c, int64_t *keys) override {
    if (c.m == 0) return;
    hipLaunchKernelGGL((k_cell_keys<T>), dim3(tblocks(c)), dim3(256), 0, c.stream, (const T *)c.pos, keys, targs(c));
  }
  void tile_sort(Ctx &c) override {
    if (c.m == 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, c.stream, c.tile_counts, 0, (int64_t)c.ntiles);
    hipLaunchKernelGGL((k_tile_count<T>), dim3(tblocks(c)), dim3(256), 0, c.stream, (const T *)c.pos, c.tile_counts,
                       targs(c));
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, c.stream, (const int *)c.tile_counts, c.tile_cursor,
                       c.ntiles);
    hipLaunchKernelGGL((k_tile_scatter<T>), dim3(tblocks(c)), dim3(256), 0, c.stream, (const T *)c.pos,
                       c.tile_cursor, c.order, targs(c));
  }
};

template <typename T>
std::unique_ptr<Pipeline> make_pipeline_t(int n) {
  switch (n) {
    case 16: return std::make_unique<PipelineImpl<T, 16>>();
    case 32: return std::make_unique<PipelineImpl<T, 32>>();
    case 64: return std::make_unique<PipelineImpl<T, 64>>();
    case 128: return std::make_unique<PipelineImpl<T, 128>>();
    case 256: return std::make_unique<PipelineImpl<T, 256>>();
    case 512: return std::make_unique<PipelineImpl<T, 512>>();
    case 1024: return std::make_unique<PipelineImpl<T, 1024>>();
    default: return nullptr;
  }
}

}  // namespace hdns
