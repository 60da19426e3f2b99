// Host-side state of one run and the launch interface of a (precision, N) pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/hdns.h"
#include "adjoint.hpp"
#include "tracers.hpp"

namespace hdns {

struct Ctx;


// Device buffers of the adjoint step (allocated on first use): the adjoints of the state and the
// register ([4][comp]: u, phi), the spectra of the physical adjoints and of the ten products, two
// pencil buffers for ten components, and one step's stage record.
struct AdjBufs {
  void *lq = nullptr, *lh = nullptr;
  void *t6 = nullptr, *t10 = nullptr;
  void *pa = nullptr, *pb = nullptr;
  void *rq[5] = {}, *rphys[5] = {}, *rphysw[5] = {};
  double *rscal = nullptr;      // [5][8]
  double *partials = nullptr;   // k_adj_pre's block sums
  double *proj = nullptr;       // [1]
  bool ready = false;
};

// The kernels of one (element type, grid size) pair, launched on the context's stream.
struct Pipeline {
  virtual ~Pipeline() = default;
  virtual void yfwd(Ctx &c) = 0;
  virtual void zf(Ctx &c, const StageArgs &sa) = 0;
  virtual void reality(Ctx &c) = 0;
  virtual void zi(Ctx &c) = 0;
  virtual void yinv(Ctx &c) = 0;
  virtual void xpass(Ctx &c, const XArgs &xa) = 0;
  // the shell spectrum's sums of the selected box: partials [nbins][SPEC_CHUNKS], then bins [nbins]
  virtual void spectrum(Ctx &c, int which, double *partials, double *bins) = 0;
  virtual void advect(Ctx &c, const AdamsArgs &aa) = 0;
  // out [3][m] of the selected box <- sum_j w_j (trail sample j), in the element type
  virtual void trail_combine(Ctx &c, const TrailCombine &tc, void *out) = 0;
  virtual void sample(Ctx &c) = 0;
  // advect then sample in one pass, the sample into the trail's newest slot (already advanced)
  virtual void advect_sample(Ctx &c, const AdamsArgs &aa) = 0;
  virtual void cell_keys(Ctx &c, int64_t *keys) = 0;
  virtual hipError_t tile_sort(Ctx &c) = 0;   // the traversal order of every box
  // the adjoint step (one box): J's adjoint from the current state, the transposed closing
  // projections, and one stage's transpose
  virtual void adj_objective(Ctx &c) = 0;
  virtual void adj_begin(Ctx &c) = 0;
  virtual void adj_stage(Ctx &c, const AdjStage &as, int stage) = 0;
  virtual size_t adj_pencil_a(const Ctx &c) const = 0;   // complex elements of the pencil buffers
  virtual size_t adj_pencil_b(const Ctx &c) const = 0;
  virtual int tile_width() const = 0;  // kx columns per y block (the widest tile)
  virtual int z_blocks(const Ctx &c) const = 0;
  // <det S> of the selected box's strain: writes one FP64 partial sum per block of its x pass into
  // partials and returns their count (the pencil buffers are overwritten: redo the deposit)
  virtual int strain_det(Ctx &c, double *partials) = 0;
  virtual int strain_det_blocks(const Ctx &c) const = 0;  // blocks of a z pass under the current schedule
  // PhaseShift: the products on the half-cell-displaced grid, transformed, into c.shifted
  virtual void shifted(Ctx &c) = 0;
};

std::unique_ptr<Pipeline> make_pipeline_f32(int n);
std::unique_ptr<Pipeline> make_pipeline_f64(int n);
// HouLi (dealias 1) and PhaseShift (2) pipelines, built for the grids WIDE_GRIDS lists
std::unique_ptr<Pipeline> make_pipeline_f32_wide(int n, int dealias);
std::unique_ptr<Pipeline> make_pipeline_f64_wide(int n, int dealias);

struct Ctx {
  hdns_config cfg{};
  Geo geo{};
  // a batch: nbox boxes of this size back to back in every array (1 for a plain context); the
  // per-box entry points address box `box` (hdns_select_box)
  int nbox = 1;
  int box = 0;
  std::vector<double> visc;   // per box: nu
  double *visc_d = nullptr, *kappa_d = nullptr;   // per box on the device: nu, nu/Sc
  size_t box_phys = 0, box_physw = 0;   // reals of phys / physw per box
  // With tracers, phys holds u and phi interleaved, (u_x, u_y, u_z, phi) at every point ([N^3][4],
  // phi 0 without a scalar), so a tracer's stencil row is one contiguous run; else [4][N^3].
  bool phys_il = false;
  bool own_stream = true;     // false inside an ensemble, which owns the shared stream
  bool in_ensemble = false;
  int fp64 = 0;
  size_t elem = 0;          // bytes of one complex element
  size_t comp = 0;          // complex elements of one compact component
  size_t buf_a = 0, buf_b = 0;
  int stages = 0;
  // launch schedule of the y passes, the z forward pass, the z inverse pass, the x pass and the y
  // radix plan (index into the NVAR[f] variants of each), set by hdns_autotune; the defaults are
  // the fastest measured at 512^3 on MI355X
  static constexpr int NFAM = 5;   // y threads, z forward, z inverse, x pass, y radix plan
  static constexpr int MAXVAR = 4;   // the widest family's variants (the autotune report's row)
  static constexpr int NVAR[NFAM] = {2, 3, 4, 2, 2};
  int sched[NFAM] = {1, 0, 0, 0, 0};
  BandSpec band{};
  BandSpec sband{};   // the scalar forcing's band
  hipStream_t stream = nullptr;
  void *uh = nullptr, *hu = nullptr, *ph = nullptr, *hp = nullptr;
  void *bufa = nullptr, *bufb = nullptr;
  void *phys = nullptr, *physw = nullptr;
  cx<double> *tw = nullptr;
  double *partials = nullptr;
  double *scalars = nullptr;   // per box [8]: A, the five reductions, the rescale factor - 1, spare
  double *rpart = nullptr;     // VarianceRescaling's block sums
  void *usrc = nullptr, *psrc = nullptr;   // StochasticOU's band force of the step, u and phi (FP64)
  unsigned long long *maxbits = nullptr;
  double *filter = nullptr;      // HouLi: the filter of each |m|^2 (FP64), 3 (N/2)^2 + 1 entries
  void *shifted = nullptr;        // PhaseShift: the displaced products' transforms [box][ncf][comp]
  double *spec_bins = nullptr;   // [nbins] shell sums, then [nbins][SPEC_CHUNKS] partials
  int nbins = 0;
  // tracers
  int64_t m = 0;
  void *pos = nullptr, *vel = nullptr, *phis = nullptr;
  void *upos = nullptr;        // the unwrapped record [box][3][cap], when tracked; else null
  int *order = nullptr;
  // The tracers' records live in slots: slot s of a box holds tracer ids[s].  An upload puts
  // tracer i in slot i; each traversal rebuild moves every record into the new traversal order
  // (k_slot_permute, into the second buffers, which then swap in), so the gathering kernels visit
  // the slots in order and read and write each record contiguously.  The host entry points map
  // slots to tracer indices through ids.
  int32_t *ids = nullptr;
  void *pos2 = nullptr, *upos2 = nullptr, *vel2 = nullptr, *phis2 = nullptr;
  int32_t *ids2 = nullptr;
  // the traversal sort's double buffers [2][box][cap] of keys and tracer indices, and its scratch
  uint32_t *skeys = nullptr;
  int32_t *svals = nullptr;
  void *sort_temp = nullptr;
  size_t sort_temp_bytes = 0;
  int sort_key_bits = 1;       // bits of the largest key, nbox * ntiles - 1
  int ntiles = 0;
  bool order_valid = false;
  // the gathering tracer kernels' launch schedule: 1 capped (128-thread blocks, four per CU), 0
  // uncapped, -1 by the precision and the tracer count (tracer_gather_capped)
  int gather_sched = -1;
  static constexpr int64_t GATHER_CAP_MIN = 1 << 16;
  // the velocity trail at the tracers (src/Particles.jl:404-526): a ring of trail_cap slots on the
  // device, [box][slot][3][cap]; the times and the ring's bookkeeping on the host
  int trail_cap = 2;           // k + 1 for ABM(k) (Heun: 2), k for AB(k)
  int trail_newest = 0;        // the newest sample's slot
  int trail_filled = 0;        // samples held
  double trail_t[MAX_TRAIL] = {0};
  bool advected = false;       // the positions moved since the newest sample
  double pending_t = 0.0;      // the time the next sample carries (the newest's plus that push's step)
  bool deposited = false;
  size_t bytes = 0;
  // optional per-kernel timing with events on the context's stream
  static constexpr int NKIND = 8;
  bool prof = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_kind;
  int prof_used = 0;
  double prof_ms[NKIND] = {0};
  long long prof_cnt[NKIND] = {0};
  // the captured step (hdns_step_graph): its kernels read dt and the substeps from step_values,
  // so one capture serves every dt; it is captured again when the schedule changes
  void *graph = nullptr, *graph_exec = nullptr;
  double *step_values = nullptr;   // [2 stage]: dt in T, [2 stage + 1]: the substep
  AdjBufs adj;
  std::unique_ptr<Pipeline> pipe;
  std::string err;
};

// By default FP64 gathers are capped from GATHER_CAP_MIN tracers on and FP32 gathers never: over the
// interleaved fields a FP32 stencil row is one 64-byte run, and the uncapped launch's extra waves
// win (2^20 tracers at 256^3: push + sample 0.257 against 0.303 ms; FP64 0.443 capped against
// 0.476 uncapped).
inline bool tracer_gather_capped(const Ctx &c) {
  return c.gather_sched == 1 ||
         (c.gather_sched < 0 && c.fp64 && c.m * (int64_t)c.nbox >= Ctx::GATHER_CAP_MIN);
}

}  // namespace hdns
