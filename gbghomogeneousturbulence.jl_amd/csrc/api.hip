// The C ABI of include/hdns.h: context lifetime, the boundary conversions between the
// reference's full half-spectrum layout and the compact retained-mode layout, and the stage
// schedule of step! + deposit! on one HIP stream.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "compare.hpp"
#include "context.hpp"
#include "forcing.hpp"
#include "observe.hpp"
#include "sort.hpp"

using namespace hdns;

struct hdns_ctx {
  Ctx c;
};

namespace {

thread_local std::string g_create_error;

hdns_status fail(hdns_ctx *ctx, hdns_status code, const std::string &msg) {
  if (ctx) ctx->c.err = msg;
  else g_create_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail(ctx, HDNS_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

hdns_status launched(hdns_ctx *ctx) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
  return HDNS_OK;
}

bool power_of_two(int n) { return n > 0 && (n & (n - 1)) == 0; }
// the transform lengths with a radix plan: 2^k (16..1024) and 3 * 2^k (48, 96, 192)
bool supported_grid(int n) { return (power_of_two(n) && n >= 16 && n <= 1024) || n == 48 || n == 96 || n == 192; }
// The grids the HouLi and PhaseShift pipelines are built for (make_pipeline_wide_t).
bool wide_grid(int n) { return (power_of_two(n) && n >= 16 && n <= 256) || n == 48; }

// The element-type value of a rational p/q: rounded once for FP64; for FP32 the nearest float
// to the extended-precision quotient.
double round_t(long double p, long double q, bool fp32) {
  const long double v = p / q;
  return fp32 ? (double)(float)v : (double)v;
}

struct Coeffs {
  int stages;
  double a[5], b[5], d[5];
};

// src/Propagator.jl:264-274 (Williamson RK3) and :304-320 (Carpenter-Kennedy RK4): exact
// rationals rounded to the element type; d are differences of the rounded abscissae formed in T.
Coeffs scheme_coefficients(int scheme, bool fp32) {
  Coeffs k{};
  if (scheme == HDNS_IF_SSP_RK3 || scheme == HDNS_IF_HEUN) {   // not 2N schemes: shu_osher_stage sets their stages
    k.stages = scheme == HDNS_IF_HEUN ? 2 : 3;
    return k;
  }
  if (scheme == HDNS_IF_LOW_STORAGE_RK3) {
    k.stages = 3;
    const long double an[3][2] = {{0, 1}, {-5, 9}, {-153, 128}};
    const long double bn[3][2] = {{1, 3}, {15, 16}, {8, 15}};
    const long double dn[3][2] = {{1, 3}, {5, 12}, {1, 4}};
    for (int i = 0; i < 3; ++i) {
      k.a[i] = round_t(an[i][0], an[i][1], fp32);
      k.b[i] = round_t(bn[i][0], bn[i][1], fp32);
      k.d[i] = round_t(dn[i][0], dn[i][1], fp32);
    }
    return k;
  }
  k.stages = 5;
  const long double an[5][2] = {{0, 1},
                                {-567301805773.0L, 1357537059087.0L},
                                {-2404267990393.0L, 2016746695238.0L},
                                {-3550918686646.0L, 2091501179385.0L},
                                {-1275806237668.0L, 842570457699.0L}};
  const long double bn[5][2] = {{1432997174477.0L, 9575080441755.0L},
                                {5161836677717.0L, 13612068292357.0L},
                                {1720146321549.0L, 2090206949498.0L},
                                {3134564353537.0L, 4481467310338.0L},
                                {2277821191437.0L, 14882151754819.0L}};
  for (int i = 0; i < 5; ++i) {
    k.a[i] = round_t(an[i][0], an[i][1], fp32);
    k.b[i] = round_t(bn[i][0], bn[i][1], fp32);
  }
  const double c2 = round_t(1432997174477.0L, 9575080441755.0L, fp32);
  const double c3 = round_t(2526269341429.0L, 6820363962896.0L, fp32);
  const double c4 = round_t(2006345519317.0L, 3224310063776.0L, fp32);
  const double c5 = round_t(2802321613138.0L, 2924317926251.0L, fp32);
  if (fp32) {
    k.d[0] = (float)c2;
    k.d[1] = (float)((float)c3 - (float)c2);
    k.d[2] = (float)((float)c4 - (float)c3);
    k.d[3] = (float)((float)c5 - (float)c4);
    k.d[4] = (float)(1.0f - (float)c5);
  } else {
    k.d[0] = c2;
    k.d[1] = c3 - c2;
    k.d[2] = c4 - c3;
    k.d[3] = c5 - c4;
    k.d[4] = 1.0 - c5;
  }
  return k;
}

// Fornberg's weights (Math. Comp. 51 (1988) 699-706) for the derivative of order `order` at 0 of
// the polynomial through values at the n nodes (src/Particles.jl:533-565): out[j] multiplies
// node j's value.
void derivative_weights(const double *x, int n, int order, double *out) {
  double c[MAX_TRAIL][MAX_TRAIL] = {};
  c[0][0] = 1.0;
  double prod_prev = 1.0;
  for (int i = 1; i < n; ++i) {
    const int top = std::min(i, order);
    double prod = 1.0;
    for (int j = 0; j < i; ++j) {
      const double gap = x[i] - x[j];
      prod *= gap;
      if (j == i - 1) {   // the new node's row, from the previous node's before that row moves
        for (int k = top; k > 0; --k) c[i][k] = prod_prev * (k * c[i - 1][k - 1] - x[i - 1] * c[i - 1][k]) / prod;
        c[i][0] = -prod_prev * x[i - 1] * c[i - 1][0] / prod;
      }
      for (int k = top; k > 0; --k) c[j][k] = (x[i] * c[j][k] - k * c[j][k - 1]) / gap;
      c[j][0] = x[i] * c[j][0] / gap;
    }
    prod_prev = prod;
  }
  for (int j = 0; j < n; ++j) out[j] = c[j][order];
}

// src/ParticlePropagator.jl:261-280: the integral over [0, h] of each Lagrange basis polynomial
// on the sample times, by 4-point Gauss-Legendre.
void adams_weights(const double *nodes, int n, double h, double *out) {
  static const double xi[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  static const double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  for (int j = 0; j < n; ++j) {
    double weight = 0.0;
    for (int g = 0; g < 4; ++g) {
      const double t = h * (xi[g] + 1) / 2;
      double basis = 1.0;
      for (int m = 0; m < n; ++m) {
        if (m == j) continue;
        basis *= (t - nodes[m]) / (nodes[j] - nodes[m]);
      }
      weight += wq[g] * h / 2 * basis;
    }
    out[j] = weight;
  }
}

int signed_index(int idx, int m, int k) { return idx <= m ? idx : idx - k; }
int bin_index(int idx, const Geo &g) { return idx <= g.m ? idx : g.n - g.k + idx; }

hdns_status alloc(hdns_ctx *ctx, void **p, size_t bytes) {
  if (bytes == 0) {
    *p = nullptr;
    return HDNS_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, HDNS_ERR_DEVICE, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  ctx->c.bytes += bytes;
  return HDNS_OK;
}

void drop_graph(Ctx &c);

void release(Ctx &c) {
  drop_graph(c);
  for (hipEvent_t e : c.prof_ev) (void)hipEventDestroy(e);
  c.prof_ev.clear();
  void *ptrs[] = {c.uh, c.hu, c.ph, c.hp, c.bufa, c.bufb, c.phys, c.physw, c.tw, c.partials, c.scalars, c.rpart,
                  c.usrc, c.psrc,
                  c.maxbits, c.spec_bins, c.pos, c.upos, c.vel, c.phis, c.order, c.skeys, c.svals, c.sort_temp,
                  c.ids, c.pos2, c.upos2, c.vel2, c.phis2, c.ids2,
                  c.visc_d, c.kappa_d, c.filter, c.shifted, c.step_values};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  AdjBufs &a = c.adj;
  void *aptrs[] = {a.lq, a.lh, a.t6, a.t10, a.pa, a.pb, a.rscal, a.partials, a.proj};
  for (void *p : aptrs)
    if (p) (void)hipFree(p);
  for (int i = 0; i < 5; ++i) {
    if (a.rq[i]) (void)hipFree(a.rq[i]);
    if (a.rphys[i]) (void)hipFree(a.rphys[i]);
    if (a.rphysw[i]) (void)hipFree(a.rphysw[i]);
  }
  c.adj = AdjBufs{};
  if (c.stream && c.own_stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// Zero-fill a device array with a kernel (no memset nodes in a captured step).
void zero_f64(Ctx &c, void *p, size_t bytes) {
  const size_t n = bytes / sizeof(double);
  hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, c.stream, (double *)p, 0.0, n);
}

// Per-kernel timing: an event before and after each launch on the context's stream, resolved
// in batches (the pool synchronises only when it is full, or on a read).
enum { K_YFWD, K_ZF, K_REALITY, K_ZI, K_FINALIZE, K_YINV, K_XPASS, K_OTHER };
constexpr int PROF_POOL = 2048;

void prof_resolve(Ctx &c) {
  if (!c.prof || c.prof_used == 0) return;
  (void)hipStreamSynchronize(c.stream);
  for (int i = 0; i < c.prof_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.prof_ev[i], c.prof_ev[i + 1]) == hipSuccess) {
      c.prof_ms[c.prof_kind[i / 2]] += ms;
      c.prof_cnt[c.prof_kind[i / 2]] += 1;
    }
  }
  c.prof_used = 0;
}

void prof_begin(Ctx &c, int kind) {
  if (!c.prof) return;
  if (c.prof_used + 2 > (int)c.prof_ev.size()) prof_resolve(c);
  c.prof_kind[c.prof_used / 2] = kind;
  (void)hipEventRecord(c.prof_ev[c.prof_used], c.stream);
}

void prof_end(Ctx &c) {
  if (!c.prof) return;
  (void)hipEventRecord(c.prof_ev[c.prof_used + 1], c.stream);
  c.prof_used += 2;
}

#define TIMED(c, kind, call) \
  do {                       \
    prof_begin(c, kind);     \
    call;                    \
    prof_end(c);             \
  } while (0)

double eps_of(const Ctx &c) { return c.fp64 ? 2.220446049250313e-16 : 1.1920928955078125e-07; }

FinalizeArgs finalize_args(const Ctx &c) {
  FinalizeArgs fa{};
  fa.nblocks = c.pipe->z_blocks(c);
  fa.inv_norm = 1.0 / pow((double)c.geo.n, 6.0);
  fa.forcing = c.cfg.velocity_forcing == HDNS_CONSTANT_POWER ? 1 : (c.cfg.velocity_forcing == HDNS_VARIANCE_CONTROLLER ? 2 : 0);
  fa.injection_rate = c.cfg.injection_rate;
  fa.eps_t = eps_of(c);
  fa.target = c.cfg.forcing_target;
  fa.gain = c.cfg.forcing_gain;
  fa.turnover_time = c.cfg.forcing_time;
  fa.visc = c.visc_d;
  return fa;
}

// modes in the cube |m| <= r (kx >= 0) of a StochasticOU band with highest shell r
size_t ou_cube(int r) { return (size_t)(2 * r + 1) * (2 * r + 1) * (r + 1); }

// StochasticOU's band force added after the stage update of each forced field.
void launch_band_source(Ctx &c, const StageArgs &sa) {
  for (int f = 0; f < 2; ++f) {
    void *src = f == 0 ? c.usrc : c.psrc;
    if (!src) continue;
    const int r = f == 0 ? c.cfg.highest_shell : c.cfg.scalar_highest_shell;
    const dim3 grid((unsigned)((ou_cube(r) + 255) / 256), 1, c.nbox);
    const BandSpec &band = f == 0 ? c.band : c.sband;
    if (c.fp64)
      hipLaunchKernelGGL(k_band_source<double>, grid, dim3(256), 0, c.stream, (cx<double> *)(f == 0 ? c.uh : c.ph),
                         (cx<double> *)(f == 0 ? c.hu : c.hp), f == 0 ? 3 : 1, c.geo, band, r,
                         (const cx<double> *)src, sa, f);
    else
      hipLaunchKernelGGL(k_band_source<float>, grid, dim3(256), 0, c.stream, (cx<float> *)(f == 0 ? c.uh : c.ph),
                         (cx<float> *)(f == 0 ? c.hu : c.hp), f == 0 ? 3 : 1, c.geo, band, r,
                         (const cx<double> *)src, sa, f);
  }
}

BandArgs band_args(const Ctx &c, int mode, double factor, int slot) {
  BandArgs ba{};
  ba.mode = mode;
  ba.factor = factor;
  ba.slot = slot;
  ba.inv_norm = 1.0 / pow((double)c.geo.n, 6.0);
  ba.eps_t = eps_of(c);
  return ba;
}

// The band sums of one field (u: 3 components at c.uh, or phi at c.ph) and the per-box factor.
void launch_band_factor(Ctx &c, bool scalar, const BandSpec &band, const BandArgs &ba) {
  const dim3 sums(RESCALE_BLOCKS, 1, c.nbox);
  const int nc = scalar ? 1 : 3;
  if (c.fp64)
    hipLaunchKernelGGL(k_band_sums<double>, sums, dim3(256), 0, c.stream,
                       (const cx<double> *)(scalar ? c.ph : c.uh), nc, c.geo, band, c.rpart);
  else
    hipLaunchKernelGGL(k_band_sums<float>, sums, dim3(256), 0, c.stream, (const cx<float> *)(scalar ? c.ph : c.uh),
                       nc, c.geo, band, c.rpart);
  hipLaunchKernelGGL(k_band_factor, dim3(c.nbox), dim3(256), 0, c.stream, (const double *)c.rpart, c.scalars, ba);
}

// VarianceRescaling's finalize! on the advanced state of every box (src/Forcing.jl:575-596): the
// velocity's factor in scalars[6], the scalar's in scalars[7].
void launch_rescale(Ctx &c, bool scalar) {
  const hdns_config &k = c.cfg;
  BandArgs ba = band_args(c, 0, scalar ? 1.0 : 0.5, scalar ? 7 : 6);
  ba.target = scalar ? k.scalar_target : k.forcing_target;
  ba.over_total = (scalar ? k.scalar_over : k.forcing_over) == HDNS_OVER_TOTAL;
  const BandSpec &band = scalar ? c.sband : c.band;
  launch_band_factor(c, scalar, band, ba);
  const int hi = scalar ? k.scalar_highest_shell : k.highest_shell;
  const int r = hi < 0 ? c.geo.m : std::min(hi, c.geo.m);
  const size_t cube = (size_t)(2 * r + 1) * (2 * r + 1) * (r + 1);
  const dim3 apply((unsigned)((cube + 255) / 256), 1, c.nbox);
  const int nc = scalar ? 1 : 3;
  if (c.fp64)
    hipLaunchKernelGGL(k_rescale_apply<double>, apply, dim3(256), 0, c.stream, (cx<double> *)(scalar ? c.ph : c.uh),
                       nc, c.geo, band, r, (const double *)c.scalars, ba.slot);
  else
    hipLaunchKernelGGL(k_rescale_apply<float>, apply, dim3(256), 0, c.stream, (cx<float> *)(scalar ? c.ph : c.uh), nc,
                       c.geo, band, r, (const double *)c.scalars, ba.slot);
}

// A ConstantPower or VarianceController scalar forcing's amplitude at the state the deposit just
// reduced, into scalars[7] (src/Forcing.jl:1041-1117, variance factor 1, diffusivity nu/Sc); the
// controller's dissipation reads the deposit's sum w |k|^2 |phi|^2 in scalars[5].
void launch_scalar_amplitude(Ctx &c) {
  const hdns_config &k = c.cfg;
  if (!c.sband.on || (k.scalar_forcing != HDNS_CONSTANT_POWER && k.scalar_forcing != HDNS_VARIANCE_CONTROLLER)) return;
  BandArgs ba = band_args(c, k.scalar_forcing == HDNS_CONSTANT_POWER ? 1 : 2, 1.0, 7);
  ba.rate = k.scalar_rate;
  ba.target = k.scalar_target;
  ba.gain = k.scalar_gain;
  ba.turnover_time = k.scalar_time;
  ba.diffusivity = c.kappa_d;
  ba.dirichlet_slot = 5;
  launch_band_factor(c, true, c.sband, ba);
}

// The inverse half of the pipeline for the current state: the field deposit!
// (src/Propagator.jl:808-842): u, w, phi in physical space, their products for the next
// stage 1, max|u|^2, and the reductions of the state.
hdns_status run_deposit(hdns_ctx *ctx, void *phys_to = nullptr, void *physw_to = nullptr) {
  Ctx &c = ctx->c;
  if (c.shifted) c.pipe->shifted(c);   // PhaseShift: the displaced products of this state, for its next stage
  c.pipe->zi(c);
  const FinalizeArgs fa = finalize_args(c);
  hipLaunchKernelGGL(k_finalize, dim3(c.nbox), dim3(256), 0, c.stream, (const double *)c.partials, c.scalars, fa);
  launch_scalar_amplitude(c);
  c.pipe->yinv(c);
  hipLaunchKernelGGL(k_fill_u64, dim3((c.nbox + 63) / 64), dim3(64), 0, c.stream, c.maxbits, 0ull, c.nbox);
  XArgs xa{};
  xa.box_phys = c.box_phys;
  xa.box_physw = c.box_physw;
  xa.store_u = (c.cfg.store_physical & 1) != 0;
  xa.store_phi = (c.cfg.store_physical & 2) != 0 && c.cfg.has_scalar;
  xa.store_w = (c.cfg.store_physical & 4) != 0;
  xa.want_max = 1;
  xa.phys = c.phys;
  xa.physw = c.physw;
  xa.interleave = c.phys_il;
  xa.max_bits = c.maxbits;
  if (phys_to) {   // the adjoint's stage record: u, phi and w into the given arrays
    xa.store_u = 1;
    xa.store_phi = c.cfg.has_scalar;
    xa.store_w = 1;
    xa.phys = phys_to;
    xa.physw = physw_to;
    xa.interleave = 0;
  }
  c.pipe->xpass(c, xa);
  c.deposited = true;
  return launched(ctx);
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// A context of nbox boxes of one size (nbox = 1: a plain context), viscosities visc[nbox] (NULL:
// the config's for every box).
// The allocation-free spectrum's scratch (hdns_spectrum_noalloc): the shell partials
// [nshells][SPEC_CHUNKS] and the shell sums [nshells] for up to NOALLOC_MAX_SHELLS shells, 272 KiB,
// in a file-scope device array that every GPU the library runs on holds until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only a stand-in for the
// symbol): each device's own address is taken once, at the first context created on that device,
// and kept in g_scratch_ptr.
constexpr int NOALLOC_MAX_SHELLS = 2048;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_SHELLS * (SPEC_CHUNKS + 1)];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// With `device` current: the device address of this device's scratch, taken once.
static hipError_t bind_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = (double *)ptr;
  return hipSuccess;
}

static double *scratch_of(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

hdns_status create_ctx(const hdns_config *config, int nbox, const double *visc, hdns_ctx **out) {
  if (!config || !out) return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "config and out must not be NULL");
  *out = nullptr;
  if (nbox < 1 || nbox > 65535) return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "a batch holds 1..65535 boxes");
  const hdns_config &k = *config;
  if (!supported_grid(k.grid_points))
    return fail(nullptr, HDNS_ERR_UNSUPPORTED,
                "grid_points must be a power of two in 16..1024, or 48, 96 or 192; got " +
                    std::to_string(k.grid_points));
  if (!(k.side_length > 0) || !isfinite(k.side_length))
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "side_length must be positive and finite");
  if (k.precision != HDNS_FP32_BULK && k.precision != HDNS_FP64)
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "precision must be HDNS_FP32_BULK or HDNS_FP64");
  if (k.scheme != HDNS_IF_LOW_STORAGE_RK3 && k.scheme != HDNS_IF_LOW_STORAGE_RK4 && k.scheme != HDNS_IF_SSP_RK3 &&
      k.scheme != HDNS_IF_HEUN)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "scheme must be IFLowStorageRK3, IFLowStorageRK4, IFSSPRK3 or IFHeun");
  if (k.dealiasing != HDNS_TWO_THIRDS && k.dealiasing != HDNS_HOU_LI && k.dealiasing != HDNS_PHASE_SHIFT)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "dealiasing must be TwoThirds, HouLi or PhaseShift");
  if (k.dealiasing != HDNS_TWO_THIRDS && !wide_grid(k.grid_points))
    return fail(nullptr, HDNS_ERR_UNSUPPORTED,
                "HouLi and PhaseShift dealiasing are built for grid_points a power of two in 16..256, or 48; got " +
                    std::to_string(k.grid_points));
  if (k.dealiasing == HDNS_HOU_LI && k.filter_order < 1)
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "the HouLi filter's `order` must be a positive integer");
  if (k.dealiasing == HDNS_HOU_LI && isinf(k.filter_strength))
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "the HouLi filter's `strength` must be finite (NaN: the default)");
  // src/Generator.jl:107-112
  for (int b = 0; b < nbox; ++b)
    if (!((visc ? visc[b] : k.viscosity) > 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`viscosity` must be positive: it sets the Reynolds number through Re = 1/nu, and a non-positive "
                  "value has no dissipation scale");
  if (!(k.schmidt > 0))
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                "`schmidt` must be positive: the scalar diffusivity is nu/Sc, so a non-positive Schmidt number makes "
                "it infinite or negative");
  if (k.velocity_forcing < HDNS_NO_FORCING || k.velocity_forcing > HDNS_STOCHASTIC_OU)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED,
                "velocity_forcing must be NoForcing, ConstantPower, Linear, VarianceController, VarianceRescaling or "
                "StochasticOU");
  // src/Forcing.jl:1315-1333: an additive force needs a finite band below the dealias cutoff
  for (int f = 0; f < 2; ++f) {
    const bool ou = (f == 0 ? k.velocity_forcing : k.scalar_forcing) == HDNS_STOCHASTIC_OU;
    const int hi = f == 0 ? k.highest_shell : k.scalar_highest_shell;
    if (ou && hi < 0)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "StochasticOU needs a finite `highest_shell`: it adds its process rather than multiplying the "
                  "field, so its band must stay below the dealias cutoff");
    if (ou && 9LL * (2 * hi + 1) * (2 * hi + 1) > 4LL * band_ball9(k.grid_points, k.dealiasing))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "the StochasticOU band reaches past the dealias cutoff; an additive force would fill modes the "
                  "dealiasing holds empty. Lower the shells, or enlarge the grid.");
  }
  if (k.velocity_forcing != HDNS_NO_FORCING) {  // src/Forcing.jl:388-409
    if (k.lowest_shell < 1)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`lowest_shell` must be at least 1: shell 0 is the mean (kappa = 0) mode, which a band forcing "
                  "cannot legitimately drive");
    if (k.highest_shell != -1 && k.highest_shell < k.lowest_shell)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`highest_shell` must be at least `lowest_shell`: the band would be empty and the forcing would "
                  "silently do nothing");
    if ((k.velocity_forcing == HDNS_CONSTANT_POWER || k.velocity_forcing == HDNS_LINEAR) && !isfinite(k.injection_rate))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  k.velocity_forcing == HDNS_LINEAR ? "the Linear rate must be finite" : "injection_rate must be finite");
  }
  if (k.velocity_forcing == HDNS_VARIANCE_CONTROLLER) {   // src/Forcing.jl:828-871
    if (!(isfinite(k.forcing_target) && k.forcing_target > 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`target` must be finite and positive: it is the variance the controller holds, and the rate "
                  "divides by the band's own variance");
    if (!(isfinite(k.forcing_gain) && k.forcing_gain >= 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`gain` must be finite and non-negative: a negative gain makes the restoring term push the "
                  "variance away from the target (gain = 0 is the bare marginal law)");
    if (!(isfinite(k.forcing_time) && k.forcing_time > 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "`turnover_time` must be finite and positive");
  }
  if (k.scalar_forcing < HDNS_NO_FORCING || k.scalar_forcing > HDNS_STOCHASTIC_OU)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED,
                "scalar_forcing must be NoForcing, ConstantPower, Linear, VarianceController, VarianceRescaling or "
                "StochasticOU");
  if (k.scalar_forcing != HDNS_NO_FORCING) {
    if (!k.has_scalar)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "a scalar forcing needs a run that carries the scalar");
    if (k.scalar_lowest_shell < 1)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`lowest_shell` must be at least 1: shell 0 is the mean (kappa = 0) mode, which a band forcing "
                  "cannot legitimately drive");
    if (k.scalar_highest_shell != -1 && k.scalar_highest_shell < k.scalar_lowest_shell)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "`highest_shell` must be at least `lowest_shell`: the band would be empty and the forcing would "
                  "silently do nothing");
    if ((k.scalar_forcing == HDNS_CONSTANT_POWER || k.scalar_forcing == HDNS_LINEAR) && !isfinite(k.scalar_rate))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "the scalar forcing's rate must be finite");
    if (k.scalar_forcing == HDNS_VARIANCE_CONTROLLER &&
        !(isfinite(k.scalar_target) && k.scalar_target > 0 && isfinite(k.scalar_gain) && k.scalar_gain >= 0 &&
          isfinite(k.scalar_time) && k.scalar_time > 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "the scalar VarianceController needs a finite positive target, a finite non-negative gain and a "
                  "finite positive turnover_time");
    if (k.scalar_forcing == HDNS_VARIANCE_RESCALING &&
        (!(isfinite(k.scalar_target) && k.scalar_target >= 0) ||
         (k.scalar_over != HDNS_OVER_BAND && k.scalar_over != HDNS_OVER_TOTAL)))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "the scalar VarianceRescaling needs a finite non-negative target and `over` band or total");
  }
  if (k.velocity_forcing == HDNS_VARIANCE_RESCALING) {    // src/Forcing.jl:555-573
    if (k.forcing_over != HDNS_OVER_BAND && k.forcing_over != HDNS_OVER_TOTAL)
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "VarianceRescaling `over` must be band or total");
    if (!(isfinite(k.forcing_target) && k.forcing_target >= 0))
      return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                  "VarianceRescaling `target` must be finite and non-negative: it is the kinetic energy the "
                  "forcing holds, and the rescale takes its square root");
  }
  if (!isfinite(k.mean_gradient)) return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "mean_gradient must be finite");
  if (k.tracer_capacity < 0) return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "tracer_capacity must be >= 0");
  if (k.unwrapped != 0 && k.unwrapped != 1) return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "unwrapped must be 0 or 1");
  if (k.unwrapped && k.tracer_capacity == 0)
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "unwrapped positions need tracers (tracer_capacity > 0)");
  if (k.tracer_capacity > 0 && k.interpolation != HDNS_CATMULL_ROM)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "tracers support CatmullRom interpolation");
  if (k.particle_scheme != HDNS_HEUN && k.particle_scheme != HDNS_ABM && k.particle_scheme != HDNS_AB)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "the particle scheme must be Heun, ABM(k) or AB(k)");
  // src/ParticlePropagator.jl:106-115, :225-234
  if (k.particle_scheme == HDNS_ABM && !(k.particle_steps >= 1 && k.particle_steps <= 7))
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                "`ABM(k)` takes a number of steps `k` from 1 to 7, got " + std::to_string(k.particle_steps) +
                    ": the four-point Gauss-Legendre rule that integrates the weights is exact only through the "
                    "degree `k = 7` needs");
  if (k.particle_scheme == HDNS_AB && !(k.particle_steps >= 1 && k.particle_steps <= 8))
    return fail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                "`AB(k)` takes a number of steps `k` from 1 to 8, got " + std::to_string(k.particle_steps) +
                    ": the four-point Gauss-Legendre rule that integrates the weights is exact only through "
                    "degree 7");
  if (k.tracer_capacity > (int64_t)INT32_MAX)
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "tracer_capacity above 2^31-1 is not supported");

  hdns_ctx *ctx = new (std::nothrow) hdns_ctx();
  if (!ctx) return fail(nullptr, HDNS_ERR_DEVICE, "out of host memory");
  Ctx &c = ctx->c;
  c.cfg = k;
  c.nbox = nbox;
  c.visc.assign(nbox, k.viscosity);
  if (visc) c.visc.assign(visc, visc + nbox);
  if (c.cfg.tracer_capacity > 0) c.cfg.store_physical |= 1 | (c.cfg.has_scalar ? 2 : 0);
  c.fp64 = k.precision == HDNS_FP64;
  const int n = k.grid_points;

  hipError_t e = hipSetDevice(k.device);
  if (e != hipSuccess) {
    delete ctx;
    return fail(nullptr, HDNS_ERR_DEVICE, std::string("hipSetDevice failed: ") + hipGetErrorString(e));
  }
  e = bind_scratch(k.device);
  if (e != hipSuccess) {
    delete ctx;
    return fail(nullptr, HDNS_ERR_DEVICE, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                                              hipGetErrorString(e));
  }
  if (k.dealiasing == HDNS_TWO_THIRDS) c.pipe = c.fp64 ? make_pipeline_f64(n) : make_pipeline_f32(n);
  else c.pipe = c.fp64 ? make_pipeline_f64_wide(n, k.dealiasing) : make_pipeline_f32_wide(n, k.dealiasing);
  if (!c.pipe) {
    delete ctx;
    return fail(nullptr, HDNS_ERR_UNSUPPORTED, "no pipeline for this grid size");
  }
  const int tx = c.pipe->tile_width();
  c.geo.n = n;
  c.geo.m = band_m(n, k.dealiasing);
  c.geo.k = band_k(n, k.dealiasing);
  c.geo.ball9 = band_ball9(n, k.dealiasing);
  c.geo.kxp = ((c.geo.m + 1 + tx - 1) / tx) * tx;
  c.geo.k0 = 2.0 * M_PI / k.side_length;
  c.elem = c.fp64 ? 16 : 8;
  const size_t realb = c.fp64 ? 8 : 4;
  c.comp = (size_t)c.geo.k * c.geo.k * c.geo.kxp;
  c.stages = k.scheme == HDNS_IF_LOW_STORAGE_RK4 ? 5 : (k.scheme == HDNS_IF_HEUN ? 2 : 3);
  // trail_capacity (src/ParticlePropagator.jl:248-249): ABM(k) k + 1, AB(k) k; Heun is ABM(1)
  c.trail_cap = k.particle_scheme == HDNS_AB ? k.particle_steps
                                             : (k.particle_scheme == HDNS_ABM ? k.particle_steps + 1 : 2);
  c.band.on = k.velocity_forcing != HDNS_NO_FORCING;
  c.band.lo4 = (2 * k.lowest_shell - 1) * (2 * k.lowest_shell - 1);
  c.band.hi4 = k.highest_shell < 0 ? -1 : (2 * k.highest_shell + 1) * (2 * k.highest_shell + 1);
  c.sband.on = k.has_scalar && k.scalar_forcing != HDNS_NO_FORCING;
  c.sband.lo4 = (2 * k.scalar_lowest_shell - 1) * (2 * k.scalar_lowest_shell - 1);
  c.sband.hi4 = k.scalar_highest_shell < 0 ? -1 : (2 * k.scalar_highest_shell + 1) * (2 * k.scalar_highest_shell + 1);
  const size_t ncf = k.has_scalar ? 6 : 3, nci = k.has_scalar ? 7 : 6;
  const size_t nn = (size_t)n * n, nk = (size_t)n * c.geo.k;
  c.buf_a = std::max(ncf * nn, nci * nk) * c.geo.kxp;
  c.buf_b = std::max(ncf * nk, nci * nn) * c.geo.kxp;
  c.geo.box_a = c.buf_a;
  c.geo.box_b = c.buf_b;
  c.geo.ncf = (int)ncf;
  c.geo.nci = (int)nci;
  const size_t vol = nn * n;
  c.phys_il = k.tracer_capacity > 0;
  c.box_phys = (k.has_scalar || c.phys_il ? 4 : 3) * vol;
  c.box_physw = 3 * vol;
  const size_t B = (size_t)nbox;

  hdns_status s = HDNS_OK;
  e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
  const int nzb = (c.geo.kxp / (tx / 2)) * c.geo.k;  // the narrowest z tile of any schedule
  c.nbins = (int)floor(sqrt(3.0) * n / 2) + 2;
  c.ntiles = (n / 8) * (n / 8) * (n / 8);
  c.m = 0;
  if (s == HDNS_OK) s = alloc(ctx, &c.uh, B * 3 * c.comp * c.elem);
  if (s == HDNS_OK) s = alloc(ctx, &c.hu, B * 3 * c.comp * c.elem);
  if (s == HDNS_OK && k.has_scalar) s = alloc(ctx, &c.ph, B * c.comp * c.elem);
  if (s == HDNS_OK && k.has_scalar) s = alloc(ctx, &c.hp, B * c.comp * c.elem);
  if (s == HDNS_OK) s = alloc(ctx, &c.bufa, B * c.buf_a * c.elem);
  if (s == HDNS_OK) s = alloc(ctx, &c.bufb, B * c.buf_b * c.elem);
  if (s == HDNS_OK && (c.cfg.store_physical & 3)) s = alloc(ctx, &c.phys, B * c.box_phys * realb);
  if (s == HDNS_OK && (c.cfg.store_physical & 4)) s = alloc(ctx, &c.physw, B * c.box_physw * realb);
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.tw, n * sizeof(cx<double>));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.partials, B * nzb * NRED * sizeof(double));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.scalars, B * 8 * sizeof(double));
  if (s == HDNS_OK && (k.velocity_forcing == HDNS_VARIANCE_RESCALING || c.sband.on))
    s = alloc(ctx, (void **)&c.rpart, B * RESCALE_BLOCKS * 2 * sizeof(double));
  for (int f = 0; f < 2 && s == HDNS_OK; ++f) {   // StochasticOU: the band force of the step, zero until uploaded
    if ((f == 0 ? k.velocity_forcing : (k.has_scalar ? k.scalar_forcing : HDNS_NO_FORCING)) != HDNS_STOCHASTIC_OU) continue;
    const size_t bytes = B * (f == 0 ? 3 : 1) * ou_cube(f == 0 ? k.highest_shell : k.scalar_highest_shell) * sizeof(cx<double>);
    void *&dst = f == 0 ? c.usrc : c.psrc;
    s = alloc(ctx, &dst, bytes);
    if (s == HDNS_OK) {
      const hipError_t z = hipMemset(dst, 0, bytes);
      if (z != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("clearing the band force: ") + hipGetErrorString(z));
    }
  }
  if (s == HDNS_OK) {   // the amplitudes and factors start at zero (a Linear scalar forcing reads A_phi = 0)
    const hipError_t z = hipMemset(c.scalars, 0, B * 8 * sizeof(double));
    if (z != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("clearing the scalars: ") + hipGetErrorString(z));
  }
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.maxbits, B * sizeof(unsigned long long));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.spec_bins, (size_t)c.nbins * (1 + SPEC_CHUNKS) * sizeof(double));
  if (s == HDNS_OK && k.dealiasing == HDNS_HOU_LI) {
    // src/SimulationDomain.jl:365-376: exp(-strength (|k| / k_Ny)^order), |k| / k_Ny = |m| / (N/2), per
    // |m|^2 up to 3 (N/2)^2; the default strength -log(eps(T)) damps the Nyquist mode to eps(T)
    const int nh = n / 2;
    const size_t len = (size_t)3 * nh * nh + 1;
    const double strength = isnan(k.filter_strength) ? -log(eps_of(c)) : k.filter_strength;
    std::vector<double> f(len);
    for (size_t m2 = 0; m2 < len; ++m2) f[m2] = exp(-strength * pow(sqrt((double)m2) / (double)nh, (double)k.filter_order));
    s = alloc(ctx, (void **)&c.filter, len * sizeof(double));
    if (s == HDNS_OK) {
      e = hipMemcpy(c.filter, f.data(), len * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("filter upload failed: ") + hipGetErrorString(e));
    }
  }
  if (s == HDNS_OK && k.dealiasing == HDNS_PHASE_SHIFT)   // the displaced products' transforms
    s = alloc(ctx, &c.shifted, B * ncf * c.comp * c.elem);
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.step_values, 16 * sizeof(double));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.visc_d, B * sizeof(double));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&c.kappa_d, B * sizeof(double));
  if (s == HDNS_OK && k.tracer_capacity > 0) {
    const size_t cap = (size_t)k.tracer_capacity;
    s = alloc(ctx, &c.pos, B * 3 * cap * realb);
    if (s == HDNS_OK && k.unwrapped) s = alloc(ctx, &c.upos, B * 3 * cap * realb);
    if (s == HDNS_OK) s = alloc(ctx, &c.vel, B * (size_t)c.trail_cap * 3 * cap * realb);
    if (s == HDNS_OK && k.has_scalar) s = alloc(ctx, &c.phis, B * cap * realb);
    if (s == HDNS_OK) s = alloc(ctx, (void **)&c.order, B * cap * sizeof(int));
    // the records' second buffers and the slot maps (every slot i holding tracer i to begin with)
    if (s == HDNS_OK) s = alloc(ctx, &c.pos2, B * 3 * cap * realb);
    if (s == HDNS_OK && k.unwrapped) s = alloc(ctx, &c.upos2, B * 3 * cap * realb);
    if (s == HDNS_OK) s = alloc(ctx, &c.vel2, B * (size_t)c.trail_cap * 3 * cap * realb);
    if (s == HDNS_OK && k.has_scalar) s = alloc(ctx, &c.phis2, B * cap * realb);
    if (s == HDNS_OK) s = alloc(ctx, (void **)&c.ids, B * cap * sizeof(int32_t));
    if (s == HDNS_OK) s = alloc(ctx, (void **)&c.ids2, B * cap * sizeof(int32_t));
    if (s == HDNS_OK) {
      std::vector<int32_t> iota(B * cap);
      for (size_t i = 0; i < iota.size(); ++i) iota[i] = (int32_t)(i % cap);
      e = hipMemcpy(c.ids, iota.data(), iota.size() * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("tracer slot map: ") + hipGetErrorString(e));
    }
    if (s == HDNS_OK) s = alloc(ctx, (void **)&c.skeys, 2 * B * cap * sizeof(uint32_t));
    if (s == HDNS_OK) s = alloc(ctx, (void **)&c.svals, 2 * B * cap * sizeof(int32_t));
    if (s == HDNS_OK) {   // the sort's scratch for every tracer slot of every box
      while ((1ull << c.sort_key_bits) < (unsigned long long)B * c.ntiles) ++c.sort_key_bits;
      e = tile_sort_pairs(nullptr, &c.sort_temp_bytes, c.skeys, c.skeys + B * cap, c.svals, c.svals + B * cap,
                          (int64_t)(B * cap), c.sort_key_bits, nullptr, nullptr);
      if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("sizing the traversal sort: ") + hipGetErrorString(e));
    }
    if (s == HDNS_OK) s = alloc(ctx, &c.sort_temp, std::max<size_t>(c.sort_temp_bytes, 256));
  }
  if (s == HDNS_OK) {
    std::vector<double> kap(nbox);
    for (int b = 0; b < nbox; ++b) kap[b] = c.visc[b] / k.schmidt;
    e = hipMemcpy(c.visc_d, c.visc.data(), B * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.kappa_d, kap.data(), B * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("viscosity upload failed: ") + hipGetErrorString(e));
  }
  if (s == HDNS_OK) {
    // W_N^m = exp(-2 pi i m / N), each entry the FP64 rounding of an extended-precision value
    std::vector<cx<double>> tw(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; ++m) {
      const long double ang = 2.0L * pi * (long double)m / (long double)n;
      tw[m] = {(double)cosl(ang), (double)(-sinl(ang))};
    }
    e = hipMemcpy(c.tw, tw.data(), n * sizeof(cx<double>), hipMemcpyHostToDevice);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("twiddle upload failed: ") + hipGetErrorString(e));
  }
  if (s == HDNS_OK) {
    // every array the step reads starts at zero (padding columns included)
    zero_f64(c, c.uh, B * 3 * c.comp * c.elem);
    zero_f64(c, c.hu, B * 3 * c.comp * c.elem);
    if (c.ph) zero_f64(c, c.ph, B * c.comp * c.elem);
    if (c.hp) zero_f64(c, c.hp, B * c.comp * c.elem);
    zero_f64(c, c.bufa, B * c.buf_a * c.elem);
    zero_f64(c, c.bufb, B * c.buf_b * c.elem);
    zero_f64(c, c.scalars, B * 8 * sizeof(double));
    if (c.phys) zero_f64(c, c.phys, B * c.box_phys * realb);
    if (c.physw) zero_f64(c, c.physw, B * c.box_physw * realb);
    if (c.shifted) zero_f64(c, c.shifted, B * ncf * c.comp * c.elem);
    s = launched(ctx);
    if (s == HDNS_OK) {
      e = hipStreamSynchronize(c.stream);
      if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("init failed: ") + hipGetErrorString(e));
    }
  }
  if (s != HDNS_OK) {
    g_create_error = c.err;
    release(c);
    delete ctx;
    return s;
  }
  *out = ctx;
  return HDNS_OK;
}

}  // namespace

extern "C" {

hdns_status hdns_ctx_create(const hdns_config *config, hdns_ctx **out) { return create_ctx(config, 1, nullptr, out); }

hdns_status hdns_ctx_create_batch(const hdns_config *config, int32_t nbox, const double *viscosities,
                                  hdns_ctx **out) {
  return create_ctx(config, nbox, viscosities, out);
}

hdns_status hdns_select_box(hdns_ctx *ctx, int32_t box) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  if (box < 0 || box >= ctx->c.nbox)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "box " + std::to_string(box) + " is outside this batch of " + std::to_string(ctx->c.nbox));
  ctx->c.box = box;
  return HDNS_OK;
}

hdns_status hdns_ctx_destroy(hdns_ctx *ctx) {
  if (!ctx) return HDNS_OK;
  if (ctx->c.in_ensemble)
    return fail(ctx, HDNS_ERR_STATE, "this context belongs to an ensemble: destroy the ensemble instead");
  (void)hipSetDevice(ctx->c.cfg.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  release(ctx->c);
  delete ctx;
  return HDNS_OK;
}

const char *hdns_last_error(const hdns_ctx *ctx) { return ctx ? ctx->c.err.c_str() : g_create_error.c_str(); }

hdns_status hdns_memory_bytes(const hdns_ctx *ctx, int64_t *bytes) {
  if (!ctx || !bytes) return HDNS_ERR_INVALID_ARGUMENT;
  *bytes = (int64_t)ctx->c.bytes;
  return HDNS_OK;
}

hdns_status hdns_scheme_coefficients(int32_t scheme, int32_t precision, double *a, double *b, double *d,
                                     int32_t *stages) {
  if (scheme != HDNS_IF_LOW_STORAGE_RK3 && scheme != HDNS_IF_LOW_STORAGE_RK4) return HDNS_ERR_INVALID_ARGUMENT;
  if (!a || !b || !d || !stages) return HDNS_ERR_INVALID_ARGUMENT;
  const Coeffs k = scheme_coefficients(scheme, precision == HDNS_FP32_BULK);
  *stages = k.stages;
  for (int i = 0; i < k.stages; ++i) {
    a[i] = k.a[i];
    b[i] = k.b[i];
    d[i] = k.d[i];
  }
  return HDNS_OK;
}

hdns_status hdns_heun_weights(double h, double *out3) {
  if (!out3 || !(h > 0) || !isfinite(h)) return HDNS_ERR_INVALID_ARGUMENT;
  const double n1[1] = {0.0};
  double b[1];
  adams_weights(n1, 1, h, b);
  const double n2[2] = {h, 0.0};
  double beta[2];
  adams_weights(n2, 2, h, beta);
  out3[0] = b[0];
  out3[1] = beta[0];
  out3[2] = beta[1] - b[0];
  return HDNS_OK;
}

hdns_status hdns_state_upload(hdns_ctx *ctx, const double *u_hat, const double *phi_hat) {
  if (!ctx || !u_hat) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "u_hat must not be NULL");
  Ctx &c = ctx->c;
  if (c.cfg.has_scalar && !phi_hat)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context evolves a scalar: phi_hat must not be NULL");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const Geo &g = c.geo;
  const int n = g.n, hx = n / 2 + 1;
  const size_t full = (size_t)n * n * hx;
  const int ncomp = c.cfg.has_scalar ? 4 : 3;
  // content outside the 2/3 ball cannot be carried by the retained-mode store
  for (int q = 0; q < ncomp; ++q) {
    const double *src = q < 3 ? u_hat + 2 * full * q : phi_hat;
    for (int iz = 0; iz < n; ++iz) {
      const int sz = iz <= n / 2 ? iz : iz - n;
      for (int iy = 0; iy < n; ++iy) {
        const int sy = iy <= n / 2 ? iy : iy - n;
        const double *row = src + 2 * (((size_t)iz * n + iy) * hx);
        for (int ix = 0; ix < hx; ++ix) {
          if (keeps(ix * ix + sy * sy + sz * sz, g)) continue;
          if (row[2 * ix] != 0.0 || row[2 * ix + 1] != 0.0) {
            char msg[400];
            snprintf(msg, sizeof msg,
                     "the %s spectrum has content outside the %s dealiasing ball (mode %d,%d,%d); the step keeps "
                     "only the retained modes, so truncate the state to the band first (truncate_to_band!)",
                     q < 3 ? "velocity" : "scalar", c.cfg.dealiasing == HDNS_PHASE_SHIFT ? "phase-shift" : "2/3",
                     ix, sy, sz);
            return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, msg);
          }
        }
      }
    }
  }
  std::vector<unsigned char> host(c.comp * c.elem);
  for (int q = 0; q < ncomp; ++q) {
    const double *src = q < 3 ? u_hat + 2 * full * q : phi_hat;
    std::fill(host.begin(), host.end(), 0);
    for (int kz = 0; kz < g.k; ++kz) {
      const int iz = bin_index(kz, g);
      for (int ky = 0; ky < g.k; ++ky) {
        const int iy = bin_index(ky, g);
        const double *row = src + 2 * (((size_t)iz * n + iy) * hx);
        const size_t at = ((size_t)kz * g.k + ky) * g.kxp;
        for (int ix = 0; ix <= g.m; ++ix) {
          if (c.fp64) {
            double *dst = (double *)host.data() + 2 * (at + ix);
            dst[0] = row[2 * ix];
            dst[1] = row[2 * ix + 1];
          } else {
            float *dst = (float *)host.data() + 2 * (at + ix);
            dst[0] = (float)row[2 * ix];
            dst[1] = (float)row[2 * ix + 1];
          }
        }
      }
    }
    void *dev = q < 3 ? (char *)c.uh + ((size_t)c.box * 3 + q) * c.comp * c.elem
                      : (char *)c.ph + (size_t)c.box * c.comp * c.elem;
    HIP_TRY(ctx, hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  hdns_status s = run_deposit(ctx);
  if (s != HDNS_OK) return s;
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return HDNS_OK;
}

hdns_status hdns_state_download(hdns_ctx *ctx, double *u_hat, double *phi_hat) {
  if (!ctx || !u_hat) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "u_hat must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const Geo &g = c.geo;
  const int n = g.n, hx = n / 2 + 1;
  const size_t full = (size_t)n * n * hx;
  const int ncomp = c.cfg.has_scalar && phi_hat ? 4 : 3;
  std::vector<unsigned char> host(c.comp * c.elem);
  for (int q = 0; q < ncomp; ++q) {
    double *dst = q < 3 ? u_hat + 2 * full * q : phi_hat;
    const void *dev = q < 3 ? (char *)c.uh + ((size_t)c.box * 3 + q) * c.comp * c.elem
                            : (char *)c.ph + (size_t)c.box * c.comp * c.elem;
    HIP_TRY(ctx, hipMemcpyAsync(host.data(), dev, host.size(), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
    memset(dst, 0, 2 * full * sizeof(double));
    for (int kz = 0; kz < g.k; ++kz) {
      const int iz = bin_index(kz, g);
      for (int ky = 0; ky < g.k; ++ky) {
        const int iy = bin_index(ky, g);
        double *row = dst + 2 * (((size_t)iz * n + iy) * hx);
        const size_t at = ((size_t)kz * g.k + ky) * g.kxp;
        for (int ix = 0; ix <= g.m; ++ix) {
          if (c.fp64) {
            const double *v = (const double *)host.data() + 2 * (at + ix);
            row[2 * ix] = v[0];
            row[2 * ix + 1] = v[1];
          } else {
            const float *v = (const float *)host.data() + 2 * (at + ix);
            row[2 * ix] = v[0];
            row[2 * ix + 1] = v[1];
          }
        }
      }
    }
  }
  return HDNS_OK;
}

}  // extern "C"

namespace {

// Every launch of one step! + its closing deposit!, on the context's stream, no host sync.
// Where a recording step keeps every stage's input, for the adjoint: the compact state
// ([4][comp]: u, phi), the normalised sums and amplitude of finalize ([8]) and the physical
// fields ([4][vol] u, phi and [3][vol] w; stage 0's are the caller's deposit).
struct StageRecord {
  void *q[5];
  double *scal[5];
  void *phys[5];
  void *physw[5];
};

// Stage i of a scheme in Shu-Osher form, q <- cb E^pb (q + dt N) + E^pa (ca h), with
// E = exp((A - nu |k|^2) dt / 2) and h the step's initial state, which stage 0 saves.
// IFSSPRK3 (src/Propagator.jl:599-669):
//   q1 = E^2 (u + dt N);  q2 = 1/4 E^-1 (q1 + dt N) + E (3/4 u);  u' = 2/3 E (q2 + dt N) + E^2 (1/3 u).
// IFHeun (src/Propagator.jl:456-517), whose u' = E u + dt/2 (E N1 + N2) is SSPRK(2,2):
//   q1 = E^2 (u + dt N);  u' = 1/2 (q1 + dt N) + E^2 (1/2 u).
// The powers are evaluated as exp(p x), never as products of E.
void shu_osher_stage(StageArgs &sa, int scheme, int i, double dt, bool fp32) {
  static const int pa3[3] = {0, 1, 2}, pb3[3] = {2, -1, 1}, pa2[2] = {0, 2}, pb2[2] = {2, 0};
  static const long double ca3[3][2] = {{0, 1}, {3, 4}, {1, 3}}, cb3[3][2] = {{1, 1}, {1, 4}, {2, 3}};
  static const long double ca2[2][2] = {{0, 1}, {1, 2}}, cb2[2][2] = {{1, 1}, {1, 2}};
  const bool heun = scheme == HDNS_IF_HEUN;
  sa.ssp = 1;
  sa.sub = dt / 2;
  sa.pa = heun ? pa2[i] : pa3[i];
  sa.pb = heun ? pb2[i] : pb3[i];
  const long double(*ca)[2] = heun ? ca2 : ca3;
  const long double(*cb)[2] = heun ? cb2 : cb3;
  sa.ca = round_t(ca[i][0], ca[i][1], fp32);
  sa.cb = round_t(cb[i][0], cb[i][1], fp32);
}

// The dt-dependent values of every stage, [2 stage] = dt in T, [2 stage + 1] = the exponent's
// substep, exactly as launch_step puts them into the stage arguments.
struct StepValues {
  double v[10];
};

static __global__ void k_set_step_values(double *dst, StepValues sv) {
  if (threadIdx.x < 10) dst[threadIdx.x] = sv.v[threadIdx.x];
}

StepValues step_values(const Ctx &c, double dt) {
  const bool fp32 = !c.fp64;
  const Coeffs k = scheme_coefficients(c.cfg.scheme, fp32);
  StepValues sv{};
  for (int i = 0; i < k.stages; ++i) {
    StageArgs sa{};
    sa.sub = k.d[i] * dt;
    if (c.cfg.scheme == HDNS_IF_SSP_RK3 || c.cfg.scheme == HDNS_IF_HEUN) shu_osher_stage(sa, c.cfg.scheme, i, dt, fp32);
    sv.v[2 * i] = fp32 ? (double)(float)dt : dt;
    sv.v[2 * i + 1] = sa.sub;
  }
  return sv;
}

// dyn: the kernels read dt and the substeps from c.step_values (a captured step that serves every
// dt); the caller writes them with k_set_step_values before each launch.
void launch_step(Ctx &c, double dt, const StageRecord *rec = nullptr, bool dyn = false) {
  const bool fp32 = !c.fp64;
  const Coeffs k = scheme_coefficients(c.cfg.scheme, fp32);
  const double dt_t = fp32 ? (double)(float)dt : dt;
  const FinalizeArgs fa = finalize_args(c);
  XArgs mid{};
  XArgs fin{};
  fin.box_phys = c.box_phys;
  fin.box_physw = c.box_physw;
  fin.store_u = (c.cfg.store_physical & 1) != 0;
  fin.store_phi = (c.cfg.store_physical & 2) != 0 && c.cfg.has_scalar;
  fin.store_w = (c.cfg.store_physical & 4) != 0;
  fin.want_max = 1;
  fin.phys = c.phys;
  fin.physw = c.physw;
  fin.interleave = c.phys_il;
  fin.max_bits = c.maxbits;
  for (int i = 0; i < k.stages; ++i) {
    const bool last = i == k.stages - 1;
    StageArgs sa{};
    sa.stage = i;
    sa.last = last;
    sa.a = k.a[i];
    sa.b = k.b[i];
    sa.dt = dt_t;
    sa.sub = k.d[i] * dt;  // d_i in T, promoted: the exponent's substep (src/Propagator.jl:375)
    sa.visc = c.visc_d;    // the exponent per |k|^2 is -D sub, D per box (nu, nu/Sc)
    sa.kappa = c.kappa_d;
    sa.mean_gradient = c.cfg.mean_gradient;
    sa.band = c.band;
    sa.scalars = c.scalars;
    sa.lin_rate = c.cfg.velocity_forcing == HDNS_LINEAR ? c.cfg.injection_rate : 0.0;
    sa.sband = c.sband;
    sa.slin_rate = c.cfg.scalar_forcing == HDNS_LINEAR ? c.cfg.scalar_rate : 0.0;
    sa.filter = c.filter;
    sa.shifted = c.shifted;
    sa.dyn = dyn ? c.step_values : nullptr;
    if (c.cfg.scheme == HDNS_IF_SSP_RK3 || c.cfg.scheme == HDNS_IF_HEUN) shu_osher_stage(sa, c.cfg.scheme, i, dt, fp32);
    XArgs xa = last ? fin : mid;
    if (rec) {
      const size_t cb = c.comp * c.elem;
      (void)hipMemcpyAsync(rec->q[i], c.uh, 3 * cb, hipMemcpyDeviceToDevice, c.stream);
      if (c.cfg.has_scalar) (void)hipMemcpyAsync((char *)rec->q[i] + 3 * cb, c.ph, cb, hipMemcpyDeviceToDevice, c.stream);
      (void)hipMemcpyAsync(rec->scal[i], c.scalars, 8 * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
      if (!last) {   // the x pass of this stage recovers the next stage's physical fields
        xa.store_u = 1;
        xa.store_phi = c.cfg.has_scalar;
        xa.store_w = 1;
        xa.phys = rec->phys[i + 1];
        xa.physw = rec->physw[i + 1];
        xa.interleave = 0;
      }
    }
    TIMED(c, K_YFWD, c.pipe->yfwd(c));
    TIMED(c, K_ZF, c.pipe->zf(c, sa));
    launch_band_source(c, sa);
    if (last) TIMED(c, K_REALITY, c.pipe->reality(c));
    if (last && c.cfg.velocity_forcing == HDNS_VARIANCE_RESCALING) launch_rescale(c, false);
    if (last && c.sband.on && c.cfg.scalar_forcing == HDNS_VARIANCE_RESCALING) launch_rescale(c, true);
    if (c.shifted) TIMED(c, K_OTHER, c.pipe->shifted(c));   // PhaseShift: the next stage's displaced products
    TIMED(c, K_ZI, c.pipe->zi(c));
    TIMED(c, K_FINALIZE,
          hipLaunchKernelGGL(k_finalize, dim3(c.nbox), dim3(256), 0, c.stream, (const double *)c.partials, c.scalars,
                             fa));
    launch_scalar_amplitude(c);
    TIMED(c, K_YINV, c.pipe->yinv(c));
    if (last) hipLaunchKernelGGL(k_fill_u64, dim3((c.nbox + 63) / 64), dim3(64), 0, c.stream, c.maxbits, 0ull, c.nbox);
    TIMED(c, K_XPASS, c.pipe->xpass(c, xa));
  }
}

hdns_status step_checks(hdns_ctx *ctx, double dt) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  if (!(isfinite(dt) && dt > 0))
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "the timestep must be positive and finite");
  if (!ctx->c.deposited)
    return fail(ctx, HDNS_ERR_STATE, "no deposit of the current state: upload a state (or deposit) before stepping");
  return HDNS_OK;
}

void drop_graph(Ctx &c) {
  if (c.graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)c.graph_exec);
  if (c.graph) (void)hipGraphDestroy((hipGraph_t)c.graph);
  c.graph_exec = nullptr;
  c.graph = nullptr;
}

}  // namespace

extern "C" {

hdns_status hdns_step(hdns_ctx *ctx, double dt) {
  hdns_status s = step_checks(ctx, dt);
  if (s != HDNS_OK) return s;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  launch_step(c, dt);
  c.deposited = true;
  return launched(ctx);
}

hdns_status hdns_step_graph(hdns_ctx *ctx, double dt) {
  hdns_status s = step_checks(ctx, dt);
  if (s != HDNS_OK) return s;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  if (!c.graph_exec) {
    const bool prof = c.prof;
    c.prof = false;  // no event records inside the captured step
    HIP_TRY(ctx, hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
    launch_step(c, dt, nullptr, true);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c.stream, &graph);
    c.prof = prof;
    if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("step capture failed: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(ctx, HDNS_ERR_DEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(e));
    }
    c.graph = graph;
    c.graph_exec = exec;
  }
  // this step's dt into the values the captured kernels read
  hipLaunchKernelGGL(k_set_step_values, dim3(1), dim3(64), 0, c.stream, c.step_values, step_values(c, dt));
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipGraphLaunch((hipGraphExec_t)c.graph_exec, c.stream));
  c.deposited = true;
  return HDNS_OK;
}

hdns_status hdns_deposit(hdns_ctx *ctx) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  HIP_TRY(ctx, hipSetDevice(ctx->c.cfg.device));
  return run_deposit(ctx);
}

hdns_status hdns_max_speed(hdns_ctx *ctx, double *out) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "max_speed has not been deposited");
  unsigned long long bits = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&bits, c.maxbits + c.box, sizeof bits, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  double v;
  memcpy(&v, &bits, sizeof v);
  *out = sqrt(v);
  return HDNS_OK;
}

hdns_status hdns_reductions(hdns_ctx *ctx, double *out) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposit of the current state");
  double s[8];
  HIP_TRY(ctx, hipMemcpyAsync(s, c.scalars + 8 * c.box, sizeof s, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  const double nu = c.visc[c.box], kappa = nu / c.cfg.schmidt;
  out[0] = 0.5 * s[1];
  out[1] = nu * s[3];
  out[2] = c.cfg.has_scalar ? s[4] : 0.0;
  out[3] = c.cfg.has_scalar ? 2.0 * kappa * s[5] : 0.0;
  out[4] = 0.5 * s[2];
  return HDNS_OK;
}

}  // extern "C"

namespace {

// sum over the velocity (ncomp = 3) or the scalar (ncomp = 1) of a's and b's selected boxes
template <typename TA, typename TB>
void launch_distance(const Ctx &ca, const Ctx &cb, bool scalar, double *partials) {
  const size_t comp_a = ca.comp, comp_b = cb.comp;
  const cx<TA> *a = scalar ? (const cx<TA> *)ca.ph + ca.box * comp_a : (const cx<TA> *)ca.uh + ca.box * 3 * comp_a;
  const cx<TB> *b = scalar ? (const cx<TB> *)cb.ph + cb.box * comp_b : (const cx<TB> *)cb.uh + cb.box * 3 * comp_b;
  hipLaunchKernelGGL((k_state_distance<TA, TB>), dim3(DIST_BLOCKS), dim3(256), 0, ca.stream, a, comp_a, ca.geo.kxp, b,
                     comp_b, cb.geo.kxp, ca.geo.k, ca.geo.m, ca.geo.n, scalar ? 1 : 3, partials);
}

double neumaier(const double *v, size_t n) {
  double s = 0.0, comp = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double t = s + v[i];
    comp += std::fabs(s) >= std::fabs(v[i]) ? (s - t) + v[i] : (v[i] - t) + s;
    s = t;
  }
  return s + comp;
}

}  // namespace

extern "C" {

hdns_status hdns_state_distance(hdns_ctx *a, hdns_ctx *b, double *out) {
  if (!a || !b || !out) return fail(a, HDNS_ERR_INVALID_ARGUMENT, "a, b and out must not be NULL");
  Ctx &ca = a->c, &cb = b->c;
  if (ca.geo.n != cb.geo.n || ca.geo.m != cb.geo.m)
    return fail(a, HDNS_ERR_INVALID_ARGUMENT, "the two contexts must have the same grid and dealiasing");
  if (ca.cfg.device != cb.cfg.device) return fail(a, HDNS_ERR_INVALID_ARGUMENT, "the two contexts must share a device");
  if (ca.cfg.has_scalar != cb.cfg.has_scalar)
    return fail(a, HDNS_ERR_INVALID_ARGUMENT, "both contexts must carry a scalar, or neither");
  HIP_TRY(a, hipSetDevice(ca.cfg.device));
  HIP_TRY(a, hipStreamSynchronize(cb.stream));   // b's last step has landed
  double *partials = nullptr;
  const size_t np = 3 * 2 * DIST_BLOCKS;
  HIP_TRY(a, hipMalloc((void **)&partials, 2 * np * sizeof(double)));
  std::vector<double> host(2 * np, 0.0);
  const int ncomp[2] = {3, ca.cfg.has_scalar ? 1 : 0};
  for (int f = 0; f < 2; ++f) {
    if (!ncomp[f]) continue;
    double *p = partials + f * np;
    const bool sc = f == 1;
    if (ca.fp64 && cb.fp64) launch_distance<double, double>(ca, cb, sc, p);
    else if (ca.fp64) launch_distance<double, float>(ca, cb, sc, p);
    else if (cb.fp64) launch_distance<float, double>(ca, cb, sc, p);
    else launch_distance<float, float>(ca, cb, sc, p);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), partials, host.size() * sizeof(double), hipMemcpyDeviceToHost, ca.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ca.stream);
  (void)hipFree(partials);
  if (e != hipSuccess) return fail(a, HDNS_ERR_DEVICE, std::string("state distance: ") + hipGetErrorString(e));
  // <f, f>_w / N^6 is the mean of |f|^2 over the grid points
  const double norm = pow((double)ca.geo.n, 6.0);
  for (int f = 0; f < 2; ++f) {
    double dd = 0.0, bb = 0.0;
    for (int c = 0; c < ncomp[f]; ++c) {
      dd += neumaier(host.data() + f * np + (size_t)(2 * c) * DIST_BLOCKS, DIST_BLOCKS);
      bb += neumaier(host.data() + f * np + (size_t)(2 * c + 1) * DIST_BLOCKS, DIST_BLOCKS);
    }
    out[2 * f] = sqrt(dd / norm);
    out[2 * f + 1] = sqrt(bb / norm);
  }
  return HDNS_OK;
}

hdns_status hdns_derivative_skewness(hdns_ctx *ctx, double *out) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  if (ctx->c.cfg.dealiasing != HDNS_TWO_THIRDS)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the derivative skewness is built for TwoThirds dealiasing only");
  Ctx &c = ctx->c;
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposit of the current state");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  double s[8];
  HIP_TRY(ctx, hipMemcpyAsync(s, c.scalars + 8 * c.box, sizeof s, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  const double ss = 0.5 * s[3];   // <S:S> = (1/2) <|w|^2> = (1/2) sum |k|^2 |u_hat|^2 / N^6
  if (!(ss > 0)) {   // a strain-free field (src/Observables.jl:472): no skewness to measure
    *out = 0.0;
    return HDNS_OK;
  }
  const int nb = c.pipe->strain_det_blocks(c);
  double *partials = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&partials, nb * sizeof(double)));
  c.pipe->strain_det(c, partials);
  std::vector<double> host(nb);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), partials, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(partials);
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("derivative skewness: ") + hipGetErrorString(e));
  const double n3 = pow((double)c.geo.n, 3.0);
  const double mean_det = neumaier(host.data(), host.size()) / n3;
  constexpr double C_ISO = (4.0 / 7.0) * 2.7386127875258306;   // (4/7) sqrt(15/2) (src/Observables.jl:406)
  *out = 3.0 * C_ISO * mean_det / pow(ss, 1.5);
  // the strain passes used the pencil buffers, which hold the next step's products
  return run_deposit(ctx);
}

hdns_status hdns_below_scale(hdns_ctx *ctx, double *scales, double *kinetic_energy, double *dissipation) {
  if (!ctx || !scales || !kinetic_energy || !dissipation)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "scales, kinetic_energy and dissipation must not be NULL");
  Ctx &c = ctx->c;
  if (c.cfg.dealiasing != HDNS_TWO_THIRDS)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the below-scale energies are built for TwoThirds dealiasing only");
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposit of the current state");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const int n = c.geo.n, m = c.geo.m;
  double s[8];
  HIP_TRY(ctx, hipMemcpyAsync(s, c.scalars + 8 * c.box, sizeof s, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  // sinc^2(pi |m| j / N) per side j = 1..N: the box mean's per-axis transfer at k_a l / 2
  std::vector<double> t((size_t)n * (m + 1));
  for (int j = 1; j <= n; ++j)
    for (int a = 0; a <= m; ++a) {
      const double x = M_PI * a * j / n;
      const double sc = a == 0 ? 1.0 : sin(x) / x;
      t[(size_t)(j - 1) * (m + 1) + a] = sc * sc;
    }
  const int groups = (n + BS_SIDES - 1) / BS_SIDES;
  double *dt = nullptr, *dp = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&dt, t.size() * sizeof(double)));
  hipError_t e = hipMalloc((void **)&dp, (size_t)n * BS_BLOCKS * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(dt, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) {
    const dim3 grid(BS_BLOCKS, groups);
    const size_t off = (size_t)c.box * 3 * c.comp;
    if (c.fp64)
      hipLaunchKernelGGL(k_below_scale<double>, grid, dim3(256), 0, c.stream, (const cx<double> *)c.uh + off, c.geo,
                         (const double *)dt, n, dp);
    else
      hipLaunchKernelGGL(k_below_scale<float>, grid, dim3(256), 0, c.stream, (const cx<float> *)c.uh + off, c.geo,
                         (const double *)dt, n, dp);
    e = hipGetLastError();
  }
  std::vector<double> host((size_t)n * BS_BLOCKS * 2);
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), dp, host.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(dt);
  (void)hipFree(dp);
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("below-scale sums: ") + hipGetErrorString(e));
  const double inv = 1.0 / pow((double)n, 6.0), nu = c.visc[c.box];
  std::vector<double> col(BS_BLOCKS);
  for (int j = 0; j < n; ++j) {
    double r[2];
    for (int q = 0; q < 2; ++q) {
      for (int b = 0; b < BS_BLOCKS; ++b) col[b] = host[((size_t)j * BS_BLOCKS + b) * 2 + q];
      r[q] = neumaier(col.data(), col.size()) * inv;
    }
    scales[j] = (j + 1) * c.cfg.side_length / n;
    kinetic_energy[j] = (s[1] - r[0]) / 2;   // k(l) = (whole - retained) / 2
    dissipation[j] = nu * (s[3] - r[1]);      // eps(l) = nu (dirichlet - retained)
  }
  return HDNS_OK;
}

// The spectrum's sums into the device buffers partials [nbins][SPEC_CHUNKS] and bins [nbins], then
// E(kappa_n) = bins / (N^6 * width) into out[0..nshells) (zeros past nbins), in place.
static hdns_status spectrum_into(hdns_ctx *ctx, int32_t which, double *partials, double *bins, double *out,
                                 int32_t nshells) {
  Ctx &c = ctx->c;
  c.pipe->spectrum(c, which, partials, bins);
  const hdns_status s = launched(ctx);
  if (s != HDNS_OK) return s;
  HIP_TRY(ctx, hipMemcpyAsync(out, bins, c.nbins * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  const double norm = pow((double)c.geo.n, 6.0);
  const double width = which == HDNS_FIELD_VELOCITY ? 2.0 * c.geo.k0 : c.geo.k0;
  for (int i = 0; i < nshells; ++i) out[i] = i < c.nbins ? out[i] / norm / width : 0.0;
  return HDNS_OK;
}

hdns_status hdns_spectrum(hdns_ctx *ctx, int32_t which, double *out, int32_t nshells) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (which != HDNS_FIELD_VELOCITY && which != HDNS_FIELD_SCALAR)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "which must be the velocity or the scalar");
  if (which == HDNS_FIELD_SCALAR && !c.cfg.has_scalar)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no scalar");
  if (nshells < c.nbins)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "nshells must be at least " + std::to_string(c.nbins));
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  return spectrum_into(ctx, which, c.spec_bins + c.nbins, c.spec_bins, out, nshells);
}

hdns_status hdns_spectrum_noalloc(hdns_ctx *ctx, int32_t which, double *out, int32_t nshells) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (which != HDNS_FIELD_VELOCITY && which != HDNS_FIELD_SCALAR)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "which must be the velocity or the scalar");
  if (which == HDNS_FIELD_SCALAR && !c.cfg.has_scalar)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no scalar");
  if (nshells < c.nbins)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "nshells must be at least " + std::to_string(c.nbins));
  if (c.nbins > NOALLOC_MAX_SHELLS)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "the allocation-free spectrum holds at most " + std::to_string(NOALLOC_MAX_SHELLS) + " shells; this "
                "grid has " + std::to_string(c.nbins) + ": use hdns_spectrum");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  double *ptr = scratch_of(c.cfg.device);
  if (!ptr) return fail(ctx, HDNS_ERR_STATE, "the scratch array of this device has no address: no context was created on it");
  // every partial and every sum is written before it is read, so nothing needs clearing
  return spectrum_into(ctx, which, ptr + NOALLOC_MAX_SHELLS, ptr, out, nshells);
}

hdns_status hdns_field_download(hdns_ctx *ctx, int32_t which, double *out) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposit of the current state");
  const size_t vol = (size_t)c.geo.n * c.geo.n * c.geo.n;
  const size_t realb = c.fp64 ? 8 : 4;
  const char *phys = (const char *)c.phys + c.box * c.box_phys * realb;
  const void *src = nullptr;
  int ncomp = 0;
  int first = 0;   // the interleaved layout's first component
  if (which == HDNS_FIELD_VELOCITY && (c.cfg.store_physical & 1)) {
    src = phys;
    ncomp = 3;
  } else if (which == HDNS_FIELD_SCALAR && (c.cfg.store_physical & 2) && c.cfg.has_scalar) {
    src = c.phys_il ? phys : phys + 3 * vol * realb;
    ncomp = 1;
    first = 3;
  } else if (which == HDNS_FIELD_VORTICITY && (c.cfg.store_physical & 4)) {
    src = (const char *)c.physw + c.box * c.box_physw * realb;
    ncomp = 3;
  } else {
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context does not keep that physical field (store_physical)");
  }
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  if (c.phys_il && which != HDNS_FIELD_VORTICITY) {   // (u_x, u_y, u_z, phi) per point -> [ncomp][N^3]
    std::vector<unsigned char> tmp(4 * vol * realb);
    HIP_TRY(ctx, hipMemcpyAsync(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
    for (int q = 0; q < ncomp; ++q)
      for (size_t i = 0; i < vol; ++i)
        out[q * vol + i] = c.fp64 ? ((const double *)tmp.data())[4 * i + first + q]
                                  : (double)((const float *)tmp.data())[4 * i + first + q];
    return HDNS_OK;
  }
  if (c.fp64) {
    HIP_TRY(ctx, hipMemcpyAsync(out, src, ncomp * vol * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  } else {
    std::vector<float> tmp(ncomp * vol);
    HIP_TRY(ctx, hipMemcpyAsync(tmp.data(), src, ncomp * vol * 4, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
    for (size_t i = 0; i < tmp.size(); ++i) out[i] = tmp[i];
  }
  return HDNS_OK;
}

// count reals from the host into the selected box's slice of a [box][3][cap] tracer array,
// rounded to the context's element type
static hdns_status copy_real_in(hdns_ctx *ctx, void *base, size_t count, const double *in) {
  Ctx &c = ctx->c;
  const size_t realb = c.fp64 ? 8 : 4;
  char *dst = (char *)base + (size_t)c.box * 3 * c.cfg.tracer_capacity * realb;
  if (c.fp64) {
    HIP_TRY(ctx, hipMemcpyAsync(dst, in, count * 8, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  } else {
    std::vector<float> tmp(count);
    for (size_t i = 0; i < count; ++i) tmp[i] = (float)in[i];
    HIP_TRY(ctx, hipMemcpyAsync(dst, tmp.data(), count * 4, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  return HDNS_OK;
}

hdns_status hdns_tracers_upload(hdns_ctx *ctx, int64_t m, const double *positions) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (m < 0 || m > c.cfg.tracer_capacity)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "tracer count must be within the context's tracer_capacity " + std::to_string(c.cfg.tracer_capacity));
  if (m > 0 && !positions) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "positions must not be NULL");
  if (c.nbox > 1 && c.m != 0 && m != c.m)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "every box of a batch carries the same number of tracers (" + std::to_string(c.m) + ")");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  c.m = m;
  c.order_valid = false;
  c.trail_filled = 0;   // new positions: the trail starts again, at t = 0
  c.advected = false;
  c.pending_t = 0.0;
  if (m == 0) return HDNS_OK;
  // tracer i in slot i
  std::vector<int32_t> iota((size_t)m);
  for (int64_t i = 0; i < m; ++i) iota[i] = (int32_t)i;
  const size_t at = (size_t)c.box * c.cfg.tracer_capacity;
  HIP_TRY(ctx, hipMemcpyAsync(c.ids + at, iota.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  const hdns_status s = copy_real_in(ctx, c.pos, 3 * (size_t)m, positions);
  if (s != HDNS_OK || !c.upos) return s;
  // the unwrapped record starts equal to the positions (src/Particles.jl:171-175)
  return copy_real_in(ctx, c.upos, 3 * (size_t)m, positions);
}

static hdns_status copy_real_out(hdns_ctx *ctx, const void *src, size_t count, double *out) {
  Ctx &c = ctx->c;
  if (count == 0) return HDNS_OK;
  if (c.fp64) {
    HIP_TRY(ctx, hipMemcpyAsync(out, src, count * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  } else {
    std::vector<float> tmp(count);
    HIP_TRY(ctx, hipMemcpyAsync(tmp.data(), src, count * 4, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
    for (size_t i = 0; i < count; ++i) out[i] = tmp[i];
  }
  return HDNS_OK;
}

// The tracer each slot of the selected box holds (the records move with every traversal rebuild).
static hdns_status slot_ids(hdns_ctx *ctx, std::vector<int32_t> &ids) {
  Ctx &c = ctx->c;
  ids.resize((size_t)c.m);
  if (c.m == 0) return HDNS_OK;
  HIP_TRY(ctx, hipMemcpyAsync(ids.data(), c.ids + (size_t)c.box * c.cfg.tracer_capacity, c.m * sizeof(int32_t),
                              hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  for (const int32_t id : ids)
    if (id < 0 || id >= c.m) return fail(ctx, HDNS_ERR_DEVICE, "the tracer slot map is corrupt");
  return HDNS_OK;
}

// comps [comps][m] blocks of slot-ordered reals from the device, into tracer order on the host.
static hdns_status copy_tracers_out(hdns_ctx *ctx, const void *src, int comps, double *out) {
  Ctx &c = ctx->c;
  const size_t m = (size_t)c.m;
  if (m == 0) return HDNS_OK;
  std::vector<int32_t> ids;
  hdns_status s = slot_ids(ctx, ids);
  if (s != HDNS_OK) return s;
  std::vector<double> tmp(comps * m);
  s = copy_real_out(ctx, src, comps * m, tmp.data());
  if (s != HDNS_OK) return s;
  for (int d = 0; d < comps; ++d)
    for (size_t k = 0; k < m; ++k) out[d * m + ids[k]] = tmp[d * m + k];
  return HDNS_OK;
}

// comps [comps][m] blocks in tracer order from the host, into the slots at dst (device).
static hdns_status copy_tracers_in(hdns_ctx *ctx, void *dst, int comps, const double *in) {
  Ctx &c = ctx->c;
  const size_t m = (size_t)c.m;
  if (m == 0) return HDNS_OK;
  std::vector<int32_t> ids;
  hdns_status s = slot_ids(ctx, ids);
  if (s != HDNS_OK) return s;
  if (c.fp64) {
    std::vector<double> tmp(comps * m);
    for (int d = 0; d < comps; ++d)
      for (size_t k = 0; k < m; ++k) tmp[d * m + k] = in[d * m + ids[k]];
    HIP_TRY(ctx, hipMemcpyAsync(dst, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  } else {
    std::vector<float> tmp(comps * m);
    for (int d = 0; d < comps; ++d)
      for (size_t k = 0; k < m; ++k) tmp[d * m + k] = (float)in[d * m + ids[k]];
    HIP_TRY(ctx, hipMemcpyAsync(dst, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  return HDNS_OK;
}

hdns_status hdns_tracers_download(hdns_ctx *ctx, double *positions) {
  if (!ctx || !positions) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "positions must not be NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->c.cfg.device));
  const Ctx &c = ctx->c;
  const size_t realb = c.fp64 ? 8 : 4;
  return copy_tracers_out(ctx, (const char *)c.pos + (size_t)c.box * 3 * c.cfg.tracer_capacity * realb, 3, positions);
}

static const char *NO_UNWRAPPED =
    "this tracer set carries no unwrapped positions: build it with `unwrapped = true` to track winding alongside "
    "the wrapped positions";

hdns_status hdns_tracers_unwrapped_download(hdns_ctx *ctx, double *positions) {
  if (!ctx || !positions) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "positions must not be NULL");
  const Ctx &c = ctx->c;
  if (!c.upos) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, NO_UNWRAPPED);
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const size_t realb = c.fp64 ? 8 : 4;
  return copy_tracers_out(ctx, (const char *)c.upos + (size_t)c.box * 3 * c.cfg.tracer_capacity * realb, 3, positions);
}

hdns_status hdns_tracers_unwrapped_upload(hdns_ctx *ctx, int64_t m, const double *positions) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (!c.upos) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, NO_UNWRAPPED);
  if (m != c.m)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "the unwrapped record holds one position per tracer: " + std::to_string(c.m) + " expected, got " +
                    std::to_string(m));
  if (m > 0 && !positions) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "positions must not be NULL");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const size_t realb = c.fp64 ? 8 : 4;
  return copy_tracers_in(ctx, (char *)c.upos + (size_t)c.box * 3 * c.cfg.tracer_capacity * realb, 3, positions);
}

hdns_status hdns_rebuild_traversal(hdns_ctx *ctx) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (c.cfg.tracer_capacity == 0) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no tracers");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const hipError_t e = c.pipe->tile_sort(c);
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("traversal sort: ") + hipGetErrorString(e));
  c.order_valid = c.m > 0;
  return launched(ctx);
}

}  // extern "C"

namespace {

// The checks and the Adams weights of a push of dt (hdns_advect, hdns_advect_deposit).
hdns_status push_args(hdns_ctx *ctx, double dt, AdamsArgs &aa) {
  Ctx &c = ctx->c;
  if (!(isfinite(dt) && dt > 0))
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "`dt` must be positive and finite: the samples a particle scheme integrates over must lie at "
                "increasing times");
  if (c.cfg.tracer_capacity == 0) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no tracers");
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposited velocity to advect the tracers with");
  if (c.trail_filled == 0)
    return fail(ctx, HDNS_ERR_STATE, "the trail holds no velocity sample at these positions: deposit the tracers first");
  if (c.advected)
    return fail(ctx, HDNS_ERR_STATE,
                "the positions moved since the trail's newest sample: deposit the tracers before the next push");
  // src/ParticlePropagator.jl:349-372: the scheme's own step once the trail holds its k samples,
  // else the predictor-corrector the samples present allow
  const int k = c.cfg.particle_scheme == HDNS_HEUN ? 1 : c.cfg.particle_steps;
  const bool full = c.trail_filled >= k;
  aa = AdamsArgs{};
  aa.J = full ? k : c.trail_filled;
  aa.corrected = !full || c.cfg.particle_scheme != HDNS_AB;
  double nodes[MAX_TRAIL + 1];
  const size_t per_slot = (size_t)3 * c.cfg.tracer_capacity;
  for (int j = 0; j < aa.J; ++j) {
    const int slot = (c.trail_newest - j + c.trail_cap) % c.trail_cap;
    aa.slot[j] = (int64_t)(slot * per_slot);
    nodes[j + 1] = c.trail_t[slot] - c.trail_t[c.trail_newest];   // 0, then negative
  }
  adams_weights(nodes + 1, aa.J, dt, aa.b);
  if (aa.corrected) {
    double beta[MAX_TRAIL + 1];
    nodes[0] = dt;
    adams_weights(nodes, aa.J + 1, dt, beta);
    aa.c[0] = beta[0];
    for (int j = 0; j < aa.J; ++j) aa.c[j + 1] = beta[j + 1] - aa.b[j];
  }
  return HDNS_OK;
}

// The ring's bookkeeping of a deposit (src/Particles.jl:454-490): a pushed set of positions takes
// the next slot, retiring the oldest sample of a full trail; unmoved positions refresh the newest
// sample in place.
void trail_push(Ctx &c) {
  if (c.trail_filled == 0) {   // the first sample: t = 0 after an upload, the stored clock after a restore
    c.trail_newest = 0;
    c.trail_t[0] = c.pending_t;
    c.trail_filled = 1;
  } else if (c.advected) {
    c.trail_newest = (c.trail_newest + 1) % c.trail_cap;
    c.trail_t[c.trail_newest] = c.pending_t;
    c.trail_filled = std::min(c.trail_filled + 1, c.trail_cap);
  }
  c.advected = false;
}

}  // namespace

extern "C" {

hdns_status hdns_advect(hdns_ctx *ctx, double dt) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  AdamsArgs aa;
  const hdns_status s = push_args(ctx, dt, aa);
  if (s != HDNS_OK) return s;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  c.pipe->advect(c, aa);
  c.advected = true;
  c.pending_t = c.trail_t[c.trail_newest] + dt;
  return launched(ctx);
}

hdns_status hdns_advect_deposit(hdns_ctx *ctx, double dt) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  AdamsArgs aa;
  const hdns_status s = push_args(ctx, dt, aa);
  if (s != HDNS_OK) return s;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  c.advected = true;
  c.pending_t = c.trail_t[c.trail_newest] + dt;
  // the slots the push reads were fixed in aa above; the ring then takes the new sample's slot
  trail_push(c);
  c.pipe->advect_sample(c, aa);
  return launched(ctx);
}

hdns_status hdns_deposit_tracers(hdns_ctx *ctx) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (c.cfg.tracer_capacity == 0) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no tracers");
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "no deposited velocity to sample at the tracers");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  trail_push(c);
  c.pipe->sample(c);
  return launched(ctx);
}

hdns_status hdns_tracer_trail(hdns_ctx *ctx, int32_t *filled, int32_t *capacity, double *times) {
  if (!ctx || !filled || !capacity) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "filled and capacity must not be NULL");
  Ctx &c = ctx->c;
  *filled = c.trail_filled;
  *capacity = c.trail_cap;
  if (times)
    for (int j = 0; j < c.trail_filled; ++j) times[j] = c.trail_t[(c.trail_newest - j + c.trail_cap) % c.trail_cap];
  return HDNS_OK;
}

hdns_status hdns_tracer_trail_restore(hdns_ctx *ctx, int32_t n, const double *times, const double *samples,
                                      double t_next) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (c.cfg.tracer_capacity == 0) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "this context carries no tracers");
  if (n < 0 || n > c.trail_cap - 1)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "a restored trail holds at most capacity - 1 = " + std::to_string(c.trail_cap - 1) +
                    " earlier samples (the next deposit measures the newest)");
  if (n > 0 && (!times || !samples)) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "times and samples must not be NULL");
  if (!isfinite(t_next)) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "a sample's time must be finite");
  for (int j = 0; j < n; ++j)   // newest first, strictly decreasing, all before t_next
    if (!isfinite(times[j]) || !(times[j] < (j == 0 ? t_next : times[j - 1])))
      return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "the restored sample times must be finite and decrease from t_next");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const size_t realb = c.fp64 ? 8 : 4, cap = (size_t)c.cfg.tracer_capacity, count = 3 * (size_t)c.m;
  // oldest in slot 0: the newest restored sample in slot n - 1, and the next deposit takes slot n
  for (int j = 0; j < n; ++j) {
    const int slot = n - 1 - j;
    char *dst = (char *)c.vel + ((size_t)c.box * c.trail_cap + slot) * 3 * cap * realb;
    const hdns_status st = copy_tracers_in(ctx, dst, 3, samples + (size_t)j * count);
    if (st != HDNS_OK) return st;
    c.trail_t[slot] = times[j];
  }
  c.trail_filled = n;
  c.trail_newest = n > 0 ? n - 1 : c.trail_cap - 1;
  c.advected = n > 0;
  c.pending_t = t_next;
  if (n == 0) c.trail_filled = 0;   // the next deposit starts the trail, at t_next
  return HDNS_OK;
}

hdns_status hdns_band_source_upload(hdns_ctx *ctx, int32_t field, const double *source) {
  if (!ctx || !source) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "source must not be NULL");
  Ctx &c = ctx->c;
  if (field != 0 && field != 1) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "field is 0 (velocity) or 1 (scalar)");
  void *dst = field == 0 ? c.usrc : c.psrc;
  if (!dst)
    return fail(ctx, HDNS_ERR_STATE, std::string("the ") + (field == 0 ? "velocity" : "scalar") +
                                         " is not driven by a StochasticOU forcing");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const int nc = field == 0 ? 3 : 1;
  const size_t bytes = nc * ou_cube(field == 0 ? c.cfg.highest_shell : c.cfg.scalar_highest_shell) * sizeof(cx<double>);
  HIP_TRY(ctx, hipMemcpyAsync((char *)dst + (size_t)c.box * bytes, source, bytes, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return HDNS_OK;
}

hdns_status hdns_tracer_trail_sample(hdns_ctx *ctx, int32_t offset, double *velocity) {
  if (!ctx || !velocity) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "velocity must not be NULL");
  Ctx &c = ctx->c;
  if (offset > 0 || offset <= -c.trail_filled)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "a trail is indexed by the offset in steps from its newest sample (0, -1, ...); it holds " +
                    std::to_string(c.trail_filled) + " sample(s), so it has no trail[" + std::to_string(offset) + "]");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const size_t realb = c.fp64 ? 8 : 4, cap = (size_t)c.cfg.tracer_capacity;
  const int slot = (c.trail_newest + offset + c.trail_cap) % c.trail_cap;
  const char *src = (const char *)c.vel + ((size_t)c.box * c.trail_cap + slot) * 3 * cap * realb;
  return copy_tracers_out(ctx, src, 3, velocity);
}

hdns_status hdns_tracer_time_derivative(hdns_ctx *ctx, int32_t order, double *out) {
  if (!ctx || !out) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "out must not be NULL");
  Ctx &c = ctx->c;
  if (order < 0) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "`order` must be non-negative, got " + std::to_string(order));
  if (c.trail_filled < order + 1)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                "a derivative of order " + std::to_string(order) + " needs at least " + std::to_string(order + 1) +
                    " samples, and the trail holds " + std::to_string(c.trail_filled));
  if (c.m == 0) return HDNS_OK;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  TrailCombine tc{};
  tc.n = c.trail_filled;
  double nodes[MAX_TRAIL];
  const size_t per_slot = (size_t)3 * c.cfg.tracer_capacity;
  for (int j = 0; j < tc.n; ++j) {
    const int slot = (c.trail_newest - j + c.trail_cap) % c.trail_cap;
    tc.slot[j] = (int64_t)(slot * per_slot);
    nodes[j] = c.trail_t[slot] - c.trail_t[c.trail_newest];
  }
  if (order == 0) {   // the newest value itself
    tc.n = 1;
    tc.w[0] = 1.0;
  } else {
    derivative_weights(nodes, tc.n, order, tc.w);
  }
  const size_t realb = c.fp64 ? 8 : 4;
  void *dev = nullptr;
  HIP_TRY(ctx, hipMalloc(&dev, 3 * (size_t)c.m * realb));
  c.pipe->trail_combine(c, tc, dev);
  const hdns_status s = copy_tracers_out(ctx, dev, 3, out);
  (void)hipFree(dev);
  return s;
}

hdns_status hdns_tracer_samples(hdns_ctx *ctx, double *velocity, double *scalar) {
  if (!ctx || !velocity) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "velocity must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const size_t realb = c.fp64 ? 8 : 4, cap = (size_t)c.cfg.tracer_capacity;
  const char *newest = (const char *)c.vel + ((size_t)c.box * c.trail_cap + c.trail_newest) * 3 * cap * realb;
  hdns_status s = copy_tracers_out(ctx, newest, 3, velocity);
  if (s != HDNS_OK || !scalar || !c.cfg.has_scalar) return s;
  return copy_tracers_out(ctx, (const char *)c.phis + c.box * cap * realb, 1, scalar);
}

hdns_status hdns_tracer_cells(hdns_ctx *ctx, int64_t *keys) {
  if (!ctx || !keys) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "keys must not be NULL");
  Ctx &c = ctx->c;
  if (c.m == 0) return HDNS_OK;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  int64_t *dev = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&dev, c.m * sizeof(int64_t)));
  c.pipe->cell_keys(c, dev);
  std::vector<int64_t> by_slot((size_t)c.m);
  hipError_t e = hipMemcpyAsync(by_slot.data(), dev, c.m * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("cell keys: ") + hipGetErrorString(e));
  std::vector<int32_t> ids;
  const hdns_status s = slot_ids(ctx, ids);
  if (s != HDNS_OK) return s;
  for (size_t k = 0; k < by_slot.size(); ++k) keys[ids[k]] = by_slot[k];
  return HDNS_OK;
}

hdns_status hdns_traversal_order(hdns_ctx *ctx, int32_t *order) {
  if (!ctx || !order) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "order must not be NULL");
  Ctx &c = ctx->c;
  if (!c.order_valid) return fail(ctx, HDNS_ERR_STATE, "no traversal order has been built for these positions");
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  HIP_TRY(ctx, hipMemcpyAsync(order, c.order + (size_t)c.box * c.cfg.tracer_capacity, c.m * sizeof(int),
                              hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return HDNS_OK;
}

hdns_status hdns_frame_tables(hdns_ctx *ctx, int32_t *mx, int32_t *my, int32_t *mz, uint8_t *mask) {
  if (!ctx || !mx || !my || !mz || !mask) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "outputs must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const int n = c.geo.n, hx = n / 2 + 1;
  const size_t total = (size_t)n * n * hx;
  int *d = nullptr;
  unsigned char *dm = nullptr;
  HIP_TRY(ctx, hipMalloc((void **)&d, (hx + 2 * n) * sizeof(int)));
  hipError_t e = hipMalloc((void **)&dm, total);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_frame_tables, dim3(1024), dim3(256), 0, c.stream, n, c.geo.ball9, d, d + hx, d + hx + n, dm);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(mx, d, hx * sizeof(int), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(my, d + hx, n * sizeof(int), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(mz, d + hx + n, n * sizeof(int), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(mask, dm, total, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(d);
  if (dm) (void)hipFree(dm);
  if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("frame tables: ") + hipGetErrorString(e));
  return HDNS_OK;
}

static_assert(Ctx::NFAM == HDNS_SCHED_FAMILIES && Ctx::MAXVAR == HDNS_SCHED_VARIANTS, "the header's schedule shape");

hdns_status hdns_autotune(hdns_ctx *ctx, int32_t reps, double *ms, int32_t *chosen) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (!c.deposited) return fail(ctx, HDNS_ERR_STATE, "autotune needs an adopted state: upload one first");
  if (reps < 1) reps = 3;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  const bool prof = c.prof;
  c.prof = false;
  hipEvent_t e0, e1;
  HIP_TRY(ctx, hipEventCreate(&e0));
  HIP_TRY(ctx, hipEventCreate(&e1));
  // A stage-shaped sequence that leaves the state bitwise unchanged: the z forward pass runs with
  // dt = 0 and a unit exponential (q <- 1 * (q + b * 0)); every other kernel writes only
  // registers the closing deposit rewrites.
  StageArgs idle{};
  idle.stage = 0;
  idle.last = 0;
  idle.b = 1.0;
  idle.sub = 0.0;   // exponent -D * 0: a unit exponential
  idle.visc = c.visc_d;
  idle.kappa = c.kappa_d;
  idle.scalars = c.scalars;
  idle.filter = c.filter;     // HouLi: the z forward kernels read the filter table
  idle.shifted = c.shifted;   // PhaseShift: and the displaced products (finite: the last deposit's)
  XArgs xa{};
  xa.phys = c.phys;
  xa.physw = c.physw;
  xa.max_bits = c.maxbits;
  xa.box_phys = c.box_phys;
  xa.box_physw = c.box_physw;
  constexpr int NF = Ctx::NFAM, MV = Ctx::MAXVAR;
  int saved[NF];
  for (int f = 0; f < NF; ++f) saved[f] = c.sched[f];
  double best[NF][MV];
  for (int f = 0; f < NF; ++f)
    for (int v = 0; v < MV; ++v) best[f][v] = -1.0;   // no such variant
  for (int fam = 0; fam < NF; ++fam) {
    for (int v = 0; v < Ctx::NVAR[fam]; ++v) {
      for (int f = 0; f < NF; ++f) c.sched[f] = saved[f];
      c.sched[fam] = v;
      double total = 0.0;
      for (int r = 0; r <= reps; ++r) {  // the first round warms up
        float t[5];
        for (int k = 0; k < 5; ++k) {
          (void)hipEventRecord(e0, c.stream);
          switch (k) {
            case 0: c.pipe->zi(c); break;
            case 1: c.pipe->yinv(c); break;
            case 2: c.pipe->xpass(c, xa); break;
            case 3: c.pipe->yfwd(c); break;
            default: c.pipe->zf(c, idle); break;
          }
          (void)hipEventRecord(e1, c.stream);
          HIP_TRY(ctx, hipEventSynchronize(e1));
          (void)hipEventElapsedTime(&t[k], e0, e1);
        }
        if (r == 0) continue;
        total += (fam == 0 || fam == 4) ? t[1] + t[3] : fam == 1 ? t[4] : fam == 2 ? t[0] : t[2];
      }
      best[fam][v] = total / reps;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int fam = 0; fam < NF; ++fam) {
    int pick = 0;
    for (int v = 1; v < Ctx::NVAR[fam]; ++v)
      if (best[fam][v] < best[fam][pick]) pick = v;
    c.sched[fam] = pick;
  }
  c.prof = prof;
  if (ms)
    for (int i = 0; i < MV * NF; ++i) ms[i] = best[i / MV][i % MV];
  if (chosen)
    for (int i = 0; i < NF; ++i) chosen[i] = c.sched[i];
  hdns_status s = launched(ctx);
  if (s != HDNS_OK) return s;
  drop_graph(c);  // a captured step holds the old schedule
  return run_deposit(ctx);
}

hdns_status hdns_schedule(hdns_ctx *ctx, const int32_t *set, int32_t *get) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (set) {
    for (int i = 0; i < Ctx::NFAM; ++i)
      if (set[i] < 0 || set[i] >= Ctx::NVAR[i])
        return fail(ctx, HDNS_ERR_INVALID_ARGUMENT,
                    "schedule entry " + std::to_string(i) + " is " + std::to_string(set[i]) + ": the family has " +
                        std::to_string(Ctx::NVAR[i]) + " variants");
    for (int i = 0; i < Ctx::NFAM; ++i) c.sched[i] = set[i];
    drop_graph(c);
  }
  if (get)
    for (int i = 0; i < Ctx::NFAM; ++i) get[i] = c.sched[i];
  return HDNS_OK;
}

hdns_status hdns_tracer_schedule(hdns_ctx *ctx, int32_t set, int32_t *get) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  if (set < -2 || set > 1) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "the tracer schedule is -1 (by count), 0 or 1");
  if (set != -2) c.gather_sched = set;
  if (get) *get = tracer_gather_capped(c) ? 1 : 0;
  return HDNS_OK;
}

hdns_status hdns_profile_enable(hdns_ctx *ctx, int32_t on) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  if (on && !c.prof) {
    c.prof_ev.resize(PROF_POOL);
    c.prof_kind.assign(PROF_POOL / 2, K_OTHER);
    for (auto &e : c.prof_ev) HIP_TRY(ctx, hipEventCreate(&e));
    c.prof_used = 0;
    c.prof = true;
  } else if (!on && c.prof) {
    prof_resolve(c);
    for (hipEvent_t e : c.prof_ev) (void)hipEventDestroy(e);
    c.prof_ev.clear();
    c.prof = false;
  }
  return HDNS_OK;
}

hdns_status hdns_profile_read(hdns_ctx *ctx, double *ms, int64_t *launches, int32_t reset) {
  if (!ctx || !ms || !launches) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "outputs must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  prof_resolve(c);
  for (int i = 0; i < Ctx::NKIND; ++i) {
    ms[i] = c.prof_ms[i];
    launches[i] = c.prof_cnt[i];
    if (reset) {
      c.prof_ms[i] = 0;
      c.prof_cnt[i] = 0;
    }
  }
  return HDNS_OK;
}

hdns_status hdns_synchronize(hdns_ctx *ctx) {
  if (!ctx) return HDNS_ERR_INVALID_ARGUMENT;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->c.stream));
  return HDNS_OK;
}

hdns_status hdns_stream(hdns_ctx *ctx, void **stream) {
  if (!ctx || !stream) return HDNS_ERR_INVALID_ARGUMENT;
  *stream = (void *)ctx->c.stream;
  return HDNS_OK;
}

// --- the adjoint step -------------------------------------------------------------------

}  // extern "C"

namespace {

hdns_status adj_alloc(hdns_ctx *ctx) {
  Ctx &c = ctx->c;
  AdjBufs &a = c.adj;
  if (a.ready) return HDNS_OK;
  if (c.nbox != 1) return fail(ctx, HDNS_ERR_UNSUPPORTED, "the adjoint runs on a single-box context");
  if (c.cfg.dealiasing != HDNS_TWO_THIRDS)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the adjoint is built for TwoThirds dealiasing only");
  if (c.cfg.velocity_forcing != HDNS_NO_FORCING && c.cfg.velocity_forcing != HDNS_CONSTANT_POWER)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the adjoint supports NoForcing and ConstantPower velocity forcing");
  if (c.cfg.scalar_forcing != HDNS_NO_FORCING)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the adjoint supports no scalar forcing (the mean gradient drives the scalar)");
  if (c.cfg.scheme == HDNS_IF_SSP_RK3 || c.cfg.scheme == HDNS_IF_HEUN)
    return fail(ctx, HDNS_ERR_UNSUPPORTED, "the adjoint supports the low-storage schemes IFLowStorageRK3 and IFLowStorageRK4");
  const size_t cb = c.comp * c.elem, realb = c.fp64 ? 8 : 4;
  const size_t vol = (size_t)c.geo.n * c.geo.n * c.geo.n;
  hdns_status s = HDNS_OK;
  if (s == HDNS_OK) s = alloc(ctx, &a.lq, 4 * cb);
  if (s == HDNS_OK) s = alloc(ctx, &a.lh, 4 * cb);
  if (s == HDNS_OK) s = alloc(ctx, &a.t6, 6 * cb);
  if (s == HDNS_OK) s = alloc(ctx, &a.t10, 10 * cb);
  if (s == HDNS_OK) s = alloc(ctx, &a.pa, c.pipe->adj_pencil_a(c) * c.elem);
  if (s == HDNS_OK) s = alloc(ctx, &a.pb, c.pipe->adj_pencil_b(c) * c.elem);
  for (int i = 0; i < c.stages && s == HDNS_OK; ++i) {
    s = alloc(ctx, &a.rq[i], 4 * cb);
    if (s == HDNS_OK) s = alloc(ctx, &a.rphys[i], 4 * vol * realb);
    if (s == HDNS_OK) s = alloc(ctx, &a.rphysw[i], 3 * vol * realb);
  }
  if (s == HDNS_OK) s = alloc(ctx, (void **)&a.rscal, 5 * 8 * sizeof(double));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&a.partials, ADJ_BLOCKS * sizeof(double));
  if (s == HDNS_OK) s = alloc(ctx, (void **)&a.proj, sizeof(double));
  if (s != HDNS_OK) return s;
  // the scalar components stay zero when the context carries no scalar
  zero_f64(c, a.lq, 4 * cb);
  zero_f64(c, a.lh, 4 * cb);
  for (int i = 0; i < c.stages; ++i) {
    zero_f64(c, a.rq[i], 4 * cb);
    zero_f64(c, a.rphys[i], 4 * vol * realb);
  }
  a.ready = true;
  return launched(ctx);
}

// copy the state ([4][comp] at src, or to dst)
void state_copy_in(Ctx &c, const void *src) {
  const size_t cb = c.comp * c.elem;
  (void)hipMemcpyAsync(c.uh, src, 3 * cb, hipMemcpyDeviceToDevice, c.stream);
  if (c.cfg.has_scalar) (void)hipMemcpyAsync(c.ph, (const char *)src + 3 * cb, cb, hipMemcpyDeviceToDevice, c.stream);
}
void state_copy_out(Ctx &c, void *dst) {
  const size_t cb = c.comp * c.elem;
  (void)hipMemcpyAsync(dst, c.uh, 3 * cb, hipMemcpyDeviceToDevice, c.stream);
  if (c.cfg.has_scalar) (void)hipMemcpyAsync((char *)dst + 3 * cb, c.ph, cb, hipMemcpyDeviceToDevice, c.stream);
}

// the transpose of one step at the current state: lam (in c.adj.lq, the output's adjoint) becomes
// the input's.  The state is recorded (advanced by one step) and restored.
// With `objective` non-null the step is the last one: its recorded forward step reaches the final
// state, whose J is written there and whose gradient seeds lam before the transpose runs.
hdns_status adj_step(hdns_ctx *ctx, double dt, double *objective = nullptr) {
  Ctx &c = ctx->c;
  AdjBufs &a = c.adj;
  StageRecord rec{};
  for (int i = 0; i < c.stages; ++i) {
    rec.q[i] = a.rq[i];
    rec.scal[i] = a.rscal + 8 * i;
    rec.phys[i] = a.rphys[i];
    rec.physw[i] = a.rphysw[i];
  }
  hdns_status s = run_deposit(ctx, a.rphys[0], a.rphysw[0]);
  if (s != HDNS_OK) return s;
  launch_step(c, dt, &rec);
  if (objective) {
    double sc[8];
    hipError_t e = hipMemcpyAsync(sc, c.scalars, sizeof sc, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return fail(ctx, HDNS_ERR_DEVICE, std::string("adjoint forward: ") + hipGetErrorString(e));
    *objective = 0.5 * sc[1] + (c.cfg.has_scalar ? sc[4] : 0.0);
    c.pipe->adj_objective(c);
  }
  state_copy_in(c, a.rq[0]);   // back to the step's input
  s = run_deposit(ctx);
  if (s != HDNS_OK) return s;
  const bool fp32 = !c.fp64;
  const Coeffs k = scheme_coefficients(c.cfg.scheme, fp32);
  c.pipe->adj_begin(c);
  for (int i = c.stages - 1; i >= 0; --i) {
    AdjStage as{};
    as.a = k.a[i];
    as.b = k.b[i];
    as.dt = fp32 ? (double)(float)dt : dt;
    as.sub = k.d[i] * dt;
    as.visc = c.visc[0];
    as.kappa = c.visc[0] / c.cfg.schmidt;
    as.mean_gradient = c.cfg.mean_gradient;
    as.last = i == c.stages - 1;
    as.scalar = c.cfg.has_scalar;
    as.band = c.band;
    as.scal = a.rscal + 8 * i;
    as.proj = a.proj;
    c.pipe->adj_stage(c, as, i);
  }
  return launched(ctx);
}

// host full layout <-> compact [4][comp] device array (components 0..2 from u, 3 from phi)
hdns_status compact_in(hdns_ctx *ctx, const double *u_full, const double *phi_full, void *dev) {
  Ctx &c = ctx->c;
  const Geo &g = c.geo;
  const int n = g.n, hx = n / 2 + 1;
  const size_t full = (size_t)n * n * hx;
  std::vector<unsigned char> host(4 * c.comp * c.elem, 0);
  for (int q = 0; q < 4; ++q) {
    const double *src = q < 3 ? u_full + 2 * full * q : phi_full;
    if (!src) continue;
    for (int kz = 0; kz < g.k; ++kz)
      for (int ky = 0; ky < g.k; ++ky) {
        const double *row = src + 2 * (((size_t)bin_index(kz, g) * n + bin_index(ky, g)) * hx);
        const size_t at = q * c.comp + ((size_t)kz * g.k + ky) * g.kxp;
        for (int ix = 0; ix <= g.m; ++ix) {
          if (c.fp64) {
            double *d = (double *)host.data() + 2 * (at + ix);
            d[0] = row[2 * ix];
            d[1] = row[2 * ix + 1];
          } else {
            float *d = (float *)host.data() + 2 * (at + ix);
            d[0] = (float)row[2 * ix];
            d[1] = (float)row[2 * ix + 1];
          }
        }
      }
  }
  HIP_TRY(ctx, hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return HDNS_OK;
}

hdns_status compact_out(hdns_ctx *ctx, const void *dev, double *u_full, double *phi_full) {
  Ctx &c = ctx->c;
  const Geo &g = c.geo;
  const int n = g.n, hx = n / 2 + 1;
  const size_t full = (size_t)n * n * hx;
  std::vector<unsigned char> host(4 * c.comp * c.elem);
  HIP_TRY(ctx, hipMemcpyAsync(host.data(), dev, host.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  for (int q = 0; q < 4; ++q) {
    double *dst = q < 3 ? u_full + 2 * full * q : phi_full;
    if (!dst) continue;
    memset(dst, 0, 2 * full * sizeof(double));
    for (int kz = 0; kz < g.k; ++kz)
      for (int ky = 0; ky < g.k; ++ky) {
        double *row = dst + 2 * (((size_t)bin_index(kz, g) * n + bin_index(ky, g)) * hx);
        const size_t at = q * c.comp + ((size_t)kz * g.k + ky) * g.kxp;
        for (int ix = 0; ix <= g.m; ++ix) {
          if (c.fp64) {
            const double *v = (const double *)host.data() + 2 * (at + ix);
            row[2 * ix] = v[0];
            row[2 * ix + 1] = v[1];
          } else {
            const float *v = (const float *)host.data() + 2 * (at + ix);
            row[2 * ix] = v[0];
            row[2 * ix + 1] = v[1];
          }
        }
      }
  }
  return HDNS_OK;
}

}  // namespace

extern "C" {

hdns_status hdns_step_vjp(hdns_ctx *ctx, double dt, const double *lam_u_out, const double *lam_phi_out,
                          double *lam_u_in, double *lam_phi_in) {
  hdns_status s = step_checks(ctx, dt);
  if (s != HDNS_OK) return s;
  if (!lam_u_out || !lam_u_in) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "the velocity adjoints must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  s = adj_alloc(ctx);
  if (s == HDNS_OK) s = compact_in(ctx, lam_u_out, c.cfg.has_scalar ? lam_phi_out : nullptr, c.adj.lq);
  if (s == HDNS_OK) s = adj_step(ctx, dt);
  if (s == HDNS_OK) s = compact_out(ctx, c.adj.lq, lam_u_in, c.cfg.has_scalar ? lam_phi_in : nullptr);
  return s;
}

hdns_status hdns_adjoint_gradient(hdns_ctx *ctx, double dt, int32_t steps, int32_t checkpoint_every,
                                  double *objective, double *grad_u, double *grad_phi) {
  hdns_status s = step_checks(ctx, dt);
  if (s != HDNS_OK) return s;
  if (steps < 1 || checkpoint_every < 1)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "steps and checkpoint_every must be at least 1");
  if (!objective || !grad_u) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "objective and grad_u must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  s = adj_alloc(ctx);
  if (s != HDNS_OK) return s;
  const size_t sb = 4 * c.comp * c.elem;   // one saved state
  const int every = std::min(checkpoint_every, (int)steps);
  const int nck = (steps + every - 1) / every;
  std::vector<void *> ck(nck, nullptr), seg(every, nullptr);
  auto cleanup = [&]() {
    for (void *p : ck)
      if (p) (void)hipFree(p);
    for (void *p : seg)
      if (p) (void)hipFree(p);
  };
  for (int i = 0; i < nck && s == HDNS_OK; ++i) {
    hipError_t e = hipMalloc(&ck[i], sb);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, "checkpoint allocation failed");
  }
  for (int i = 0; i < every && s == HDNS_OK; ++i) {
    hipError_t e = hipMalloc(&seg[i], sb);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, "segment allocation failed");
  }
  if (s != HDNS_OK) {
    cleanup();
    return s;
  }
  // forward, keeping a checkpoint every `every` steps (checkpoint 0 is the initial state)
  for (int t = 0; t < steps; ++t) {
    if (t % every == 0) state_copy_out(c, ck[t / every]);
    launch_step(c, dt);
  }
  double sc[8];
  hipError_t e = hipMemcpyAsync(sc, c.scalars, sizeof sc, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) {
    cleanup();
    return fail(ctx, HDNS_ERR_DEVICE, std::string("adjoint forward: ") + hipGetErrorString(e));
  }
  *objective = 0.5 * sc[1] + (c.cfg.has_scalar ? sc[4] : 0.0);
  c.pipe->adj_objective(c);
  // reverse sweep: each segment is recomputed from its checkpoint, then walked backwards
  for (int sgi = nck - 1; sgi >= 0 && s == HDNS_OK; --sgi) {
    const int t0 = sgi * every, t1 = std::min((int)steps, t0 + every);
    state_copy_in(c, ck[sgi]);
    s = run_deposit(ctx);   // a step starts from its state's products
    for (int t = t0; t < t1 && s == HDNS_OK; ++t) {
      state_copy_out(c, seg[t - t0]);
      if (t + 1 < t1) launch_step(c, dt);
    }
    for (int t = t1 - 1; t >= t0 && s == HDNS_OK; --t) {
      state_copy_in(c, seg[t - t0]);
      s = adj_step(ctx, dt);
    }
  }
  if (s == HDNS_OK) s = compact_out(ctx, c.adj.lq, grad_u, c.cfg.has_scalar ? grad_phi : nullptr);
  // the context is left at the initial state
  if (s == HDNS_OK) {
    state_copy_in(c, ck[0]);
    s = run_deposit(ctx);
  }
  if (s == HDNS_OK) {
    e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("adjoint: ") + hipGetErrorString(e));
  }
  cleanup();
  return s;
}

}  // extern "C"

namespace {

// Binomial checkpointing (A. Griewank, "Achieving logarithmic growth of temporal and spatial
// complexity in reverse automatic differentiation", Optim. Methods Softw. 1, 1992; the "revolve"
// schedule of Griewank and Walther, ACM TOMS 26, 2000).  With s state slots, one of them holding
// state l, the steps [l, l+n) are reversed by advancing from l to a split l+m, storing the state
// there, reversing [l+m, l+n) with s-1 slots, restoring state l and reversing [l, l+m) with s
// slots; with one slot every step is recomputed from state l.  beta(s, r) = C(s+r, s) steps can be
// reversed with s slots when each step runs forward at most r times.  For the smallest such r the
// split is m = max(n - beta(s-1, r), beta(s, r-2)): the right part then runs its steps r times and
// the left part r-1 times besides the advance to l+m, which balances the two (a step moved from
// one part to the other does not lower the count), and the total number of forward steps is
// Griewank's minimum r n - beta(s+1, r-1).  The adjoint of a step records its own forward step, so
// the last step's, the first turn, also yields the final state, whose objective seeds the adjoint.
enum { RV_ADVANCE = 0, RV_SNAPSHOT = 1, RV_RESTORE = 2, RV_FIRST_TURN = 3, RV_ADJOINT = 4 };
struct RvAction {
  int32_t op, step, slot;   // advance to `step`; store / restore state `step` in / from `slot`; the
                            // adjoint of step `step` (first turn: with the final state's seed)
};

int64_t binomial_reach(int s, int r) {   // C(s + r, s), saturated at 2^62
  int64_t b = 1;
  for (int i = 1; i <= r; ++i) {
    const __int128 p = (__int128)b * (s + i) / i;
    if (p > ((__int128)1 << 62)) return (int64_t)1 << 62;
    b = (int64_t)p;
  }
  return b;
}

void revolve_gen(int l, int n, int s, int slot, bool rightmost, std::vector<RvAction> &out) {
  // on entry the context holds state l, which slot `slot` also holds; s slots from `slot` on are ours
  if (n == 1) {
    out.push_back({rightmost ? RV_FIRST_TURN : RV_ADJOINT, l, slot});
    return;
  }
  if (s == 1) {   // no free slot: each step is recomputed from state l
    for (int t = l + n - 1; t >= l; --t) {
      if (t < l + n - 1) out.push_back({RV_RESTORE, l, slot});
      if (t > l) out.push_back({RV_ADVANCE, t, -1});
      out.push_back({rightmost && t == l + n - 1 ? RV_FIRST_TURN : RV_ADJOINT, t, slot});
    }
    return;
  }
  int r = 1;
  while (binomial_reach(s, r) < n) ++r;
  int64_t m = std::max<int64_t>(n - binomial_reach(s - 1, r), r >= 2 ? binomial_reach(s, r - 2) : 0);
  m = std::min<int64_t>(std::max<int64_t>(m, 1), n - 1);
  out.push_back({RV_ADVANCE, l + (int)m, -1});
  out.push_back({RV_SNAPSHOT, l + (int)m, slot + 1});
  revolve_gen(l + (int)m, n - (int)m, s - 1, slot + 1, rightmost, out);
  out.push_back({RV_RESTORE, l, slot});
  revolve_gen(l, (int)m, s, slot, false, out);
}

std::vector<RvAction> revolve_schedule(int steps, int snapshots) {
  std::vector<RvAction> out;
  out.push_back({RV_SNAPSHOT, 0, 0});
  revolve_gen(0, steps, std::min(snapshots, steps), 0, true, out);
  return out;
}

}  // namespace

extern "C" {

hdns_status hdns_revolve_schedule(int32_t steps, int32_t snapshots, int32_t *actions, int64_t capacity,
                                  int64_t *count) {
  if (steps < 1 || snapshots < 1 || !count) return HDNS_ERR_INVALID_ARGUMENT;
  const std::vector<RvAction> sched = revolve_schedule(steps, snapshots);
  *count = (int64_t)sched.size();
  if (!actions) return HDNS_OK;
  if (capacity < (int64_t)sched.size()) return HDNS_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < sched.size(); ++i) {
    actions[3 * i] = sched[i].op;
    actions[3 * i + 1] = sched[i].step;
    actions[3 * i + 2] = sched[i].slot;
  }
  return HDNS_OK;
}

hdns_status hdns_adjoint_gradient_revolve(hdns_ctx *ctx, double dt, int32_t steps, int32_t snapshots,
                                          double *objective, double *grad_u, double *grad_phi) {
  hdns_status s = step_checks(ctx, dt);
  if (s != HDNS_OK) return s;
  if (steps < 1 || snapshots < 1)
    return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "steps and snapshots must be at least 1");
  if (!objective || !grad_u) return fail(ctx, HDNS_ERR_INVALID_ARGUMENT, "objective and grad_u must not be NULL");
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipSetDevice(c.cfg.device));
  s = adj_alloc(ctx);
  if (s != HDNS_OK) return s;
  const std::vector<RvAction> sched = revolve_schedule(steps, snapshots);
  int nslot = 0;
  for (const RvAction &a : sched) nslot = std::max(nslot, a.slot + 1);
  const size_t sb = 4 * c.comp * c.elem;   // one saved state
  std::vector<void *> slot(nslot, nullptr);
  auto cleanup = [&]() {
    for (void *p : slot)
      if (p) (void)hipFree(p);
  };
  for (int i = 0; i < nslot && s == HDNS_OK; ++i) {
    const hipError_t e = hipMalloc(&slot[i], sb);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, "snapshot allocation failed");
  }
  int cur = 0;                 // the step whose state the context holds
  bool deposited = true;       // the context's products belong to its state
  for (size_t i = 0; i < sched.size() && s == HDNS_OK; ++i) {
    const RvAction &a = sched[i];
    switch (a.op) {
      case RV_ADVANCE:
        if (!deposited) s = run_deposit(ctx);
        for (; cur < a.step && s == HDNS_OK; ++cur) launch_step(c, dt);
        deposited = true;
        break;
      case RV_SNAPSHOT:
        state_copy_out(c, slot[a.slot]);
        break;
      case RV_RESTORE:
        state_copy_in(c, slot[a.slot]);
        cur = a.step;
        deposited = false;
        break;
      case RV_FIRST_TURN:   // the last step: its recorded forward step reaches the final state
        s = adj_step(ctx, dt, objective);
        deposited = true;
        break;
      case RV_ADJOINT:
        s = adj_step(ctx, dt);   // the context holds state a.step, as the schedule guarantees
        deposited = true;
        break;
    }
  }
  if (s == HDNS_OK) s = compact_out(ctx, c.adj.lq, grad_u, c.cfg.has_scalar ? grad_phi : nullptr);
  // the context is left at the initial state
  if (s == HDNS_OK) {
    state_copy_in(c, slot[0]);
    s = run_deposit(ctx);
  }
  if (s == HDNS_OK) {
    const hipError_t e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) s = fail(ctx, HDNS_ERR_DEVICE, std::string("adjoint: ") + hipGetErrorString(e));
  }
  cleanup();
  return s;
}

// --- ensembles ---------------------------------------------------------------------------

}  // extern "C"

struct hdns_ensemble {
  std::vector<hdns_ctx *> classes;          // one batch context per grid size
  std::vector<std::pair<int, int>> where;   // box -> (class, slot)
  hipStream_t stream = nullptr;
  // the captured step runs every batch after the first on a side stream of its own, forked from and
  // joined back to `stream` by events, so the graph holds one branch per grid size and the small
  // boxes' kernels overlap (default-priority streams; a batch's own kernels stay in order)
  std::vector<hipStream_t> side;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<int> graph_sched;             // the schedules the captured step holds (dt is written per step)
  std::string err;
};

namespace {

hdns_status efail(hdns_ensemble *e, hdns_status code, const std::string &msg) {
  if (e) e->err = msg;
  else g_create_error = msg;
  return code;
}

void drop_ensemble_graph(hdns_ensemble &e) {
  if (e.graph_exec) (void)hipGraphExecDestroy(e.graph_exec);
  if (e.graph) (void)hipGraphDestroy(e.graph);
  e.graph_exec = nullptr;
  e.graph = nullptr;
}

void release_ensemble(hdns_ensemble *e) {
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_ensemble_graph(*e);
  for (hdns_ctx *c : e->classes) {
    c->c.in_ensemble = false;
    hdns_ctx_destroy(c);
  }
  e->classes.clear();
  for (hipStream_t st : e->side)
    if (st) (void)hipStreamDestroy(st);
  e->side.clear();
  for (hipEvent_t ev : e->join)
    if (ev) (void)hipEventDestroy(ev);
  e->join.clear();
  if (e->fork) (void)hipEventDestroy(e->fork);
  e->fork = nullptr;
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->stream = nullptr;
}

std::vector<int> ensemble_schedules(const hdns_ensemble &e) {
  std::vector<int> s;
  for (hdns_ctx *c : e.classes) s.insert(s.end(), c->c.sched, c->c.sched + Ctx::NFAM);
  return s;
}

hdns_status ensemble_step_checks(hdns_ensemble *e, double dt) {
  if (!e) return HDNS_ERR_INVALID_ARGUMENT;
  if (!(isfinite(dt) && dt > 0)) return efail(e, HDNS_ERR_INVALID_ARGUMENT, "the timestep must be positive and finite");
  for (hdns_ctx *c : e->classes)
    if (!c->c.deposited) return efail(e, HDNS_ERR_STATE, "upload a state for every box before stepping the ensemble");
  return HDNS_OK;
}

}  // namespace

extern "C" {

hdns_status hdns_ensemble_create(const hdns_config *configs, int32_t nbox, hdns_ensemble **out) {
  if (!configs || !out || nbox < 1) return efail(nullptr, HDNS_ERR_INVALID_ARGUMENT, "configs, nbox >= 1 and out");
  *out = nullptr;
  const hdns_config &k0 = configs[0];
  for (int b = 1; b < nbox; ++b) {
    hdns_config k = configs[b];
    k.grid_points = k0.grid_points;
    k.viscosity = k0.viscosity;
    if (memcmp(&k, &k0, sizeof k) != 0)
      return efail(nullptr, HDNS_ERR_INVALID_ARGUMENT,
                   "the boxes of an ensemble may differ in grid_points and viscosity only (box " + std::to_string(b) +
                       " differs in another field)");
  }
  hdns_ensemble *e = new (std::nothrow) hdns_ensemble();
  if (!e) return efail(nullptr, HDNS_ERR_DEVICE, "out of host memory");
  e->device = k0.device;
  e->where.resize(nbox);
  std::vector<int> sizes;
  for (int b = 0; b < nbox; ++b)
    if (std::find(sizes.begin(), sizes.end(), configs[b].grid_points) == sizes.end()) sizes.push_back(configs[b].grid_points);
  hipError_t he = hipSetDevice(e->device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete e;
    return efail(nullptr, HDNS_ERR_DEVICE, std::string("ensemble stream: ") + hipGetErrorString(he));
  }
  for (size_t ci = 0; ci < sizes.size(); ++ci) {
    std::vector<double> visc;
    for (int b = 0; b < nbox; ++b)
      if (configs[b].grid_points == sizes[ci]) {
        e->where[b] = {(int)ci, (int)visc.size()};
        visc.push_back(configs[b].viscosity);
      }
    hdns_config k = k0;
    k.grid_points = sizes[ci];
    k.viscosity = visc[0];
    hdns_ctx *c = nullptr;
    hdns_status s = create_ctx(&k, (int)visc.size(), visc.data(), &c);
    if (s != HDNS_OK) {
      const std::string msg = g_create_error;
      release_ensemble(e);
      delete e;
      return efail(nullptr, s, "ensemble batch of " + std::to_string(sizes[ci]) + "^3 boxes: " + msg);
    }
    // the batch joins the ensemble's stream (its own is idle: creation synchronised it)
    (void)hipStreamDestroy(c->c.stream);
    c->c.stream = e->stream;
    c->c.own_stream = false;
    c->c.in_ensemble = true;
    e->classes.push_back(c);
  }
  // the side streams and events of the captured step's branches
  he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  e->side.assign(e->classes.size(), nullptr);
  e->join.assign(e->classes.size(), nullptr);
  for (size_t ci = 1; ci < e->classes.size() && he == hipSuccess; ++ci) {
    he = hipStreamCreateWithFlags(&e->side[ci], hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->join[ci], hipEventDisableTiming);
  }
  if (he != hipSuccess) {
    release_ensemble(e);
    delete e;
    return efail(nullptr, HDNS_ERR_DEVICE, std::string("ensemble streams: ") + hipGetErrorString(he));
  }
  *out = e;
  return HDNS_OK;
}

hdns_status hdns_ensemble_destroy(hdns_ensemble *ens) {
  if (!ens) return HDNS_OK;
  (void)hipSetDevice(ens->device);
  release_ensemble(ens);
  delete ens;
  return HDNS_OK;
}

const char *hdns_ensemble_last_error(const hdns_ensemble *ens) { return ens ? ens->err.c_str() : g_create_error.c_str(); }

hdns_status hdns_ensemble_member(hdns_ensemble *ens, int32_t box, hdns_ctx **ctx) {
  if (!ens || !ctx) return HDNS_ERR_INVALID_ARGUMENT;
  if (box < 0 || box >= (int)ens->where.size())
    return efail(ens, HDNS_ERR_INVALID_ARGUMENT, "box " + std::to_string(box) + " is outside the ensemble");
  hdns_ctx *c = ens->classes[ens->where[box].first];
  c->c.box = ens->where[box].second;
  *ctx = c;
  return HDNS_OK;
}

hdns_status hdns_ensemble_step(hdns_ensemble *ens, double dt) {
  hdns_status s = ensemble_step_checks(ens, dt);
  if (s != HDNS_OK) return s;
  if (hipSetDevice(ens->device) != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, "hipSetDevice failed");
  for (hdns_ctx *c : ens->classes) launch_step(c->c, dt);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
  return HDNS_OK;
}

hdns_status hdns_ensemble_step_graph(hdns_ensemble *ens, double dt) {
  hdns_status s = ensemble_step_checks(ens, dt);
  if (s != HDNS_OK) return s;
  if (hipSetDevice(ens->device) != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, "hipSetDevice failed");
  const std::vector<int> sched = ensemble_schedules(*ens);
  if (!ens->graph_exec || ens->graph_sched != sched) {
    drop_ensemble_graph(*ens);
    std::vector<bool> prof;
    for (hdns_ctx *c : ens->classes) {
      prof.push_back(c->c.prof);
      c->c.prof = false;  // no event records inside the captured step
    }
    hipError_t e = hipStreamBeginCapture(ens->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, std::string("capture: ") + hipGetErrorString(e));
    // fork: batch i >= 1 captures on its side stream, which waits for the origin stream
    const size_t nc = ens->classes.size();
    if (nc > 1) e = hipEventRecord(ens->fork, ens->stream);
    for (size_t i = 1; i < nc && e == hipSuccess; ++i) {
      e = hipStreamWaitEvent(ens->side[i], ens->fork, 0);
      ens->classes[i]->c.stream = ens->side[i];
    }
    if (e == hipSuccess)
      for (hdns_ctx *c : ens->classes) launch_step(c->c, dt, nullptr, true);
    // join: the origin stream waits for every branch
    for (size_t i = 1; i < nc; ++i) {
      if (e == hipSuccess) e = hipEventRecord(ens->join[i], ens->side[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(ens->stream, ens->join[i], 0);
      ens->classes[i]->c.stream = ens->stream;
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ens->stream, &graph);
    if (e == hipSuccess) e = ec;
    else if (graph) (void)hipGraphDestroy(graph);
    for (size_t i = 0; i < ens->classes.size(); ++i) ens->classes[i]->c.prof = prof[i];
    if (e != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, std::string("step capture failed: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return efail(ens, HDNS_ERR_DEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(e));
    }
    ens->graph = graph;
    ens->graph_exec = exec;
    ens->graph_sched = sched;
  }
  for (hdns_ctx *c : ens->classes)   // this step's dt into the values the captured kernels read
    hipLaunchKernelGGL(k_set_step_values, dim3(1), dim3(64), 0, ens->stream, c->c.step_values, step_values(c->c, dt));
  const hipError_t e = hipGraphLaunch(ens->graph_exec, ens->stream);
  if (e != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
  return HDNS_OK;
}

hdns_status hdns_ensemble_advance_tracers(hdns_ensemble *ens, double dt) {
  if (!ens) return HDNS_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(ens->device) != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, "hipSetDevice failed");
  // the batches' rebuilds and pushes are independent: batch i >= 1 runs on its side stream, forked
  // from the ensemble's stream and joined back to it, as in the captured field step
  const size_t nc = ens->classes.size();
  hipError_t he = nc > 1 ? hipEventRecord(ens->fork, ens->stream) : hipSuccess;
  for (size_t i = 1; i < nc && he == hipSuccess; ++i) {
    he = hipStreamWaitEvent(ens->side[i], ens->fork, 0);
    if (he == hipSuccess) ens->classes[i]->c.stream = ens->side[i];
  }
  hdns_status s = he == hipSuccess ? HDNS_OK : HDNS_ERR_DEVICE;
  std::string msg = he == hipSuccess ? std::string() : std::string("tracer streams: ") + hipGetErrorString(he);
  for (size_t i = 0; i < nc && s == HDNS_OK; ++i) {
    hdns_ctx *c = ens->classes[i];
    s = hdns_rebuild_traversal(c);
    if (s == HDNS_OK) s = hdns_advect_deposit(c, dt);
    if (s != HDNS_OK) msg = c->c.err;
  }
  for (size_t i = 1; i < nc; ++i) {
    if (ens->classes[i]->c.stream == ens->side[i]) {
      if (hipEventRecord(ens->join[i], ens->side[i]) != hipSuccess ||
          hipStreamWaitEvent(ens->stream, ens->join[i], 0) != hipSuccess) {
        (void)hipStreamSynchronize(ens->side[i]);   // the join must hold even if its event failed
        if (s == HDNS_OK) {
          s = HDNS_ERR_DEVICE;
          msg = "tracer streams: the join failed";
        }
      }
      ens->classes[i]->c.stream = ens->stream;
    }
  }
  if (s != HDNS_OK) return efail(ens, s, msg);
  return HDNS_OK;
}

hdns_status hdns_ensemble_synchronize(hdns_ensemble *ens) {
  if (!ens) return HDNS_ERR_INVALID_ARGUMENT;
  const hipError_t e = hipStreamSynchronize(ens->stream);
  if (e != hipSuccess) return efail(ens, HDNS_ERR_DEVICE, std::string("synchronize: ") + hipGetErrorString(e));
  return HDNS_OK;
}

}  // extern "C"
