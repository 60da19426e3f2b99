// Tracer kernels: the Catmull-Rom gather, the Adams pushes (Heun = ABM(1), ABM(k), AB(k)) with the
// periodic wrap, the closing sample into the velocity trail, the trail's time derivative, the
// canonical cell keys, the keys of the traversal order's stable sort, and the move of every
// tracer's records into that order.
//
// Positions and samples are stored [3][M] (the reference's (M, 3) column-major PointSet matrix).
// Grid point (i, j, k) sits at (i, j, k) * L/N; the fields the gathers read are interleaved,
// [z][y][x][4] with (u_x, u_y, u_z, phi) at every point.  The cell of x is
// floor(x * (N/L)) with N/L rounded once to the element type and the product formed in it, the
// contract the tests hold bit-exact.  Periodic indices wrap modulo N (a mask when N is a power of
// two); any integer, including the cast of a non-finite coordinate, lands inside the grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xcd.hpp"

namespace hdns {

__device__ __forceinline__ int wrap_index(int i, int n) {
  if ((n & (n - 1)) == 0) return i & (n - 1);
  const int r = i % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ T wrap_box(T x, T L) {
  // test/regressions/02-api-contracts-02.jl:1-6: fmod, lift negatives by L, and an equality
  // test at L (NaN and Inf fall through both comparisons untouched)
  T r = fmod(x, L);
  if (r < T(0)) r += L;
  if (r == L) r = T(0);
  return r;
}

template <typename T>
__device__ __forceinline__ void catmull_rom(T t, T *w) {
  w[0] = t * ((T(2) - t) * t - T(1)) * T(0.5);
  w[1] = (t * t * (T(3) * t - T(5)) + T(2)) * T(0.5);
  w[2] = t * ((T(4) - T(3) * t) * t + T(1)) * T(0.5);
  w[3] = t * t * (t - T(1)) * T(0.5);
}

struct Stencil {
  int i, j, k;
};

template <typename T>
__device__ __forceinline__ Stencil stencil_of(T x, T y, T z, T inv_dx, int n, T *wx, T *wy, T *wz) {
  const T sx = x * inv_dx, sy = y * inv_dx, sz = z * inv_dx;
  const T fx = floor(sx), fy = floor(sy), fz = floor(sz);
  catmull_rom(sx - fx, wx);
  catmull_rom(sy - fy, wy);
  catmull_rom(sz - fz, wz);
  return {wrap_index((int)fx, n), wrap_index((int)fy, n), wrap_index((int)fz, n)};
}

// One point of the interleaved physical fields, (u_x, u_y, u_z, phi), from a 16-byte-aligned
// address: one 16-byte load (FP32) or two (FP64).
typedef float point4_f __attribute__((ext_vector_type(4)));
typedef double point2_d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load_point4(const float *p, float *v) {
  const point4_f r = *reinterpret_cast<const point4_f *>(p);
  v[0] = r.x;
  v[1] = r.y;
  v[2] = r.z;
  v[3] = r.w;
}
__device__ __forceinline__ void load_point4(const double *p, double *v) {
  const point2_d a = reinterpret_cast<const point2_d *>(p)[0], b = reinterpret_cast<const point2_d *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

// a * b rounded on its own: the product carries no contraction flag, so it is never fused into a
// following addition
template <typename T>
__device__ __forceinline__ T rounded_product(T a, T b) {
#pragma clang fp contract(off)
  return a * b;
}

// The 4^3 gather of one tracer by a group of TRACER_LANES = 4 lanes from the interleaved fields f4
// ([N][N][N][4] per box: u_x, u_y, u_z, phi at every point): lane a takes the stencil's x point a
// (periodic) on all 16 (y, z) rows, so in every load the group reads the 64 contiguous bytes (FP32)
// of one stencil row, and sums wz[c] wy[b] f over the rows with fused multiply-adds in a fixed
// order; the group then adds wx[a] times its sums by butterflies (lanes l and l xor 1, then l xor
// 2), which every lane forms alike, so all of them hold the same values.  The first NF components
// are gathered: 3 (u) or 4 (u and phi).
constexpr int TRACER_LANES = 4;
template <typename T, int NF>
__device__ __forceinline__ void gather_points(const T *__restrict__ f4, int n, const Stencil &s, int l, const T *wx,
                                              const T *wy, const T *wz, T *out) {
  static_assert(TRACER_LANES == 4, "one x point of the stencil per lane");
  const int a = l & 3;
  const size_t ii = (size_t)wrap_index(s.i + a - 1, n);
  T acc[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) acc[q] = T(0);
  // FP32: all 16 points in flight (64 registers); FP64 points are twice as wide, and one plane at a
  // time keeps the gathering kernels inside their registers
#pragma unroll(sizeof(T) == 4 ? 4 : 1)
  for (int c = 0; c < 4; ++c) {
    const size_t kk = (size_t)wrap_index(s.k + c - 1, n);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const size_t jj = (size_t)wrap_index(s.j + b - 1, n);
      T v[4];
      load_point4(f4 + 4 * ((kk * n + jj) * n + ii), v);
      const T w = wz[c] * wy[b];
#pragma unroll
      for (int q = 0; q < NF; ++q) acc[q] = fma(w, v[q], acc[q]);
    }
  }
  // the product rounded on its own (never fused into the first butterfly's addition, which would
  // round the lanes' sums differently)
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    out[q] = rounded_product(wx[a], acc[q]);
#pragma unroll
    for (int m = 1; m < TRACER_LANES; m <<= 1) out[q] = out[q] + __shfl_xor(out[q], m, TRACER_LANES);
  }
}

struct TracerArgs {
  int64_t m;          // tracers per box
  int n;
  double side_length;
  int64_t cap;        // tracer slots per box: box b's arrays start at b * cap (positions [3][m] inside)
  size_t box_phys;    // reals of the physical fields per box
  int ntiles;         // traversal tiles per box
  int64_t vstride;    // reals of the velocity trail per box: [slot][3][cap]
};

constexpr int MAX_TRAIL = 8;   // samples an Adams step integrates over, and the longest trail

// One Adams step over the trail's newest J samples f_j (newest first), with weights already
// formed on the host from the samples' times (src/ParticlePropagator.jl:287-316):
//   X* = wrap(X + sum_j b_j f_j);  when corrected, f* = u(X*) and
//   X <- wrap(X* + c_0 f* + sum_j c_{j+1} f_j),  c_0 = beta_*, c_{j+1} = beta_{j+1} - b_j.
// Each sum is added left to right in the order written (one fused displacement each).  The
// unwrapped record U, when tracked (upos non-null), takes the same terms in the same order and is
// never reduced to the box (src/ParticlePropagator.jl:299,313).
struct AdamsArgs {
  int J;
  int corrected;
  int64_t slot[MAX_TRAIL];   // reals from a box's trail start to sample j: slot * 3 * cap
  double b[MAX_TRAIL];
  double c[MAX_TRAIL + 1];
};

// One tracer's Adams step; g3(stencil, wx, wy, wz, out) gathers u at a point.  The new (wrapped)
// position, as stored, is also left in xn.  Every lane of the tracer's group runs it (the gather
// needs all of them); only the writer stores.
template <typename T, typename G3>
__device__ __forceinline__ void advect_one(int64_t p, T *__restrict__ pos, T *__restrict__ upos,
                                           const T *__restrict__ trail, const TracerArgs &ta, const AdamsArgs &aa,
                                           G3 g3, T (&xn)[3], bool writer) {
  const T L = T(ta.side_length), inv_dx = T(ta.n / ta.side_length);
  T F[MAX_TRAIL][3], acc[3], Xs[3], ua[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int d = 0; d < 3; ++d) acc[d] = pos[d * ta.m + p];
  if (upos) {
#pragma unroll
    for (int d = 0; d < 3; ++d) ua[d] = upos[d * ta.m + p];
  }
#pragma unroll
  for (int j = 0; j < MAX_TRAIL; ++j) {
    if (j >= aa.J) break;
    const T bj = T(aa.b[j]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      F[j][d] = trail[aa.slot[j] + d * ta.m + p];
      acc[d] = acc[d] + bj * F[j][d];
      ua[d] = ua[d] + bj * F[j][d];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) Xs[d] = wrap_box(acc[d], L);
  if (!aa.corrected) {
#pragma unroll
    for (int d = 0; d < 3; ++d) xn[d] = Xs[d];
    if (!writer) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) pos[d * ta.m + p] = Xs[d];
    if (upos) {
#pragma unroll
      for (int d = 0; d < 3; ++d) upos[d * ta.m + p] = ua[d];
    }
    return;
  }
  T wx[4], wy[4], wz[4], fs[3];
  const Stencil s = stencil_of(Xs[0], Xs[1], Xs[2], inv_dx, ta.n, wx, wy, wz);
  g3(s, wx, wy, wz, fs);
  const T c0 = T(aa.c[0]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    acc[d] = Xs[d] + c0 * fs[d];
    ua[d] = ua[d] + c0 * fs[d];
  }
#pragma unroll
  for (int j = 0; j < MAX_TRAIL; ++j) {
    if (j >= aa.J) break;
    const T cj = T(aa.c[j + 1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[d] = acc[d] + cj * F[j][d];
      ua[d] = ua[d] + cj * F[j][d];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) xn[d] = wrap_box(acc[d], L);
  if (!writer) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) pos[d * ta.m + p] = xn[d];
  if (upos) {
#pragma unroll
    for (int d = 0; d < 3; ++d) upos[d * ta.m + p] = ua[d];
  }
}

// The thread of a gathering kernel: tracer t (in traversal order) of its box, lane c of the
// tracer's group of TRACER_LANES.  Blocks of one XCD take consecutive chunks of the order, so the
// field lines that neighbouring tiles share stay in that XCD's L2.
struct GroupThread {
  int64_t t;
  int c;
};
__device__ __forceinline__ GroupThread group_thread() {
  const int64_t g = xcd_chunk(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x;
  return {g / TRACER_LANES, (int)(g % TRACER_LANES)};
}

// The push in tracer (or traversal) order, a group of TRACER_LANES lanes per tracer.
template <typename T>
__global__ void k_advect(T *__restrict__ pos, T *__restrict__ upos, const T *__restrict__ trail, const T *__restrict__ u,
                         const int *__restrict__ order, TracerArgs ta, AdamsArgs aa) {
  const GroupThread qt = group_thread();
  const int64_t t = qt.t;
  if (t >= ta.m) return;   // the whole group leaves together
  const int box = blockIdx.y;
  pos += box * 3 * ta.cap;
  if (upos) upos += box * 3 * ta.cap;
  trail += box * ta.vstride;
  u += box * ta.box_phys;
  if (order) order += box * ta.cap;
  const int64_t p = order ? order[t] : t;
  T xn[3];
  advect_one(
      p, pos, upos, trail, ta, aa,
      [&](const Stencil &s, const T *wx, const T *wy, const T *wz, T *out) {
        gather_points<T, 3>(u, ta.n, s, qt.c, wx, wy, wz, out);
      },
      xn, qt.c == 0);
}

// sum_j w_j f_j over n trail samples of the selected box, added left to right (time_derivative!,
// src/Particles.jl:597-631); out [3][m].
struct TrailCombine {
  int n;
  int64_t slot[MAX_TRAIL];
  double w[MAX_TRAIL];
};

template <typename T>
__global__ void k_trail_combine(const T *__restrict__ trail, TracerArgs ta, TrailCombine tc, T *__restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ta.m) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T acc = T(tc.w[0]) * trail[tc.slot[0] + d * ta.m + t];
    for (int j = 1; j < tc.n; ++j) acc = acc + T(tc.w[j]) * trail[tc.slot[j] + d * ta.m + t];
    out[d * ta.m + t] = acc;
  }
}

// The particle deposit! (src/ParticlePropagator.jl:378-386): u (and phi) at the positions, into
// the trail slot `vel` points at (inside box 0's trail; box b's is b * vstride further).  The
// physical fields are [4][N^3] per box (u, then phi), so phi is u's fourth component.
template <typename T, typename G>
__device__ __forceinline__ void sample_at(int64_t p, T x, T y, T z, T *__restrict__ vel, T *__restrict__ phis,
                                          const TracerArgs &ta, G g, bool writer) {
  const T inv_dx = T(ta.n / ta.side_length);
  T wx[4], wy[4], wz[4], v[4];
  const Stencil s = stencil_of(x, y, z, inv_dx, ta.n, wx, wy, wz);
  g(s, wx, wy, wz, v, phis != nullptr);   // u, and phi when sampled
  if (!writer) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) vel[d * ta.m + p] = v[d];
  if (phis) phis[p] = v[3];
}

template <typename T>
__global__ void k_sample(const T *__restrict__ pos, T *__restrict__ vel, T *__restrict__ phis,
                         const T *__restrict__ u, const int *__restrict__ order, TracerArgs ta) {
  const GroupThread qt = group_thread();
  const int64_t t = qt.t;
  if (t >= ta.m) return;   // the whole group leaves together
  const int box = blockIdx.y;
  pos += box * 3 * ta.cap;
  vel += box * ta.vstride;
  if (phis) phis += box * ta.cap;
  u += box * ta.box_phys;
  if (order) order += box * ta.cap;
  const int64_t p = order ? order[t] : t;
  sample_at(
      p, pos[p], pos[ta.m + p], pos[2 * ta.m + p], vel, phis, ta,
      [&](const Stencil &s, const T *wx, const T *wy, const T *wz, T *out, bool with_phi) {
        if (with_phi) gather_points<T, 4>(u, ta.n, s, qt.c, wx, wy, wz, out);
        else gather_points<T, 3>(u, ta.n, s, qt.c, wx, wy, wz, out);
      },
      qt.c == 0);
}

// The push and the closing sample in one pass (an iteration's advect! then deposit!, both on the
// new field): the sample is taken at the new position the push leaves in registers, whose 4^3
// stencil mostly overlaps the one the corrector just gathered.  Each value is the one k_advect then
// k_sample would produce, bit for bit.  The sample goes into the trail slot `vel` points at; when
// that slot is also one the push reads (a full AB(k) ring), the group reads its tracer's samples
// there before its writer stores the new one, and no other group touches that tracer's entries.
template <typename T>
__global__ void k_advect_sample(T *__restrict__ pos, T *__restrict__ upos, const T *trail, T *vel,
                                T *__restrict__ phis, const T *__restrict__ u, const int *__restrict__ order,
                                TracerArgs ta, AdamsArgs aa) {
  const GroupThread qt = group_thread();
  const int64_t t = qt.t;
  if (t >= ta.m) return;   // the whole group leaves together
  const int box = blockIdx.y;
  pos += box * 3 * ta.cap;
  if (upos) upos += box * 3 * ta.cap;
  trail += box * ta.vstride;
  vel += box * ta.vstride;
  if (phis) phis += box * ta.cap;
  u += box * ta.box_phys;
  if (order) order += box * ta.cap;
  const int64_t p = order ? order[t] : t;
  T xn[3];
  advect_one(
      p, pos, upos, trail, ta, aa,
      [&](const Stencil &s, const T *wx, const T *wy, const T *wz, T *out) {
        gather_points<T, 3>(u, ta.n, s, qt.c, wx, wy, wz, out);
      },
      xn, qt.c == 0);
  sample_at(
      p, xn[0], xn[1], xn[2], vel, phis, ta,
      [&](const Stencil &s, const T *wx, const T *wy, const T *wz, T *out, bool with_phi) {
        if (with_phi) gather_points<T, 4>(u, ta.n, s, qt.c, wx, wy, wz, out);
        else gather_points<T, 3>(u, ta.n, s, qt.c, wx, wy, wz, out);
      },
      qt.c == 0);
}

template <typename T>
__global__ void k_cell_keys(const T *__restrict__ pos, int64_t *__restrict__ keys, TracerArgs ta) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= ta.m) return;   // one box per launch: the caller passes that box's arrays
  const T inv_dx = T(ta.n / ta.side_length);
  const int i = wrap_index((int)floor(pos[p] * inv_dx), ta.n);
  const int j = wrap_index((int)floor(pos[ta.m + p] * inv_dx), ta.n);
  const int k = wrap_index((int)floor(pos[2 * ta.m + p] * inv_dx), ta.n);
  keys[p] = ((int64_t)k * ta.n + j) * ta.n + i;
}

// Traversal order (src/Particles.jl:220-271): the tracers sorted by 8^3-cell tile, and within a
// tile by index.  The sort key of tracer p of box b is b * ntiles + tile, so one stable sort of the
// (key, p) pairs of every box leaves box b's order in its own run of m entries.  The order only
// reorders the visits, so every gather gives the same values with or without it.
template <typename T>
__device__ __forceinline__ int tile_of(const T *pos, int64_t m, int64_t p, int n, T inv_dx) {
  const int nt = n >> 3;
  const int i = wrap_index((int)floor(pos[p] * inv_dx), n);
  const int j = wrap_index((int)floor(pos[m + p] * inv_dx), n);
  const int k = wrap_index((int)floor(pos[2 * m + p] * inv_dx), n);
  return ((k >> 3) * nt + (j >> 3)) * nt + (i >> 3);
}

// Slot s holds tracer ids[s]: its (key, slot) pair goes to entry box * m + ids[s], so the sort's
// input is in tracer-index order whatever the slots' arrangement (a rebuilt order equals a fresh
// one), and the sorted values name the slot each tracer of the new order comes from.
template <typename T>
__global__ void k_tile_keys(const T *__restrict__ pos, const int32_t *__restrict__ ids, uint32_t *__restrict__ keys,
                            int32_t *__restrict__ vals, TracerArgs ta) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= ta.m) return;
  const int box = blockIdx.y;
  pos += box * 3 * ta.cap;
  const int32_t id = ids[box * ta.cap + s];
  if (id < 0 || id >= ta.m) return;   // ids is a permutation of 0 .. m-1
  const size_t i = (size_t)box * ta.m + id;
  keys[i] = (uint32_t)box * (uint32_t)ta.ntiles + (uint32_t)tile_of(pos, ta.m, s, ta.n, T(ta.n / ta.side_length));
  vals[i] = (int32_t)s;
}

// The records into the traversal order: slot t of box b receives the tracer of slot from[t] (box
// b's run of the sorted values), with everything it owns -- its position, its unwrapped position,
// its samples in every slot of the trail's ring, its scalar sample -- and its index, which is also
// entry t of the traversal order.  The outputs are the second buffers, which the caller swaps in.
template <typename T>
__global__ void k_slot_permute(const int32_t *__restrict__ from, const int32_t *__restrict__ ids,
                               int32_t *__restrict__ ids_out, int *__restrict__ order, const T *__restrict__ pos,
                               T *__restrict__ pos_out, const T *__restrict__ upos, T *__restrict__ upos_out,
                               const T *__restrict__ vel, T *__restrict__ vel_out, int ring, const T *__restrict__ phis,
                               T *__restrict__ phis_out, TracerArgs ta) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ta.m) return;
  const int box = blockIdx.y;
  const int64_t cap = ta.cap, m = ta.m;
  const int64_t s = from[box * m + t];
  if (s < 0 || s >= m) return;
  const int32_t id = ids[box * cap + s];
  ids_out[box * cap + t] = id;
  order[box * cap + t] = id;
  const size_t pb = (size_t)box * 3 * cap;
#pragma unroll
  for (int d = 0; d < 3; ++d) pos_out[pb + d * m + t] = pos[pb + d * m + s];
  if (upos) {
#pragma unroll
    for (int d = 0; d < 3; ++d) upos_out[pb + d * m + t] = upos[pb + d * m + s];
  }
  const size_t vb = (size_t)box * ta.vstride;
  for (int r = 0; r < ring; ++r)
#pragma unroll
    for (int d = 0; d < 3; ++d) vel_out[vb + (size_t)r * 3 * cap + d * m + t] = vel[vb + (size_t)r * 3 * cap + d * m + s];
  if (phis) phis_out[box * cap + t] = phis[box * cap + s];
}

}  // namespace hdns
