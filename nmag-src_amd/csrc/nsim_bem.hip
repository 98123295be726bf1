// Dense Fredkin-Koehler boundary matrix: device assembly (Lindholm weights) and
// the per-evaluation GEMV phi2|_b = B phi1|_b (DM*V BEM, nmag_lam.py:749).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "nsim_dev_common.h"
#include "nsim_device.h"

namespace nsim {

namespace {

using namespace dev;

// Lindholm weight of vertex slot `a` of triangle (p0,p1,p2) seen from observer
// o (src/bem3d_stubs.c:413-610, identity lattice transform, min_dist 0).
__device__ double lindholm_weight(const double* o, const double* p0, const double* p1,
                                  const double* p2, const double* nout, double geo_eps,
                                  int a) {
  double s0[3], s1[3], s2[3], r0[3], r1[3], r2[3];
  for (int d = 0; d < 3; d++) {
    s0[d] = p2[d] - p1[d];
    s1[d] = p0[d] - p2[d];
    s2[d] = p1[d] - p0[d];
    r0[d] = p0[d] - o[d];
    r1[d] = p1[d] - o[d];
    r2[d] = p2[d] - o[d];
  }
  const double r0l = sqrt(dot3(r0, r0)), r1l = sqrt(dot3(r1, r1)), r2l = sqrt(dot3(r2, r2));
  if (r0l <= 0.0 || r1l <= 0.0 || r2l <= 0.0) return 0.0;  // observer on a vertex
  const double s0l = sqrt(dot3(s0, s0)), s1l = sqrt(dot3(s1, s1)), s2l = sqrt(dot3(s2, s2));
  double zv[3];
  cross(s1, s2, zv);
  const double zl = sqrt(dot3(zv, zv));
  if (zl == 0.0 || s0l == 0.0 || s1l == 0.0 || s2l == 0.0) return 0.0;
  for (int d = 0; d < 3; d++) zv[d] /= zl;
  const double area = 0.5 * zl;
  double xi0[3], xi1[3], xi2[3];
  for (int d = 0; d < 3; d++) {
    xi0[d] = s0[d] / s0l;
    xi1[d] = s1[d] / s1l;
    xi2[d] = s2[d] / s2l;
  }
  const double c01 = dot3(xi0, xi1), c12 = dot3(xi1, xi2), c20 = dot3(xi2, xi0);
  const double lg0 = log((r1l + r2l + s0l) / (r1l + r2l - s0l));
  const double lg1 = log((r2l + r0l + s1l) / (r2l + r0l - s1l));
  const double lg2 = log((r0l + r1l + s2l) / (r0l + r1l - s2l));
  const double sp01 = dot3(r0, r1), sp12 = dot3(r1, r2), sp20 = dot3(r2, r0);
  const double num = r0l * r1l * r2l + r0l * sp12 + r1l * sp20 + r2l * sp01;
  const double den =
      sqrt(2.0 * (r1l * r2l + sp12) * (r2l * r0l + sp20) * (r0l * r1l + sp01));
  if (den == 0.0) return 0.0;
  // The reference's 2*acos(num/den)*sign(zeta) equals 2*atan2(r0.(r1 x r2), num)
  // (den^2 = num^2 + triple^2, zeta ~ triple); atan2 keeps the small angles of
  // distant triangles to full precision where acos near 1 loses half the digits.
  double r1xr2[3];
  cross(r1, r2, r1xr2);
  const double angle = 2.0 * atan2(dot3(r0, r1xr2), num);
  const double zeta = dot3(zv, r0);
  double f = 1.0 / (8.0 * kPi * area);
  if (dot3(zv, nout) < 0.0) f = -f;
  if (fabs(zeta) < geo_eps) f = 0.0;
  double eta[3], res;
  if (a == 0) {
    cross(zv, xi0, eta);
    res = f * s0l * (angle * dot3(eta, r1) - zeta * (lg0 + c01 * lg1 + c20 * lg2));
  } else if (a == 1) {
    cross(zv, xi1, eta);
    res = f * s1l * (angle * dot3(eta, r2) - zeta * (c01 * lg0 + lg1 + c12 * lg2));
  } else {
    cross(zv, xi2, eta);
    res = f * s2l * (angle * dot3(eta, r0) - zeta * (c20 * lg0 + c12 * lg1 + lg2));
  }
  return res;
}

// One thread per matrix entry (i, j): B_ij = sum over the triangles incident to
// boundary node j of j's Lindholm weight seen from node i, plus the diagonal
// Omega_i/(4 pi) - 1 (src/bem3d.ml:938-1046).  Deterministic; writes coalesced.
template <typename TB>
__global__ __launch_bounds__(256) void k_bem_assemble(int nb, int ldb, int col_blocks,
                                                      const double* __restrict__ x,
                                                      const int* __restrict__ bnode,
                                                      const int* __restrict__ tri_node,
                                                      const double* __restrict__ tnorm,
                                                      const int* __restrict__ inc_rp,
                                                      const int* __restrict__ inc_tri,
                                                      const double* __restrict__ omega,
                                                      double geo_eps, TB* __restrict__ B) {
  const int i = blockIdx.x / col_blocks;
  const int j = (blockIdx.x % col_blocks) * 256 + threadIdx.x;
  if (i >= nb || j >= ldb) return;
  double v = 0.0;
  if (j < nb) {
    const int oi = bnode[i];
    const double o[3] = {x[3 * oi], x[3 * oi + 1], x[3 * oi + 2]};
    for (int k = inc_rp[j]; k < inc_rp[j + 1]; k++) {
      const int t = inc_tri[k] / 3, a = inc_tri[k] % 3;
      const int v0 = tri_node[3 * t], v1 = tri_node[3 * t + 1], v2 = tri_node[3 * t + 2];
      const double p0[3] = {x[3 * v0], x[3 * v0 + 1], x[3 * v0 + 2]};
      const double p1[3] = {x[3 * v1], x[3 * v1 + 1], x[3 * v1 + 2]};
      const double p2[3] = {x[3 * v2], x[3 * v2 + 1], x[3 * v2 + 2]};
      v += lindholm_weight(o, p0, p1, p2, tnorm + 3 * t, geo_eps, a);
    }
    if (i == j) v += omega[i] / (4.0 * kPi) - 1.0;
  }
  store_b<TB>(B + (size_t)i * ldb + j, v);
}

// Storage type TB of B, type TX of the LDS-staged input vector, accumulation
// type TA.  FP32 storage is accumulated in FP64 (the kernel is HBM-bound, the
// FP64 FMAs are free; an FP32 input, which would halve the LDS footprint,
// measured no faster at the 20k-node slab and is not used); BF16 storage uses
// FP32 inputs and accumulation.
template <typename TB>
struct BemVec;
template <>
struct BemVec<double> {
  static constexpr int V = 2;  // 16 B per lane per load
  using X = double;
  using A = double;
};
template <>
struct BemVec<float> {
  static constexpr int V = 4;
  using X = double;
  using A = double;
};
template <>
struct BemVec<uint16_t> {
  static constexpr int V = 8;
  using X = float;
  using A = float;
};
template <>
struct BemVec<F24> {
  static constexpr int V = 8;  // one 16-byte and one 8-byte load per lane
  using X = double;
  using A = double;
};

// 16-byte row chunks of a dense operator.  The operator is streamed once per
// launch with no reuse, so the loads are non-temporal: on MI355X this lifts a
// plain streaming read from about 6.4 to 7.1 TB/s, and the GEMV with it.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

template <typename TB, typename X>
__device__ inline void load_row_chunk(const TB* p, X* out);
template <>
__device__ inline void load_row_chunk<double, double>(const double* p, double* out) {
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
  out[0] = v.x;
  out[1] = v.y;
}
template <>
__device__ inline void load_row_chunk<float, double>(const float* p, double* out) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}
template <>
__device__ inline void load_row_chunk<uint16_t, float>(const uint16_t* p, float* out) {
  const u4v v = __builtin_nontemporal_load(reinterpret_cast<const u4v*>(p));
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    out[2 * k] = __uint_as_float(w[k] << 16);
    out[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// Entries c .. c+V-1 of a row of length ldb.
template <typename TB, typename X>
__device__ inline void load_row(const TB* row, int c, int ldb, X* out) {
  load_row_chunk<TB, X>(row + c, out);
}
template <>
__device__ inline void load_row<F24, double>(const F24* row, int c, int ldb, double* out) {
  f24_load8(row, c, ldb, out);
}

// xs[j] = x[j] for j < nb, 0 up to ldb, by T threads: eight loads in flight
// per thread before any LDS store (a plain loop issues one load at a time
// and waits on L2 latency before every LDS store).
template <typename X, int T>
__device__ inline void stage_x(X* xs, const double* __restrict__ x, int nb, int ldb) {
  constexpr int U = 8;
  for (int j0 = threadIdx.x; j0 < ldb; j0 += T * U) {
    X v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + u * T;
      v[u] = j < nb ? (X)x[j] : (X)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + u * T;
      if (j < ldb) xs[j] = v[u];
    }
  }
}

// NR consecutive rows rb.. of a row-major segment (stride ldb) against the
// staged x: y[map[rb + k]] = row . x.  Every lane loads V entries of each
// row per step (16 bytes; 24 for 24-bit storage); the partial sums are
// reduced across the wavefront.
template <typename TB, int NR, typename X>
__device__ inline void gemv_rows(const TB* __restrict__ B, int ldb, int rb, const X* xs,
                                 int lane, const int* __restrict__ map, double* __restrict__ y) {
  using A = typename BemVec<TB>::A;
  constexpr int V = BemVec<TB>::V;
  const TB* rp[NR];
  A acc[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    rp[k] = B + (size_t)(rb + k) * ldb;
    acc[k] = 0;
  }
#pragma unroll 4
  for (int c = lane * V; c < ldb; c += 64 * V) {
    X xv[V];
#pragma unroll
    for (int e = 0; e < V; e++) xv[e] = xs[c + e];
#pragma unroll
    for (int k = 0; k < NR; k++) {
      X bv[V];
      load_row<TB, X>(rp[k], c, ldb, bv);
#pragma unroll
      for (int e = 0; e < V; e++) acc[k] += (A)bv[e] * (A)xv[e];
    }
  }
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double t = wave_sum((double)acc[k]);
    if (lane == 0) y[map[rb + k]] = t;
  }
}

// y[map[off + r]] = sum_j A_rj x[off + j] over a block-diagonal set of dense
// matrices (one per independent magnet of a batch; one for a single magnet):
// the BEM GEMV phi2|_b = B phi1|_b and the dense potential phi = M rho.
// A table maps every workgroup to a segment and a run of its rows.  The
// segment's input vector is staged once per workgroup in LDS (up to 160 KiB
// on gfx950); each wavefront streams ROWS rows at a time with 16-byte loads
// per lane, then its remaining rows one at a time, so a wavefront's row count
// need not be a multiple of ROWS (no row is read twice).  KIND only names the
// instantiation after its operator (0: BEM, 1: potential) so that profiles
// report the two GEMVs apart.
template <typename TB, int ROWS, int WAVES, int KIND>
__global__ __launch_bounds__(64 * WAVES) void k_dense_gemv(
    const TB* __restrict__ Abase, const long long* __restrict__ seg_aoff,
    const int* __restrict__ seg_ld, const int* __restrict__ seg_n,
    const int* __restrict__ seg_off, const int* __restrict__ wg_seg,
    const int* __restrict__ wg_row0, int rows_per_wave, const double* __restrict__ xall,
    const int* __restrict__ mapall, double* __restrict__ y, const int* __restrict__ halt) {
  if (halt && *halt) return;  // queued by a finished integrator run
  using X = typename BemVec<TB>::X;
  constexpr int T = 64 * WAVES;
  extern __shared__ __align__(16) unsigned char smem[];
  X* xs = reinterpret_cast<X*>(smem);
  const int sg = wg_seg[blockIdx.x];
  const int ldb = seg_ld[sg], nb = seg_n[sg], boff = seg_off[sg];
  const TB* B = Abase + seg_aoff[sg];
  const double* xb = xall + boff;
  const int* map = mapall + boff;
  stage_x<X, T>(xs, xb, nb, ldb);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r0 = wg_row0[blockIdx.x] + (threadIdx.x >> 6) * rows_per_wave;
  const int r1 = min(r0 + rows_per_wave, nb);
  int rb = r0;
  for (; rb + ROWS <= r1; rb += ROWS) gemv_rows<TB, ROWS, X>(B, ldb, rb, xs, lane, map, y);
  for (; rb < r1; rb++) gemv_rows<TB, 1, X>(B, ldb, rb, xs, lane, map, y);
}

// Single-segment variant with a few workgroups per CU: x is staged once per
// workgroup, then every wavefront claims rows from a device counter, the
// claim of its next row issued before the current row is streamed.  Each
// wavefront makes exactly one claim past the last row, so the counter ends at
// n + total_waves; the last wavefront out (counted in ctr[1]) resets both
// words, and launches need no host-side reset (graph replays included).
template <typename TB, int WAVES, int KIND>
__global__ __launch_bounds__(64 * WAVES) void k_dense_gemv_dyn(
    const TB* __restrict__ B, int ldb, int nb, const double* __restrict__ x,
    const int* __restrict__ map, double* __restrict__ y, unsigned* ctr, unsigned total_waves,
    const int* __restrict__ halt) {
  if (halt && *halt) return;  // every wavefront: the counter stays untouched
  using X = typename BemVec<TB>::X;
  constexpr int T = 64 * WAVES;
  extern __shared__ __align__(16) unsigned char smem[];
  X* xs = reinterpret_cast<X*>(smem);
  stage_x<X, T>(xs, x, nb, ldb);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  unsigned r = 0;
  if (lane == 0) r = atomicAdd(&ctr[0], 1u);
  r = __shfl(r, 0, 64);
  while (r < (unsigned)nb) {
    unsigned next = 0;
    if (lane == 0) next = atomicAdd(&ctr[0], 1u);
    gemv_rows<TB, 1, X>(B, ldb, (int)r, xs, lane, map, y);
    r = __shfl(next, 0, 64);
  }
  if (lane == 0 && atomicAdd(&ctr[1], 1u) == total_waves - 1) {
    __hip_atomic_store(&ctr[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctr[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

constexpr size_t kMaxLds = 160 * 1024;  // LDS per workgroup on gfx950

// Raises a kernel's dynamic-LDS limit above the 64 KiB default once per
// kernel, device and size, so that a launch makes no runtime call besides
// the launch itself.
inline hipError_t allow_lds(const void* k, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  struct Done {
    const void* k;
    int dev;
    size_t lds;
  };
  static std::mutex mu;
  static std::vector<Done> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  for (const Done& d : done)
    if (d.k == k && d.dev == dev && d.lds >= lds) return hipSuccess;
  e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) done.push_back(Done{k, dev, lds});
  return e;
}

template <typename TB, int ROWS, int WAVES, int KIND>
hipError_t launch_shape(const DenseOp& op, size_t lds, const double* x, double* y,
                        hipStream_t st) {
  const void* k = reinterpret_cast<const void*>(&k_dense_gemv<TB, ROWS, WAVES, KIND>);
  const hipError_t e = allow_lds(k, lds);
  if (e != hipSuccess) return e;
  k_dense_gemv<TB, ROWS, WAVES, KIND><<<op.grid, 64 * WAVES, lds, st>>>(
      (const TB*)op.A, op.aoff, op.ld, op.n, op.off, op.wg_seg, op.wg_row0, op.rpw, x, op.map,
      y, op.halt);
  return hipGetLastError();
}

template <typename TB, int WAVES, int KIND>
hipError_t launch_dyn_shape(const DenseOp& op, size_t lds, const double* x, double* y,
                            hipStream_t st) {
  const void* k = reinterpret_cast<const void*>(&k_dense_gemv_dyn<TB, WAVES, KIND>);
  const hipError_t e = allow_lds(k, lds);
  if (e != hipSuccess) return e;
  k_dense_gemv_dyn<TB, WAVES, KIND><<<op.grid, 64 * WAVES, lds, st>>>(
      (const TB*)op.A, op.h_ld[0], op.h_n[0], x + op.h_off[0], op.map + op.h_off[0], y, op.ctr,
      (unsigned)(op.grid * WAVES), op.halt);
  return hipGetLastError();
}

template <typename TB, int KIND>
hipError_t launch_dyn(const DenseOp& op, size_t lds, const double* x, double* y,
                      hipStream_t st) {
  if (!op.ctr || op.h_n.size() != 1 || op.h_aoff[0] != 0) return hipErrorInvalidConfiguration;
  if (op.waves == 4) return launch_dyn_shape<TB, 4, KIND>(op, lds, x, y, st);
  if (op.waves == 16) return launch_dyn_shape<TB, 16, KIND>(op, lds, x, y, st);
  return hipErrorInvalidConfiguration;
}

template <typename TB, int ROWS, int KIND>
hipError_t launch_waves(const DenseOp& op, size_t lds, const double* x, double* y,
                        hipStream_t st) {
  if (op.waves == 4) return launch_shape<TB, ROWS, 4, KIND>(op, lds, x, y, st);
  if (op.waves == 16) return launch_shape<TB, ROWS, 16, KIND>(op, lds, x, y, st);
  return hipErrorInvalidConfiguration;
}

template <typename TB, int ROWS>
hipError_t launch_rows(const DenseOp& op, size_t lds, const double* x, double* y,
                       hipStream_t st) {
  return op.kind == 1 ? launch_waves<TB, ROWS, 1>(op, lds, x, y, st)
                      : launch_waves<TB, ROWS, 0>(op, lds, x, y, st);
}

template <typename TB>
hipError_t gemv_typed(const DenseOp& op, const double* x, double* y, hipStream_t st) {
  using X = typename BemVec<TB>::X;
  const size_t lds = (size_t)op.max_ld * sizeof(X);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  switch (op.rows) {
    case 0:
      return op.kind == 1 ? launch_dyn<TB, 1>(op, lds, x, y, st)
                          : launch_dyn<TB, 0>(op, lds, x, y, st);
    case 1: return launch_rows<TB, 1>(op, lds, x, y, st);
    case 2: return launch_rows<TB, 2>(op, lds, x, y, st);
    case 4: return launch_rows<TB, 4>(op, lds, x, y, st);
    default: return hipErrorInvalidConfiguration;
  }
}

template <typename TB>
hipError_t assemble_segment(const DeviceOps& d, int s, TB* out, int ld, hipStream_t st) {
  const DenseOp& b = d.bem;
  const int nb = b.h_n[s], bo = b.h_off[s];
  const int cb = (ld + 255) / 256;
  const long long nblk = (long long)cb * nb;
  if (nblk <= 0) return hipSuccess;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  k_bem_assemble<TB><<<(int)nblk, 256, 0, st>>>(nb, ld, cb, d.coords, d.bnode + bo, d.tri_node,
                                                d.tnorm, d.inc_rp + bo, d.inc_tri, d.omega + bo,
                                                d.h_bem_geo_eps[s], out);
  return hipGetLastError();
}

template <typename TB>
hipError_t assemble_typed(const DeviceOps& d, hipStream_t st) {
  for (size_t s = 0; s < d.bem.h_n.size(); s++) {
    const hipError_t e = assemble_segment<TB>(d, (int)s, (TB*)d.bem.A + d.bem.h_aoff[s],
                                              d.bem.h_ld[s], st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

int dense_gemv_max_order(int type) {
  const size_t x = type == kBemBF16 ? sizeof(float) : sizeof(double);
  return (int)(kMaxLds / x);
}

bool plan_dense_gemv(DenseOp& op, int num_cus, int rows_per_pass, int wg_per_cu, int waves,
                     std::vector<int>* wg_seg, std::vector<int>* wg_row0) {
  if (rows_per_pass != 0 && rows_per_pass != 1 && rows_per_pass != 2 && rows_per_pass != 4)
    return false;
  if (wg_per_cu < 1 || wg_per_cu > 8) return false;
  if (waves != 4 && waves != 16) return false;
  if (rows_per_pass == 0) {
    // dynamic row claims: one segment, wg_per_cu workgroups per CU
    if (op.h_n.size() != 1 || op.h_aoff[0] != 0) return false;
    op.max_ld = op.h_ld[0];
    op.rpw = 0;
    op.rows = 0;
    op.wg_per_cu = wg_per_cu;
    op.waves = waves;
    op.grid = wg_per_cu * (num_cus > 0 ? num_cus : 256);
    wg_seg->assign(1, 0);
    wg_row0->assign(1, 0);
    return true;
  }
  // wg_per_cu workgroups of `waves` wavefronts per CU; every wavefront gets
  // the same number of rows, every workgroup one segment
  long long rows = 0;
  for (int nb : op.h_n) rows += nb;
  const long long nw = (long long)waves * wg_per_cu * (num_cus > 0 ? num_cus : 256);
  const int R = rows_per_pass;
  const int rpw = std::max(1, (int)((rows + nw - 1) / nw));
  const int rows_per_wg = waves * rpw;
  wg_seg->clear();
  wg_row0->clear();
  op.max_ld = 0;
  for (size_t s = 0; s < op.h_n.size(); s++) {
    op.max_ld = std::max(op.max_ld, op.h_ld[s]);
    for (int r = 0; r < op.h_n[s]; r += rows_per_wg) {
      wg_seg->push_back((int)s);
      wg_row0->push_back(r);
    }
  }
  op.rpw = rpw;
  op.rows = R;
  op.wg_per_cu = wg_per_cu;
  op.waves = waves;
  op.grid = (int)wg_seg->size();
  return true;
}

hipError_t launch_bem_assemble(const DeviceOps& d, hipStream_t st) {
  if (d.bem.type == kBemF64) return assemble_typed<double>(d, st);
  if (d.bem.type == kBemF32) return assemble_typed<float>(d, st);
  return assemble_typed<uint16_t>(d, st);
}

hipError_t launch_bem_assemble_f64(const DeviceOps& d, int s, double* out, hipStream_t st) {
  return assemble_segment<double>(d, s, out, d.bem.h_ld[s], st);
}

hipError_t launch_dense_gemv(const DenseOp& op, const double* x, double* y, hipStream_t st) {
  if (!op.A || op.grid == 0) return hipSuccess;
  if (op.type == kBemF64) return gemv_typed<double>(op, x, y, st);
  if (op.type == kBemF32) return gemv_typed<float>(op, x, y, st);
  if (op.type == kBemF24) return gemv_typed<F24>(op, x, y, st);
  return gemv_typed<uint16_t>(op, x, y, st);
}

}  // namespace nsim
