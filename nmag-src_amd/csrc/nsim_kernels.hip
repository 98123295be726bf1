// HIP kernels of the LLG right-hand-side path for CDNA4 (gfx950).
//
// One field evaluation (the reference's update_dmdt program,
// interface/nmag/nmag_lam.py:734-832) is five launches on one stream:
//   1. k_exch_div   H_exch = negJ K m / V and rho = Ms D.m       (SELL-64 SpMV, 3 rhs)
//   2. k_cg         Neumann solve K phi1 = rho                    (persistent pipelined PCG)
//   3. k_bem_gemv   phi2|_b = B phi1|_b                           (LDS-staged dense GEMV)
//   4. k_cg         Dirichlet solve for phi2 in the interior      (same kernel, masked rows)
//   5. k_demag_llg  H_demag = G phi, H_total, LLG dm/dt           (fused gather + pointwise)
// Operators use a SELL format with slice height 64 (one wavefront): the rows
// of a wavefront share one loop bound and every slot load is coalesced.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "nsim_device.h"

namespace nsim {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  using T = double;
};
template <>
struct Acc<float> {
  using T = float;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a 256-thread block; every thread gets the same (deterministic) value.
__device__ inline void block_sum2(double& a, double& b, double* sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[2 * w] = a;
    sh[2 * w + 1] = b;
  }
  __syncthreads();
  a = (sh[0] + sh[2]) + (sh[4] + sh[6]);
  b = (sh[1] + sh[3]) + (sh[5] + sh[7]);
}

__device__ inline void cross(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// ------------------------------------------------------------ BEM assembly

__device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Lindholm weight of vertex slot `a` of triangle (p0,p1,p2) for observer o
// (src/bem3d_stubs.c:413-610, identity lattice transform, min_dist 0).
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
  const double q = fmin(1.0, fmax(-1.0, num / den));
  double angle = 2.0 * acos(q);
  const double zeta = dot3(zv, r0);
  if (zeta < 0.0) angle = -angle;
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

__device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <typename TB>
__device__ inline void store_b(TB* p, double v);
template <>
__device__ inline void store_b<double>(double* p, double v) {
  *p = v;
}
template <>
__device__ inline void store_b<float>(float* p, double v) {
  *p = (float)v;
}
template <>
__device__ inline void store_b<uint16_t>(uint16_t* p, double v) {
  *p = f32_to_bf16((float)v);
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

// -------------------------------------------------------------- SELL SpMV

template <typename TV>
__device__ inline typename Acc<TV>::T sell_dot(const int* __restrict__ off,
                                               const int* __restrict__ col,
                                               const TV* __restrict__ val,
                                               const double* __restrict__ v, int row) {
  using A = typename Acc<TV>::T;
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  A acc = 0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    acc += (A)val[e] * (A)v[col[e]];
  }
  return acc;
}

// H_exch = negJ/V_i sum_j K_ij m_j (set_H_exch, nmag_lam.py:734-739) and
// rho_i = Ms sum_j D_ij . m_j (op_div_m + SCALE -1, nmag_lam.py:742-743).
template <typename TV>
__global__ __launch_bounds__(256) void k_exch_div(int n, const int* __restrict__ off,
                                                  const int* __restrict__ col,
                                                  const TV* __restrict__ K,
                                                  const TV* __restrict__ D,
                                                  const double* __restrict__ Vinv,
                                                  double negJ, double Ms,
                                                  const double* __restrict__ m,
                                                  double* __restrict__ Hexch,
                                                  double* __restrict__ rho, int want_rho) {
  using A = typename Acc<TV>::T;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  A ex = 0, ey = 0, ez = 0, r = 0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    const int j = col[e];
    const A kv = (A)K[e];
    const A mx = (A)m[3 * j], my = (A)m[3 * j + 1], mz = (A)m[3 * j + 2];
    ex += kv * mx;
    ey += kv * my;
    ez += kv * mz;
    if (want_rho) r += (A)D[3 * e] * mx + (A)D[3 * e + 1] * my + (A)D[3 * e + 2] * mz;
  }
  const double f = negJ * Vinv[row];
  Hexch[3 * row] = f * (double)ex;
  Hexch[3 * row + 1] = f * (double)ey;
  Hexch[3 * row + 2] = f * (double)ez;
  if (want_rho) rho[row] = Ms * (double)r;
}

// --------------------------------------------------------- persistent PCG

template <typename TV>
struct CgArgs {
  int n, gmax, maxit;
  const int* off;
  const int* col;
  const TV* val;
  const double* dinv;  // 0 marks a fixed (Dirichlet) row
  const double* b;     // may be null (zero load)
  double* x;           // in: warm start (+ Dirichlet values), out: solution
  double* pub;
  double* part;
  unsigned* sync;
  double rtol2;
  const double* xadd;  // optional: xsum = x + xadd
  double* xsum;
  double* rout;        // optional: rout[blocal[i]] = x_i on boundary nodes
  const int* blocal;
};

constexpr unsigned kSpinLimit = 1u << 24;

// Grid-wide barrier for a cooperative launch.  The spin is bounded: on
// timeout (or when another workgroup has already timed out) the error flag is
// set and every workgroup leaves the solver, so the grid always drains.
__device__ bool grid_barrier(unsigned* sync, unsigned target) {
  __shared__ int ok;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&sync[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    int good = 1;
    while (__hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&sync[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++spins > kSpinLimit) {
        __hip_atomic_store(&sync[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// Sum of the per-workgroup partials of slot q written before barrier `par`.
__device__ inline void reduce_partials(const double* part, int gmax, int par, int G,
                                       double& a, double& b, int qa, int qb, double* sh) {
  double sa = 0.0, sb = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) {
    sa += part[(par * 4 + qa) * gmax + g];
    sb += part[(par * 4 + qb) * gmax + g];
  }
  block_sum2(sa, sb, sh);
  a = sa;
  b = sb;
}

// Pipelined Jacobi-preconditioned CG (Ghysels & Vanroose 2014) with one grid
// barrier per iteration: the two inner products of iteration i and the
// preconditioned vector m_i that the next SpMV gathers are published together.
// Each thread owns RPT rows and keeps their 8 recurrence vectors in registers;
// only m_i travels through memory.  Stops when (r,u) <= rtol^2 (b_eff,D^-1 b_eff)
// with b_eff = b - K x_fixed (the reference's KSP rtol against the load vector).
template <typename TV, int RPT>
__global__ __launch_bounds__(256) void k_cg(CgArgs<TV> a) {
  using A = typename Acc<TV>::T;
  __shared__ double sh[8];
  const int G = gridDim.x, stride = G * 256, n = a.n;
  double x[RPT], r[RPT], u[RPT], w[RPT], z[RPT], q[RPT], s[RPT], p[RPT], di[RPT], mk[RPT];
  double lbn = 0.0;
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
    x[k] = r[k] = u[k] = w[k] = z[k] = q[k] = s[k] = p[k] = di[k] = mk[k] = 0.0;
    if (row < n) {
      di[k] = a.dinv[row];
      x[k] = a.x[row];
      if (di[k] != 0.0) {
        const int sl = row >> 6, lane = row & 63;
        const int bb = a.off[sl], ww = (a.off[sl + 1] - bb) >> 6;
        A ax = 0, af = 0;
        for (int t = 0; t < ww; t++) {
          const int e = bb + (t << 6) + lane;
          const int j = a.col[e];
          const A pr = (A)a.val[e] * (A)a.x[j];
          ax += pr;
          if (a.dinv[j] == 0.0) af += pr;
        }
        const double bi = a.b ? a.b[row] : 0.0;
        r[k] = bi - (double)ax;
        const double be = bi - (double)af;
        lbn += be * be * di[k];
      }
      u[k] = di[k] * r[k];
      a.pub[row] = u[k];
    }
  }
  double dummy = 0.0;
  block_sum2(lbn, dummy, sh);
  if (threadIdx.x == 0) a.part[(0 * 4 + 2) * a.gmax + blockIdx.x] = lbn;
  unsigned nbar = 1;
  bool alive = grid_barrier(a.sync, nbar * G);
  double bn = 0.0;
  int it = 0;
  if (alive) {
    reduce_partials(a.part, a.gmax, 0, G, bn, dummy, 2, 2, sh);
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
      if (row < n && di[k] != 0.0) w[k] = (double)sell_dot<TV>(a.off, a.col, a.val, a.pub, row);
    }
    double gam_old = 1.0, alpha_old = 1.0;
    const double stop = a.rtol2 * bn;
    for (;; it++) {
      const int par = (it + 1) & 1;
      double lg = 0.0, ld = 0.0;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
        lg += r[k] * u[k];
        ld += w[k] * u[k];
        mk[k] = di[k] * w[k];
        if (row < n) a.pub[par * n + row] = mk[k];
      }
      block_sum2(lg, ld, sh);
      if (threadIdx.x == 0) {
        a.part[(par * 4 + 0) * a.gmax + blockIdx.x] = lg;
        a.part[(par * 4 + 1) * a.gmax + blockIdx.x] = ld;
      }
      nbar++;
      if (!grid_barrier(a.sync, nbar * G)) break;
      double gam, del;
      reduce_partials(a.part, a.gmax, par, G, gam, del, 0, 1, sh);
      if (!(gam > stop) || gam <= 1e-300 || it >= a.maxit) break;
      if (!isfinite(gam) || !isfinite(del)) {
        if (threadIdx.x == 0 && blockIdx.x == 0)
          __hip_atomic_store(&a.sync[2], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      const double beta = it > 0 ? gam / gam_old : 0.0;
      const double alpha = it > 0 ? gam / (del - beta * gam / alpha_old) : gam / del;
      const double* mv = a.pub + par * n;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
        double nk = 0.0;
        if (row < n && di[k] != 0.0) nk = (double)sell_dot<TV>(a.off, a.col, a.val, mv, row);
        z[k] = nk + beta * z[k];
        q[k] = mk[k] + beta * q[k];
        s[k] = w[k] + beta * s[k];
        p[k] = u[k] + beta * p[k];
        x[k] += alpha * p[k];
        r[k] -= alpha * s[k];
        u[k] -= alpha * q[k];
        w[k] -= alpha * z[k];
      }
      gam_old = gam;
      alpha_old = alpha;
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
    if (row < n) {
      a.x[row] = x[k];
      if (a.xsum) a.xsum[row] = x[k] + a.xadd[row];
      if (a.rout) {
        const int lb = a.blocal[row];
        if (lb >= 0) a.rout[lb] = x[k];
      }
    }
  }
  // Self-cleaning: the last workgroup out resets the barrier counters.
  __syncthreads();
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0) {
      a.sync[3] = (unsigned)it;
      a.sync[4] += (unsigned)it;
      a.sync[5] += 1u;
    }
    const unsigned prev =
        __hip_atomic_fetch_add(&a.sync[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == (unsigned)G - 1) {
      __hip_atomic_store(&a.sync[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&a.sync[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// --------------------------------------------------------------- BEM GEMV

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
  using X = float;
  using A = float;
};
template <>
struct BemVec<uint16_t> {
  static constexpr int V = 8;
  using X = float;
  using A = float;
};

template <typename TB>
__device__ inline void load_row_chunk(const TB* p, typename BemVec<TB>::X* out);
template <>
__device__ inline void load_row_chunk<double>(const double* p, double* out) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  out[0] = v.x;
  out[1] = v.y;
}
template <>
__device__ inline void load_row_chunk<float>(const float* p, float* out) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}
template <>
__device__ inline void load_row_chunk<uint16_t>(const uint16_t* p, float* out) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    out[2 * k] = __uint_as_float(w[k] << 16);
    out[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// phi2[bnode[i]] = sum_j B_ij phi1b_j  (DM*V BEM, nmag_lam.py:749).  The input
// vector is staged once per workgroup in LDS; each wavefront streams ROWS rows
// of B at a time with 16-byte loads per lane and reduces with cross-lane adds.
// Rows are split evenly over all wavefronts of the grid.
template <typename TB, int ROWS>
__global__ __launch_bounds__(256) void k_bem_gemv(const TB* __restrict__ B, int ldb, int nb,
                                                  int rows_per_wave,
                                                  const double* __restrict__ xb,
                                                  const int* __restrict__ bnode,
                                                  double* __restrict__ y) {
  using X = typename BemVec<TB>::X;
  using A = typename BemVec<TB>::A;
  constexpr int V = BemVec<TB>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  X* xs = reinterpret_cast<X*>(smem);
  for (int j = threadIdx.x; j < ldb; j += 256) xs[j] = j < nb ? (X)xb[j] : (X)0;
  __syncthreads();
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r0 = wave * rows_per_wave;
  const int r1 = min(r0 + rows_per_wave, nb);
  for (int rb = r0; rb < r1; rb += ROWS) {
    const TB* rp[ROWS];
    A acc[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      rp[k] = B + (size_t)min(rb + k, nb - 1) * ldb;
      acc[k] = 0;
    }
#pragma unroll 4
    for (int c = lane * V; c < ldb; c += 64 * V) {
      X xv[V];
#pragma unroll
      for (int e = 0; e < V; e++) xv[e] = xs[c + e];
#pragma unroll
      for (int k = 0; k < ROWS; k++) {
        X bv[V];
        load_row_chunk<TB>(rp[k] + c, bv);
#pragma unroll
        for (int e = 0; e < V; e++) acc[k] += (A)bv[e] * (A)xv[e];
      }
    }
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const double t = wave_sum((double)acc[k]);
      if (lane == 0 && rb + k < r1) y[bnode[rb + k]] = t;
    }
  }
}

// ------------------------------------------- fused H_demag + H_total + LLG

// H_demag_i = sum_j G_ij phi_j (op_grad_phi + CFBOX, nmag_lam.py:755-759);
// H_anis from the uniaxial tangential gradient (src/nsim_anisotropy.ml:167-234);
// H_total (nmag_lam.py:377-411); dm/dt (nmag_lam.py:417-430) times pin.
template <typename TV>
__global__ __launch_bounds__(256) void k_demag_llg(int n, const int* __restrict__ off,
                                                   const int* __restrict__ col,
                                                   const TV* __restrict__ G, Coeffs c,
                                                   FieldArgs f, FieldBufs bf,
                                                   double* __restrict__ dmdt, int do_demag) {
  using A = typename Acc<TV>::T;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  double hd[3] = {0.0, 0.0, 0.0};
  if (do_demag) {
    const int s = row >> 6, lane = row & 63;
    const int b = off[s], w = (off[s + 1] - b) >> 6;
    A gx = 0, gy = 0, gz = 0;
    for (int k = 0; k < w; k++) {
      const int e = b + (k << 6) + lane;
      const A ph = (A)bf.phi[col[e]];
      gx += (A)G[3 * e] * ph;
      gy += (A)G[3 * e + 1] * ph;
      gz += (A)G[3 * e + 2] * ph;
    }
    hd[0] = gx;
    hd[1] = gy;
    hd[2] = gz;
    bf.Hdemag[3 * row] = hd[0];
    bf.Hdemag[3 * row + 1] = hd[1];
    bf.Hdemag[3 * row + 2] = hd[2];
  }
  const double m[3] = {f.m[3 * row], f.m[3 * row + 1], f.m[3 * row + 2]};
  double ha[3] = {0.0, 0.0, 0.0};
  if (f.k1) {
    const double ax[3] = {f.axis[3 * row], f.axis[3 * row + 1], f.axis[3 * row + 2]};
    const double uu = dot3(ax, m);
    const double k2 = f.k2 ? f.k2[row] : 0.0;
    const double sc = (2.0 * f.k1[row] * uu + 4.0 * k2 * uu * uu * uu) / (c.mu0 * c.Ms);
    for (int d = 0; d < 3; d++) ha[d] = sc * (ax[d] - uu * m[d]);
    bf.Hanis[3 * row] = ha[0];
    bf.Hanis[3 * row + 1] = ha[1];
    bf.Hanis[3 * row + 2] = ha[2];
  }
  double H[3];
  for (int d = 0; d < 3; d++)
    H[d] = (f.Hext ? f.Hext[3 * row + d] : 0.0) + bf.Hexch[3 * row + d] + ha[d] + hd[d];
  bf.Htotal[3 * row] = H[0];
  bf.Htotal[3 * row + 1] = H[1];
  bf.Htotal[3 * row + 2] = H[2];
  double mxH[3], mxmxH[3];
  cross(m, H, mxH);
  cross(m, mxH, mxmxH);
  const double mm = dot3(m, m);
  const double pin = f.pin ? f.pin[row] : 1.0;
  for (int d = 0; d < 3; d++)
    dmdt[3 * row + d] = pin * (c.c1 * mxH[d] + c.c2 * mxmxH[d] + c.c3 * (1.0 - mm) * m[d]);
}

// Jv = c1 (v x H + m x h) + c2 (v x (m x H) + m x (v x H) + m x (m x h)),
// h = negJ K v / V: the reference's jacobi_ts1 linearisation (nmag_lam.py:432-436,
// src/nsim.ml:1004-1042) with H_total frozen except for its exchange part.
template <typename TV>
__global__ __launch_bounds__(256) void k_jtimes(int n, const int* __restrict__ off,
                                                const int* __restrict__ col,
                                                const TV* __restrict__ K,
                                                const double* __restrict__ Vinv, Coeffs c,
                                                const double* __restrict__ m,
                                                const double* __restrict__ Ht,
                                                const double* __restrict__ v,
                                                double* __restrict__ Jv) {
  using A = typename Acc<TV>::T;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  A hx = 0, hy = 0, hz = 0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    const int j = col[e];
    const A kv = (A)K[e];
    hx += kv * (A)v[3 * j];
    hy += kv * (A)v[3 * j + 1];
    hz += kv * (A)v[3 * j + 2];
  }
  const double fct = c.negJ * Vinv[row];
  const double h[3] = {fct * hx, fct * hy, fct * hz};
  const double mi[3] = {m[3 * row], m[3 * row + 1], m[3 * row + 2]};
  const double vi[3] = {v[3 * row], v[3 * row + 1], v[3 * row + 2]};
  const double H[3] = {Ht[3 * row], Ht[3 * row + 1], Ht[3 * row + 2]};
  double vxH[3], mxh[3], mxH[3], vxmxH[3], mxvxH[3], mxmxh[3];
  cross(vi, H, vxH);
  cross(mi, h, mxh);
  cross(mi, H, mxH);
  cross(vi, mxH, vxmxH);
  cross(mi, vxH, mxvxH);
  cross(mi, mxh, mxmxh);
  for (int d = 0; d < 3; d++)
    Jv[3 * row + d] = c.c1 * (vxH[d] + mxh[d]) + c.c2 * (vxmxH[d] + mxvxH[d] + mxmxh[d]);
}

inline int blocks_for(int n) { return (n + 255) / 256; }

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_bem_assemble(const DeviceOps& d, hipStream_t st) {
  const int cb = (d.ldb + 255) / 256;
  const long long nblk = (long long)cb * d.nb;
  if (nblk <= 0) return hipSuccess;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  if (d.btype == kBemF64)
    k_bem_assemble<double><<<(int)nblk, 256, 0, st>>>(
        d.nb, d.ldb, cb, d.coords, d.bnode, d.tri_node, d.tnorm, d.inc_rp, d.inc_tri,
        d.omega, d.geo_eps, (double*)d.B);
  else if (d.btype == kBemF32)
    k_bem_assemble<float><<<(int)nblk, 256, 0, st>>>(
        d.nb, d.ldb, cb, d.coords, d.bnode, d.tri_node, d.tnorm, d.inc_rp, d.inc_tri,
        d.omega, d.geo_eps, (float*)d.B);
  else
    k_bem_assemble<uint16_t><<<(int)nblk, 256, 0, st>>>(
        d.nb, d.ldb, cb, d.coords, d.bnode, d.tri_node, d.tnorm, d.inc_rp, d.inc_tri,
        d.omega, d.geo_eps, (uint16_t*)d.B);
  return hipGetLastError();
}

hipError_t launch_exch_div(const DeviceOps& d, bool mixed, const Coeffs& c, const double* m,
                           double* Hexch, double* rho, bool want_rho, hipStream_t st) {
  if (mixed)
    k_exch_div<float><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Kf,
                                                         d.Df, d.Vinv, c.negJ, c.Ms, m, Hexch,
                                                         rho, want_rho ? 1 : 0);
  else
    k_exch_div<double><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.K,
                                                          d.D, d.Vinv, c.negJ, c.Ms, m, Hexch,
                                                          rho, want_rho ? 1 : 0);
  return hipGetLastError();
}

namespace {
template <typename TV, int RPT>
hipError_t coop_launch_cg(CgArgs<TV>& a, int grid, hipStream_t st) {
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_cg<TV, RPT>),
                                    dim3(grid), dim3(256), args, 0, st);
}
template <typename TV>
hipError_t coop_launch_cg_rpt(CgArgs<TV>& a, int grid, int rpt, hipStream_t st) {
  switch (rpt) {
    case 1: return coop_launch_cg<TV, 1>(a, grid, st);
    case 2: return coop_launch_cg<TV, 2>(a, grid, st);
    case 4: return coop_launch_cg<TV, 4>(a, grid, st);
    case 8: return coop_launch_cg<TV, 8>(a, grid, st);
    default: return hipErrorInvalidValue;
  }
}
template <typename TV>
int max_wg(int rpt, int device) {
  int per_cu = 0, cus = 0;
  hipError_t e = hipSuccess;
  switch (rpt) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<TV, 1>, 256, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<TV, 2>, 256, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<TV, 4>, 256, 0); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<TV, 8>, 256, 0); break;
    default: return 0;
  }
  if (e != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * cus;
}
}  // namespace

int cg_max_workgroups(bool mixed, int rpt, int device) {
  return mixed ? max_wg<float>(rpt, device) : max_wg<double>(rpt, device);
}

hipError_t launch_cg(const DeviceOps& d, bool mixed, const CgWork& w, int grid, int rpt,
                     const double* dinv, const double* b, double* x, double rtol, int maxit,
                     const double* xadd, double* xsum, double* restrict_out,
                     hipStream_t st) {
  if (grid <= 0 || grid > w.gmax || (long long)grid * 256 * rpt < d.n)
    return hipErrorInvalidConfiguration;
  if (mixed) {
    CgArgs<float> a{d.n, w.gmax, maxit, d.sell_off, d.sell_col, d.Kf, dinv, b, x, w.pub,
                    w.partial, w.sync, rtol * rtol, xadd, xsum, restrict_out, d.blocal};
    return coop_launch_cg_rpt<float>(a, grid, rpt, st);
  }
  CgArgs<double> a{d.n, w.gmax, maxit, d.sell_off, d.sell_col, d.K, dinv, b, x, w.pub,
                   w.partial, w.sync, rtol * rtol, xadd, xsum, restrict_out, d.blocal};
  return coop_launch_cg_rpt<double>(a, grid, rpt, st);
}

namespace {
template <typename TB>
hipError_t gemv_typed(const DeviceOps& d, const double* phi1b, double* phi2, hipStream_t st,
                      int num_cus) {
  using X = typename BemVec<TB>::X;
  const size_t lds = (size_t)d.ldb * sizeof(X);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  constexpr int ROWS = 2;
  // two workgroups per CU (LDS-limited at the largest sizes), rows split evenly
  int waves = 4 * 2 * (num_cus > 0 ? num_cus : 256);
  int rpw = (d.nb + waves - 1) / waves;
  rpw = ((rpw + ROWS - 1) / ROWS) * ROWS;
  const int used_waves = (d.nb + rpw - 1) / rpw;
  const int grid = (used_waves + 3) / 4;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bem_gemv<TB, ROWS>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  k_bem_gemv<TB, ROWS><<<grid, 256, lds, st>>>((const TB*)d.B, d.ldb, d.nb, rpw, phi1b,
                                               d.bnode, phi2);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_bem_gemv(const DeviceOps& d, const double* phi1b, double* phi2,
                           hipStream_t st, int num_cus) {
  if (d.nb == 0) return hipSuccess;
  if (d.btype == kBemF64) return gemv_typed<double>(d, phi1b, phi2, st, num_cus);
  if (d.btype == kBemF32) return gemv_typed<float>(d, phi1b, phi2, st, num_cus);
  return gemv_typed<uint16_t>(d, phi1b, phi2, st, num_cus);
}

hipError_t launch_demag_llg(const DeviceOps& d, bool mixed, const Coeffs& c,
                            const FieldArgs& f, const FieldBufs& b, double* dmdt,
                            bool do_demag, hipStream_t st) {
  if (mixed)
    k_demag_llg<float><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Gf, c,
                                                          f, b, dmdt, do_demag ? 1 : 0);
  else
    k_demag_llg<double><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.G, c,
                                                           f, b, dmdt, do_demag ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_jtimes(const DeviceOps& d, bool mixed, const Coeffs& c, const double* m,
                         const double* Htotal, const double* v, double* Jv, hipStream_t st) {
  if (mixed)
    k_jtimes<float><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Kf,
                                                       d.Vinv, c, m, Htotal, v, Jv);
  else
    k_jtimes<double><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.K,
                                                        d.Vinv, c, m, Htotal, v, Jv);
  return hipGetLastError();
}

}  // namespace nsim
