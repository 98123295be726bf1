// Persistent pipelined Jacobi-preconditioned CG for the two Poisson solves of
// the Fredkin-Koehler demag (solve_neg_laplace_phi and solve_laplace_DBC,
// interface/nmag/nmag_lam.py:614-633, 744-754).
//
// One cooperative launch runs the whole solve.  Workgroups exchange data once
// per iteration: each publishes its slice of the preconditioned vector that the
// next SpMV gathers, its two partial inner products, and then a tag; everyone
// polls the tags of all workgroups.  Published data uses agent-coherent stores
// and loads, ordered by an explicit wait before the tag store, so no L2
// write-back, invalidation or contended atomic is on the critical path.
#include <hip/hip_runtime.h>

#include "nsim_dev_common.h"
#include "nsim_device.h"

namespace nsim {

namespace {

using namespace dev;

template <typename TV>
struct CgArgs {
  int n, gmax, maxit;
  const int* off;
  const int* col;
  const TV* val;
  const double* dinv;  // 0 marks a fixed (Dirichlet) row
  const double* b;     // may be null (zero load)
  double* x;           // in: warm start (+ Dirichlet values), out: solution
  double* pub;         // 2n published vector (by phase parity)
  double* pa;          // 2*gmax partial sums (by phase parity)
  double* pb;
  unsigned long long* tag;  // 2*gmax phase tags
  unsigned* sync;           // [2] error flag, [3] last its, [4] total its, [5] solves
  unsigned long long epoch; // unique per launch
  double rtol2;
  const double* xadd;       // optional: xsum = x + xadd
  double* xsum;
  double* rout;             // optional: rout[blocal[i]] = x_i on boundary nodes
  const int* blocal;
};

constexpr unsigned kSpinLimit = 1u << 20;

template <typename TV>
__device__ inline double sell_dot_coh(const int* __restrict__ off, const int* __restrict__ col,
                                      const TV* __restrict__ val, const double* v, int row) {
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  double acc = 0.0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    acc += (double)val[e] * ld_coh(v + col[e]);
  }
  return acc;
}

// Publishes this workgroup's partial sums (s1, s2) of phase `ph` and replaces
// them with the grid totals.  Every thread must have issued its agent-coherent
// stores of the phase's published vector before the call.  Returns false if
// the exchange timed out or another workgroup reported an error; the spin is
// bounded so that every workgroup leaves the solver and the grid drains.
template <typename TV>
__device__ bool exchange(const CgArgs<TV>& a, unsigned ph, double& s1, double& s2,
                         double* sh) {
  wait_all_memory();
  block_sum2(s1, s2, sh);
  const int par = ph & 1, G = gridDim.x;
  const unsigned long long tg = (a.epoch << 24) | ph;
  if (threadIdx.x == 0) {
    st_coh(&a.pa[par * a.gmax + blockIdx.x], s1);
    st_coh(&a.pb[par * a.gmax + blockIdx.x], s2);
    wait_all_memory();
    st_coh(&a.tag[par * a.gmax + blockIdx.x], tg);
  }
  double v1 = 0.0, v2 = 0.0;
  int bad = 0;
  for (int t = threadIdx.x; t < G && !bad; t += 256) {
    unsigned spins = 0;
    while (ld_coh(&a.tag[par * a.gmax + t]) != tg) {
      if ((++spins & 255u) == 0u && (ld_coh(&a.sync[2]) != 0u || spins > kSpinLimit)) {
        st_coh(&a.sync[2], 1u);
        bad = 1;
        break;
      }
    }
    if (!bad) {
      v1 += ld_coh(&a.pa[par * a.gmax + t]);
      v2 += ld_coh(&a.pb[par * a.gmax + t]);
    }
  }
  if (__syncthreads_or(bad)) return false;
  block_sum2(v1, v2, sh);
  s1 = v1;
  s2 = v2;
  return true;
}

// Pipelined Jacobi-PCG (Ghysels & Vanroose 2014): the inner products of
// iteration i and the vector m_i = D^-1 w_i gathered by the next SpMV travel in
// the same exchange, so an iteration costs one grid-wide round trip.  Each
// thread owns RPT rows and keeps their recurrence vectors in registers.  Stops
// when (r,u) <= rtol^2 (b_eff, D^-1 b_eff), b_eff = b - K x_fixed: the KSP rtol
// measured against the load vector.
template <typename TV, int RPT>
__global__ __launch_bounds__(256) void k_cg(CgArgs<TV> a) {
  __shared__ double sh[8];
  const int G = gridDim.x, stride = G * 256, n = a.n;
  double x[RPT], r[RPT], u[RPT], w[RPT], z[RPT], q[RPT], s[RPT], p[RPT], di[RPT], mk[RPT];
  double lbn = 0.0, dummy = 0.0;
  // phase 1: r0 = b - K x0, u0 = D^-1 r0 (published), |b_eff|
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
        double ax = 0.0, af = 0.0;
        for (int t = 0; t < ww; t++) {
          const int e = bb + (t << 6) + lane;
          const int j = a.col[e];
          const double pr = (double)a.val[e] * a.x[j];
          ax += pr;
          if (a.dinv[j] == 0.0) af += pr;
        }
        const double bi = a.b ? a.b[row] : 0.0;
        r[k] = bi - ax;
        const double be = bi - af;
        lbn += be * be * di[k];
      }
      u[k] = di[k] * r[k];
      st_coh(&a.pub[1 * n + row], u[k]);
    }
  }
  int it = 0;
  bool alive = exchange(a, 1u, lbn, dummy, sh);
  if (alive) {
    const double stop = a.rtol2 * lbn;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
      if (row < n && di[k] != 0.0) w[k] = sell_dot_coh<TV>(a.off, a.col, a.val, a.pub + n, row);
    }
    double gam_old = 1.0, alpha_old = 1.0;
    for (;; it++) {
      const unsigned ph = (unsigned)it + 2u;
      const int par = ph & 1;
      double gam = 0.0, del = 0.0;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
        gam += r[k] * u[k];
        del += w[k] * u[k];
        mk[k] = di[k] * w[k];
        if (row < n) st_coh(&a.pub[par * n + row], mk[k]);
      }
      if (!exchange(a, ph, gam, del, sh)) break;
      if (!(gam > stop) || gam <= 1e-300 || it >= a.maxit) break;
      if (!isfinite(gam) || !isfinite(del)) {
        if (threadIdx.x == 0) st_coh(&a.sync[2], 2u);
        break;
      }
      const double beta = it > 0 ? gam / gam_old : 0.0;
      const double alpha = it > 0 ? gam / (del - beta * gam / alpha_old) : gam / del;
      const double* mv = a.pub + par * n;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
        double nk = 0.0;
        if (row < n && di[k] != 0.0) nk = sell_dot_coh<TV>(a.off, a.col, a.val, mv, row);
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
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sync[3] = (unsigned)it;
    a.sync[4] += (unsigned)it;
    a.sync[5] += 1u;
  }
}

template <typename TV, int RPT>
hipError_t coop_launch(CgArgs<TV>& a, int grid, hipStream_t st) {
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_cg<TV, RPT>), dim3(grid),
                                    dim3(256), args, 0, st);
}

template <typename TV>
hipError_t coop_launch_rpt(CgArgs<TV>& a, int grid, int rpt, hipStream_t st) {
  switch (rpt) {
    case 1: return coop_launch<TV, 1>(a, grid, st);
    case 2: return coop_launch<TV, 2>(a, grid, st);
    case 4: return coop_launch<TV, 4>(a, grid, st);
    case 8: return coop_launch<TV, 8>(a, grid, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename TV, int RPT>
int occupancy() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<TV, RPT>, 256, 0) !=
      hipSuccess)
    return 0;
  return per_cu;
}

template <typename TV>
int max_wg(int rpt, int device) {
  int per_cu = 0, cus = 0;
  switch (rpt) {
    case 1: per_cu = occupancy<TV, 1>(); break;
    case 2: per_cu = occupancy<TV, 2>(); break;
    case 4: per_cu = occupancy<TV, 4>(); break;
    case 8: per_cu = occupancy<TV, 8>(); break;
    default: return 0;
  }
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * cus;
}

}  // namespace

int cg_max_workgroups(bool mixed, int rpt, int device) {
  return mixed ? max_wg<float>(rpt, device) : max_wg<double>(rpt, device);
}

hipError_t launch_cg(const DeviceOps& d, bool mixed, CgWork& w, int grid, int rpt,
                     const double* dinv, const double* b, double* x, double rtol, int maxit,
                     const double* xadd, double* xsum, double* restrict_out,
                     hipStream_t st) {
  if (grid <= 0 || grid > w.gmax || (long long)grid * 256 * rpt < d.n || maxit >= (1 << 23))
    return hipErrorInvalidConfiguration;
  const unsigned long long epoch = ++w.epoch;
  if (mixed) {
    CgArgs<float> a{d.n,    w.gmax,      maxit,  d.sell_off, d.sell_col, d.Kf,
                    dinv,   b,           x,      w.pub,      w.pa,       w.pb,
                    w.tag,  w.sync,      epoch,  rtol * rtol, xadd,      xsum,
                    restrict_out, d.blocal};
    return coop_launch_rpt<float>(a, grid, rpt, st);
  }
  CgArgs<double> a{d.n,    w.gmax,      maxit,  d.sell_off, d.sell_col, d.K,
                   dinv,   b,           x,      w.pub,      w.pa,       w.pb,
                   w.tag,  w.sync,      epoch,  rtol * rtol, xadd,      xsum,
                   restrict_out, d.blocal};
  return coop_launch_rpt<double>(a, grid, rpt, st);
}

}  // namespace nsim
