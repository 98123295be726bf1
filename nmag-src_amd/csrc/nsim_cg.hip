// Persistent pipelined Jacobi-preconditioned CG for the two Poisson solves of
// the Fredkin-Koehler demag (solve_neg_laplace_phi and solve_laplace_DBC,
// interface/nmag/nmag_lam.py:614-633, 744-754).
//
// One cooperative launch runs the whole solve.  Workgroups exchange data once
// per iteration: each publishes its slice of the preconditioned vector that the
// next SpMV gathers and its two partial inner products, waits for those stores
// to complete, then publishes a tag; everyone polls the tags of all
// workgroups.  Exchanged data is written with agent-coherent (sc1) stores and
// read with sc1 buffer loads, so no L2 write-back or invalidation and no
// contended atomic is on the critical path, and the gathers of one row are all
// in flight at once.  Each row's operator slots stay in registers for the
// whole solve.
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
  double* part;        // 2 * gmax * 2 partial sums (by phase parity)
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
constexpr int kBufferConfig = 0x00020000;  // raw buffer, gfx9 family
constexpr int kSc1 = 16;                   // cache policy: device-coherent

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           kBufferConfig);
}

// Device-coherent load of element `idx` (out-of-range indices read 0).
__device__ inline double load_coh(__amdgpu_buffer_rsrc_t r, int idx) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, kSc1));
}

// Row slots of one thread: kept in registers (MAXW > 0) or streamed (MAXW == 0).
template <typename TV, int MAXW>
struct RowSlots {
  int col[MAXW > 0 ? MAXW : 1];
  TV val[MAXW > 0 ? MAXW : 1];
  int base, width, lane;

  __device__ void load(const CgArgs<TV>& a, int row) {
    const int s = row >> 6;
    lane = row & 63;
    base = a.off[s];
    width = (a.off[s + 1] - base) >> 6;
    if constexpr (MAXW > 0) {
#pragma unroll
      for (int k = 0; k < MAXW; k++) {
        if (k < width) {
          const int e = base + (k << 6) + lane;
          col[k] = a.col[e];
          val[k] = a.val[e];
        } else {
          col[k] = row;
          val[k] = (TV)0;
        }
      }
    }
  }

  // sum_j K_ij v_j with v read device-coherently from r at offset `shift`.
  __device__ double gather(const CgArgs<TV>& a, __amdgpu_buffer_rsrc_t r, int shift) const {
    double acc = 0.0;
    if constexpr (MAXW > 0) {
      double v[MAXW];
#pragma unroll
      for (int k = 0; k < MAXW; k++) v[k] = load_coh(r, shift + col[k]);
#pragma unroll
      for (int k = 0; k < MAXW; k++) acc += (double)val[k] * v[k];
    } else {
      for (int k0 = 0; k0 < width; k0 += 8) {
        int c[8];
        double w8[8], v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const bool in = k0 + k < width;
          const int e = base + ((k0 + k) << 6) + lane;
          c[k] = in ? a.col[e] : 0;
          w8[k] = in ? (double)a.val[e] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = load_coh(r, shift + c[k]);
#pragma unroll
        for (int k = 0; k < 8; k++) acc += w8[k] * v[k];
      }
    }
    return acc;
  }
};

// Publishes this workgroup's partial sums (s1, s2) of phase `ph` and replaces
// them with the grid totals.  Every thread must have issued its sc1 stores of
// the phase's published vector before the call.  Returns false if the exchange
// timed out or another workgroup reported an error; the spin is bounded so that
// every workgroup leaves the solver and the grid drains.
template <typename TV>
__device__ bool exchange(const CgArgs<TV>& a, __amdgpu_buffer_rsrc_t rpart, unsigned ph,
                         double& s1, double& s2, double* sh) {
  wait_all_memory();
  block_sum2(s1, s2, sh);
  const int par = ph & 1, G = gridDim.x;
  const unsigned long long tg = (a.epoch << 24) | ph;
  if (threadIdx.x == 0) {
    st_coh(&a.part[2 * (par * a.gmax + blockIdx.x)], s1);
    st_coh(&a.part[2 * (par * a.gmax + blockIdx.x) + 1], s2);
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
      const int i = 2 * (par * a.gmax + t);
      const double p1 = load_coh(rpart, i), p2 = load_coh(rpart, i + 1);
      v1 += p1;
      v2 += p2;
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
template <typename TV, int RPT, int MAXW>
__global__ __launch_bounds__(256) void k_cg(CgArgs<TV> a) {
  __shared__ double sh[8];
  const int G = gridDim.x, stride = G * 256, n = a.n;
  const __amdgpu_buffer_rsrc_t rpub = make_rsrc(a.pub, 2u * (unsigned)n * 8u);
  const __amdgpu_buffer_rsrc_t rpart = make_rsrc(a.part, 4u * (unsigned)a.gmax * 8u);
  RowSlots<TV, MAXW> rs[RPT];
  double x[RPT], r[RPT], u[RPT], w[RPT], z[RPT], q[RPT], s[RPT], p[RPT], di[RPT], mk[RPT];
  double lbn = 0.0, dummy = 0.0;
  // phase 1: r0 = b - K x0, u0 = D^-1 r0 (published), |b_eff|
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
    x[k] = r[k] = u[k] = w[k] = z[k] = q[k] = s[k] = p[k] = di[k] = mk[k] = 0.0;
    rs[k].load(a, row < n ? row : n - 1);
    if (row < n) {
      di[k] = a.dinv[row];
      x[k] = a.x[row];
      if (di[k] != 0.0) {
        double ax = 0.0, af = 0.0;
        for (int t = 0; t < rs[k].width; t++) {
          const int e = rs[k].base + (t << 6) + rs[k].lane;
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
  const bool alive = exchange(a, rpart, 1u, lbn, dummy, sh);
  if (alive) {
    const double stop = a.rtol2 * lbn;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
      if (row < n && di[k] != 0.0) w[k] = rs[k].gather(a, rpub, n);
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
      if (!exchange(a, rpart, ph, gam, del, sh)) break;
      if (!(gam > stop) || gam <= 1e-300 || it >= a.maxit) break;
      if (!isfinite(gam) || !isfinite(del)) {
        if (threadIdx.x == 0) st_coh(&a.sync[2], 2u);
        break;
      }
      const double beta = it > 0 ? gam / gam_old : 0.0;
      const double alpha = it > 0 ? gam / (del - beta * gam / alpha_old) : gam / del;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = blockIdx.x * 256 + threadIdx.x + k * stride;
        double nk = 0.0;
        if (row < n && di[k] != 0.0) nk = rs[k].gather(a, rpub, par * n);
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

// The kernel variants: (rows per thread, register slots).  Register slots need
// every slice of the operator to be at most MAXW wide.
template <typename TV, int RPT, int MAXW>
const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_cg<TV, RPT, MAXW>);
}

template <typename TV>
const void* select_kernel(int rpt, int maxw) {
  if (maxw == 16) {
    if (rpt == 1) return kernel_ptr<TV, 1, 16>();
    if (rpt == 2) return kernel_ptr<TV, 2, 16>();
  } else if (maxw == 32) {
    if (rpt == 1) return kernel_ptr<TV, 1, 32>();
    if (rpt == 2) return kernel_ptr<TV, 2, 32>();
  } else if (maxw == 0) {
    if (rpt == 1) return kernel_ptr<TV, 1, 0>();
    if (rpt == 2) return kernel_ptr<TV, 2, 0>();
    if (rpt == 4) return kernel_ptr<TV, 4, 0>();
    if (rpt == 8) return kernel_ptr<TV, 8, 0>();
  }
  return nullptr;
}

}  // namespace

// Register-slot width used for an operator whose widest slice has `width` slots.
int cg_slot_variant(int width, int rpt) {
  if (rpt <= 2 && width <= 16) return 16;
  if (rpt <= 2 && width <= 32) return 32;
  return 0;
}

int cg_max_workgroups(bool mixed, int rpt, int maxw, int device) {
  const void* k = mixed ? select_kernel<float>(rpt, maxw) : select_kernel<double>(rpt, maxw);
  if (!k) return 0;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * cus;
}

hipError_t launch_cg(const DeviceOps& d, bool mixed, CgWork& w, int grid, int rpt, int maxw,
                     const double* dinv, const double* b, double* x, double rtol, int maxit,
                     const double* xadd, double* xsum, double* restrict_out,
                     hipStream_t st) {
  if (grid <= 0 || grid > w.gmax || (long long)grid * 256 * rpt < d.n || maxit >= (1 << 23) ||
      (maxw > 0 && d.max_width > maxw) || 2LL * d.n * 8 >= 0x7fffffffLL)
    return hipErrorInvalidConfiguration;
  const unsigned long long epoch = ++w.epoch;
  const void* k = mixed ? select_kernel<float>(rpt, maxw) : select_kernel<double>(rpt, maxw);
  if (!k) return hipErrorInvalidConfiguration;
  if (mixed) {
    CgArgs<float> a{d.n,   w.gmax, maxit, d.sell_off, d.sell_col, d.Kf,        dinv,
                    b,     x,      w.pub, w.part,     w.tag,      w.sync,      epoch,
                    rtol * rtol, xadd, xsum, restrict_out, d.blocal};
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), args, 0, st);
  }
  CgArgs<double> a{d.n,   w.gmax, maxit, d.sell_off, d.sell_col, d.K,         dinv,
                   b,     x,      w.pub, w.part,     w.tag,      w.sync,      epoch,
                   rtol * rtol, xadd, xsum, restrict_out, d.blocal};
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), args, 0, st);
}

}  // namespace nsim
