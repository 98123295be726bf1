// Persistent GMRES for the Newton systems (I - gamma J) x = r of the BDF
// integrator: the work of CVODE's SPGMR (src/sundials_sp_stubs.c:700-1003
// registers it with the jacobi_ts1 J.v and the preconditioner) in a single
// launch per linear solve.
//
// Right-preconditioned GMRES in the space scaled by the error weights (so the
// Euclidean residual is sqrt(n) times its WRMS norm), classical Gram-Schmidt
// applied twice, Givens rotations on the Hessenberg columns.  Each workgroup
// owns whole SELL slices (64 nodes each) and every vector element of their
// nodes; everything it computes on them is local except the exchange part of
// J.z, which gathers the preconditioned direction z at neighbouring nodes.
// Grid-wide data meet in three exchanges per iteration: the two Gram-Schmidt
// coefficient columns and the new vector's norm, the last one carrying the
// next direction z' = M^-1 (v / ewt) (still unnormalised; J is linear, so the
// gather scales by 1/norm).  An exchange publishes per-workgroup partial sums
// with agent-coherent stores, then a phase tag; every workgroup polls all
// tags and sums the partials in the same fixed order, so all of them hold
// bit-identical totals, take the same Givens rotations and stop at the same
// iteration without further communication.  Workgroup 0 keeps the triangular
// factor, solves the least-squares problem at the end and broadcasts its
// solution; the final correction x and sum (x ewt)^2 (the Newton update's
// WRMS norm) are formed in the same launch.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_gmres.h"

namespace nsim {

namespace {

using namespace dev;

constexpr int kSplit = 4;                 // wavefronts per slice in the J.z gather
constexpr unsigned kSpinLimit = 1u << 20;  // polls before an exchange gives up
constexpr int kMaxGrid = 1024;

template <typename TV>
struct GmresArgs {
  int n, nslices, spw, maxl, gmax, pstride;
  const int* off;
  const int* col;
  const TV* K;
  const double* Vinv;
  const Coeffs* ct;
  const int* inst;
  const double* m;
  const double* Ht;
  const double* ewt;
  const double* Minv;  // null: no preconditioner
  double gamma, tol;
  const double* r;
  double* x;
  double* V;
  size_t ld;
  double* zp;
  double* part;
  unsigned long long* tag;
  double* R;
  double* y;
  double* status;
  unsigned* err;
  unsigned long long epoch;
};

struct GShared {
  double cs[kGmresMaxl], sn[kGmresMaxl], g[kGmresMaxl + 1];
  double col[kGmresMaxl + 1];  // Hessenberg column of this iteration
  double loc[kGmresMaxl + 1];  // this workgroup's partial sums
  double tot[kGmresMaxl + 1];  // grid totals
  double gp[3 * kSplit][64];   // J.z gather partials
  double red[8];
};

// z = M^-1 t on one node (3x3 block), t = v / ewt.
__device__ inline void pc_node(const double* __restrict__ Minv, int i, const double* t,
                               double* z) {
  if (!Minv) {
    z[0] = t[0];
    z[1] = t[1];
    z[2] = t[2];
    return;
  }
  const double* o = Minv + 9 * i;
  z[0] = o[0] * t[0] + o[1] * t[1] + o[2] * t[2];
  z[1] = o[3] * t[0] + o[4] * t[1] + o[5] * t[2];
  z[2] = o[6] * t[0] + o[7] * t[1] + o[8] * t[2];
}

// Publishes sm.loc[0, cnt) for phase `ph` and, unless this workgroup is not
// the root of a root_only exchange, sums every workgroup's values into
// sm.tot[0, cnt): wavefront w takes the entries w, w + 4, ..., lane l the
// workgroups l, l + 64, ..., then a fixed butterfly; the same order in every
// workgroup.  Returns false if the exchange timed out or another workgroup
// reported an error (the spin is bounded, so every workgroup leaves and the
// grid drains).
template <typename TV>
__device__ bool exchange(const GmresArgs<TV>& a, GShared& sm, unsigned ph, int cnt,
                         bool root_only) {
  const int par = ph & 1, G = gridDim.x;
  const unsigned long long tg = (a.epoch << 24) | ph;
  __syncthreads();  // sm.loc complete
  double* mine = a.part + ((size_t)par * a.gmax + blockIdx.x) * a.pstride;
  for (int i = threadIdx.x; i < cnt; i += 256) st_coh(mine + i, sm.loc[i]);
  wait_all_memory();  // this wave's exchanged stores (z', partials) completed
  __syncthreads();
  if (threadIdx.x == 0) st_coh(&a.tag[par * a.gmax + blockIdx.x], tg);
  if (root_only && blockIdx.x != 0) return true;
  int bad = 0;
  for (int t = threadIdx.x; t < G && !bad; t += 256) {
    unsigned spins = 0;
    while (ld_coh(&a.tag[par * a.gmax + t]) != tg) {
      if ((++spins & 255u) == 0u && (ld_coh(a.err) != 0u || spins > kSpinLimit)) {
        st_coh(a.err, 1u);
        bad = 1;
        break;
      }
    }
  }
  if (__syncthreads_or(bad)) return false;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const double* base = a.part + (size_t)par * a.gmax * a.pstride;
  for (int i = w; i < cnt; i += 4) {
    double s = 0.0;
    for (int t = l; t < G; t += 64) s += ld_coh(base + (size_t)t * a.pstride + i);
    s = wave_sum(s);
    if (l == 0) sm.tot[i] = s;
  }
  __syncthreads();
  return true;
}

// Waits for workgroup 0's tag of phase `ph` (a broadcast).
template <typename TV>
__device__ bool wait_root(const GmresArgs<TV>& a, unsigned ph) {
  const unsigned long long tg = (a.epoch << 24) | ph;
  int bad = 0;
  if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (ld_coh(&a.tag[(ph & 1) * a.gmax]) != tg) {
      if ((++spins & 255u) == 0u && (ld_coh(a.err) != 0u || spins > kSpinLimit)) {
        st_coh(a.err, 1u);
        bad = 1;
        break;
      }
    }
  }
  return !__syncthreads_or(bad);
}

// Sum over the workgroup into sm.loc[0] (fixed order).
__device__ inline void block_total(double s, GShared& sm) {
  double dummy = 0.0;
  block_sum2(s, dummy, sm.red);
  if (threadIdx.x == 0) sm.loc[0] = s;
}

// z' = M^-1 (src / ewt) on this workgroup's nodes, published coherently; with
// src_is_t the source is already t = v / ewt.
template <typename TV>
__device__ void publish_direction(const GmresArgs<TV>& a, const double* src, bool src_is_t,
                                  int r0, int r1) {
  for (int i = r0 + threadIdx.x; i < r1; i += 256) {
    double t[3], z[3];
    for (int d = 0; d < 3; d++)
      t[d] = src_is_t ? src[3 * i + d] : src[3 * i + d] / a.ewt[3 * i + d];
    pc_node(a.Minv, i, t, z);
    for (int d = 0; d < 3; d++) st_coh(a.zp + 3 * i + d, z[d]);
  }
}

template <typename TV>
__global__ __launch_bounds__(256) void k_gmres(GmresArgs<TV> a) {
  __shared__ GShared sm;
  const int s0 = blockIdx.x * a.spw, s1 = min(a.nslices, s0 + a.spw);
  const int r0 = 64 * s0, r1 = min(a.n, 64 * s1);
  const int e0 = 3 * r0, e1 = 3 * r1;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const double tol = a.tol;
  unsigned ph = 0;
  bool ok = true;
  int k = 0;

  // v0 = r ewt, beta = |v0|; z'_0 = M^-1 r
  {
    double s = 0.0;
    for (int e = e0 + tid; e < e1; e += 256) {
      const double v = a.r[e] * a.ewt[e];
      a.V[e] = v;
      s += v * v;
    }
    publish_direction(a, a.r, true, r0, r1);
    block_total(s, sm);
    ok = exchange(a, sm, ++ph, 1, false);
  }
  const double beta = ok ? sqrt(sm.tot[0]) : 0.0;
  if (ok && !(beta > tol)) {  // already solved: x = 0
    for (int e = e0 + tid; e < e1; e += 256) a.x[e] = 0.0;
    if (blockIdx.x == 0 && tid == 0) {
      a.status[0] = 0.0;
      a.status[1] = 0.0;
      a.status[2] = 0.0;
    }
    return;
  }
  double scale = ok ? 1.0 / beta : 0.0;
  if (ok) {
    for (int e = e0 + tid; e < e1; e += 256) a.V[e] *= scale;
    if (tid == 0) sm.g[0] = beta;
  }
  for (int j = 0; ok && j < a.maxl; j++) {
    double* Vn = a.V + (size_t)(j + 1) * a.ld;
    // Vn = (z - gamma J z) ewt, z = scale z'
    for (int s = s0; s < s1; s++) {
      const int row = 64 * s + lane;
      const bool live = row < a.n;
      const int b = a.off[s], w = (a.off[s + 1] - b) >> 6;
      double hv[3] = {0.0, 0.0, 0.0};
      if (live)
        for (int q = wv; q < w; q += kSplit) {
          const int e = b + (q << 6) + lane;
          const int c = a.col[e];
          const double kv = (double)a.K[e];
          hv[0] += kv * ld_coh(a.zp + 3 * c);
          hv[1] += kv * ld_coh(a.zp + 3 * c + 1);
          hv[2] += kv * ld_coh(a.zp + 3 * c + 2);
        }
      for (int q = 0; q < 3; q++) sm.gp[wv * 3 + q][lane] = hv[q];
      __syncthreads();
      if (wv == 0 && live) {
        for (int q = 0; q < 3; q++) {
          double acc = sm.gp[q][lane];
          for (int u = 1; u < kSplit; u++) acc += sm.gp[u * 3 + q][lane];
          hv[q] = acc;
        }
        const Coeffs& cf = a.ct[a.inst ? a.inst[row] : 0];
        const double fct = cf.negJ * a.Vinv[row] * scale;
        const double h[3] = {fct * hv[0], fct * hv[1], fct * hv[2]};
        const double mi[3] = {a.m[3 * row], a.m[3 * row + 1], a.m[3 * row + 2]};
        const double Hi[3] = {a.Ht[3 * row], a.Ht[3 * row + 1], a.Ht[3 * row + 2]};
        double zi[3], jv[3];
        for (int d = 0; d < 3; d++) zi[d] = scale * ld_coh(a.zp + 3 * row + d);
        llg_jv_row(cf.c1, cf.c2, h, mi, zi, Hi, jv);
        for (int d = 0; d < 3; d++)
          Vn[3 * row + d] = (zi[d] - a.gamma * jv[d]) * a.ewt[3 * row + d];
      }
      __syncthreads();
    }
    // classical Gram-Schmidt, twice
    for (int pass = 0; pass < 2 && ok; pass++) {
      for (int i = wv; i <= j; i += 4) {
        const double* Vi = a.V + (size_t)i * a.ld;
        double s = 0.0;
        for (int e = e0 + lane; e < e1; e += 64) s += Vi[e] * Vn[e];
        s = wave_sum(s);
        if (lane == 0) sm.loc[i] = s;
      }
      ok = exchange(a, sm, ++ph, j + 1, false);
      if (!ok) break;
      for (int e = e0 + tid; e < e1; e += 256) {
        double v = Vn[e];
        for (int i = 0; i <= j; i++) v -= sm.tot[i] * a.V[(size_t)i * a.ld + e];
        Vn[e] = v;
      }
      for (int i = tid; i <= j; i += 256) sm.col[i] = pass == 0 ? sm.tot[i] : sm.col[i] + sm.tot[i];
      __syncthreads();
    }
    if (!ok) break;
    // norm of the new vector; its direction z' travels in the same exchange
    {
      double s = 0.0;
      for (int e = e0 + tid; e < e1; e += 256) s += Vn[e] * Vn[e];
      publish_direction(a, Vn, false, r0, r1);
      block_total(s, sm);
      ok = exchange(a, sm, ++ph, 1, false);
      if (!ok) break;
    }
    const double hn = sqrt(sm.tot[0]);
    if (hn > 0.0) {
      scale = 1.0 / hn;
      for (int e = e0 + tid; e < e1; e += 256) Vn[e] *= scale;
    }
    if (tid == 0) {
      double* c = sm.col;
      c[j + 1] = hn;
      for (int i = 0; i < j; i++) {  // previous rotations
        const double p = c[i], q = c[i + 1];
        c[i] = sm.cs[i] * p + sm.sn[i] * q;
        c[i + 1] = -sm.sn[i] * p + sm.cs[i] * q;
      }
      const double p = c[j], q = c[j + 1];
      const double rr = hypot(p, q);
      sm.cs[j] = rr > 0.0 ? p / rr : 1.0;
      sm.sn[j] = rr > 0.0 ? q / rr : 0.0;
      c[j] = rr;
      sm.g[j + 1] = -sm.sn[j] * sm.g[j];
      sm.g[j] = sm.cs[j] * sm.g[j];
      if (blockIdx.x == 0)
        for (int i = 0; i <= j; i++) a.R[(size_t)i * a.maxl + j] = c[i];
    }
    __syncthreads();
    k = j + 1;
    if (fabs(sm.g[j + 1]) <= tol || !(hn > 0.0) || !isfinite(hn)) break;
  }
  // y = R^-1 g (workgroup 0), broadcast; x = M^-1 (sum_i y_i V_i / ewt)
  if (ok) {
    ++ph;
    if (blockIdx.x == 0) {
      if (tid == 0) {
        for (int i = k - 1; i >= 0; i--) {
          double s = sm.g[i];
          for (int c = i + 1; c < k; c++) s -= a.R[(size_t)i * a.maxl + c] * sm.tot[c];
          sm.tot[i] = s / a.R[(size_t)i * a.maxl + i];
          st_coh(a.y + i, sm.tot[i]);
        }
        wait_all_memory();
        st_coh(&a.tag[(ph & 1) * a.gmax], (a.epoch << 24) | ph);
      }
      __syncthreads();
    } else {
      ok = wait_root(a, ph);
      if (ok) {
        for (int i = tid; i < k; i += 256) sm.tot[i] = ld_coh(a.y + i);
        __syncthreads();
      }
    }
  }
  if (ok) {
    for (int e = e0 + tid; e < e1; e += 256) {
      double v = 0.0;
      for (int i = 0; i < k; i++) v += sm.tot[i] * a.V[(size_t)i * a.ld + e];
      a.x[e] = v / a.ewt[e];
    }
    __syncthreads();
    double s = 0.0;
    for (int i = r0 + tid; i < r1; i += 256) {
      const double t[3] = {a.x[3 * i], a.x[3 * i + 1], a.x[3 * i + 2]};
      double z[3];
      pc_node(a.Minv, i, t, z);
      for (int d = 0; d < 3; d++) {
        a.x[3 * i + d] = z[d];
        const double u = z[d] * a.ewt[3 * i + d];
        s += u * u;
      }
    }
    block_total(s, sm);
    ok = exchange(a, sm, ++ph, 1, true);
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.status[0] = (double)k;
    a.status[1] = ok ? sm.tot[0] : NAN;
    a.status[2] = ok ? 0.0 : 1.0;
  }
}

template <typename TV>
const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_gmres<TV>);
}

}  // namespace

bool gmres_grid(int nslices, int maxl, int device, int gmax, int* grid, int* spw) {
  if (nslices <= 0 || maxl < 1 || maxl > kGmresMaxl) return false;
  int per_cu = 0, cus = 0;
  // the FP32 and FP64 variants differ only in one load; take the smaller
  int pf = 0, pd = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pf, kernel_ptr<float>(), 256, 0) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&pd, kernel_ptr<double>(), 256, 0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return false;
  per_cu = pf < pd ? pf : pd;
  int cap = per_cu * cus;
  if (cap > kMaxGrid) cap = kMaxGrid;
  if (cap > gmax) cap = gmax;
  if (cap < 1) return false;
  const int s = (nslices + cap - 1) / cap;
  *spw = s;
  *grid = (nslices + s - 1) / s;
  return true;
}

hipError_t launch_gmres(const DeviceOps& d, bool mixed, GmresWork& w, int grid, int spw,
                        int maxl, double gamma, double tol, const double* m, const double* Ht,
                        const double* ewt, const double* Minv, const double* r, double* x,
                        hipStream_t st) {
  if (grid < 1 || grid > w.gmax || maxl < 1 || maxl > kGmresMaxl || w.pstride < maxl + 1 ||
      (long long)grid * spw < d.nslices || (long long)(grid - 1) * spw >= d.nslices)
    return hipErrorInvalidConfiguration;
  const unsigned long long epoch = ++w.epoch;
  const size_t ld = 3 * (size_t)d.n;
  if (mixed) {
    GmresArgs<float> a{d.n, d.nslices, spw, maxl, w.gmax, w.pstride, d.sell_off, d.sell_col,
                       d.Kf, d.Vinv, d.coefs, d.inst, m, Ht, ewt, Minv, gamma, tol, r, x,
                       w.V, ld, w.zp, w.part, w.tag, w.R, w.y, w.status, w.err, epoch};
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(kernel_ptr<float>(), dim3(grid), dim3(256), args, 0, st);
  }
  GmresArgs<double> a{d.n, d.nslices, spw, maxl, w.gmax, w.pstride, d.sell_off, d.sell_col,
                      d.K, d.Vinv, d.coefs, d.inst, m, Ht, ewt, Minv, gamma, tol, r, x,
                      w.V, ld, w.zp, w.part, w.tag, w.R, w.y, w.status, w.err, epoch};
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(kernel_ptr<double>(), dim3(grid), dim3(256), args, 0, st);
}

}  // namespace nsim
