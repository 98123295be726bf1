// Persistent pipelined two-level preconditioned CG for the two Poisson solves
// of the Fredkin-Koehler demag (solve_neg_laplace_phi and solve_laplace_DBC,
// interface/nmag/nmag_lam.py:614-633, 744-754).
//
// Rows are reordered so that each workgroup owns one spatially compact part of
// the mesh (recursive coordinate bisection).  The coarse aggregates are two
// or four sub-parts of every part (the part bisected again; at most 320, each
// workgroup evaluating only the rows of the coarse inverse its rows touch),
// or, where that does not fit, runs of consecutive parts (at most 64).  The
// preconditioner is additive two-level:
//     M = D^-1 + P A_c^-1 P^T,   A_c = P^T K P,
// with P the piecewise-constant prolongation from the aggregates.  P^T w is a
// per-workgroup partial sum, so it travels with the inner products: one
// launch runs the whole solve and every iteration costs a single grid-wide
// exchange.  Each workgroup publishes its slice of D^-1 w (which the next SpMV
// gathers) and its partial sums, waits for those stores to complete, then
// publishes a tag; everyone polls the tags of all workgroups.  Exchanged data
// is written with agent-coherent (sc1) stores and read with sc1 buffer loads,
// so no L2 write-back or invalidation and no contended atomic is on the
// critical path, and all gathers of a row are in flight at once.
//
// The operator is stored in FP64 at every precision setting.  Rounding the
// singular Neumann stiffness to FP32 moves its null eigenvalue off zero (on a
// 6x4x2 Kuhn mesh to -1.2e-9 of the largest: the matrix becomes indefinite)
// and the load picks up a component along it, so tight solves stall; the
// solve is bound by exchange latency, not operator bytes, so FP32 storage
// would buy no time either.
#include <hip/hip_runtime.h>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_internal.h"

namespace nsim {

namespace {

using namespace dev;

constexpr int kMaxAggs = 64;
constexpr int kMaxParts = 1024;
constexpr int kSubs = 4;  // aggregates per part at most (fine coarse spaces)
static_assert(2 + 2 * kSubs <= kCgRecord, "workgroup record");
// restriction slots per workgroup: fixed per variant so that they stay in
// registers (indexed by compile-time constants only)
template <bool FINE>
constexpr int kSlots = FINE ? kSubs : 1;

template <typename TV>
struct CgArgs {
  int n;           // padded rows (parts * cap)
  int gmax, maxit;
  int cap_shift;   // log2(rows per part)
  int agg_shift;   // log2(parts per aggregate)
  int sub_shift;   // log2(aggregates per part); then agg_shift is 0
  int naggs;       // 0: Jacobi only
  const int* off;
  const int* col;
  const TV* val;
  const int* p2o;      // row -> original node, -1 on padding rows
  const double* dinv;  // 0 on padding and fixed (Dirichlet) rows
  const double* ainv;  // naggs x naggs coarse inverse
  const float* ainvf;  // the same in FP32 (fine coarse spaces)
  const int* need;     // fine coarse spaces: per part, kCgNeed aggregate slots
  const int* need_n;   // and how many of them are used
  int need_lg;         // log2 threads per row of the coarse inverse
  const double* b;     // original order, may be null (zero load)
  double* x;           // original order; in: warm start (+ Dirichlet values), out: solution
  double* pub;         // 2n published vector (by phase parity)
  double* part;        // 2 * gmax records of kCgRecord partial sums (by phase parity)
  unsigned long long* tag;  // 2*gmax phase tags
  unsigned* sync;           // [2] error flag, [3] last its, [4] total its, [5] solves,
                            // [6] solves stopped at maxit
  unsigned long long epoch; // unique per launch
  double rtol2;
  const double* xadd;       // optional: xsum = x + xadd (original order)
  double* xsum;
  double* rout;             // optional: rout[blocal[i]] = x_i on boundary nodes
  const int* blocal;
  const int* halt;          // nonzero word: every workgroup returns at once
};

// LDS of a workgroup: the FP64 coarse inverse of at most 64 aggregates (three
// workgroups per CU), or, with a fine coarse space, the rows of the FP32
// inverse of at most kCgFineAggs sub-part aggregates that the workgroup's
// rows touch (at most kCgNeed of them; one workgroup per CU).
template <bool FINE>
struct Shared;
template <>
struct Shared<false> {
  double red[8];
  double wred[4][kCgRecord];
  double tot[kCgRecord];
  double ainv[kMaxAggs * kMaxAggs];
  double cpart[kMaxParts];
  double cin[kMaxAggs];
  double zc[kMaxAggs];
};
template <>
struct Shared<true> {
  double red[8];
  double wred[4][kCgRecord];
  double tot[kCgRecord];
  float ainv[kCgNeed * kCgFineAggs];  // row r: the inverse's row need[r]
  double cin[kCgFineAggs];
  double zc[kCgFineAggs];             // valid on need[0, nneed) only
  int need[kCgNeed];
  int nneed;
};
// three workgroups per CU (one with a fine coarse space), with 1 KiB for the
// compiler's own LDS use
static_assert(3 * lds_granules((int)sizeof(Shared<false>) + 1024) * kLdsGranule <= kLdsPerCu,
              "CG workgroup LDS too large for three workgroups per CU");
static_assert(lds_granules((int)sizeof(Shared<true>) + 1024) * kLdsGranule <= kLdsPerCu,
              "CG workgroup LDS (fine coarse space) too large for a CU");

constexpr unsigned kSpinLimit = 1u << 20;

// Row slots of one thread: kept in registers (MAXW > 0) or streamed (MAXW == 0).
// `free_mask` bit k marks a slot whose column is an unknown (receives the
// coarse correction).
template <typename TV, int MAXW>
struct RowSlots {
  int col[MAXW > 0 ? MAXW : 1];
  TV val[MAXW > 0 ? MAXW : 1];
  unsigned long long free_mask;
  int base, width, lane;

  __device__ void load(const CgArgs<TV>& a, int row) {
    const int s = row >> 6;
    lane = row & 63;
    base = a.off[s];
    width = (a.off[s + 1] - base) >> 6;
    free_mask = 0;
    if constexpr (MAXW > 0) {
      // every slot's loads issued at once (slots past the width re-read the
      // last one, then become the row itself with value 0); a slice of
      // padding rows only has no slots
      if (width > 0) {
        double dv[MAXW];
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
          const int e = base + (min(k, width - 1) << 6) + lane;
          col[k] = a.col[e];
          val[k] = a.val[e];
        }
#pragma unroll
        for (int k = 0; k < MAXW; k++) dv[k] = a.dinv[col[k]];
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
          if (k >= width) {
            col[k] = row;
            val[k] = (TV)0;
          } else if (dv[k] != 0.0) {
            free_mask |= 1ull << k;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
          col[k] = row;
          val[k] = (TV)0;
        }
      }
    }
  }

  // sum_j K_ij x_j over all columns (ax) and over the fixed ones (af), x in
  // the original order (the warm start)
  __device__ void residual_terms(const CgArgs<TV>& a, double& ax, double& af) const {
    ax = af = 0.0;
    if constexpr (MAXW > 0) {
      int o[MAXW];
      double xv[MAXW];
#pragma unroll
      for (int k = 0; k < MAXW; k++) o[k] = a.p2o[col[k]];
#pragma unroll
      for (int k = 0; k < MAXW; k++) xv[k] = a.x[max(o[k], 0)];
#pragma unroll
      for (int k = 0; k < MAXW; k++) {
        if (k < width) {
          const double pr = (double)val[k] * xv[k];
          ax += pr;
          if (!((free_mask >> k) & 1ull)) af += pr;
        }
      }
    } else {
      for (int k0 = 0; k0 < width; k0 += 8) {
        int c[8], o[8];
        double w8[8], xv[8], dv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int e = base + (min(k0 + k, width - 1) << 6) + lane;
          c[k] = a.col[e];
          w8[k] = (double)a.val[e];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
          o[k] = a.p2o[c[k]];
          dv[k] = a.dinv[c[k]];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) xv[k] = a.x[max(o[k], 0)];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          if (k0 + k < width) {
            const double pr = w8[k] * xv[k];
            ax += pr;
            if (dv[k] == 0.0) af += pr;
          }
        }
      }
    }
  }

  // sum_j K_ij (v_j + [j free] zc[agg(j)]), v read device-coherently at `shift`
  __device__ double gather(const CgArgs<TV>& a, __amdgpu_buffer_rsrc_t r, int shift,
                           const double* zc, bool coarse) const {
    const int sh = a.cap_shift + a.agg_shift - a.sub_shift;  // row -> aggregate
    double acc = 0.0;
    if constexpr (MAXW > 0) {
      double v[MAXW];
#pragma unroll
      for (int k = 0; k < MAXW; k++) v[k] = load_coh(r, shift + col[k]);
#pragma unroll
      for (int k = 0; k < MAXW; k++) {
        double t = v[k];
        if (coarse && ((free_mask >> k) & 1ull)) t += zc[col[k] >> sh];
        acc += (double)val[k] * t;
      }
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
        for (int k = 0; k < 8; k++) {
          double t = v[k];
          if (coarse && w8[k] != 0.0 && a.dinv[c[k]] != 0.0) t += zc[c[k] >> sh];
          acc += w8[k] * t;
        }
      }
    }
    return acc;
  }
};

// dst[i] = src(i) for i < total over the workgroup, eight loads in flight per
// thread (src is called on in-range indices only)
template <typename T, typename Src>
__device__ void lds_fill(T* dst, int total, Src src) {
  for (int i0 = 0; i0 < total; i0 += 8 * 256) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = src(min(i0 + (k << 8) + (int)threadIdx.x, total - 1));
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = i0 + (k << 8) + threadIdx.x;
      if (i < total) dst[i] = v[k];
    }
  }
}

// Sums the entries v[i], i < Q with used(i), over the workgroup into
// sm.tot[i]: a butterfly per wavefront, then the four wavefronts as
// (w0 + w1) + (w2 + w3), the order of block_sum2, so every workgroup forms
// its totals the same way.  `used` is uniform over the workgroup.
template <int Q, bool FINE, typename Used>
__device__ void block_sums(const double (&v)[Q], Used used, Shared<FINE>& sm) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    if (!used(i)) continue;
    const double t = wave_sum(v[i]);
    if (l == 0) sm.wred[w][i] = t;
  }
  __syncthreads();
  if (threadIdx.x < Q && used((int)threadIdx.x)) {
    const int i = threadIdx.x;
    sm.tot[i] = (sm.wred[0][i] + sm.wred[1][i]) + (sm.wred[2][i] + sm.wred[3][i]);
  }
  __syncthreads();
}

// Where an exchange leaves the restrictions it reads, one slot per workgroup
// and aggregate of it: sm.cpart (aggregates of whole parts, summed per
// aggregate by gather_cin) or, with a fine coarse space, sm.cin directly.
__device__ inline double* coarse_in(Shared<false>& sm) { return sm.cpart; }
__device__ inline double* coarse_in(Shared<true>& sm) { return sm.cin; }

// sm.cin = the aggregates' restrictions, from the per-part ones in sm.cpart
// (aggregates of whole parts).
template <typename TV>
__device__ void gather_cin(const CgArgs<TV>& a, Shared<false>& sm, int G) {
  const int na = a.naggs;
  if (threadIdx.x < na) {
    double sacc = 0.0;
    const int g0 = threadIdx.x << a.agg_shift;
    const int g1 = min(G, (threadIdx.x + 1) << a.agg_shift);
    for (int g = g0; g < g1; g++) sacc += sm.cpart[g];
    sm.cin[threadIdx.x] = sacc;
  }
  __syncthreads();
}

// (A_c^-1 sm.cin)_i for aggregate i (threads i < naggs)
template <typename TV>
__device__ double coarse_row(const CgArgs<TV>& a, const Shared<false>& sm, int i) {
  const int na = a.naggs;
  double z = 0.0;
  for (int b = 0; b < na; b++) z += sm.ainv[i * na + b] * sm.cin[b];
  return z;
}

// Fine coarse space: (A_c^-1 sm.cin)_{need[r]} for r < nr, with
// T = 2^need_lg threads per row r = threadIdx.x / T (T is the same in every
// workgroup, at most 64, and T times the longest need list is at most 256).
// Every lane must call: the T partial sums are combined by lane shuffles in
// an order that depends on the row alone, so every workgroup gets the same
// value for an aggregate.
template <typename TV>
__device__ double coarse_need_row(const CgArgs<TV>& a, const Shared<true>& sm, int nr) {
  const int na = a.naggs, T = 1 << a.need_lg;
  const int r = threadIdx.x >> a.need_lg, q = threadIdx.x & (T - 1);
  double z = 0.0;
  if (r < nr) {
    const float* row = &sm.ainv[r * na];
    for (int b = q; b < na; b += T) z += (double)row[b] * sm.cin[b];
  }
  for (int o = 1; o < T; o <<= 1) z += __shfl_xor(z, o, 64);
  return z;
}

// sm.zc = A_c^-1 cin over the aggregates (cin gathered as above); with a
// fine coarse space, on the aggregates this workgroup's rows touch.
template <typename TV, bool FINE>
__device__ void coarse_solve(const CgArgs<TV>& a, Shared<FINE>& sm, int G) {
  if constexpr (FINE) {
    const int nr = sm.nneed, r = threadIdx.x >> a.need_lg;
    const double z = coarse_need_row(a, sm, nr);
    if (r < nr && (threadIdx.x & ((1 << a.need_lg) - 1)) == 0) sm.zc[sm.need[r]] = z;
  } else {
    gather_cin(a, sm, G);
    if (threadIdx.x < a.naggs) sm.zc[threadIdx.x] = coarse_row(a, sm, threadIdx.x);
  }
  __syncthreads();
}

// Publishes this workgroup's record of phase `ph`, [s1, s2, c_0 .. c_{L-1},
// cb_0 .. cb_{L-1}] (L = kSlots<FINE>) with c and cb the coarse restrictions
// of its S <= L aggregates (one per part without a fine coarse space; slots
// from S on are not written or read), and, if `with3`, s3 after them;
// returns the grid totals of s1, s2 (and s3) and leaves in sm.zc the coarse
// correction A_c^-1 (the restrictions c gathered per aggregate; with a fine
// coarse space they are read straight into sm.cin, one slot per aggregate,
// and zc is formed on the aggregates this workgroup's rows touch).  Every
// thread must have issued its
// sc1 stores of the phase's published vector before the call.  Returns false
// if the exchange timed out or another workgroup reported an error; the spin
// is bounded so that every workgroup leaves the solver and the grid drains.
template <typename TV, bool FINE>
__device__ bool exchange(const CgArgs<TV>& a, __amdgpu_buffer_rsrc_t rpart, unsigned ph,
                         double& s1, double& s2, const double (&c)[kSlots<FINE>],
                         const double (&cb)[kSlots<FINE>], double& s3, bool with3,
                         Shared<FINE>& sm) {
  wait_all_memory();
  constexpr int L = kSlots<FINE>, Q = 2 + 2 * L;
  static_assert(Q + 1 <= kCgRecord, "workgroup record");
  const int S = 1 << a.sub_shift;
  const auto used = [S, with3](int i) {
    return i < 2 || (i < Q && (i - 2) % L < S) || (i == Q && with3);
  };
  double v[Q + 1];
  v[0] = s1;
  v[1] = s2;
#pragma unroll
  for (int i = 0; i < L; i++) {
    v[2 + i] = c[i];
    v[2 + L + i] = cb[i];
  }
  v[Q] = s3;
  block_sums(v, used, sm);
  const int par = ph & 1, G = gridDim.x;
  const unsigned long long tg = (a.epoch << 24) | ph;
  if (threadIdx.x == 0) {
    double* slot = &a.part[kCgRecord * (par * a.gmax + blockIdx.x)];
    for (int i = 0; i <= Q; i++)
      if (used(i)) st_coh(slot + i, sm.tot[i]);
    wait_all_memory();
    st_coh(&a.tag[tag_slot(par, a.gmax, blockIdx.x)], tg);
  }
  double v1 = 0.0, v2 = 0.0, v3 = 0.0;
  double* const cdst = coarse_in(sm);
  int bad = 0;
  for (int t = threadIdx.x; t < G && !bad; t += 256) {
    unsigned spins = 0;
    while (ld_coh(&a.tag[tag_slot(par, a.gmax, t)]) != tg) {
      if ((++spins & 255u) == 0u && (ld_coh(&a.sync[2]) != 0u || spins > kSpinLimit)) {
        st_coh(&a.sync[2], 1u);
        bad = 1;
        break;
      }
    }
    if (!bad) {
      const int i = kCgRecord * (par * a.gmax + t);
      const double p1 = load_coh(rpart, i), p2 = load_coh(rpart, i + 1);
      for (int k = 0; k < S; k++) cdst[t * S + k] = load_coh(rpart, i + 2 + k);
      if (with3) v3 += load_coh(rpart, i + Q);
      v1 += p1;
      v2 += p2;
    }
  }
  if (__syncthreads_or(bad)) return false;
  block_sum2(v1, v2, sm.red);
  s1 = v1;
  s2 = v2;
  if (with3) {
    double none = 0.0;
    block_sum2(v3, none, sm.red);
    s3 = v3;
  }
  if (a.naggs > 0) coarse_solve(a, sm, G);
  return true;
}

// Reads the load's coarse restrictions cb of phase 1 into sm.cin and returns
// cb^T A_c^-1 cb (sm.zc, the correction of the exchange, is left alone).
// With a fine coarse space, only this workgroup's own aggregates' terms
// (need[0, S)): the grid sums them in the next exchange.
template <typename TV, bool FINE>
__device__ double load_norm_coarse(const CgArgs<TV>& a, __amdgpu_buffer_rsrc_t rpart,
                                   Shared<FINE>& sm) {
  constexpr int L = kSlots<FINE>;
  const int G = gridDim.x, na = a.naggs, S = 1 << a.sub_shift;
  double* const dst = coarse_in(sm);
  for (int t = threadIdx.x; t < G; t += 256)
    for (int k = 0; k < S; k++)
      dst[t * S + k] = load_coh(rpart, kCgRecord * (1 * a.gmax + t) + 2 + L + k);
  __syncthreads();
  double q = 0.0, dummy = 0.0;
  if constexpr (FINE) {
    const int r = threadIdx.x >> a.need_lg;
    const double z = coarse_need_row(a, sm, S);
    if (r < S && (threadIdx.x & ((1 << a.need_lg) - 1)) == 0) q = z * sm.cin[sm.need[r]];
  } else {
    gather_cin(a, sm, G);
    if (threadIdx.x < na) q = coarse_row(a, sm, threadIdx.x) * sm.cin[threadIdx.x];
  }
  block_sum2(q, dummy, sm.red);
  return q;
}

// Pipelined PCG (Ghysels & Vanroose 2014) with the two-level M: the inner
// products of iteration i and the vector m_i = M w_i gathered by the next SpMV
// travel in the same exchange.  Each thread owns RPT rows of its workgroup's
// part and keeps their recurrence vectors in registers.  Stops when
// (r, M r) <= rtol^2 (b_eff, M b_eff), b_eff = b - K x_fixed: the KSP rtol
// measured against the load vector.  FINE: the coarse space of sub-part
// aggregates (several per workgroup, FP32 inverse in LDS).
template <typename TV, int RPT, int MAXW, bool FINE>
__global__ __launch_bounds__(256) void k_cg(CgArgs<TV> a) {
  if (a.halt && *a.halt) return;  // queued by a finished integrator run
  __shared__ Shared<FINE> sm;
  const int n = a.n;
  const bool coarse = a.naggs > 0;
  const __amdgpu_buffer_rsrc_t rpub = make_rsrc(a.pub, 2u * (unsigned)n * 8u);
  const __amdgpu_buffer_rsrc_t rpart =
      make_rsrc(a.part, 2u * kCgRecord * (unsigned)a.gmax * 8u);
  if constexpr (FINE) {
    // the rows of the coarse inverse for the aggregates this part touches
    const int nr = a.naggs > 0 ? a.need_n[blockIdx.x] : 0, na = a.naggs;
    const int* nd = a.need + (size_t)blockIdx.x * kCgNeed;
    if (threadIdx.x < kCgNeed) sm.need[threadIdx.x] = (int)threadIdx.x < nr ? nd[threadIdx.x] : 0;
    if (threadIdx.x == 0) sm.nneed = nr;
    __syncthreads();
    lds_fill(sm.ainv, nr * na, [&](int i) {
      const int r = i / na;
      return a.ainvf[(size_t)sm.need[r] * na + (i - r * na)];
    });
  } else {
    lds_fill(sm.ainv, a.naggs * a.naggs, [&](int i) { return a.ainv[i]; });
  }
  // row -> aggregate, and each of this thread's rows' aggregate within its
  // part (the restrictions a workgroup publishes are per such aggregate)
  const int ash = a.cap_shift + a.agg_shift - a.sub_shift;
  int sub[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++)
    sub[k] = (((k << 8) + (int)threadIdx.x) >> (a.cap_shift - a.sub_shift)) &
             ((1 << a.sub_shift) - 1);
  RowSlots<TV, MAXW> rs[RPT];
  double x[RPT], r[RPT], u[RPT], w[RPT], z[RPT], q[RPT], s[RPT], p[RPT], di[RPT], mk[RPT];
  int orig[RPT];
  constexpr int L = kSlots<FINE>;
  double lbn = 0.0, dummy = 0.0, cr[L] = {}, cb[L] = {};
  // phase 1: r0 = b - K x0, publish D^-1 r0 and the coarse restrictions of r0, b_eff
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int row = (blockIdx.x * RPT + k) * 256 + threadIdx.x;
    x[k] = r[k] = u[k] = w[k] = z[k] = q[k] = s[k] = p[k] = di[k] = mk[k] = 0.0;
    orig[k] = row < n ? a.p2o[row] : -1;
    rs[k].load(a, row < n ? row : n - 1);
    if (orig[k] >= 0) {
      di[k] = a.dinv[row];
      x[k] = a.x[orig[k]];
      if (di[k] != 0.0) {
        double ax, af;
        rs[k].residual_terms(a, ax, af);
        const double bi = a.b ? a.b[orig[k]] : 0.0;
        r[k] = bi - ax;
        const double be = bi - af;
        lbn += be * be * di[k];
#pragma unroll
        for (int j = 0; j < L; j++) {
          cr[j] += sub[k] == j ? r[k] : 0.0;
          cb[j] += sub[k] == j ? be : 0.0;
        }
      }
    }
    if (row < n) st_coh(&a.pub[1 * n + row], di[k] * r[k]);
  }
  int it = 0;
  unsigned unconverged = 0u;
  double bq = 0.0;  // the load's coarse norm (a fine space's own terms until summed)
  const bool alive = exchange(a, rpart, 1u, lbn, dummy, cr, cb, bq, false, sm);
  if (alive) {
    if (coarse) bq = load_norm_coarse(a, rpart, sm);
    // with a fine coarse space the first exchange of the loop completes bq
    double stop = a.rtol2 * (lbn + bq);
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int row = (blockIdx.x * RPT + k) * 256 + threadIdx.x;
      if (di[k] != 0.0) {
        u[k] = di[k] * r[k] + (coarse ? sm.zc[row >> ash] : 0.0);
        w[k] = rs[k].gather(a, rpub, n, sm.zc, coarse);
      }
    }
    double gam_old = 1.0, alpha_old = 1.0;
    for (;; it++) {
      const unsigned ph = (unsigned)it + 2u;
      const int par = ph & 1;
      double gam = 0.0, del = 0.0, cw[L] = {};
      const double none[L] = {};
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = (blockIdx.x * RPT + k) * 256 + threadIdx.x;
        gam += r[k] * u[k];
        del += w[k] * u[k];
#pragma unroll
        for (int j = 0; j < L; j++) cw[j] += di[k] != 0.0 && sub[k] == j ? w[k] : 0.0;
        if (row < n) st_coh(&a.pub[par * n + row], di[k] * w[k]);
      }
      const bool sum_bq = FINE && coarse && it == 0;
      if (!exchange(a, rpart, ph, gam, del, cw, none, bq, sum_bq, sm)) break;
      if (sum_bq) stop = a.rtol2 * (lbn + bq);
      if (!(gam > stop) || gam <= 1e-300) break;
      if (it >= a.maxit) {
        unconverged = 1u;
        break;
      }
      if (!isfinite(gam) || !isfinite(del)) {
        if (threadIdx.x == 0) st_coh(&a.sync[2], 2u);
        break;
      }
      const double beta = it > 0 ? gam / gam_old : 0.0;
      const double alpha = it > 0 ? gam / (del - beta * gam / alpha_old) : gam / del;
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int row = (blockIdx.x * RPT + k) * 256 + threadIdx.x;
        double nk = 0.0;
        if (di[k] != 0.0) {
          mk[k] = di[k] * w[k] + (coarse ? sm.zc[row >> ash] : 0.0);
          nk = rs[k].gather(a, rpub, par * n, sm.zc, coarse);
        } else {
          mk[k] = 0.0;
        }
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
    const int o = orig[k];
    if (o >= 0) {
      a.x[o] = x[k];
      if (a.xsum) a.xsum[o] = x[k] + a.xadd[o];
      if (a.rout) {
        const int lb = a.blocal[o];
        if (lb >= 0) a.rout[lb] = x[k];
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sync[3] = (unsigned)it;
    a.sync[4] += (unsigned)it;
    a.sync[5] += 1u;
    a.sync[6] += unconverged;
  }
}

template <typename TV, int RPT, int MAXW, bool FINE>
const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_cg<TV, RPT, MAXW, FINE>);
}

template <typename TV, bool FINE>
const void* select_kernel_f(int rpt, int maxw) {
  if (maxw == 16) {
    if (rpt == 1) return kernel_ptr<TV, 1, 16, FINE>();
    if (rpt == 2) return kernel_ptr<TV, 2, 16, FINE>();
  } else if (maxw == 32) {
    if (rpt == 1) return kernel_ptr<TV, 1, 32, FINE>();
    if (rpt == 2) return kernel_ptr<TV, 2, 32, FINE>();
  } else if (maxw == 0) {
    if (rpt == 1) return kernel_ptr<TV, 1, 0, FINE>();
    if (rpt == 2) return kernel_ptr<TV, 2, 0, FINE>();
    if (rpt == 4) return kernel_ptr<TV, 4, 0, FINE>();
  }
  return nullptr;
}

template <typename TV>
const void* select_kernel(int rpt, int maxw, bool fine) {
  return fine ? select_kernel_f<TV, true>(rpt, maxw) : select_kernel_f<TV, false>(rpt, maxw);
}

}  // namespace

int cg_slot_variant(int width, int rpt) {
  if (rpt <= 2 && width <= 16) return 16;
  if (rpt <= 2 && width <= 32) return 32;
  return 0;
}

int resident_per_cu(const void* kernel, int threads, int device) {
  int occ = 0, lds_cu = 0;
  hipFuncAttributes fa;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess ||
      hipFuncGetAttributes(&fa, kernel) != hipSuccess ||
      hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                            device) != hipSuccess)
    return 0;
  if (lds_cu <= 0 || lds_cu > kLdsPerCu) lds_cu = kLdsPerCu;
  if (fa.sharedSizeBytes > 0) {
    const int per = lds_granules((int)fa.sharedSizeBytes) * kLdsGranule;
    const int by_lds = lds_cu / per;
    if (by_lds < occ) occ = by_lds;
  }
  return occ > 0 ? occ : 0;
}

int cg_max_workgroups(int rpt, int maxw, int device, bool fine) {
  const void* k = select_kernel<double>(rpt, maxw, fine);
  if (!k) return 0;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return resident_per_cu(k, 256, device) * cus;
}

hipError_t launch_cg(const CgDevice& c, CgWork& w, bool dirichlet, bool coarse,
                     const double* b, double* x, double rtol, int maxit, const double* xadd,
                     double* xsum, double* restrict_out, const int* blocal,
                     hipStream_t st) {
  const int grid = c.parts;
  // a plain launch above the resident count would leave workgroups waiting
  // for ones that cannot start
  if (grid <= 0 || grid > w.gmax || grid > kMaxParts || grid > c.resident ||
      maxit >= (1 << 23) ||
      (c.maxw > 0 && c.max_width > c.maxw) ||
      c.naggs > (c.fine ? kCgFineAggs : kMaxAggs) || (c.fine && grid > kCgFineParts) ||
      (c.fine ? c.agg_shift != 0 || c.sub_shift < 1 || c.sub_shift > 2 || !c.ainvf_nbc ||
                    !c.ainvf_dbc || !c.need || !c.need_n || c.naggs != grid << c.sub_shift ||
                    c.need_lg < 2 || c.need_lg > 6 || c.need_max > (256 >> c.need_lg)
              : c.sub_shift != 0) ||
      2LL * c.npad * 8 >= 0x7fffffffLL || (256 << 0) * c.rpt != (1 << c.cap_shift))
    return hipErrorInvalidConfiguration;
  const unsigned long long epoch = ++w.epoch;
  const void* k = select_kernel<double>(c.rpt, c.maxw, c.fine);
  if (!k) return hipErrorInvalidConfiguration;
  const double* dinv = dirichlet ? c.dinv_dbc : c.dinv_nbc;
  const double* ainv = dirichlet ? c.ainv_dbc : c.ainv_nbc;
  const float* ainvf = dirichlet ? c.ainvf_dbc : c.ainvf_nbc;
  const int naggs = coarse ? c.naggs : 0;
  CgArgs<double> a{c.npad,  w.gmax, maxit, c.cap_shift, c.agg_shift, c.sub_shift, naggs,
                   c.sell_off, c.sell_col, c.K, c.p2o, dinv, ainv, ainvf, c.need, c.need_n,
                   c.need_lg, b, x, w.pub, w.part,
                   w.tag,   w.sync, epoch, rtol * rtol, xadd, xsum, restrict_out, blocal,
                   w.halt};
  void* args[] = {&a};
  if (w.cooperative) return hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), args, 0, st);
  return hipLaunchKernel(k, dim3(grid), dim3(256), args, 0, st);
}

}  // namespace nsim
