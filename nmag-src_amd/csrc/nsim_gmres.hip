// Persistent Newton iteration for the BDF integrator: residual, block-Jacobi
// set-up, the GMRES solve of (I - gamma J) x = r, the update and its norms in
// one launch.  It does the work of CVODE's SPGMR with the jacobi_ts1 J.v and
// the preconditioner callbacks (src/sundials_sp_stubs.c:700-1003,
// src/nsim.ml:2273-2484).
//
// GMRES runs right-preconditioned in the space scaled by the error weights
// (the Euclidean residual is sqrt(n) times its WRMS norm) in its flexible
// form: the directions z_j are stored, the Krylov basis V is orthonormalised
// by classical Gram-Schmidt applied twice, and x = Z y.  Each workgroup owns
// whole SELL slices (64 nodes each) and every vector element of their nodes;
// everything it computes on them is local except the exchange part of J.z,
// which gathers z at neighbouring nodes.  Grid-wide sums meet in exchanges:
// each workgroup publishes its partial sums with agent-coherent stores, then
// a phase tag; every workgroup polls all tags and sums the partials in the
// same fixed order, so all of them hold bit-identical totals, take the same
// Givens rotations and stop at the same iteration without further
// communication.
//
// The first kOneExMax Arnoldi steps take one exchange each: the second
// Gram-Schmidt pass of a basis vector travels with the first pass of the next
// (delayed reorthogonalisation), the next direction is formed from the
// operator's output before the exchange (a Newton basis, z_{j+1} = (M^-1 A -
// I) z_j: flexible GMRES admits any directions, and these span the same
// Krylov space), and the norm of the next basis vector is estimated ahead
// (|w|^2 - |h|^2), so a converged solve stops at the exchange that shows it.
// Later steps take two exchanges: the first Gram-Schmidt column, then the
// second with |v'|^2 of the once-projected v', whose direction z_{j+1} = M^-1
// (v' / ewt) is published at the same time (the new norm follows as |v'|^2 -
// |h2|^2, measured directly when that difference loses more than two
// digits).  Both keep the Arnoldi relation A Z = V H for the z actually used.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_gmres.h"

namespace nsim {

namespace {

using namespace dev;

constexpr int kSplit = 4;                  // wavefronts of a workgroup (J.z gather partials)
constexpr unsigned kSpinLimit = 1u << 20;  // polls before an exchange gives up
constexpr int kMaxGrid = 1024;
constexpr int kPcNodes = 256;              // largest node block of the local sweeps
// Krylov iterations whose R every workgroup keeps in LDS (what fits beside
// the rest while three workgroups share a CU's 160 KiB)
constexpr int kRtCap = 32;
// Arnoldi steps taken with one exchange each before the two-exchange steps
// take over (the Newton basis of those steps stays well conditioned below it)
constexpr int kOneExMax = kGmresOneExMax;
constexpr int kSumChunk = 4;  // exchanged entries a wavefront sums per round trip

template <typename TV>
struct GmresArgs {
  int n, nslices, spw, maxl, gmax, pstride;
  int sweeps;          // local Jacobi sweeps of the preconditioner (0: block Jacobi)
  int onex;            // Arnoldi steps with one exchange each (at most kOneExMax)
  const int* off;
  const int* col;
  const TV* K;
  const double* Vinv;
  const Coeffs* ct;
  const int* inst;
  double* m;           // Newton iterate ycur; receives ycur + x
  const double* Ht;    // H_total of the rhs at m
  const double* ewt;
  const double* kappa;
  double* Minv;        // block-Jacobi inverses, set up here (null: none)
  double gamma, tol;
  const double* avec;  // residual r = avec + gamma f - m
  const double* f;
  const double* yp;    // predictor (for the local error sum)
  double* x;           // the correction
  double* V;
  double* Z;           // maxl directions, published coherently
  size_t ld;
  double* part;
  unsigned long long* tag;
  double* R;
  double* y;
  double* status;
  unsigned* err;
  unsigned long long epoch;
  const double* gamma_p;  // if set, gamma is read here (set by the device controller)
  const int* halt;        // nonzero word: every workgroup returns at once
  double* const* slots;  // the controller's history ring (read with ts)
  // If set, workgroup 0 runs the step controller on the results at the end.
  // Every workgroup reads gamma and halt (fields of this record) at its start,
  // and workgroup 0 reaches the end only after the last exchange, to which
  // every workgroup contributes; so no read races the controller's write.
  TsState* ts;
};

struct GShared {
  double cs[kGmresMaxl], sn[kGmresMaxl], g[kGmresMaxl + 1];
  double zs[kGmresMaxl];       // z_j = zs[j] * Z_j
  double col[kGmresMaxl + 1];  // Hessenberg column of this iteration
  double colp[kOneExMax + 1];  // first-pass entries of the pending column (one-exchange steps)
  double loc[kGmresMaxl + 2];  // this workgroup's partial sums
  double tot[kGmresMaxl + 2];  // grid totals
  double gp[3 * kSplit][64];   // J.z gather partials
  double pt[3 * kPcNodes];     // preconditioner: right-hand side on own nodes
  double pz[2][3 * kPcNodes];  //                 sweep iterates
  double red[8];
  double hn2, vn2;  // |v|^2 from |v'|^2 - |h2|^2, and |v'|^2
  int stop;         // the look-ahead column met the tolerance
  // the first kRtCap columns of R, upper triangle packed column by column
  // (R[i][j] at j (j + 1) / 2 + i): every workgroup holds the same R, so up
  // to kRtCap iterations each solves R y = g itself, with no broadcast
  double Rt[kRtCap * (kRtCap + 1) / 2];
};
// Three workgroups per CU (the VGPR budget allows three), each rounded up to
// whole LDS granules, with 1 KiB for the compiler's own LDS use.  A record
// that rounded past a third of the CU would let only two start per CU while
// the occupancy query still reports three.
static_assert(3 * lds_granules((int)sizeof(GShared) + 1024) * kLdsGranule <= kLdsPerCu,
              "GMRES workgroup LDS too large for three workgroups per CU");

// A workgroup takes its slices [s0, s1) in rounds of up to four, all of a
// round's slices at once: with ns slices in the round, 4 / ns wavefronts (one
// when ns is 3) share a slice, wavefront wv taking slice sb + wv / per and the
// column slots part, part + per, ... of its rows.  One slice per workgroup
// splits every row over the four wavefronts.
struct SliceRound {
  int s, part, per;
  bool active;
};
__device__ inline SliceRound slice_round(int sb, int s1, int wv) {
  const int ns = min(4, s1 - sb);
  const int per = ns == 1 ? 4 : ns == 2 ? 2 : 1;
  return SliceRound{sb + wv / per, wv % per, per, wv / per < ns};
}

// z = M^-1 t on one node (3x3 block).
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

template <typename TV>
__device__ inline __amdgpu_buffer_rsrc_t z_rsrc(const GmresArgs<TV>& a, int j) {
  return make_rsrc(a.Z + (size_t)j * a.ld, (unsigned)(a.ld * 8));
}

// Publishes sm.loc[0, cnt) for phase `ph` and, unless this workgroup is not
// the root of a root_only exchange, sums every workgroup's values into
// sm.tot[0, cnt).  The values are stored entry-major.  Thread t polls the
// tags of workgroups t, t + 256, ... with their loads in flight together.
// Then wavefront w sums the entries
// w, w + 4, ..., kSumChunk of them per pass with all their loads in flight:
// lane l over the workgroups l, l + 64, ... in increasing order, then a fixed
// butterfly.  The order is the same in every workgroup, so all of them hold
// bit-identical totals.  Every wave's earlier coherent stores (the published
// direction) complete before the tag.  Returns false if the exchange timed
// out or another workgroup reported an error (the spin is bounded, so every
// workgroup leaves and the grid drains).
template <typename TV>
__device__ bool exchange(const GmresArgs<TV>& a, GShared& sm, unsigned ph, int cnt,
                         bool root_only) {
  constexpr int Q = kMaxGrid / 256;  // workgroups per polling thread
  const int par = ph & 1, G = gridDim.x;
  const unsigned long long tg = (a.epoch << 24) | ph;
  __syncthreads();  // sm.loc complete
  // entry-major: entry i of every workgroup is contiguous, so a wavefront's
  // reads of one entry coalesce into whole lines
  double* part = a.part + (size_t)par * a.pstride * a.gmax;
  for (int i = threadIdx.x; i < cnt; i += 256) st_coh(part + (size_t)i * a.gmax + blockIdx.x, sm.loc[i]);
  wait_all_memory();
  __syncthreads();
  if (threadIdx.x == 0) st_coh(&a.tag[tag_slot(par, a.gmax, blockIdx.x)], tg);
  if (root_only && blockIdx.x != 0) return true;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const __amdgpu_buffer_rsrc_t rt = make_rsrc(a.tag, (unsigned)(2 * a.gmax * kTagStride * 8));
  unsigned need = 0u;
#pragma unroll
  for (int q = 0; q < Q; q++)
    if (tid + 256 * q < G) need |= 1u << q;
  int bad = 0;
  unsigned spins = 0;
  while (need) {
    unsigned long long v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++)
      v[q] = (need >> q) & 1u ? load_tag(rt, tag_slot(par, a.gmax, tid + 256 * q)) : tg;
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (v[q] == tg) need &= ~(1u << q);
    if (need && (++spins & 255u) == 0u && (ld_coh(a.err) != 0u || spins > kSpinLimit)) {
      st_coh(a.err, 1u);
      bad = 1;
      break;
    }
  }
  if (__syncthreads_or(bad)) return false;
  const __amdgpu_buffer_rsrc_t rp = make_rsrc(part, (unsigned)(a.gmax * a.pstride * 8));
  for (int i0 = w; i0 < cnt; i0 += 4 * kSumChunk) {
    double sv[kSumChunk];
#pragma unroll
    for (int c = 0; c < kSumChunk; c++) sv[c] = 0.0;
#pragma unroll 4
    for (int t = l; t < G; t += 64) {
#pragma unroll
      for (int c = 0; c < kSumChunk; c++)
        if (i0 + 4 * c < cnt) sv[c] += load_coh(rp, (i0 + 4 * c) * a.gmax + t);
    }
#pragma unroll
    for (int c = 0; c < kSumChunk; c++) {
      const double s = wave_sum(sv[c]);
      if (l == 0 && i0 + 4 * c < cnt) sm.tot[i0 + 4 * c] = s;
    }
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
    while (ld_coh(&a.tag[tag_slot(ph & 1, a.gmax, 0)]) != tg) {
      if ((++spins & 255u) == 0u && (ld_coh(a.err) != 0u || spins > kSpinLimit)) {
        st_coh(a.err, 1u);
        bad = 1;
        break;
      }
    }
  }
  return !__syncthreads_or(bad);
}

// Sum over the workgroup into sm.loc[slot] (fixed order).
__device__ inline void block_total(double s, GShared& sm, int slot) {
  double dummy = 0.0;
  block_sum2(s, dummy, sm.red);
  if (threadIdx.x == 0) sm.loc[slot] = s;
}

// Z_j = M^-1 t - sub Z_jsub on this workgroup's nodes (no subtraction when
// jsub < 0), published coherently, with t = v / ewt (or, for the first
// direction, t = the Newton residual); returns this thread's part of |Z_j|^2.  M is
// the 3x3 block-Jacobi splitting A = D - N of A = I - gamma J, improved by
// `sweeps` Jacobi sweeps z <- D^-1 (t + N z) over the couplings inside the
// workgroup's own nodes (an additive-Schwarz block solve; couplings to other
// workgroups' nodes are dropped, so M stays local and fixed).  N z at node i
// is gamma (c1 m x h + c2 m x (m x h)) with h = negJ/V_i sum_{j != i} K_ij z_j.
template <typename TV>
__device__ double publish_direction(const GmresArgs<TV>& a, GShared& sm, const double* v,
                                    bool residual, int j, int s0, int s1, int r0, int r1,
                                    int jsub, double sub) {
  double* Zj = a.Z + (size_t)j * a.ld;
  const bool local = a.Minv && a.sweeps > 0 && r1 - r0 <= kPcNodes;
  const __amdgpu_buffer_rsrc_t rs = z_rsrc(a, jsub >= 0 ? jsub : 0);
  double zz = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += 256) {
    double t[3], z[3];
    for (int d = 0; d < 3; d++) {
      if (residual) {
        double r = 1.0 * a.avec[3 * i + d];
        r += a.gamma * a.f[3 * i + d];
        r += -1.0 * a.m[3 * i + d];
        t[d] = r;
      } else {
        t[d] = v[3 * i + d] / a.ewt[3 * i + d];
      }
    }
    pc_node(a.Minv, i, t, z);
    if (local) {
      for (int d = 0; d < 3; d++) {
        sm.pt[3 * (i - r0) + d] = t[d];
        sm.pz[0][3 * (i - r0) + d] = z[d];
      }
    } else {
      for (int d = 0; d < 3; d++) {
        const double zd = jsub >= 0 ? z[d] - sub * load_coh(rs, 3 * i + d) : z[d];
        zz += zd * zd;
        st_coh(Zj + 3 * i + d, zd);
      }
    }
  }
  if (!local) return zz;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int cur = 0;
  for (int sw = 0; sw < a.sweeps; sw++) {
    __syncthreads();
    for (int sb = s0; sb < s1; sb += 4) {
      const SliceRound sr = slice_round(sb, s1, wv);
      const int row = 64 * sr.s + lane;
      const bool live = sr.active && row < a.n;
      double hv[3] = {0.0, 0.0, 0.0};
      if (live) {
        const int b = a.off[sr.s], w = (a.off[sr.s + 1] - b) >> 6;
        for (int q = sr.part; q < w; q += sr.per) {
          const int e = b + (q << 6) + lane;
          const int c = a.col[e];
          if (c == row || c < r0 || c >= r1) continue;
          const double kv = (double)a.K[e];
          hv[0] += kv * sm.pz[cur][3 * (c - r0)];
          hv[1] += kv * sm.pz[cur][3 * (c - r0) + 1];
          hv[2] += kv * sm.pz[cur][3 * (c - r0) + 2];
        }
      }
      for (int q = 0; q < 3; q++) sm.gp[wv * 3 + q][lane] = hv[q];
      __syncthreads();
      if (sr.part == 0 && live) {
        for (int q = 0; q < 3; q++) {
          double acc = sm.gp[wv * 3 + q][lane];
          for (int u = 1; u < sr.per; u++) acc += sm.gp[(wv + u) * 3 + q][lane];
          hv[q] = acc;
        }
        const Coeffs& cf = a.ct[a.inst ? a.inst[row] : 0];
        const double fct = cf.negJ * a.Vinv[row];
        const double h[3] = {fct * hv[0], fct * hv[1], fct * hv[2]};
        const double mi[3] = {a.m[3 * row], a.m[3 * row + 1], a.m[3 * row + 2]};
        const double zero[3] = {0.0, 0.0, 0.0};
        double nz[3], t[3], z[3];
        llg_jv_row(cf.c1, cf.c2, h, mi, zero, zero, nz);
        for (int d = 0; d < 3; d++) t[d] = sm.pt[3 * (row - r0) + d] + a.gamma * nz[d];
        pc_node(a.Minv, row, t, z);
        for (int d = 0; d < 3; d++) sm.pz[cur ^ 1][3 * (row - r0) + d] = z[d];
      }
      __syncthreads();
    }
    cur ^= 1;
  }
  for (int i = r0 + threadIdx.x; i < r1; i += 256)
    for (int d = 0; d < 3; d++) {
      const double zp = sm.pz[cur][3 * (i - r0) + d];
      const double zd = jsub >= 0 ? zp - sub * load_coh(rs, 3 * i + d) : zp;
      zz += zd * zd;
      st_coh(Zj + 3 * i + d, zd);
    }
  return zz;
}

// sm.loc[i] = this workgroup's part of V_i . v for i <= j.
template <typename TV>
__device__ void local_dots(const GmresArgs<TV>& a, GShared& sm, const double* v, int j, int e0,
                           int e1) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wv; i <= j; i += 4) {
    const double* Vi = a.V + (size_t)i * a.ld;
    double s = 0.0;
    for (int e = e0 + lane; e < e1; e += 64) s += Vi[e] * v[e];
    s = wave_sum(s);
    if (lane == 0) sm.loc[i] = s;
  }
}

// Vn = (z - gamma J z) ewt on this workgroup's rows, z = zs[j] Z_j: the
// exchange part of J z gathers z at neighbouring nodes (published by other
// workgroups, read device-coherently).
template <typename TV>
__device__ void apply_op(const GmresArgs<TV>& a, GShared& sm, int j, double* Vn, int s0,
                         int s1) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const __amdgpu_buffer_rsrc_t rz = z_rsrc(a, j);
  const double sc = sm.zs[j];
  for (int sb = s0; sb < s1; sb += 4) {
    const SliceRound sr = slice_round(sb, s1, wv);
    const int row = 64 * sr.s + lane;
    const bool live = sr.active && row < a.n;
    double hv[3] = {0.0, 0.0, 0.0};
    if (live) {
      const int b = a.off[sr.s], w = (a.off[sr.s + 1] - b) >> 6;
      for (int q = sr.part; q < w; q += sr.per) {
        const int e = b + (q << 6) + lane;
        const int c = a.col[e];
        const double kv = (double)a.K[e];
        hv[0] += kv * load_coh(rz, 3 * c);
        hv[1] += kv * load_coh(rz, 3 * c + 1);
        hv[2] += kv * load_coh(rz, 3 * c + 2);
      }
    }
    for (int q = 0; q < 3; q++) sm.gp[wv * 3 + q][lane] = hv[q];
    __syncthreads();
    if (sr.part == 0 && live) {
      for (int q = 0; q < 3; q++) {
        double acc = sm.gp[wv * 3 + q][lane];
        for (int u = 1; u < sr.per; u++) acc += sm.gp[(wv + u) * 3 + q][lane];
        hv[q] = acc;
      }
      const Coeffs& cf = a.ct[a.inst ? a.inst[row] : 0];
      const double fct = cf.negJ * a.Vinv[row] * sc;
      const double h[3] = {fct * hv[0], fct * hv[1], fct * hv[2]};
      const double mi[3] = {a.m[3 * row], a.m[3 * row + 1], a.m[3 * row + 2]};
      const double Hi[3] = {a.Ht[3 * row], a.Ht[3 * row + 1], a.Ht[3 * row + 2]};
      double zi[3], jv[3];
      for (int d = 0; d < 3; d++) zi[d] = sc * load_coh(rz, 3 * row + d);
      llg_jv_row(cf.c1, cf.c2, h, mi, zi, Hi, jv);
      for (int d = 0; d < 3; d++)
        Vn[3 * row + d] = (zi[d] - a.gamma * jv[d]) * a.ewt[3 * row + d];
    }
    __syncthreads();
  }
}

// The sums of a one-exchange step over this workgroup's elements, in one
// pass: sm.loc[i] = V_i . u and sm.loc[j + i] = V_i . w for i < j, then
// sm.loc[2j] = u . u, [2j + 1] = u . w, [2j + 2] = the threads' zz (their
// parts of |z_{j+1}|^2), [2j + 3] = w . w; the w terms are zero when w is
// null.  Thread t takes the elements t, t + 256, ..., the basis vectors go
// eight at a time with all their loads in flight, and each sum is a
// wavefront butterfly followed by the four wavefronts in a fixed order (so
// every workgroup forms its parts the same way).  sm.gp holds the wavefronts'
// partials (4 (2j + 4) <= 4 (2 kOneExMax + 4) doubles).
template <typename TV>
__device__ void step_dots(const GmresArgs<TV>& a, GShared& sm, const double* u, const double* w,
                          int j, int e0, int e1, double zz) {
  constexpr int C = 8;
  static_assert(4 * (2 * kOneExMax + 4) <= 3 * kSplit * 64, "wavefront partials");
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int cnt = 2 * j + 4;
  double* part = &sm.gp[0][0] + wv * cnt;
  {
    double uu = 0.0, cc = 0.0, ww = 0.0;
    for (int e = e0 + tid; e < e1; e += 256) {
      const double ue = u[e];
      uu += ue * ue;
      if (w) {
        const double we = w[e];
        cc += ue * we;
        ww += we * we;
      }
    }
    uu = wave_sum(uu);
    cc = wave_sum(cc);
    zz = wave_sum(zz);
    ww = wave_sum(ww);
    if (lane == 0) {
      part[2 * j] = uu;
      part[2 * j + 1] = cc;
      part[2 * j + 2] = zz;
      part[2 * j + 3] = ww;
    }
  }
  for (int i0 = 0; i0 < j; i0 += C) {
    double sa[C], sb[C];
#pragma unroll
    for (int c = 0; c < C; c++) sa[c] = sb[c] = 0.0;
    for (int e = e0 + tid; e < e1; e += 256) {
      double vi[C];
#pragma unroll
      for (int c = 0; c < C; c++) vi[c] = i0 + c < j ? a.V[(size_t)(i0 + c) * a.ld + e] : 0.0;
      const double ue = u[e], we = w ? w[e] : 0.0;
#pragma unroll
      for (int c = 0; c < C; c++) {
        sa[c] += vi[c] * ue;
        sb[c] += vi[c] * we;
      }
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
      if (i0 + c < j) {  // uniform over the workgroup
        const double ta = wave_sum(sa[c]), tb = wave_sum(sb[c]);
        if (lane == 0) {
          part[i0 + c] = ta;
          part[j + i0 + c] = tb;
        }
      }
    }
  }
  __syncthreads();
  const double* p = &sm.gp[0][0];
  for (int i = tid; i < cnt; i += 256)
    sm.loc[i] = (p[i] + p[cnt + i]) + (p[2 * cnt + i] + p[3 * cnt + i]);
}

// Thread 0: Hessenberg column j (sm.col[0 .. j + 1]) through the earlier
// Givens rotations and a new one, the residual vector g updated; the column
// of R kept in LDS (j < kRtCap) and, by workgroup 0, in R.
template <typename TV>
__device__ void rotate_column(const GmresArgs<TV>& a, GShared& sm, int j) {
  double* c = sm.col;
  for (int i = 0; i < j; i++) {
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
  if (j < kRtCap)
    for (int i = 0; i <= j; i++) sm.Rt[j * (j + 1) / 2 + i] = c[i];
  if (blockIdx.x == 0)
    for (int i = 0; i <= j; i++) a.R[(size_t)i * a.maxl + j] = c[i];
}

// The step controller as a call on the record in place: inlined, or on a
// register copy of the record, its registers would count against the whole
// kernel.  The kernel's occupancy target (three waves per SIMD) makes the
// controller spill what does not fit instead; it runs on one thread.
// (Measured at the slab: in place 88.3 us per Newton iteration, a copy in LDS
// 88.7-90.8 us, a separate one-thread launch 83.8 + 9.9 us.)
__device__ __attribute__((noinline)) void run_controller(TsState* s, double its, double xw2,
                                                         double bad, double errsq,
                                                         double errsq_m, double errsq_p,
                                                         int n3) {
  ts_control_at(*s, its, xw2, bad, errsq, errsq_m, errsq_p, n3);
}

template <typename TV>
// Three waves per SIMD: what the LDS allows (see GShared) and what the
// exchanges want; the step controller called at the end must fit the same
// register budget.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_gmres(
    GmresArgs<TV> a) {
  if (a.halt && *a.halt) return;  // queued by a finished integrator run
  if (a.gamma_p) a.gamma = *a.gamma_p;
  __shared__ GShared sm;
  const int s0 = blockIdx.x * a.spw, s1 = min(a.nslices, s0 + a.spw);
  const int r0 = 64 * s0, r1 = min(a.n, 64 * s1);
  const int e0 = 3 * r0, e1 = 3 * r1;
  const int tid = threadIdx.x;
  const double tol = a.tol;
  unsigned ph = 0;
  bool ok = true;
  int k = 0;

  // Newton residual r = avec + gamma f - ycur and this workgroup's
  // block-Jacobi inverses at the current iterate (its own nodes only);
  // v0 = r ewt, beta = |v0|; Z_0 = M^-1 r
  {
    if (a.Minv)
      for (int i = r0 + tid; i < r1; i += 256) {
        const Coeffs& cf = a.ct[a.inst ? a.inst[i] : 0];
        const double mi[3] = {a.m[3 * i], a.m[3 * i + 1], a.m[3 * i + 2]};
        const double Hi[3] = {a.Ht[3 * i], a.Ht[3 * i + 1], a.Ht[3 * i + 2]};
        bjac_block(cf.c1, cf.c2, a.gamma, a.kappa[i], mi, Hi, a.Minv + 9 * i);
      }
    publish_direction(a, sm, nullptr, true, 0, s0, s1, r0, r1, -1, 0.0);
    double s = 0.0;
    for (int e = e0 + tid; e < e1; e += 256) {
      double r = 1.0 * a.avec[e];
      r += a.gamma * a.f[e];
      r += -1.0 * a.m[e];
      const double v = r * a.ewt[e];
      a.V[e] = v;
      s += v * v;
    }
    block_total(s, sm, 0);
    ok = exchange(a, sm, ++ph, 1, false);
  }
  const double beta = ok ? sqrt(sm.tot[0]) : 0.0;
  const bool solved = !(beta > tol);  // then x = 0 (k stays 0)
  if (ok && !solved) {
    const double sc = 1.0 / beta;
    for (int e = e0 + tid; e < e1; e += 256) a.V[e] *= sc;
    if (tid == 0) {
      sm.g[0] = beta;
      sm.zs[0] = sc;
    }
    __syncthreads();
  }
  int j = 0;
  bool conv = !ok || solved;  // no (further) Arnoldi steps
  // One exchange per Arnoldi step while j < kOneExMax: step j applies the
  // operator to z_j, publishes the next direction
  //   z_{j+1} = M^-1 (w_j / ewt) - z_j,  w_j = (z_j - gamma J z_j) ewt,
  // and in one exchange gets the second Gram-Schmidt pass of the pending
  // u_j (the once-projected w_{j-1}): a = V^T u_j and |u_j|^2, together with
  // the first pass of w_j: b = V^T w_j, u_j . w_j and |w_j|^2.  Then v_j =
  // (u_j - V a) / alpha finishes Hessenberg column j - 1 (its first-pass
  // entries plus a, and alpha below them), whose Givens rotation gives the
  // residual after j steps, and u_{j+1} = w_j - V b - h_jj v_j waits for the
  // next step.  The directions form a Newton basis of the same Krylov space
  // as the two-exchange steps below (with M fixed, exactly the iterates of
  // right-preconditioned GMRES), whose conditioning grows with the step
  // count, so after a.onex (at most kOneExMax) steps the two-exchange steps
  // take over.
  if (a.onex > 0 && !conv) {
    for (;; j++) {
      const bool pend = j > 0;  // u_j waits in V_j for its second pass
      const bool havew = j < a.onex && j < a.maxl;
      double* Vj = a.V + (size_t)j * a.ld;
      double* Vn = a.V + (size_t)(j + 1) * a.ld;  // w_j, then u_{j+1}
      double zz = 0.0;
      if (havew) {
        apply_op(a, sm, j, Vn, s0, s1);
        if (j + 1 < a.maxl)
          zz = publish_direction(a, sm, Vn, false, j + 1, s0, s1, r0, r1, j, sm.zs[j]);
      }
      step_dots(a, sm, Vj, havew ? Vn : nullptr, j, e0, e1, zz);
      ok = exchange(a, sm, ++ph, 2 * j + 4, false);
      if (!ok) break;
      // everything needed from the totals, before another exchange may
      // overwrite them: column j - 1 = pending entries + a, the new pending
      // entries b, and a.a, a.b, b.b
      const double uu = sm.tot[2 * j], cc = sm.tot[2 * j + 1], zt = sm.tot[2 * j + 2];
      const double ww = sm.tot[2 * j + 3];
      double aa = 0.0, ab = 0.0, bb = 0.0;
      for (int i = 0; i < j; i++) {
        aa += sm.tot[i] * sm.tot[i];
        ab += sm.tot[i] * sm.tot[j + i];
        bb += sm.tot[j + i] * sm.tot[j + i];
      }
      for (int i = tid; i < j; i += 256) sm.col[i] = sm.colp[i] + sm.tot[i];
      __syncthreads();
      for (int i = tid; i < j; i += 256) sm.colp[i] = sm.tot[j + i];
      double alpha = 1.0;
      bool vdone = false;  // V_j already holds u_j - V a
      if (pend) {
        double al2 = uu - aa;
        if (!(al2 > 1e-2 * uu)) {  // cancellation: form u_j - V a and measure it directly
          for (int e = e0 + tid; e < e1; e += 256) {
            double v = Vj[e];
            for (int i = 0; i < j; i++) v -= sm.tot[i] * a.V[(size_t)i * a.ld + e];
            Vj[e] = v;
          }
          vdone = true;
          double sq = 0.0;
          for (int e = e0 + tid; e < e1; e += 256) sq += Vj[e] * Vj[e];
          block_total(sq, sm, 0);
          ok = exchange(a, sm, ++ph, 1, false);
          if (!ok) break;
          al2 = sm.tot[0];
        }
        alpha = sqrt(fmax(al2, 0.0));
        if (tid == 0) {
          sm.col[j] = alpha;
          rotate_column(a, sm, j - 1);
        }
        __syncthreads();
        k = j;
        if (fabs(sm.g[j]) <= tol || !(alpha > 0.0) || !isfinite(alpha)) {
          conv = true;
          break;
        }
      }
      const double inv = pend ? 1.0 / alpha : 1.0;
      const double hjj = pend ? (cc - ab) / alpha : cc;
      if (havew) {
        // look-ahead: |u_{j+1}|^2 = |w_j|^2 - |b|^2 - h_jj^2 (v_0 .. v_j
        // orthonormal) completes column j provisionally.  While the
        // difference keeps 8 of the 16 digits, the estimate is good to about
        // 1e-8 relative, far finer than the stopping test needs; if that
        // column's rotation already meets the tolerance, the solve stops here
        // with it (its second Gram-Schmidt pass would change it at rounding
        // level), and the step that would finish it is not taken.
        const double un2 = ww - bb - hjj * hjj;
        if (tid == 0) {
          sm.stop = 0;
          if (un2 > 1e-8 * ww) {
            double* c = sm.col;
            for (int i = 0; i < j; i++) c[i] = sm.colp[i];
            c[j] = hjj;
            c[j + 1] = sqrt(un2);
            // row j of the column after the earlier rotations, without
            // applying them yet
            double p = c[0];
            for (int i = 0; i < j; i++) p = -sm.sn[i] * p + sm.cs[i] * c[i + 1];
            const double rr = hypot(p, c[j + 1]);
            if (rr > 0.0 && fabs(c[j + 1] / rr * sm.g[j]) <= tol) {
              rotate_column(a, sm, j);
              sm.stop = 1;
            }
          }
        }
        __syncthreads();
        if (sm.stop) {
          k = j + 1;
          conv = true;
          break;
        }
      }
      // v_j = (u_j - V a) / alpha and, while the steps go on, u_{j+1} =
      // w_j - V b - h_jj v_j, in one pass over V_0 .. V_{j-1}
      for (int e = e0 + tid; e < e1; e += 256) {
        double v = Vj[e], w = havew ? Vn[e] : 0.0;
        for (int i = 0; i < j; i++) {
          const double vi = a.V[(size_t)i * a.ld + e];
          if (!vdone) v -= sm.tot[i] * vi;
          if (havew) w -= sm.colp[i] * vi;
        }
        v *= inv;
        if (pend) Vj[e] = v;
        if (havew) Vn[e] = w - hjj * v;
      }
      if (!havew) {  // the last column is finished, v_j normalised for what follows
        __syncthreads();
        break;
      }
      if (tid == 0) {
        sm.colp[j] = hjj;
        sm.zs[j + 1] = zt > 0.0 ? 1.0 / sqrt(zt) : 0.0;
      }
      __syncthreads();
    }
    if (j >= a.maxl) conv = true;
  }
  // two exchanges per Arnoldi step (classical Gram-Schmidt applied twice)
  for (; ok && !conv && j < a.maxl; j++) {
    double* Vn = a.V + (size_t)(j + 1) * a.ld;
    apply_op(a, sm, j, Vn, s0, s1);
    // Gram-Schmidt pass 1
    local_dots(a, sm, Vn, j, e0, e1);
    ok = exchange(a, sm, ++ph, j + 1, false);
    if (!ok) break;
    for (int e = e0 + tid; e < e1; e += 256) {
      double v = Vn[e];
      for (int i = 0; i <= j; i++) v -= sm.tot[i] * a.V[(size_t)i * a.ld + e];
      Vn[e] = v;
    }
    for (int i = tid; i <= j; i += 256) sm.col[i] = sm.tot[i];
    __syncthreads();
    // pass 2 with |v'|^2 and the next direction in the same exchange
    local_dots(a, sm, Vn, j, e0, e1);
    {
      double s = 0.0;
      for (int e = e0 + tid; e < e1; e += 256) s += Vn[e] * Vn[e];
      if (j + 1 < a.maxl) publish_direction(a, sm, Vn, false, j + 1, s0, s1, r0, r1, -1, 0.0);
      block_total(s, sm, j + 1);
    }
    ok = exchange(a, sm, ++ph, j + 2, false);
    if (!ok) break;
    for (int e = e0 + tid; e < e1; e += 256) {
      double v = Vn[e];
      for (int i = 0; i <= j; i++) v -= sm.tot[i] * a.V[(size_t)i * a.ld + e];
      Vn[e] = v;
    }
    if (tid == 0) {
      double h2 = 0.0;
      for (int i = 0; i <= j; i++) {
        sm.col[i] += sm.tot[i];
        h2 += sm.tot[i] * sm.tot[i];
      }
      sm.hn2 = sm.tot[j + 1] - h2;
      sm.vn2 = sm.tot[j + 1];
    }
    __syncthreads();
    double hn2 = sm.hn2;
    if (!(hn2 > 1e-2 * sm.vn2)) {  // cancellation: measure |v| directly
      double s = 0.0;
      for (int e = e0 + tid; e < e1; e += 256) s += Vn[e] * Vn[e];
      block_total(s, sm, 0);
      ok = exchange(a, sm, ++ph, 1, false);
      if (!ok) break;
      hn2 = sm.tot[0];
    }
    const double hn = sqrt(fmax(hn2, 0.0));
    if (hn > 0.0) {
      const double inv = 1.0 / hn;
      for (int e = e0 + tid; e < e1; e += 256) Vn[e] *= inv;
    }
    if (tid == 0) {
      sm.col[j + 1] = hn;
      if (j + 1 < a.maxl) sm.zs[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
      rotate_column(a, sm, j);
    }
    __syncthreads();
    k = j + 1;
    if (fabs(sm.g[j + 1]) <= tol || !(hn > 0.0) || !isfinite(hn)) break;
  }
  // y = R^-1 g into sm.tot as y_i zs_i: in every workgroup from its own copy
  // of R when k <= kRtCap (the same k everywhere, so all take this branch),
  // else by workgroup 0 and broadcast
  if (ok && k > 0 && k <= kRtCap) {
    if (tid == 0)
      for (int i = k - 1; i >= 0; i--) {
        double s = sm.g[i];
        for (int c = i + 1; c < k; c++) s -= sm.Rt[c * (c + 1) / 2 + i] * sm.tot[c];
        sm.tot[i] = s / sm.Rt[i * (i + 1) / 2 + i];
      }
    __syncthreads();
    for (int i = tid; i < k; i += 256) sm.tot[i] *= sm.zs[i];
    __syncthreads();
  } else if (ok && k > 0) {
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
        st_coh(&a.tag[tag_slot(ph & 1, a.gmax, 0)], (a.epoch << 24) | ph);
      }
      __syncthreads();
    } else {
      ok = wait_root(a, ph);
      if (ok) {
        const __amdgpu_buffer_rsrc_t ry = make_rsrc(a.y, (unsigned)(a.maxl * 8));
        for (int i = tid; i < k; i += 256) sm.tot[i] = load_coh(ry, i);
        __syncthreads();
      }
    }
    if (ok) {
      for (int i = tid; i < k; i += 256) sm.tot[i] *= sm.zs[i];
      __syncthreads();
    }
  }
  if (ok) {
    // x = Z (y zs); the Newton update ycur += x, with the weighted sums of
    // squares of the correction, of the distance from the predictor and of
    // the controller's order q - 1 and q + 1 error estimates (combinations of
    // the new iterate and the history, weights set by the controller)
    const int em = a.ts ? a.ts->em : 0, ep = a.ts ? a.ts->ep : 0;
    double sx = 0.0, se = 0.0, sem = 0.0, sep = 0.0;
    for (int e = e0 + tid; e < e1; e += 256) {
      double xv = 0.0;
      for (int i = 0; i < k; i++) xv += sm.tot[i] * load_coh(z_rsrc(a, i), e);
      a.x[e] = xv;
      const double w = a.ewt[e];
      const double u = xv * w;
      sx += u * u;
      double yv = 1.0 * a.m[e];
      yv += 1.0 * xv;
      a.m[e] = yv;
      double dy = 1.0 * yv;
      dy += -1.0 * a.yp[e];
      const double v = dy * w;
      se += v * v;
      if (em > 0) {
        double t = a.ts->Em[0] * yv;
        for (int i = 1; i <= em; i++) t += a.ts->Em[i] * a.slots[a.ts->sh[i - 1]][e];
        t *= w;
        sem += t * t;
      }
      if (ep > 0) {
        double t = a.ts->Ep[0] * yv;
        for (int i = 1; i <= ep; i++) t += a.ts->Ep[i] * a.slots[a.ts->sh[i - 1]][e];
        t *= w;
        sep += t * t;
      }
    }
    block_sum2(sx, se, sm.red);
    block_sum2(sem, sep, sm.red);
    if (tid == 0) {
      sm.loc[0] = sx;
      sm.loc[1] = se;
      sm.loc[2] = sem;
      sm.loc[3] = sep;
    }
    ok = exchange(a, sm, ++ph, 4, true);
  }
  if (blockIdx.x == 0 && tid == 0) {
    const double xw2 = ok ? sm.tot[0] : NAN, errsq = ok ? sm.tot[1] : NAN;
    const double errsq_m = ok ? sm.tot[2] : NAN, errsq_p = ok ? sm.tot[3] : NAN;
    a.status[0] = (double)k;
    a.status[1] = xw2;
    a.status[2] = ok ? 0.0 : 1.0;
    a.status[3] = errsq;
    a.status[4] = errsq_m;
    a.status[5] = errsq_p;
    if (a.ts)
      run_controller(a.ts, (double)k, xw2, ok ? 0.0 : 1.0, errsq, errsq_m, errsq_p,
                     (int)a.ld);
  }
}

template <typename TV>
const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_gmres<TV>);
}

template <typename TV>
hipError_t launch_typed(const DeviceOps& d, const TV* K, GmresWork& w, int grid, int spw,
                        int maxl, const NewtonSystem& s, hipStream_t st) {
  GmresArgs<TV> a{d.n,     d.nslices, spw,      maxl,     w.gmax,  w.pstride, s.sweeps,
                  s.onex_steps,
                  d.sell_off, d.sell_col, K,  d.Vinv,   d.coefs,  d.inst,  s.ycur,    s.Ht,
                  s.ewt,   s.kappa,   s.Minv,   s.gamma,  s.tol,   s.avec,    s.f,
                  s.yp,    s.x,       w.V,      w.Z,      3 * (size_t)d.n,    w.part,
                  w.tag,   w.R,       w.y,      w.status, w.err,   ++w.epoch,
                  s.gamma_p, s.halt, s.slots, s.ts};
  void* args[] = {&a};
  if (w.cooperative)
    return hipLaunchCooperativeKernel(kernel_ptr<TV>(), dim3(grid), dim3(256), args, 0, st);
  return hipLaunchKernel(kernel_ptr<TV>(), dim3(grid), dim3(256), args, 0, st);
}

}  // namespace

bool gmres_grid(int nslices, int maxl, int device, int gmax, int min_spw, int* grid,
                int* spw, int* resident) {
  if (nslices <= 0 || maxl < 1 || maxl > kGmresMaxl) return false;
  int cus = 0;
  // the FP32 and FP64 variants differ only in one load; take the smaller
  const int pf = resident_per_cu(kernel_ptr<float>(), 256, device);
  const int pd = resident_per_cu(kernel_ptr<double>(), 256, device);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return false;
  int cap = (pf < pd ? pf : pd) * cus;
  *resident = cap;
  // one slot in sixteen is left free, so that a few workgroups of another
  // queue's kernel on the device cannot keep part of the grid from starting
  cap -= cap / 16;
  if (cap > kMaxGrid) cap = kMaxGrid;
  if (cap > gmax) cap = gmax;
  if (cap < 1) return false;
  int s = (nslices + cap - 1) / cap;
  if (s < min_spw) s = min_spw;
  *spw = s;
  *grid = (nslices + s - 1) / s;
  return true;
}

hipError_t launch_newton_gmres(const DeviceOps& d, bool mixed, GmresWork& w, int grid, int spw,
                               int maxl, const NewtonSystem& s, hipStream_t st) {
  if (grid < 1 || grid > w.gmax || grid > w.resident || maxl < 1 || maxl > kGmresMaxl ||
      s.onex_steps < 0 || s.onex_steps > kGmresOneExMax ||
      w.pstride < gmres_pstride(maxl) ||
      (long long)grid * spw < d.nslices || (long long)(grid - 1) * spw >= d.nslices ||
      3LL * d.n * 8 >= (1LL << 32) || (long long)w.gmax * w.pstride * 8 >= (1LL << 32))
    return hipErrorInvalidConfiguration;
  if (mixed) return launch_typed<float>(d, d.Kf, w, grid, spw, maxl, s, st);
  return launch_typed<double>(d, d.K, w, grid, spw, maxl, s, st);
}

}  // namespace nsim
