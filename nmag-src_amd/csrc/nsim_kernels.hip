// Sparse and pointwise kernels of the LLG right-hand-side path (gfx950).
//
// One field evaluation (the reference's update_dmdt program,
// interface/nmag/nmag_lam.py:734-832) is five launches on one stream:
//   1. k_exch_div   H_exch = negJ K m / V and rho = Ms D.m     (this file)
//   2. k_cg         Neumann solve K phi1 = rho                  (nsim_cg.hip)
//   3. k_bem_gemv   phi2|_b = B phi1|_b                         (nsim_bem.hip)
//   4. k_cg         Dirichlet solve for phi2 in the interior    (nsim_cg.hip)
//   5. k_demag_llg  H_demag = G phi, H_total, LLG dm/dt         (this file)
// Operators use a SELL format with slice height 64 (one wavefront): the rows of
// a wavefront share one loop bound and every slot load is coalesced.  Values
// are stored in FP64, or in FP32 at mixed precision for the exchange stiffness
// and the gradient (the divergence stays FP64); accumulation is always FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_anis.h"

namespace nsim {

namespace {

using namespace dev;

// Coefficients of the magnet that owns node `row`: the table has one entry per
// magnet of a batch; `inst` is null for a single magnet.
__device__ inline const Coeffs& coef_of(const Coeffs* __restrict__ ct,
                                        const int* __restrict__ inst, int row) {
  return ct[inst ? inst[row] : 0];
}

// Work split of the sparse kernels below: a workgroup of four wavefronts owns
// one SELL slice (64 rows, one per lane); wavefront t takes the slots
// k = t, t + 4, ... of every row, so even a 20k-node mesh keeps 1200+
// wavefronts in flight, and partial sums meet in LDS in a fixed order.
constexpr int kSplit = 4;

// Sums the kSplit partials of Q quantities into wavefront 0's registers.
template <int Q>
__device__ inline void gather_partials(double (&v)[Q], double (*sh)[64], int t, int lane) {
#pragma unroll
  for (int q = 0; q < Q; q++) sh[t * Q + q][lane] = v[q];
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
      double s = sh[q][lane];
      for (int u = 1; u < kSplit; u++) s += sh[u * Q + q][lane];
      v[q] = s;
    }
  }
}

// H_exch = negJ/V_i sum_j K_ij m_j (set_H_exch, nmag_lam.py:734-739) and
// rho_i = Ms sum_j D_ij . m_j (op_div_m then SCALE -1, nmag_lam.py:742-743).
// The magnets of a batch share no node, so row and columns share one Ms.
// The stiffness rows sum to zero (K 1 = 0 for P1 elements), so the exchange
// sum is taken as sum_j K_ij (m_j - m_i): identical in exact arithmetic, and
// with FP32-stored K a uniform m still gives exactly zero field (the rounded
// rows no longer sum to zero, the differences do not care).  D is always FP64:
// rho is the load of the singular Neumann problem, whose solvability
// (sum_i rho_i = 0) must hold to FP64 roundoff.
template <typename TV>
__global__ __launch_bounds__(256) void k_exch_div(int n, const int* __restrict__ off,
                                                  const int* __restrict__ col,
                                                  const TV* __restrict__ K,
                                                  const double* __restrict__ D,
                                                  const double* __restrict__ Vinv,
                                                  const Coeffs* __restrict__ ct,
                                                  const int* __restrict__ inst,
                                                  const double* __restrict__ m,
                                                  double* __restrict__ Hexch,
                                                  double* __restrict__ rho, int want_rho,
                                                  const int* __restrict__ halt,
                                                  const int* __restrict__ rho_pos,
                                                  double* __restrict__ rho_perm) {
  if (halt && *halt) return;  // queued by a finished integrator run
  __shared__ double sh[4 * kSplit][64];
  const int s = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = s * 64 + lane;
  const bool live = row < n;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (live) {
    const double mix = m[3 * row], miy = m[3 * row + 1], miz = m[3 * row + 2];
    for (int k = t; k < w; k += kSplit) {
      const int e = b + (k << 6) + lane;
      const int j = col[e];
      const double kv = (double)K[e];
      const double mx = m[3 * j], my = m[3 * j + 1], mz = m[3 * j + 2];
      v[0] += kv * (mx - mix);
      v[1] += kv * (my - miy);
      v[2] += kv * (mz - miz);
      if (want_rho) v[3] += D[3 * e] * mx + D[3 * e + 1] * my + D[3 * e + 2] * mz;
    }
  }
  gather_partials<4>(v, sh, t, lane);
  if (t != 0 || !live) return;
  const Coeffs& c = coef_of(ct, inst, row);
  const double f = c.negJ * Vinv[row];
  Hexch[3 * row] = f * v[0];
  Hexch[3 * row + 1] = f * v[1];
  Hexch[3 * row + 2] = f * v[2];
  if (want_rho) {
    rho[row] = c.Ms * v[3];
    // the H-matrix route reads rho in its cluster order
    if (rho_perm) rho_perm[rho_pos[row]] = c.Ms * v[3];
  }
}

// H_demag_i = sum_j G_ij phi_j (op_grad_phi + CFBOX, nmag_lam.py:755-759),
// taken as sum_j G_ij (phi_j - phi_i): the rows of G sum to zero (the
// barycentric gradients of a tetrahedron sum to zero), so this is the same
// field, and with FP32-stored G it stays blind to the arbitrary constant of
// the Neumann potential;
// H_anis from the uniaxial tangential gradient (src/nsim_anisotropy.ml:167-234)
// and the polynomial anisotropy of the magnet (nsim_anis.h);
// H_total (nmag_lam.py:377-411); dm/dt (nmag_lam.py:417-430) times pin.
template <typename TV>
__global__ __launch_bounds__(256) void k_demag_llg(int n, const int* __restrict__ off,
                                                   const int* __restrict__ col,
                                                   const TV* __restrict__ G,
                                                   const Coeffs* __restrict__ ct,
                                                   const int* __restrict__ inst, FieldArgs f,
                                                   FieldBufs bf, double* __restrict__ dmdt,
                                                   int do_demag, const int* __restrict__ halt) {
  if (halt && *halt) return;  // queued by a finished integrator run
  __shared__ double sh[3 * kSplit][64];
  const int s = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = s * 64 + lane;
  const bool live = row < n;
  double hd[3] = {0.0, 0.0, 0.0};
  if (do_demag) {
    const int b = off[s], w = (off[s + 1] - b) >> 6;
    if (live) {
      const double phi_i = bf.phi[row];
      for (int k = t; k < w; k += kSplit) {
        const int e = b + (k << 6) + lane;
        const double ph = bf.phi[col[e]] - phi_i;
        hd[0] += (double)G[3 * e] * ph;
        hd[1] += (double)G[3 * e + 1] * ph;
        hd[2] += (double)G[3 * e + 2] * ph;
      }
    }
    gather_partials<3>(hd, sh, t, lane);
  }
  if (t != 0 || !live) return;
  const Coeffs& c = coef_of(ct, inst, row);
  if (do_demag) {
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
  }
  if (f.poly) poly_field(f, inst ? inst[row] : 0, m, 1.0 / (c.mu0 * c.Ms), ha);
  if (f.k1 || f.poly) {
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
                                                const double* __restrict__ Vinv,
                                                const Coeffs* __restrict__ ct,
                                                const int* __restrict__ inst,
                                                const double* __restrict__ m,
                                                const double* __restrict__ Ht,
                                                const double* __restrict__ v,
                                                double* __restrict__ Jv) {
  __shared__ double sh[3 * kSplit][64];
  const int s = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = s * 64 + lane;
  const bool live = row < n;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  double hv[3] = {0.0, 0.0, 0.0};
  if (live)
    for (int k = t; k < w; k += kSplit) {
      const int e = b + (k << 6) + lane;
      const int j = col[e];
      const double kv = (double)K[e];
      hv[0] += kv * v[3 * j];
      hv[1] += kv * v[3 * j + 1];
      hv[2] += kv * v[3 * j + 2];
    }
  gather_partials<3>(hv, sh, t, lane);
  if (t != 0 || !live) return;
  const Coeffs& c = coef_of(ct, inst, row);
  const double fct = c.negJ * Vinv[row];
  const double h[3] = {fct * hv[0], fct * hv[1], fct * hv[2]};
  const double mi[3] = {m[3 * row], m[3 * row + 1], m[3 * row + 2]};
  const double vi[3] = {v[3 * row], v[3 * row + 1], v[3 * row + 2]};
  const double H[3] = {Ht[3 * row], Ht[3 * row + 1], Ht[3 * row + 2]};
  double out[3];
  llg_jv_row(c.c1, c.c2, h, mi, vi, H, out);
  for (int d = 0; d < 3; d++) Jv[3 * row + d] = out[d];
}

}  // namespace

hipError_t launch_exch_div(const DeviceOps& d, bool mixed, const double* m, double* Hexch,
                           double* rho, bool want_rho, hipStream_t st) {
  if (mixed)
    k_exch_div<float><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Kf,
                                                         d.D, d.Vinv, d.coefs, d.inst, m,
                                                         Hexch, rho, want_rho ? 1 : 0, d.halt,
                                                         d.hm.pos, d.hm.xp);
  else
    k_exch_div<double><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.K,
                                                          d.D, d.Vinv, d.coefs, d.inst, m,
                                                          Hexch, rho, want_rho ? 1 : 0, d.halt,
                                                          d.hm.pos, d.hm.xp);
  return hipGetLastError();
}

hipError_t launch_demag_llg(const DeviceOps& d, bool mixed, const FieldArgs& f,
                            const FieldBufs& b, double* dmdt, bool do_demag, hipStream_t st) {
  if (mixed)
    k_demag_llg<float><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Gf,
                                                          d.coefs, d.inst, f, b, dmdt,
                                                          do_demag ? 1 : 0, d.halt);
  else
    k_demag_llg<double><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.G,
                                                           d.coefs, d.inst, f, b, dmdt,
                                                           do_demag ? 1 : 0, d.halt);
  return hipGetLastError();
}

hipError_t launch_jtimes(const DeviceOps& d, bool mixed, const double* m, const double* Htotal,
                         const double* v, double* Jv, hipStream_t st) {
  if (mixed)
    k_jtimes<float><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Kf,
                                                       d.Vinv, d.coefs, d.inst, m, Htotal, v,
                                                       Jv);
  else
    k_jtimes<double><<<d.nslices, 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.K,
                                                        d.Vinv, d.coefs, d.inst, m, Htotal, v,
                                                        Jv);
  return hipGetLastError();
}

}  // namespace nsim
