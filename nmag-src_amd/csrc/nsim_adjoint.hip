// Transposed Jacobian of the LLG right-hand side (vector-Jacobian products),
// the building block of the discrete adjoint.  The reference has no adjoint;
// these kernels transpose the forward operators of this library one by one:
//
//   f = pin [c1 m x H + c2 m x (m x H) + c3 (1 - |m|^2) m]         (site-local)
//   H = H_ext + negJ V^-1 A m + H_anis(m) + G (I + E B R) K^+ Ms D m
//
// with A the (difference-form) stiffness, D the divergence load, K^+ the
// Neumann solve, R the restriction to the boundary, B the BEM matrix, E the
// harmonic (Dirichlet) extension and G the lumped gradient.  The adjoint of
// the demag chain uses two identities of the P1 operators, G^T g = -D (g/V)
// and D^T w = -V (G w) (both from G_ij = -D_ji / V_i), so no transposed copy
// of D or G is stored; K is symmetric; only B needs a transposed product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "nsim_adjoint.h"
#include "nsim_anis.h"
#include "nsim_dev_common.h"
#include "nsim_device.h"

namespace nsim {

namespace {

using namespace dev;

__device__ inline const Coeffs& coef_at(const Coeffs* __restrict__ ct,
                                        const int* __restrict__ inst, int row) {
  return ct[inst ? inst[row] : 0];
}

inline int blocks_for(int n) { return (n + 255) / 256; }

__global__ __launch_bounds__(256) void k_adj_local(int n, const Coeffs* __restrict__ ct,
                                                   const int* __restrict__ inst, FieldArgs f,
                                                   FieldBufs bf, int have_demag,
                                                   const double* __restrict__ Vinv,
                                                   const double* __restrict__ lam,
                                                   double* __restrict__ g,
                                                   double* __restrict__ x,
                                                   double* __restrict__ a, AdjAccum acc) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const Coeffs& c = coef_at(ct, inst, row);
  const double m[3] = {f.m[3 * row], f.m[3 * row + 1], f.m[3 * row + 2]};
  const double H[3] = {bf.Htotal[3 * row], bf.Htotal[3 * row + 1], bf.Htotal[3 * row + 2]};
  const double pin = f.pin ? f.pin[row] : 1.0;
  const double l[3] = {pin * lam[3 * row], pin * lam[3 * row + 1], pin * lam[3 * row + 2]};
  const double mm = dot3(m, m), lm = dot3(l, m), mH = dot3(m, H), lH = dot3(l, H);
  double lxm[3], Hxl[3];
  cross(l, m, lxm);
  cross(H, l, Hxl);
  double gv[3], av[3];
  for (int d = 0; d < 3; d++) {
    // (df/dH)^T l: c1 l x m + c2 ((l.m) m - (m.m) l)
    gv[d] = c.c1 * lxm[d] + c.c2 * (lm * m[d] - mm * l[d]);
    // (df/dm|_H)^T l: c1 H x l + c2 ((m.H) l + (l.m) H - 2 (l.H) m)
    //                 + c3 ((1 - m.m) l - 2 (l.m) m)
    av[d] = c.c1 * Hxl[d] + c.c2 * (mH * l[d] + lm * H[d] - 2.0 * lH * m[d]) +
            c.c3 * ((1.0 - mm) * l[d] - 2.0 * lm * m[d]);
  }
  double ha[3] = {0.0, 0.0, 0.0};
  if (f.k1) {
    // H_anis = q(u) (a - u m), u = a.m, q = (2 K1 u + 4 K2 u^3) / (mu0 Ms)
    const double ax[3] = {f.axis[3 * row], f.axis[3 * row + 1], f.axis[3 * row + 2]};
    const double u = dot3(ax, m), K1 = f.k1[row], K2 = f.k2 ? f.k2[row] : 0.0;
    const double inv = 1.0 / (c.mu0 * c.Ms);
    const double q = (2.0 * K1 * u + 4.0 * K2 * u * u * u) * inv;
    const double qp = (2.0 * K1 + 12.0 * K2 * u * u) * inv;
    const double gm = dot3(gv, m), gt = dot3(gv, ax) - u * gm;  // g.(a - u m)
    for (int d = 0; d < 3; d++) av[d] += (qp * gt - q * gm) * ax[d] - q * u * gv[d];
    if (acc.k1bar) acc.k1bar[row] += 2.0 * u * inv * gt;
  }
  if (f.poly) poly_field_vjp(f, inst ? inst[row] : 0, m, 1.0 / (c.mu0 * c.Ms), gv, av);
  if (f.k1 || f.poly)
    for (int d = 0; d < 3; d++) ha[d] = bf.Hanis[3 * row + d];
  if (acc.msbar) {
    // H_exch and H_anis scale as 1/Ms, H_demag as Ms
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
      const double hd = have_demag ? bf.Hdemag[3 * row + d] : 0.0;
      s += gv[d] * (hd - bf.Hexch[3 * row + d] - ha[d]);
    }
    acc.msbar[row] += s / c.Ms;
  }
  if (acc.hbar) acc.hbar[row] += dot3(gv, acc.dir);
  const double vi = Vinv[row];
  for (int d = 0; d < 3; d++) {
    g[3 * row + d] = gv[d];
    x[3 * row + d] = gv[d] * vi;
    a[3 * row + d] = av[d];
  }
}

// a_j += negJ sum_i K_ji (x_i - x_j);  psi_j = -sum_i D_ji . x_i
template <typename TV>
__global__ __launch_bounds__(256) void k_adj_exch_div(int n, const int* __restrict__ off,
                                                      const int* __restrict__ col,
                                                      const TV* __restrict__ K,
                                                      const double* __restrict__ D,
                                                      const Coeffs* __restrict__ ct,
                                                      const int* __restrict__ inst,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ a,
                                                      double* __restrict__ psi) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  const double xj[3] = {x[3 * row], x[3 * row + 1], x[3 * row + 2]};
  double e0 = 0.0, e1 = 0.0, e2 = 0.0, r = 0.0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    const int i = col[e];
    const double kv = (double)K[e];
    const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    e0 += kv * (x0 - xj[0]);
    e1 += kv * (x1 - xj[1]);
    e2 += kv * (x2 - xj[2]);
    r += D[3 * e] * x0 + D[3 * e + 1] * x1 + D[3 * e + 2] * x2;
  }
  const double nj = coef_at(ct, inst, row).negJ;
  a[3 * row] += nj * e0;
  a[3 * row + 1] += nj * e1;
  a[3 * row + 2] += nj * e2;
  psi[row] = -r;
}

__global__ __launch_bounds__(256) void k_adj_boundary(int nb, const int* __restrict__ bnode,
                                                      const int* __restrict__ off,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ K,
                                                      const double* __restrict__ psi,
                                                      const double* __restrict__ z,
                                                      double* __restrict__ e) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  if (l >= nb) return;
  const int i = bnode[l];
  const int s = i >> 6, lane = i & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  double acc = 0.0;
  for (int k = 0; k < w; k++) {
    const int q = b + (k << 6) + lane;
    acc += K[q] * z[col[q]];
  }
  e[l] = psi[i] - acc;
}

template <typename TB>
struct TVec;
// load(row, j, ld, o): entries j .. j+V-1 of a row of length ld
template <>
struct TVec<double> {
  static constexpr int V = 2;
  __device__ static void load(const double* row, int j, int, double* o) {
    const double2 v = *reinterpret_cast<const double2*>(row + j);
    o[0] = v.x;
    o[1] = v.y;
  }
};
template <>
struct TVec<float> {
  static constexpr int V = 4;
  __device__ static void load(const float* row, int j, int, double* o) {
    const float4 v = *reinterpret_cast<const float4*>(row + j);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
};
template <>
struct TVec<F24> {
  static constexpr int V = 8;
  __device__ static void load(const F24* row, int j, int ld, double* o) {
    f24_load8(row, j, ld, o);
  }
};
template <>
struct TVec<uint16_t> {
  static constexpr int V = 8;
  __device__ static void load(const uint16_t* row, int j, int, double* o) {
    const uint4 v = *reinterpret_cast<const uint4*>(row + j);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (int k = 0; k < 4; k++) {
      o[2 * k] = __uint_as_float(w[k] << 16);
      o[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
    }
  }
};

// part[chunk][j] = sum over the chunk's rows i of B_ij e_i; each thread owns
// V consecutive columns (one 16-byte load per row), rows of a chunk are
// streamed in order (eight rows' loads in flight), so a wavefront reads
// 1 KiB of a row per load.
template <typename TB>
__global__ __launch_bounds__(256) void k_bem_gemv_t(int nb, int ldb, int rows_per_chunk,
                                                    const TB* __restrict__ B,
                                                    const double* __restrict__ e,
                                                    double* __restrict__ part) {
  constexpr int V = TVec<TB>::V;
  const int j0 = (blockIdx.x * 256 + threadIdx.x) * V;
  if (j0 >= ldb) return;
  const int r0 = blockIdx.y * rows_per_chunk;
  const int r1 = min(r0 + rows_per_chunk, nb);
  double acc[V];
  for (int v = 0; v < V; v++) acc[v] = 0.0;
#pragma unroll 8
  for (int i = r0; i < r1; i++) {
    double bv[V];
    TVec<TB>::load(B + (size_t)i * ldb, j0, ldb, bv);
    const double ei = e[i];
    for (int v = 0; v < V; v++) acc[v] += bv[v] * ei;
  }
  for (int v = 0; v < V; v++) part[(size_t)blockIdx.y * ldb + j0 + v] = acc[v];
}

// chi[bnode[j]] (+)= sum over the chunks c of part[c][j], 64 columns per
// workgroup: wavefront w sums the chunks w, w + 16, ... in increasing order
// (eight loads in flight), then the 16 wavefront sums meet in LDS in a fixed
// order, so the result does not depend on timing.
__global__ __launch_bounds__(1024) void k_bem_t_reduce(int nb, int ldb, int chunks,
                                                       const double* __restrict__ part,
                                                       const int* __restrict__ bnode,
                                                       double* __restrict__ chi, int accumulate) {
  __shared__ double sh[16][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int j = blockIdx.x * 64 + l;
  double s = 0.0;
  if (j < nb) {
#pragma unroll 8
    for (int c = w; c < chunks; c += 16) s += part[(size_t)c * ldb + j];
  }
  sh[w][l] = s;
  __syncthreads();
  if (w == 0 && j < nb) {
    double t = 0.0;
    for (int q = 0; q < 16; q++) t += sh[q][l];
    if (accumulate) t += chi[bnode[j]];
    chi[bnode[j]] = t;
  }
}

__global__ __launch_bounds__(1024) void k_mean_project(int n, double* __restrict__ x) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += x[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < 16; w++) t += sh[w];
  const double mean = t / n;
  for (int i = threadIdx.x; i < n; i += 1024) x[i] -= mean;
}

// a_i -= Ms V_i sum_j G_ij (phi_j - phi_i)
template <typename TV>
__global__ __launch_bounds__(256) void k_adj_grad(int n, const int* __restrict__ off,
                                                  const int* __restrict__ col,
                                                  const TV* __restrict__ G,
                                                  const Coeffs* __restrict__ ct,
                                                  const int* __restrict__ inst,
                                                  const double* __restrict__ Vinv,
                                                  const double* __restrict__ phi,
                                                  double* __restrict__ a) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  const double pi = phi[row];
  double h0 = 0.0, h1 = 0.0, h2 = 0.0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    const double ph = phi[col[e]] - pi;
    h0 += (double)G[3 * e] * ph;
    h1 += (double)G[3 * e + 1] * ph;
    h2 += (double)G[3 * e + 2] * ph;
  }
  const double f = coef_at(ct, inst, row).Ms / Vinv[row];
  a[3 * row] -= f * h0;
  a[3 * row + 1] -= f * h1;
  a[3 * row + 2] -= f * h2;
}

template <typename TB>
hipError_t gemv_t_typed(const DenseOp& op, const double* e, double* part, double* chi,
                        bool accumulate, hipStream_t st) {
  constexpr int V = TVec<TB>::V;
  const int ldb = op.h_ld[0], nb = op.h_n[0];
  const int chunks = dense_t_chunks(op);
  const int rpc = (nb + chunks - 1) / chunks;
  const int cols = (ldb / V + 255) / 256;
  k_bem_gemv_t<TB><<<dim3(cols, chunks), 256, 0, st>>>(nb, ldb, rpc, (const TB*)op.A, e, part);
  k_bem_t_reduce<<<(nb + 63) / 64, 1024, 0, st>>>(nb, ldb, chunks, part, op.map, chi,
                                                   accumulate ? 1 : 0);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_adj_local(const DeviceOps& d, const FieldArgs& f, const FieldBufs& b,
                            const double* lam, double* g, double* x, double* a,
                            const AdjAccum& acc, hipStream_t st) {
  k_adj_local<<<blocks_for(d.n), 256, 0, st>>>(d.n, d.coefs, d.inst, f, b,
                                               (b.Hdemag && d.nb > 0) ? 1 : 0, d.Vinv, lam, g,
                                               x, a, acc);
  return hipGetLastError();
}

hipError_t launch_adj_exch_div(const DeviceOps& d, bool mixed, const double* x, double* a,
                               double* psi, hipStream_t st) {
  if (mixed)
    k_adj_exch_div<float><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Kf,
                                                             d.D, d.coefs, d.inst, x, a, psi);
  else
    k_adj_exch_div<double><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.K,
                                                              d.D, d.coefs, d.inst, x, a, psi);
  return hipGetLastError();
}

hipError_t launch_adj_boundary(const DeviceOps& d, const double* psi, const double* z,
                               double* e, hipStream_t st) {
  if (d.nb == 0) return hipSuccess;
  k_adj_boundary<<<blocks_for(d.nb), 256, 0, st>>>(d.nb, d.bnode, d.sell_off, d.sell_col, d.K,
                                                   psi, z, e);
  return hipGetLastError();
}

int dense_t_chunks(const DenseOp& op) {
  if (op.h_n.empty() || op.h_n[0] == 0) return 1;
  // about 512 workgroups in all (each thread then streams a few tens of rows
  // with its loads in flight together), each chunk at least 8 rows
  const int cols = (op.h_ld[0] / 2 + 255) / 256;
  int chunks = std::max(1, 512 / std::max(1, cols));
  chunks = std::min(chunks, std::max(1, op.h_n[0] / 8));
  return chunks;
}

hipError_t launch_dense_gemv_t(const DenseOp& op, const double* e, double* part, double* chi,
                               bool accumulate, hipStream_t st) {
  if (!op.A || op.h_n.empty() || op.h_n[0] == 0) return hipSuccess;
  if (op.h_n.size() != 1) return hipErrorInvalidConfiguration;
  if (op.type == kBemF64) return gemv_t_typed<double>(op, e, part, chi, accumulate, st);
  if (op.type == kBemF32) return gemv_t_typed<float>(op, e, part, chi, accumulate, st);
  if (op.type == kBemF24) return gemv_t_typed<F24>(op, e, part, chi, accumulate, st);
  return gemv_t_typed<uint16_t>(op, e, part, chi, accumulate, st);
}

hipError_t launch_mean_project(int n, double* x, hipStream_t st) {
  k_mean_project<<<1, 1024, 0, st>>>(n, x);
  return hipGetLastError();
}

hipError_t launch_adj_grad(const DeviceOps& d, bool mixed, const double* phi, double* a,
                           hipStream_t st) {
  if (mixed)
    k_adj_grad<float><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.Gf,
                                                         d.coefs, d.inst, d.Vinv, phi, a);
  else
    k_adj_grad<double><<<blocks_for(d.n), 256, 0, st>>>(d.n, d.sell_off, d.sell_col, d.G,
                                                          d.coefs, d.inst, d.Vinv, phi, a);
  return hipGetLastError();
}

}  // namespace nsim
