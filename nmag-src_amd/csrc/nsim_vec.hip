// N_Vector operations of the device-resident time integrator, and the 3x3
// block-Jacobi preconditioner of I - gamma J.
//
// Reductions write one partial per workgroup (fixed grid, fixed order) and the
// host adds them, so every decision of the step controller is reproducible.
// CVODE's parallel N_Vector (N_VLinearSum, N_VWrmsNorm, N_VDotProd ...) is what
// these replace for the reference's timestepper (src/sundials_sp_stubs.c).
#include <hip/hip_runtime.h>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_vec.h"

namespace nsim {

namespace {

using namespace dev;

__device__ inline double block_sum(double v, double* sh) {
  double dummy = 0.0;
  block_sum2(v, dummy, sh);
  return v;
}

__global__ __launch_bounds__(256) void k_lincomb(int n, double c0, const double* x0, double c1,
                                                 const double* x1, double c2, const double* x2,
                                                 double c3, const double* x3, double* y) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double v = c0 * x0[i];
    if (x1) v += c1 * x1[i];
    if (x2) v += c2 * x2[i];
    if (x3) v += c3 * x3[i];
    y[i] = v;
  }
}

// y_k = c_k x for four outputs in one pass.
__global__ __launch_bounds__(256) void k_scale4(int n, const double* x, double c0, double* y0,
                                                double c1, double* y1, double c2, double* y2,
                                                double c3, double* y3) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double v = x[i];
    y0[i] = c0 * v;
    y1[i] = c1 * v;
    y2[i] = c2 * v;
    y3[i] = c3 * v;
  }
}

__global__ __launch_bounds__(256) void k_copy4(int n, const double* s0, const double* s1,
                                               const double* s2, const double* s3,
                                               double* dst) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    if (s0) dst[i] = s0[i];
    if (s1) dst[n + i] = s1[i];
    if (s2) dst[2 * (size_t)n + i] = s2[i];
    if (s3) dst[3 * (size_t)n + i] = s3[i];
  }
}

// y1 += a1 x and y2 += a2 x in one pass.
__global__ __launch_bounds__(256) void k_update2(int n, const double* x, double a1, double* y1,
                                                 double a2, double* y2) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double v = x[i];
    y1[i] += a1 * v;
    y2[i] += a2 * v;
  }
}

// y = x * w (inverse: y = x / w)
__global__ __launch_bounds__(256) void k_vmul(int n, const double* x, const double* w,
                                              int inverse, double* y) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    y[i] = inverse ? x[i] / w[i] : x[i] * w[i];
}

__global__ __launch_bounds__(256) void k_ewt(int n, const double* y, double rtol, double atol,
                                             double* ewt) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    ewt[i] = 1.0 / (rtol * fabs(y[i]) + atol);
}

// part[b] = sum over this block's elements of (x_i * w_i)^2 (w may be null: 1)
__global__ __launch_bounds__(256) void k_wsq(int n, const double* x, const double* w,
                                             double* part) {
  __shared__ double sh[8];
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double v = w ? x[i] * w[i] : x[i];
    s += v * v;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// part[j * G + b] = sum_i V_j[i] * y[i] for j < k (k <= 8), V_j = V + j*ld
__global__ __launch_bounds__(256) void k_multidot(int n, int k, const double* V, size_t ld,
                                                  const double* y, double* part) {
  __shared__ double sh[8];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double yi = y[i];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j < k) s[j] += V[j * ld + i] * yi;
  }
  for (int j = 0; j < k; j++) {
    const double t = block_sum(s[j], sh);
    if (threadIdx.x == 0) part[j * gridDim.x + blockIdx.x] = t;
  }
}

// y += sum_j c[j] V_j  (c in device memory, k <= 64)
__global__ __launch_bounds__(256) void k_axpy_multi(int n, int k, const double* V, size_t ld,
                                                    const double* c, double* y) {
  __shared__ double cs[64];
  if (threadIdx.x < k) cs[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double v = y[i];
    for (int j = 0; j < k; j++) v += cs[j] * V[j * ld + i];
    y[i] = v;
  }
}

// part[b] = max over this block's nodes of |a_i - b_i| (3-vectors)
__global__ __launch_bounds__(256) void k_maxdist(int nn, const double* a, const double* b,
                                                 double* part) {
  __shared__ double sh[4];
  double mx = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256) {
    const double dx = a[3 * i] - b[3 * i], dy = a[3 * i + 1] - b[3 * i + 1],
                 dz = a[3 * i + 2] - b[3 * i + 2];
    mx = fmax(mx, sqrt(dx * dx + dy * dy + dz * dz));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// Per node: M_i = I - gamma J_ii with J_ii the 3x3 diagonal block of the
// jacobi_ts1 linearisation (H frozen, dH_i/dm_i = kappa_i I, kappa_i =
// negJ K_ii / V_i); stores M_i^-1 (row-major, 9 doubles).
__global__ __launch_bounds__(256) void k_bjac_setup(int nn, const Coeffs* __restrict__ ct,
                                                    const int* __restrict__ inst, double gamma,
                                                    const double* m, const double* H,
                                                    const double* kappa, double* Minv) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const Coeffs& c = ct[inst ? inst[i] : 0];
  const double mi[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
  const double Hi[3] = {H[3 * i], H[3 * i + 1], H[3 * i + 2]};
  bjac_block(c.c1, c.c2, gamma, kappa[i], mi, Hi, Minv + 9 * i);
}

__global__ __launch_bounds__(256) void k_bjac_solve(int nn, const double* Minv, const double* r,
                                                    double* z) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const double* o = Minv + 9 * i;
  const double a = r[3 * i], b = r[3 * i + 1], d = r[3 * i + 2];
  z[3 * i] = o[0] * a + o[1] * b + o[2] * d;
  z[3 * i + 1] = o[3] * a + o[4] * b + o[5] * d;
  z[3 * i + 2] = o[6] * a + o[7] * b + o[8] * d;
}

// z = Minv^T r per node (the transposed block preconditioner of the adjoint)
__global__ __launch_bounds__(256) void k_bjac_solve_t(int nn, const double* Minv,
                                                      const double* r, double* z) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const double* o = Minv + 9 * i;
  const double a = r[3 * i], b = r[3 * i + 1], d = r[3 * i + 2];
  z[3 * i] = o[0] * a + o[3] * b + o[6] * d;
  z[3 * i + 1] = o[1] * a + o[4] * b + o[7] * d;
  z[3 * i + 2] = o[2] * a + o[5] * b + o[8] * d;
}

inline int grid_for(int n) {
  const int g = (n + 255) / 256;
  return g < kReduceBlocks ? g : kReduceBlocks;
}

// out[j] = sum_b part[j * np + b] in a fixed order, neg[j] = -out[j] (neg may
// be null): the device-side end of a reduction whose result feeds further
// device work without a host round trip.
__global__ __launch_bounds__(256) void k_reduce_parts(int np, int k, const double* part,
                                                      double* out, double* neg) {
  for (int j = threadIdx.x; j < k; j += 256) {
    double s = 0.0;
    for (int b = 0; b < np; b++) s += part[(size_t)j * np + b];
    out[j] = s;
    if (neg) neg[j] = -s;
  }
}

}  // namespace

int vec_reduce_grid(int n) { return grid_for(n); }

hipError_t vec_reduce_parts(hipStream_t st, int np, int k, const double* part, double* out,
                            double* neg) {
  if (k <= 0) return hipSuccess;
  k_reduce_parts<<<1, 256, 0, st>>>(np, k, part, out, neg);
  return hipGetLastError();
}

hipError_t vec_lincomb(hipStream_t st, int n, double c0, const double* x0, double c1,
                       const double* x1, double c2, const double* x2, double c3,
                       const double* x3, double* y) {
  k_lincomb<<<grid_for(n), 256, 0, st>>>(n, c0, x0, c1, x1, c2, x2, c3, x3, y);
  return hipGetLastError();
}

hipError_t vec_scale4(hipStream_t st, int n, const double* x, double c0, double* y0, double c1,
                      double* y1, double c2, double* y2, double c3, double* y3) {
  k_scale4<<<grid_for(n), 256, 0, st>>>(n, x, c0, y0, c1, y1, c2, y2, c3, y3);
  return hipGetLastError();
}

hipError_t vec_copy4(hipStream_t st, int n, const double* s0, const double* s1,
                     const double* s2, const double* s3, double* dst) {
  k_copy4<<<grid_for(n), 256, 0, st>>>(n, s0, s1, s2, s3, dst);
  return hipGetLastError();
}

hipError_t vec_update2(hipStream_t st, int n, const double* x, double a1, double* y1, double a2,
                       double* y2) {
  k_update2<<<grid_for(n), 256, 0, st>>>(n, x, a1, y1, a2, y2);
  return hipGetLastError();
}

hipError_t vec_mul(hipStream_t st, int n, const double* x, const double* w, bool inverse,
                   double* y) {
  k_vmul<<<grid_for(n), 256, 0, st>>>(n, x, w, inverse ? 1 : 0, y);
  return hipGetLastError();
}

hipError_t vec_ewt(hipStream_t st, int n, const double* y, double rtol, double atol,
                   double* ewt) {
  k_ewt<<<grid_for(n), 256, 0, st>>>(n, y, rtol, atol, ewt);
  return hipGetLastError();
}

hipError_t vec_wsq_partials(hipStream_t st, int n, const double* x, const double* w,
                            double* part, int* nparts) {
  *nparts = grid_for(n);
  k_wsq<<<*nparts, 256, 0, st>>>(n, x, w, part);
  return hipGetLastError();
}

hipError_t vec_multidot_partials(hipStream_t st, int n, int k, const double* V, size_t ld,
                                 const double* y, double* part, int* nparts) {
  if (k < 1 || k > 8) return hipErrorInvalidValue;
  *nparts = grid_for(n);
  k_multidot<<<*nparts, 256, 0, st>>>(n, k, V, ld, y, part);
  return hipGetLastError();
}

hipError_t vec_axpy_multi(hipStream_t st, int n, int k, const double* V, size_t ld,
                          const double* c_dev, double* y) {
  if (k < 1 || k > 64) return hipErrorInvalidValue;
  k_axpy_multi<<<grid_for(n), 256, 0, st>>>(n, k, V, ld, c_dev, y);
  return hipGetLastError();
}

hipError_t vec_maxdist_partials(hipStream_t st, int nn, const double* a, const double* b,
                                double* part, int* nparts) {
  *nparts = grid_for(nn);
  k_maxdist<<<*nparts, 256, 0, st>>>(nn, a, b, part);
  return hipGetLastError();
}

hipError_t bjac_setup(hipStream_t st, int nn, const Coeffs* ct, const int* inst, double gamma,
                      const double* m, const double* H, const double* kappa, double* Minv) {
  k_bjac_setup<<<(nn + 255) / 256, 256, 0, st>>>(nn, ct, inst, gamma, m, H, kappa, Minv);
  return hipGetLastError();
}

hipError_t bjac_solve(hipStream_t st, int nn, const double* Minv, const double* r, double* z) {
  k_bjac_solve<<<(nn + 255) / 256, 256, 0, st>>>(nn, Minv, r, z);
  return hipGetLastError();
}

hipError_t bjac_solve_t(hipStream_t st, int nn, const double* Minv, const double* r, double* z) {
  k_bjac_solve_t<<<(nn + 255) / 256, 256, 0, st>>>(nn, Minv, r, z);
  return hipGetLastError();
}

}  // namespace nsim
