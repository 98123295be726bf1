// Launchers of the integrator's vector kernels (nsim_vec.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "nsim_device.h"

namespace nsim {

constexpr int kReduceBlocks = 240;  // upper bound on partials of one reduction

hipError_t vec_lincomb(hipStream_t st, int n, double c0, const double* x0, double c1,
                       const double* x1, double c2, const double* x2, double c3,
                       const double* x3, double* y);
hipError_t vec_mul(hipStream_t st, int n, const double* x, const double* w, bool inverse,
                   double* y);
// y_k = c_k x, k = 0..3, in one launch.
hipError_t vec_scale4(hipStream_t st, int n, const double* x, double c0, double* y0, double c1,
                      double* y1, double c2, double* y2, double c3, double* y3);
// dst[k * n ...] = src_k (k = 0..3; a null source leaves its slice alone),
// in one launch.
hipError_t vec_copy4(hipStream_t st, int n, const double* s0, const double* s1,
                     const double* s2, const double* s3, double* dst);
// y1 += a1 x; y2 += a2 x, in one launch.
hipError_t vec_update2(hipStream_t st, int n, const double* x, double a1, double* y1, double a2,
                       double* y2);
hipError_t vec_ewt(hipStream_t st, int n, const double* y, double rtol, double atol,
                   double* ewt);
hipError_t vec_wsq_partials(hipStream_t st, int n, const double* x, const double* w,
                            double* part, int* nparts);
hipError_t vec_multidot_partials(hipStream_t st, int n, int k, const double* V, size_t ld,
                                 const double* y, double* part, int* nparts);
hipError_t vec_axpy_multi(hipStream_t st, int n, int k, const double* V, size_t ld,
                          const double* c_dev, double* y);
// Number of partials the *_partials reductions produce at length n.
int vec_reduce_grid(int n);
// Device-side sum of k reductions' partials (part[j * np + b]) into out[j],
// and -out[j] into neg (may be null).
hipError_t vec_reduce_parts(hipStream_t st, int np, int k, const double* part, double* out,
                            double* neg);
hipError_t vec_maxdist_partials(hipStream_t st, int nn, const double* a, const double* b,
                                double* part, int* nparts);
// ct: per-magnet coefficient table, inst: node -> magnet (null: one magnet)
hipError_t bjac_setup(hipStream_t st, int nn, const Coeffs* ct, const int* inst, double gamma,
                      const double* m, const double* H, const double* kappa, double* Minv);
hipError_t bjac_solve(hipStream_t st, int nn, const double* Minv, const double* r, double* z);
// z = Minv^T r per node
hipError_t bjac_solve_t(hipStream_t st, int nn, const double* Minv, const double* r, double* z);

}  // namespace nsim
