// Launcher of the persistent GMRES kernel (nsim_gmres.hip) that solves the
// integrator's Newton systems in one launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "nsim_device.h"

namespace nsim {

constexpr int kGmresMaxl = 512;  // largest Krylov dimension the kernel holds

// Device buffers of the solver (owned by the caller).
struct GmresWork {
  double* V = nullptr;        // (maxl + 1) * n3 Krylov basis, stride n3
  double* zp = nullptr;       // n3: the preconditioned direction, exchanged
  double* part = nullptr;     // 2 * gmax * pstride partial sums
  unsigned long long* tag = nullptr;  // 2 * gmax phase tags
  double* R = nullptr;        // maxl * maxl triangular factor (workgroup 0)
  double* y = nullptr;        // maxl: the least-squares solution, broadcast
  double* status = nullptr;   // [0] iterations [1] sum (x ewt)^2 [2] error flag
  unsigned* err = nullptr;    // exchange error flag
  int gmax = 0, pstride = 0;
  unsigned long long epoch = 0;  // launch counter; makes every phase tag unique
};

// Grid of the kernel for a handle with `nslices` SELL slices: workgroups
// (all co-resident) and slices per workgroup.  Returns false if no grid fits.
bool gmres_grid(int nslices, int maxl, int device, int gmax, int* grid, int* spw);

// x = (I - gamma J)^-1 r by right-preconditioned GMRES in the space scaled by
// ewt, J the jacobi_ts1 linearisation at (m, Ht); Minv: 3x3 block-Jacobi
// inverse per node, or null.  Stops at |residual| <= tol (scaled Euclidean) or
// after maxl iterations.  Writes status[0] = iterations, status[1] =
// sum_i (x_i ewt_i)^2; status[2] != 0 if an exchange timed out.
hipError_t launch_gmres(const DeviceOps& d, bool mixed, GmresWork& w, int grid, int spw,
                        int maxl, double gamma, double tol, const double* m, const double* Ht,
                        const double* ewt, const double* Minv, const double* r, double* x,
                        hipStream_t st);

}  // namespace nsim
