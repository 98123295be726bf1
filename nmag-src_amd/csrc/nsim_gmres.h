// Launcher of the persistent GMRES kernel (nsim_gmres.hip) that solves the
// integrator's Newton systems in one launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_tsctl.h"

namespace nsim {

// Largest Krylov dimension the kernel holds (the reference's default maxl is
// 300).  Larger bases run the host-driven GMRES.  The per-workgroup arrays of
// this size are what keep three workgroups' LDS within one CU.
constexpr int kGmresMaxl = 320;
constexpr int kGmresTagStride = dev::kTagStride;  // tag words per workgroup and parity
// Arnoldi steps the kernel takes with one grid exchange each (the later ones
// take two).
constexpr int kGmresOneExMax = 32;

// Partial sums per workgroup and exchange: up to maxl + 2 for a two-exchange
// step, 2 j + 4 for one-exchange step j; rounded to whole 64-byte lines.
inline int gmres_pstride(int maxl) {
  const int one = 2 * (maxl < kGmresOneExMax ? maxl : kGmresOneExMax) + 4;
  const int need = maxl + 2 > one ? maxl + 2 : one;
  return (need + 7) & ~7;
}

// Device buffers of the solver (owned by the caller).
struct GmresWork {
  double* V = nullptr;        // (maxl + 1) * n3 Krylov basis, stride n3
  double* Z = nullptr;        // maxl * n3 preconditioned directions, exchanged
  double* part = nullptr;     // 2 * gmax * pstride partial sums
  unsigned long long* tag = nullptr;  // 2 * gmax phase tags, kGmresTagStride apart
  double* R = nullptr;        // maxl * maxl triangular factor (workgroup 0)
  double* y = nullptr;        // maxl: the least-squares solution, broadcast
  double* status = nullptr;   // [0] iterations [1] sum (x ewt)^2 [2] error flag, [3..5] sums
  unsigned* err = nullptr;    // exchange error flag
  int gmax = 0, pstride = 0;
  int resident = 0;              // workgroups the device holds at once (gmres_grid)
  unsigned long long epoch = 0;  // launch counter; makes every phase tag unique
  // The grid never exceeds what the device holds at once (gmres_grid), and
  // the launches before and after it on the stream wait for it, so its
  // workgroups are co-resident with a plain launch; a cooperative launch
  // guarantees the same at about 20 us more dispatch cost per launch
  // (measured on MI355X).  Every exchange wait is bounded either way.
  bool cooperative = false;
};

// Grid of the kernel for a handle with `nslices` SELL slices: workgroups
// (all co-resident, at most gmax) and slices per workgroup (at least
// min_spw); `resident` receives the number of workgroups the device holds at
// once.  Returns false if no grid fits or maxl exceeds kGmresMaxl.
bool gmres_grid(int nslices, int maxl, int device, int gmax, int min_spw, int* grid, int* spw,
                int* resident);

// One Newton iteration of a BDF step, y - gamma f(y) - avec = 0.
struct NewtonSystem {
  double* ycur;         // iterate; receives ycur + x
  const double* Ht;     // H_total of the rhs evaluated at ycur
  const double* f;      // that rhs
  const double* avec;
  const double* yp;     // the step's predictor
  const double* ewt;    // error weights
  const double* kappa;  // negJ K_ii / V_i per node
  double* Minv;         // 9 per node: block-Jacobi inverses set up here, or null
  double* x;            // the correction
  double gamma, tol;
  int sweeps = 0;       // local Jacobi sweeps over each workgroup's own nodes
  int onex_steps = kGmresOneExMax;  // Arnoldi steps with one grid exchange each
                                    // (0 .. kGmresOneExMax; the rest take two)
  const double* gamma_p = nullptr;  // device word holding gamma (overrides gamma)
  const int* halt = nullptr;        // nonzero device word: the launch does nothing
  // step controller record: when set, the launch's last act is the
  // controller's step on this Newton iteration's results (ts_control), so no
  // separate controller launch is needed
  TsState* ts = nullptr;
  double* const* slots = nullptr;  // device table of the controller's history ring
};

// Solves (I - gamma J) x = avec + gamma f - ycur by right-preconditioned
// GMRES in the space scaled by ewt (J the jacobi_ts1 linearisation at ycur)
// to |residual| <= tol (scaled Euclidean) or maxl iterations, then sets
// ycur += x.  Writes w.status: [0] iterations, [1] sum (x ewt)^2,
// [2] != 0 if an exchange timed out, [3] sum ((ycur - yp) ewt)^2,
// [4], [5] the sums of squares of the controller's order q - 1 and q + 1
// error estimates (0 unless ts asks for them).
// Refuses a grid above w.resident.
hipError_t launch_newton_gmres(const DeviceOps& d, bool mixed, GmresWork& w, int grid, int spw,
                               int maxl, const NewtonSystem& s, hipStream_t st);

}  // namespace nsim
