// Launchers of the adjoint kernels (nsim_adjoint.hip): the transposed
// Jacobian of the right-hand side, piece by piece.
#pragma once

#include <hip/hip_runtime.h>

#include "nsim_device.h"

namespace nsim {

// Accumulators of parameter gradients (device, may each be null).
struct AdjAccum {
  double* k1bar;  // n: d/dK1_i
  double* msbar;  // n: per-node share of d/dMs of the node's magnet
  double* hbar;   // n: d/d(pulse amplitude) per node, along `dir`
  double dir[3];
};

// Site-local part: g = (df/dH)^T lam, x = g / V, a = (df/dm|_H)^T lam +
// (dH_anis/dm)^T g, and the parameter accumulations.
hipError_t launch_adj_local(const DeviceOps& d, const FieldArgs& f, const FieldBufs& b,
                            const double* lam, double* g, double* x, double* a,
                            const AdjAccum& acc, hipStream_t st);
// a += negJ A x (exchange, A the difference-form stiffness) and
// psi = -D x (= G^T g, the adjoint of the gradient).
hipError_t launch_adj_exch_div(const DeviceOps& d, bool mixed, const double* x, double* a,
                               double* psi, hipStream_t st);
// e_l = psi_i - (K z)_i on the boundary nodes i = bnode[l] (E^T psi with z the
// interior solve of psi).
hipError_t launch_adj_boundary(const DeviceOps& d, const double* psi, const double* z,
                               double* e, hipStream_t st);
// chi[map[j]] += (A^T e)_j (accumulate) or = (A^T e)_j for a single-segment
// DenseOp (the BEM: e on the boundary, map = boundary nodes; the potential
// operator: e and chi on all nodes); part: scratch of dense_t_chunks(op) * ld
// doubles.
int dense_t_chunks(const DenseOp& op);
hipError_t launch_dense_gemv_t(const DenseOp& op, const double* e, double* part, double* chi,
                               bool accumulate, hipStream_t st);
// x -= mean(x) (one workgroup)
hipError_t launch_mean_project(int n, double* x, hipStream_t st);
// a -= Ms V_i (G phi)_i (the adjoint of the divergence load: D^T = -V G)
hipError_t launch_adj_grad(const DeviceOps& d, bool mixed, const double* phi, double* a,
                           hipStream_t st);

}  // namespace nsim
