// The handle behind the C ABI: one magnet (or a batch of independent magnets)
// with its operators in HBM, its field vectors, one HIP stream and the
// persistent-CG communication buffers.  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "nsim_device.h"
#include "nsim_integrator.h"
#include "nsim_internal.h"

namespace nsim {

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  long count = 0;
};

// One dense-route field evaluation captured as a HIP graph (exchange and
// divergence, potential GEMV, demag + LLG), keyed by everything its kernel
// arguments depend on.
struct RhsGraph {
  const double* m = nullptr;
  double* dmdt = nullptr;
  FieldArgs f{};
  bool demag = false;
  hipGraphExec_t exec = nullptr;
};

struct AdjointState;  // buffers of the adjoint (nsim_traj.cpp)
void destroy_adjoint(AdjointState* a);

}  // namespace nsim

struct nsim_handle {
  nsim::HostOperators host;
  nsim::DeviceOps d;
  nsim::CgWork cgw;
  unsigned* sync_nbc = nullptr;
  unsigned* sync_dbc = nullptr;
  nsim::Coeffs coef{};           // the first magnet's coefficients
  std::vector<nsim::Coeffs> coefs;  // host copy of the per-magnet table
  std::vector<int> node_off;     // magnet k owns nodes [node_off[k], node_off[k+1])
  std::vector<int> bnd_off;      // and boundary nodes [bnd_off[k], bnd_off[k+1])
  bool use_potential = false;    // demag through the potential operator M
  bool use_hmatrix = false;      // M compressed as an H-matrix (else dense)
  double tuned_bem_us = 0.0, tuned_cg_us = 0.0, tuned_pot_us = 0.0;  // autotuner's best
  nsim_options opt{};
  bool mixed = false;
  int device = 0, num_cus = 0;
  hipStream_t stream = nullptr;
  nsim::CgDevice cg;             // the CG's reordered operator
  std::vector<void*> cg_allocs;  // its device buffers
  bool cg_coarse = true;         // two-level preconditioner on
  bool cg_fine = true;           // its coarse space of sub-part aggregates when it fits
  // field vectors (site-major 3-vectors or scalars)
  double *m = nullptr, *Hext = nullptr, *k1 = nullptr, *k2 = nullptr, *axis = nullptr,
         *pin = nullptr;
  double *Hexch = nullptr, *Hdemag = nullptr, *Hanis = nullptr, *Htotal = nullptr,
         *dmdt = nullptr, *rho = nullptr, *phi1 = nullptr, *phi2 = nullptr, *phi = nullptr,
         *phi1b = nullptr;
  double *io_a = nullptr, *io_b = nullptr, *io_c = nullptr;  // staging for host-pointer calls
  bool have_hext = false, have_anis = false, have_k2 = false, have_pin = false;
  // polynomial anisotropy (nsim_set_anisotropy): each magnet's terms, and
  // their device table with per-magnet offsets (null: none)
  std::vector<std::vector<nsim::AnisTerm>> poly_terms;
  nsim::AnisTerm* poly = nullptr;
  int* poly_off = nullptr;
  long rhs_evals = 0;
  bool rhs_graph = true;               // replay captured dense-route evaluations
  std::vector<nsim::RhsGraph> graphs;  // at most kMaxRhsGraphs, oldest first
  std::vector<nsim::RhsGraph> seen;    // keys evaluated once, not captured yet
  long graph_launches = 0, graph_captures = 0;
  bool timers_on = false;
  std::map<std::string, nsim::Timer> timers;
  nsim::Integrator* ts = nullptr;
  nsim::TsOptions ts_opt;
  double* kappa = nullptr;    // negJ K_ii / V_i: dH_exch_i/dm_i
  // CVODE preconditioner callbacks (nsim_psolve.hip): the Jacobian frozen at
  // the last re-Jacobi (linearisation point m and its H_total), M = I - gamma J
  // with its 3x3 node-block inverses, psolve's KSP settings and workspace
  double* pc_Minv = nullptr;
  double *pc_m = nullptr, *pc_H = nullptr;
  bool pc_have_J = false;
  double pc_gamma = 0.0;
  double pc_rtol = 1e-2, pc_atol = 1e-7;  // interface/nmag/main.py:728-729
  double pc_dtol = 1e5;                   // PETSc's defaults
  int pc_maxit = 10000;
  bool pc_full = true;  // 0: psolve applies the block preconditioner only
  double* pc_ws = nullptr;
  long pc_solves = 0, pc_iters_total = 0, pc_unconverged = 0, pc_jac_builds = 0;
  int pc_iters_last = 0;
  nsim::AdjointState* adj = nullptr;
  std::vector<void*> allocations;

  void track(void* p) {
    if (p) allocations.push_back(p);
  }
};

namespace nsim {

// Memory the dense potential operators of a handle may take.
constexpr double kPotentialBudgetBytes = 32.0 * 1024 * 1024 * 1024;

// Polynomial anisotropy: terms per magnet and the largest power of a
// component (order 16 is the reference's practical limit, ~153 terms).
constexpr int kMaxAnisTerms = 256;
constexpr int kMaxAnisPower = 16;

// Captured evaluations kept per handle (one per (m, dmdt) buffer pair in use).
// A pair is captured on its second evaluation, so callers that stream through
// many buffers once each (a sweep over stored states) stay on eager launches.
constexpr size_t kMaxRhsGraphs = 8;
constexpr size_t kMaxRhsSeen = 16;

// Shared by the C-ABI translation units.
int device_rhs(nsim_handle* h, const double* m, double* dmdt);
// The field inputs of an evaluation at m (device) from the handle's fields.
FieldArgs field_args(nsim_handle* h, const double* m);
// Destroys the captured evaluations; called whenever a launch shape changes.
void drop_rhs_graphs(nsim_handle* h);
// Builds the dense demag potential operator of every magnet (nsim_dense.cpp).
int build_potential(nsim_handle* h);
// Re-uploads the per-magnet coefficient table and the exchange diagonal kappa
// from h->coefs.
int refresh_coefficients(nsim_handle* h);

}  // namespace nsim
