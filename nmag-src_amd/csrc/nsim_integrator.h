// Device-resident time integrator: variable-step BDF (orders 1 to 5) with a
// Newton iteration whose linear systems (I - gamma J) delta = -G are solved by
// scaled GMRES with the analytic jacobi_ts1 J.v and a block-Jacobi right
// preconditioner.  It stands in for CVODE (BDF + Newton + SPGMR) behind the
// reference's timestepper (src/nsim.ml:2485-2580, src/sundials_sp_stubs.c:700-1308).
//
// The step controller (nsim_tsctl.h) runs on the device after every Newton
// iteration, so an advance call queues whole Newton cycles -- start of
// attempt, field evaluation, persistent GMRES, controller -- in batches and
// the host only watches for the end.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "nsim_gmres.h"
#include "nsim_tsctl.h"

struct nsim_handle;

namespace nsim {

// Entry points of the rhs path used by the integrator.
int device_rhs(nsim_handle* h, const double* m, double* dmdt);
int device_jtimes(nsim_handle* h, const double* m, const double* v, double* Jv);
hipStream_t handle_stream(nsim_handle* h);
int handle_n(nsim_handle* h);
double* handle_m(nsim_handle* h);
const double* handle_Htotal(nsim_handle* h);
const double* handle_kappa(nsim_handle* h);
struct Coeffs;
const Coeffs* handle_coef_table(nsim_handle* h);  // device, one entry per magnet
const int* handle_inst(nsim_handle* h);           // device node -> magnet, or null
const DeviceOps& handle_ops(nsim_handle* h);
bool handle_mixed(nsim_handle* h);
int handle_device(nsim_handle* h);
// Points the field kernels of the handle at a halt word (null: none).
void handle_set_halt(nsim_handle* h, const int* halt);
int handle_error_flags(nsim_handle* h, unsigned* out);
void set_error(const std::string& msg);

struct TsOptions {
  double rtol = 1e-6, atol = 1e-6;  // scalar tolerances (main.py:724-725)
  int max_order = 2;                // timestepper_max_order (main.py:838)
  int krylov_max = 300;             // timestepper_krylov_max (main.py:839)
  int precond = 1;                  // 0 none, 1 block Jacobi of I - gamma J
  int device_gmres = 1;             // 1: one persistent GMRES launch per Newton
                                    // iteration, queued without host round trips;
                                    // 0: host-driven Arnoldi steps
  int gmres_spw = 1;                // least SELL slices per GMRES workgroup
  int gmres_pc_sweeps = 1;          // local Jacobi sweeps of the persistent GMRES's
                                    // preconditioner (0: 3x3 block Jacobi only)
  int gmres_onex_steps = 32;        // persistent GMRES: Arnoldi steps with one
                                    // grid exchange each (0 .. 32; the rest take two)
  int max_batch = 16;               // Newton cycles queued per batch
  bool fixed_order = false;         // orders 3-5: rise to max_order as the history
                                    // allows instead of choosing the order
  bool cooperative = false;         // persistent GMRES as a cooperative launch
};

struct TsStats {
  long steps = 0, rhs = 0, newton = 0, linear = 0, err_fails = 0, conv_fails = 0;
  long linear_max = 0;  // most Krylov iterations of one Newton iteration
  double h_last = 0.0, t = 0.0, max_dm = 0.0;
  int q_last = 0;
  long long order_steps[kMaxOrder + 1] = {};  // accepted steps per order
};

// Called after every controller step of a host-driven run with the state of
// the attempt that ran and the state the controller left; a non-zero return
// ends the run with that status.
using CycleHook = std::function<int(const TsState& attempt, const TsState& next)>;

class Integrator {
 public:
  Integrator(nsim_handle* h, const TsOptions& o);
  ~Integrator();
  // Takes y(t0) from the handle's m; forgets all history (CVodeReInit).
  int init(double t0);
  // Integrates towards t_final (at most max_steps steps if max_steps > 0).
  // exact_tstop: never step past t_final; else step past and interpolate.
  // The handle's m receives y at *t_reached.
  int advance(double t_final, int max_steps, bool exact_tstop, double* t_reached);
  const TsStats& stats() const { return st_; }
  TsOptions& options() { return opt_; }
  // workgroups of the persistent GMRES launch (0: host-driven GMRES only)
  int gmres_grid() const { return ggrid_; }
  // With a hook set, advance runs the host-driven cycles and calls it after
  // each; iterate() is the Newton iterate (the accepted y_{n+1} after an
  // accepting step).
  void set_cycle_hook(CycleHook f) { hook_ = std::move(f); }
  // Points the handle's field kernels at this integrator's halt word again
  // (after another integrator on the same handle has gone).
  void rebind_halt() {
    if (ok_) handle_set_halt(h_, &ds_->halt);
  }
  const double* iterate() const { return ycur_; }

 private:
  int run_device_cycles();
  int run_host_cycles();
  int enqueue_cycle();
  int upload_state();
  int download_state();
  int gmres(const double* r, double* x, int* iters);
  int apply_A(const double* z, double* w);
  double wrms(const double* x);
  int reduce(int nparts, int k, std::vector<double>* out);
  int estimate(int cnt, const double* W, double* out);
  int check(hipError_t e, const char* what);

  nsim_handle* h_;
  TsOptions opt_;
  TsStats st_;
  hipStream_t s_;
  int n3_ = 0, nn_ = 0;
  bool ok_ = false;
  TsState hs_{};              // controller state, host copy (current between calls)
  TsState* ds_ = nullptr;     // its device copy, which the controller kernel updates
  TsState* pin_ = nullptr;    // pinned: [0], [1] batch mirrors, [2] upload staging
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double* slot_[kHist] = {};     // ring of y_n, y_{n-1}, ..., y_{n-5}
  double** slots_dev_ = nullptr;                   // the same pointers on the device
  double *fn_ = nullptr;    // f(y_n) (first step's predictor)
  double *yp_ = nullptr, *ycur_ = nullptr, *avec_ = nullptr, *f_ = nullptr, *res_ = nullptr,
         *del_ = nullptr, *tmp_ = nullptr, *tmp2_ = nullptr, *ewt_ = nullptr, *ystart_ = nullptr;
  double* status_ = nullptr;  // Newton-iteration results for the controller (host GMRES)
  double* V_ = nullptr;     // (maxl+1) Krylov vectors
  double* Minv_ = nullptr;  // 9 per node
  double* coef_ = nullptr;  // device coefficients for multi-axpy
  double* part_ = nullptr;  // reduction partials
  double* hdev_ = nullptr;  // device-reduced Gram-Schmidt coefficients and norm
  double* hhost_ = nullptr; // their pinned host copy
  int np_ = 0;              // partials per reduction at this vector length
  GmresWork gw_;            // persistent GMRES buffers (V shared with V_)
  int ggrid_ = 0, gspw_ = 0;  // its grid; 0: not available
  std::vector<double> host_part_;
  double gamma_ = 0.0;      // Newton matrix coefficient of the host-driven GMRES
  std::vector<void*> allocs_;
  CycleHook hook_;
};

}  // namespace nsim
