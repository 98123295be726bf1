// Differentiation form: the vector-Jacobian product of the right-hand side,
// a fixed-step classical Runge-Kutta (RK4) trajectory, and its discrete
// adjoint with checkpointing (states kept every k steps, the steps in between
// recomputed in the backward sweep).
//
// The reference has no adjoint.  The time stepper here is explicit and
// fixed-step on purpose: its discrete adjoint is exact for the discrete
// forward map (gradients match finite differences of the same forward to the
// Poisson solves' tolerance), where the adaptive BDF's step selection is not
// differentiable.  The step must sit inside RK4's stability region for the
// exchange spectrum (|lambda dt| < 2.8 on the imaginary axis).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/nsim.h"
#include "nsim_adjoint.h"
#include "nsim_device.h"
#include "nsim_handle.h"
#include "nsim_vec.h"

namespace nsim {

struct AdjointState {
  int n = 0, nb = 0;
  // vector-Jacobian product scratch
  double *g = nullptr, *x = nullptr, *a = nullptr, *psi = nullptr, *z = nullptr,
         *e = nullptr, *part = nullptr, *phia = nullptr, *lam = nullptr, *out = nullptr;
  // parameter-gradient accumulators
  double *k1bar = nullptr, *msbar = nullptr, *hbar = nullptr;
  int hbar_windows = 0;
  // trajectory
  double *y = nullptr, *tmp = nullptr, *ybar = nullptr, *kv[4] = {}, *kb[4] = {};
  double *hstatic = nullptr, *hdir = nullptr;
  double* ckpt = nullptr;
  int ckpt_count = 0;
  double* stages = nullptr;  // checkpoint_every x 4 stage inputs
  int stage_steps = 0;
  bool forward_done = false;
  nsim_traj last{};
  std::vector<double> last_amps;
  bool had_hext = false;
  std::vector<void*> allocs;
};

void destroy_adjoint(AdjointState* a) {
  if (!a) return;
  for (void* p : a->allocs) (void)hipFree(p);
  delete a;
}

namespace {

int fail(const std::string& msg) {
  set_error(msg);
  return -1;
}

#define HIPCHK(expr)                                                  \
  do {                                                                \
    hipError_t _e = (expr);                                           \
    if (_e != hipSuccess)                                             \
      return fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)
#define TRY(expr)          \
  do {                     \
    int _rc = (expr);      \
    if (_rc != 0) return _rc; \
  } while (0)

int dalloc(AdjointState* A, double** p, size_t count) {
  HIPCHK(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(double)));
  A->allocs.push_back(*p);
  HIPCHK(hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(double)));
  return 0;
}

int ensure_vjp(nsim_handle* h) {
  if (h->d.ninst != 1)
    return fail("the adjoint supports a handle with a single magnet");
  if (!h->adj) {
    AdjointState* A = new AdjointState();
    h->adj = A;
    const size_t n = h->d.n, n3 = 3 * n;
    A->n = (int)n;
    A->nb = h->d.nb;
    for (double** p : {&A->g, &A->x, &A->a, &A->lam, &A->out}) TRY(dalloc(A, p, n3));
    for (double** p : {&A->psi, &A->z, &A->phia, &A->k1bar, &A->msbar}) TRY(dalloc(A, p, n));
    TRY(dalloc(A, &A->e, std::max(1, h->d.nb)));
    const DenseOp& op = h->d.pot.A ? h->d.pot : h->d.bem;
    const size_t ld = op.h_ld.empty() ? 1 : op.h_ld[0];
    TRY(dalloc(A, &A->part, (size_t)dense_t_chunks(op) * ld));
  }
  return 0;
}

FieldArgs field_args_of(nsim_handle* h, const double* m) { return field_args(h, m); }

// out = J(y)^T lam, plus the parameter accumulations of `acc`.  Evaluates the
// right-hand side at y first, so the handle's fields are those of y.
int vjp(nsim_handle* h, const double* y, const double* lam, double* out, const AdjAccum& acc) {
  AdjointState* A = h->adj;
  const DeviceOps& d = h->d;
  hipStream_t st = h->stream;
  TRY(device_rhs(h, y, h->dmdt));
  const bool demag = h->opt.do_demag != 0 && d.nb > 0;
  FieldBufs b{h->Hexch, demag ? h->Hdemag : nullptr, h->Hanis, h->Htotal, h->rho,
              h->phi1,  h->phi2, h->phi, h->phi1b};
  HIPCHK(launch_adj_local(d, field_args_of(h, y), b, lam, A->g, A->x, out, acc, st));
  HIPCHK(launch_adj_exch_div(d, h->mixed, A->x, out, A->psi, st));
  if (demag && d.pot.A) {
    // dense route phi = M rho: the adjoint potential is M^T psi (written
    // whole: the operator's map covers every node)
    HIPCHK(launch_dense_gemv_t(d.pot, A->psi, A->part, A->phia, false, st));
    HIPCHK(launch_adj_grad(d, h->mixed, A->phia, out, st));
  } else if (demag) {
    CgWork& w = h->cgw;
    // E^T psi: interior solve K_II z = psi_I (z = 0 on the boundary), then
    // e = psi_B - K_BI z_I
    w.sync = h->sync_dbc;
    HIPCHK(launch_cg(h->cg, w, true, h->cg_coarse, A->psi, A->z, h->opt.dbc_rtol,
                     h->opt.dbc_maxits, nullptr, nullptr, nullptr, d.blocal, st));
    HIPCHK(launch_adj_boundary(d, A->psi, A->z, A->e, st));
    // chi = psi + R^T B^T e, then K^+ chi (the pseudo-inverse projects out the
    // constants first)
    HIPCHK(launch_dense_gemv_t(d.bem, A->e, A->part, A->psi, true, st));
    HIPCHK(launch_mean_project(d.n, A->psi, st));
    w.sync = h->sync_nbc;
    HIPCHK(launch_cg(h->cg, w, false, h->cg_coarse, A->psi, A->phia, h->opt.nbc_rtol,
                     h->opt.nbc_maxits, nullptr, nullptr, nullptr, d.blocal, st));
    HIPCHK(launch_adj_grad(d, h->mixed, A->phia, out, st));
  }
  return 0;
}

int check_traj(nsim_handle* h, const nsim_traj* tr) {
  if (!tr) return fail("null trajectory description");
  if (!(tr->dt > 0.0) || tr->n_steps <= 0 || tr->checkpoint_every <= 0)
    return fail("trajectory needs dt > 0, n_steps > 0, checkpoint_every > 0");
  if (tr->n_pulses < 0) return fail("n_pulses < 0");
  if (tr->n_pulses > 0) {
    const double r = tr->pulse_len / tr->dt;
    if (!(tr->pulse_len > 0.0) || std::fabs(r - std::round(r)) > 1e-9 * std::max(1.0, r))
      return fail("pulse_len must be a positive multiple of dt (pulse edges on step edges)");
  }
  (void)h;
  return 0;
}

int window_of(const nsim_traj* tr, int step) {
  if (tr->n_pulses == 0) return -1;
  const int w = (int)std::floor((step + 0.5) * tr->dt / tr->pulse_len);
  return w < tr->n_pulses ? w : -1;
}

int ensure_traj(nsim_handle* h, const nsim_traj* tr) {
  TRY(ensure_vjp(h));
  AdjointState* A = h->adj;
  const size_t n3 = 3 * (size_t)A->n;
  if (!A->y) {
    for (double** p : {&A->y, &A->tmp, &A->ybar, &A->kv[0], &A->kv[1], &A->kv[2], &A->kv[3],
                       &A->kb[0], &A->kb[1], &A->kb[2], &A->kb[3], &A->hstatic, &A->hdir})
      TRY(dalloc(A, p, n3));
  }
  const int nck = (tr->n_steps + tr->checkpoint_every - 1) / tr->checkpoint_every;
  if (nck > A->ckpt_count) {
    TRY(dalloc(A, &A->ckpt, (size_t)nck * n3));
    A->ckpt_count = nck;
  }
  if (tr->checkpoint_every > A->stage_steps) {
    TRY(dalloc(A, &A->stages, (size_t)tr->checkpoint_every * 4 * n3));
    A->stage_steps = tr->checkpoint_every;
  }
  if (tr->n_pulses > A->hbar_windows) {
    TRY(dalloc(A, &A->hbar, (size_t)tr->n_pulses * A->n));
    A->hbar_windows = tr->n_pulses;
  }
  std::vector<double> dir(n3);
  for (size_t i = 0; i < n3; i++) dir[i] = tr->pulse_dir[i % 3];
  HIPCHK(hipMemcpyAsync(A->hdir, dir.data(), n3 * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// H_ext = static part + amplitude of the step's pulse window along the pulse
// direction (all four stages of a step see the same window).
int set_window(nsim_handle* h, const nsim_traj* tr, const double* amps, int step) {
  AdjointState* A = h->adj;
  const int w = window_of(tr, step);
  const double amp = w >= 0 ? amps[w] : 0.0;
  const int n3 = 3 * A->n;
  HIPCHK(vec_lincomb(h->stream, n3, 1.0, A->hstatic, amp, A->hdir, 0.0, nullptr, 0.0, nullptr,
                     h->Hext));
  h->have_hext = true;
  return 0;
}

int restore_hext(nsim_handle* h) {
  AdjointState* A = h->adj;
  HIPCHK(hipMemcpyAsync(h->Hext, A->hstatic, 3 * (size_t)A->n * sizeof(double),
                        hipMemcpyDeviceToDevice, h->stream));
  h->have_hext = A->had_hext;
  return 0;
}

// One RK4 step y -> y (stage inputs copied to `stage` when not null).
int rk4_step(nsim_handle* h, double dt, double* y, double* stage) {
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const int n3 = 3 * A->n;
  const size_t bytes = (size_t)n3 * sizeof(double);
  double** k = A->kv;
  if (stage) HIPCHK(hipMemcpyAsync(stage, y, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, y, k[0]));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, 0.5 * dt, k[0], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[1]));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, 0.5 * dt, k[1], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + 2 * n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[2]));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, dt, k[2], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + 3 * n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[3]));
  HIPCHK(vec_lincomb(st, n3, 1.0, k[0], 2.0, k[1], 2.0, k[2], 1.0, k[3], A->tmp));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, dt / 6.0, A->tmp, 0, nullptr, 0, nullptr, y));
  return 0;
}

bool same_traj(const nsim_traj& a, const nsim_traj& b) {
  return a.dt == b.dt && a.n_steps == b.n_steps && a.checkpoint_every == b.checkpoint_every &&
         a.n_pulses == b.n_pulses && a.pulse_len == b.pulse_len &&
         a.pulse_dir[0] == b.pulse_dir[0] && a.pulse_dir[1] == b.pulse_dir[1] &&
         a.pulse_dir[2] == b.pulse_dir[2];
}

unsigned cg_flags(nsim_handle* h) {
  unsigned a[8] = {0}, b[8] = {0};
  if (hipMemcpy(a, h->sync_nbc, sizeof(a), hipMemcpyDeviceToHost) != hipSuccess) return ~0u;
  if (hipMemcpy(b, h->sync_dbc, sizeof(b), hipMemcpyDeviceToHost) != hipSuccess) return ~0u;
  return a[2] | b[2];
}

}  // namespace

}  // namespace nsim

using namespace nsim;

extern "C" {

int nsim_rhs_vjp(nsim_handle* h, double t, const double* y, const double* lam, double* y_bar,
                 double* K1_bar, double* Ms_bar, double* Hext_bar) {
  (void)t;
  if (!h || !y || !lam || !y_bar) return fail("nsim_rhs_vjp: null argument");
  TRY(ensure_vjp(h));
  AdjointState* A = h->adj;
  const size_t n = A->n, n3 = 3 * n;
  hipStream_t st = h->stream;
  HIPCHK(hipMemcpyAsync(h->io_a, y, n3 * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(A->lam, lam, n3 * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(A->k1bar, 0, n * sizeof(double), st));
  HIPCHK(hipMemsetAsync(A->msbar, 0, n * sizeof(double), st));
  AdjAccum acc{A->k1bar, A->msbar, nullptr, {0.0, 0.0, 0.0}};
  TRY(vjp(h, h->io_a, A->lam, A->out, acc));
  HIPCHK(hipMemcpyAsync(y_bar, A->out, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
  std::vector<double> ms(n);
  if (K1_bar)
    HIPCHK(hipMemcpyAsync(K1_bar, A->k1bar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (Ms_bar)
    HIPCHK(hipMemcpyAsync(ms.data(), A->msbar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (Hext_bar)
    HIPCHK(hipMemcpyAsync(Hext_bar, A->g, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (Ms_bar) {
    double s = 0.0;
    for (double v : ms) s += v;
    *Ms_bar = s;
  }
  if (cg_flags(h) != 0) return fail("nsim_rhs_vjp: a Poisson solve reported an error");
  return 0;
}

int nsim_traj_forward(nsim_handle* h, const nsim_traj* tr, const double* amps) {
  if (!h) return fail("nsim_traj_forward: null handle");
  TRY(check_traj(h, tr));
  if (tr->n_pulses > 0 && !amps) return fail("nsim_traj_forward: null pulse amplitudes");
  TRY(ensure_traj(h, tr));
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const size_t n3 = 3 * (size_t)A->n, bytes = n3 * sizeof(double);
  A->had_hext = h->have_hext;
  if (h->have_hext)
    HIPCHK(hipMemcpyAsync(A->hstatic, h->Hext, bytes, hipMemcpyDeviceToDevice, st));
  else
    HIPCHK(hipMemsetAsync(A->hstatic, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(A->y, h->m, bytes, hipMemcpyDeviceToDevice, st));
  for (int s = 0; s < tr->n_steps; s++) {
    if (s % tr->checkpoint_every == 0)
      HIPCHK(hipMemcpyAsync(A->ckpt + (size_t)(s / tr->checkpoint_every) * n3, A->y, bytes,
                            hipMemcpyDeviceToDevice, st));
    TRY(set_window(h, tr, amps, s));
    TRY(rk4_step(h, tr->dt, A->y, nullptr));
  }
  HIPCHK(hipMemcpyAsync(h->m, A->y, bytes, hipMemcpyDeviceToDevice, st));
  TRY(restore_hext(h));
  HIPCHK(hipStreamSynchronize(st));
  if (cg_flags(h) != 0) return fail("nsim_traj_forward: a Poisson solve reported an error");
  A->forward_done = true;
  A->last = *tr;
  A->last_amps.assign(amps ? amps : nullptr, amps ? amps + tr->n_pulses : nullptr);
  return 0;
}

int nsim_traj_adjoint(nsim_handle* h, const nsim_traj* tr, const double* amps,
                      const double* m_bar_final, double* m_bar_initial, double* K1_bar,
                      double* Ms_bar, double* amps_bar) {
  if (!h || !m_bar_final) return fail("nsim_traj_adjoint: null argument");
  TRY(check_traj(h, tr));
  AdjointState* A = h->adj;
  if (!A || !A->forward_done || !same_traj(A->last, *tr))
    return fail("nsim_traj_adjoint: run nsim_traj_forward with the same description first");
  if (tr->n_pulses > 0 &&
      (!amps || !std::equal(amps, amps + tr->n_pulses, A->last_amps.begin())))
    return fail("nsim_traj_adjoint: pulse amplitudes differ from the forward run");
  hipStream_t st = h->stream;
  const int n3 = 3 * A->n;
  const size_t n = A->n, bytes = (size_t)n3 * sizeof(double);
  const double dt = tr->dt;
  HIPCHK(hipMemcpyAsync(A->ybar, m_bar_final, bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(A->k1bar, 0, n * sizeof(double), st));
  HIPCHK(hipMemsetAsync(A->msbar, 0, n * sizeof(double), st));
  if (tr->n_pulses > 0)
    HIPCHK(hipMemsetAsync(A->hbar, 0, (size_t)tr->n_pulses * n * sizeof(double), st));
  const int k = tr->checkpoint_every;
  const int nseg = (tr->n_steps + k - 1) / k;
  double** kb = A->kb;
  for (int seg = nseg - 1; seg >= 0; seg--) {
    const int s0 = seg * k, s1 = std::min(s0 + k, tr->n_steps);
    // recompute the segment's stage inputs from its checkpoint
    HIPCHK(hipMemcpyAsync(A->y, A->ckpt + (size_t)seg * n3, bytes, hipMemcpyDeviceToDevice, st));
    for (int s = s0; s < s1; s++) {
      TRY(set_window(h, tr, amps, s));
      TRY(rk4_step(h, dt, A->y, A->stages + (size_t)(s - s0) * 4 * n3));
    }
    // backward through the segment
    for (int s = s1 - 1; s >= s0; s--) {
      TRY(set_window(h, tr, amps, s));
      const int w = window_of(tr, s);
      AdjAccum acc{A->k1bar, A->msbar, w >= 0 ? A->hbar + (size_t)w * n : nullptr,
                   {tr->pulse_dir[0], tr->pulse_dir[1], tr->pulse_dir[2]}};
      const double* Y = A->stages + (size_t)(s - s0) * 4 * n3;
      HIPCHK(vec_scale4(st, n3, A->ybar, dt / 6.0, kb[0], dt / 3.0, kb[1], dt / 3.0, kb[2],
                        dt / 6.0, kb[3]));
      // stage 4: Y4 = y + dt k3
      TRY(vjp(h, Y + 3 * (size_t)n3, kb[3], A->out, acc));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, dt, kb[2]));
      // stage 3: Y3 = y + dt/2 k2
      TRY(vjp(h, Y + 2 * (size_t)n3, kb[2], A->out, acc));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, 0.5 * dt, kb[1]));
      // stage 2: Y2 = y + dt/2 k1
      TRY(vjp(h, Y + (size_t)n3, kb[1], A->out, acc));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, 0.5 * dt, kb[0]));
      // stage 1: Y1 = y
      TRY(vjp(h, Y, kb[0], A->out, acc));
      HIPCHK(vec_lincomb(st, n3, 1.0, A->ybar, 1.0, A->out, 0, nullptr, 0, nullptr, A->ybar));
    }
  }
  TRY(restore_hext(h));
  if (m_bar_initial)
    HIPCHK(hipMemcpyAsync(m_bar_initial, A->ybar, bytes, hipMemcpyDeviceToHost, st));
  if (K1_bar) HIPCHK(hipMemcpyAsync(K1_bar, A->k1bar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  std::vector<double> ms(n), hb((size_t)tr->n_pulses * n);
  HIPCHK(hipMemcpyAsync(ms.data(), A->msbar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (tr->n_pulses > 0)
    HIPCHK(hipMemcpyAsync(hb.data(), A->hbar, hb.size() * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  HIPCHK(hipStreamSynchronize(st));
  if (Ms_bar) {
    double s = 0.0;
    for (double v : ms) s += v;
    *Ms_bar = s;
  }
  if (amps_bar)
    for (int w = 0; w < tr->n_pulses; w++) {
      double s = 0.0;
      for (size_t i = 0; i < n; i++) s += hb[(size_t)w * n + i];
      amps_bar[w] = s;
    }
  if (cg_flags(h) != 0) return fail("nsim_traj_adjoint: a Poisson solve reported an error");
  return 0;
}

}  // extern "C"
