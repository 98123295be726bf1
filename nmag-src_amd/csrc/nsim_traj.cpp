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
#include "nsim_integrator.h"
#include "nsim_tsctl.h"
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
  // checkpoint_every x 4 stages x the fields the vector-Jacobian product reads
  // there (H_exch, H_demag, H_anis, H_total), kept by the recompute pass
  double* sfields = nullptr;
  int stage_steps = 0;
  bool forward_done = false;
  nsim_traj last{};
  std::vector<double> last_amps;
  bool had_hext = false;
  // BDF trajectory (nsim_bdf_*): the accepted steps of the recorded run and
  // every state Y_0 .. Y_K, the adjoint's accumulators (a ring over six
  // consecutive steps) and the GMRES work space of both solves
  struct BdfStep {
    double t1, h, gamma;
    int q, window;
    double A[kMaxOrder];  // the step's history term sum_j A[j] Y_{k-j}
  };
  nsim_bdf_traj bdf_tr{};
  std::vector<double> bdf_amps;
  std::vector<BdfStep> bdf_steps;
  double* bdf_states = nullptr;
  int bdf_cap = 0;  // states the buffer holds
  double* bdf_ring[kHist] = {};
  double *bdf_V = nullptr, *bdf_w = nullptr, *bdf_r = nullptr, *bdf_mu = nullptr,
         *bdf_avec = nullptr, *bdf_Minv = nullptr, *bdf_coef = nullptr, *bdf_part = nullptr;
  std::vector<double> bdf_host;
  bool bdf_ready = false;
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
// saved: the fields of y (H_exch, H_demag, H_anis, H_total, 3n each) kept
// when y's right-hand side was evaluated; then it is not evaluated again.
int vjp(nsim_handle* h, const double* y, const double* lam, double* out, const AdjAccum& acc,
        bool eval_rhs = true, const double* saved = nullptr) {
  AdjointState* A = h->adj;
  const DeviceOps& d = h->d;
  hipStream_t st = h->stream;
  // eval_rhs false: the handle's fields are those of y already
  if (eval_rhs && !saved) TRY(device_rhs(h, y, h->dmdt));
  const bool demag = h->opt.do_demag != 0 && d.nb > 0;
  const size_t n3 = 3 * (size_t)d.n;
  FieldBufs b{h->Hexch, demag ? h->Hdemag : nullptr, h->Hanis, h->Htotal, h->rho,
              h->phi1,  h->phi2, h->phi, h->phi1b};
  if (saved) {
    b.Hexch = const_cast<double*>(saved);
    b.Hdemag = demag ? const_cast<double*>(saved + n3) : nullptr;
    b.Hanis = const_cast<double*>(saved + 2 * n3);
    b.Htotal = const_cast<double*>(saved + 3 * n3);
  }
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
    TRY(dalloc(A, &A->sfields, (size_t)tr->checkpoint_every * 16 * n3));
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

// One RK4 step y -> y (stage inputs copied to `stage` and the fields of each
// stage's right-hand side to `fields`, 4 x 3n per stage, when not null).
int rk4_step(nsim_handle* h, double dt, double* y, double* stage, double* fields = nullptr) {
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const int n3 = 3 * A->n;
  const size_t bytes = (size_t)n3 * sizeof(double);
  double** k = A->kv;
  const bool demag = h->opt.do_demag != 0 && h->d.nb > 0;
  auto keep = [&](int i) -> int {
    if (fields)
      HIPCHK(vec_copy4(st, n3, h->Hexch, demag ? h->Hdemag : nullptr, h->Hanis, h->Htotal,
                       fields + (size_t)i * 4 * n3));
    return 0;
  };
  if (stage) HIPCHK(hipMemcpyAsync(stage, y, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, y, k[0]));
  TRY(keep(0));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, 0.5 * dt, k[0], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[1]));
  TRY(keep(1));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, 0.5 * dt, k[1], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + 2 * n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[2]));
  TRY(keep(2));
  HIPCHK(vec_lincomb(st, n3, 1.0, y, dt, k[2], 0, nullptr, 0, nullptr, A->tmp));
  if (stage) HIPCHK(hipMemcpyAsync(stage + 3 * n3, A->tmp, bytes, hipMemcpyDeviceToDevice, st));
  TRY(device_rhs(h, A->tmp, k[3]));
  TRY(keep(3));
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

// ---- BDF trajectory and its discrete adjoint ----

constexpr int kBdfKrylov = 40;  // restart length of the two GMRES solves

using BdfStep = AdjointState::BdfStep;

// H_ext = the static part + amps[w] along the pulse direction (w < 0: none)
int bdf_set_window(nsim_handle* h, const double* amps, int w) {
  AdjointState* A = h->adj;
  const double amp = w >= 0 ? amps[w] : 0.0;
  HIPCHK(vec_lincomb(h->stream, 3 * A->n, 1.0, A->hstatic, amp, A->hdir, 0.0, nullptr, 0.0,
                     nullptr, h->Hext));
  h->have_hext = true;
  return 0;
}

int ensure_bdf(nsim_handle* h, const nsim_bdf_traj* tr) {
  TRY(ensure_vjp(h));
  AdjointState* A = h->adj;
  const size_t n = A->n, n3 = 3 * n;
  if (!A->hstatic) TRY(dalloc(A, &A->hstatic, n3));
  if (!A->hdir) TRY(dalloc(A, &A->hdir, n3));
  if (!A->tmp) TRY(dalloc(A, &A->tmp, n3));
  if (!A->ybar) TRY(dalloc(A, &A->ybar, n3));
  if (!A->bdf_V) {
    for (double*& r : A->bdf_ring) TRY(dalloc(A, &r, n3));
    TRY(dalloc(A, &A->bdf_V, (kBdfKrylov + 1) * n3));
    for (double** p : {&A->bdf_w, &A->bdf_r, &A->bdf_mu, &A->bdf_avec}) TRY(dalloc(A, p, n3));
    TRY(dalloc(A, &A->bdf_Minv, 9 * n));
    TRY(dalloc(A, &A->bdf_coef, kBdfKrylov + 8));
    TRY(dalloc(A, &A->bdf_part, (size_t)kReduceBlocks * 8));
    A->bdf_host.resize((size_t)kReduceBlocks * 8);
  }
  if (tr->max_steps + 1 > A->bdf_cap) {
    TRY(dalloc(A, &A->bdf_states, (size_t)(tr->max_steps + 1) * n3));
    A->bdf_cap = tr->max_steps + 1;
  }
  if (tr->n_pulses > A->hbar_windows) {
    TRY(dalloc(A, &A->hbar, (size_t)tr->n_pulses * n));
    A->hbar_windows = tr->n_pulses;
  }
  std::vector<double> dir(n3);
  for (size_t i = 0; i < n3; i++) dir[i] = tr->pulse_dir[i % 3];
  HIPCHK(hipMemcpyAsync(A->hdir, dir.data(), n3 * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// out[i] = V_i . y for i < k (V with stride ld), on the host
int bdf_dots(nsim_handle* h, int k, const double* V, size_t ld, const double* y,
             std::vector<double>* out) {
  AdjointState* A = h->adj;
  const int n3 = 3 * A->n;
  out->assign(k, 0.0);
  for (int c0 = 0; c0 < k; c0 += 8) {
    const int kc = std::min(8, k - c0);
    int np = 0;
    HIPCHK(vec_multidot_partials(h->stream, n3, kc, V + c0 * ld, ld, y, A->bdf_part, &np));
    HIPCHK(hipMemcpyAsync(A->bdf_host.data(), A->bdf_part, sizeof(double) * np * kc,
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int j = 0; j < kc; j++)
      for (int b = 0; b < np; b++) (*out)[c0 + j] += A->bdf_host[(size_t)j * np + b];
  }
  return 0;
}

int bdf_norm(nsim_handle* h, const double* x, double* out) {
  std::vector<double> d;
  TRY(bdf_dots(h, 1, x, 0, x, &d));
  *out = std::sqrt(std::max(d[0], 0.0));
  return 0;
}

// y += sum_i c[i] V_i (c on the host)
int bdf_axpy(nsim_handle* h, int k, const double* V, size_t ld, const double* c, double* y) {
  AdjointState* A = h->adj;
  for (int c0 = 0; c0 < k; c0 += 32) {
    const int kc = std::min(32, k - c0);
    HIPCHK(hipMemcpyAsync(A->bdf_coef, c + c0, sizeof(double) * kc, hipMemcpyHostToDevice,
                          h->stream));
    HIPCHK(vec_axpy_multi(h->stream, 3 * A->n, kc, V + c0 * ld, ld, A->bdf_coef, y));
  }
  return 0;
}

// x = Op^-1 b by restarted, right-preconditioned GMRES (classical
// Gram-Schmidt applied twice), from x = 0, to |b - Op x| <= rtol |b| or
// maxit iterations.  apply(v, out) and precond(v, out) work on device
// vectors; out never aliases v.  Returns the iterations in *its.
template <typename Op, typename Pc>
int bdf_gmres(nsim_handle* h, Op apply, Pc precond, const double* b, double* x, double rtol,
              int maxit, int* its) {
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const int n3 = 3 * A->n;
  const size_t ld = n3, bytes = n3 * sizeof(double);
  double* V = A->bdf_V;
  double *w = A->bdf_w, *r = A->bdf_r;
  HIPCHK(hipMemsetAsync(x, 0, bytes, st));
  double bnorm = 0.0;
  TRY(bdf_norm(h, b, &bnorm));
  *its = 0;
  if (!(bnorm > 0.0)) return 0;
  HIPCHK(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, st));
  double beta = bnorm;
  while (*its < maxit) {
    HIPCHK(vec_lincomb(st, n3, 1.0 / beta, r, 0, nullptr, 0, nullptr, 0, nullptr, V));
    const int m = kBdfKrylov;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), g(m + 1, 0.0), cs(m), sn(m);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m && *its < maxit; j++) {
      double* Vn = V + (j + 1) * ld;
      TRY(precond(V + j * ld, w));
      TRY(apply(w, Vn));
      std::vector<double> c, c2;
      TRY(bdf_dots(h, j + 1, V, ld, Vn, &c));
      for (double& v : c) v = -v;
      TRY(bdf_axpy(h, j + 1, V, ld, c.data(), Vn));
      TRY(bdf_dots(h, j + 1, V, ld, Vn, &c2));
      for (int i = 0; i <= j; i++) H[i * m + j] = -c[i] + c2[i];
      for (double& v : c2) v = -v;
      TRY(bdf_axpy(h, j + 1, V, ld, c2.data(), Vn));
      double hn = 0.0;
      TRY(bdf_norm(h, Vn, &hn));
      H[(j + 1) * m + j] = hn;
      if (hn > 0.0)
        HIPCHK(vec_lincomb(st, n3, 1.0 / hn, Vn, 0, nullptr, 0, nullptr, 0, nullptr, Vn));
      for (int i = 0; i < j; i++) {
        const double a = H[i * m + j], q = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * a + sn[i] * q;
        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * q;
      }
      const double a = H[j * m + j], q = H[(j + 1) * m + j];
      const double rr = std::hypot(a, q);
      cs[j] = rr > 0.0 ? a / rr : 1.0;
      sn[j] = rr > 0.0 ? q / rr : 0.0;
      H[j * m + j] = rr;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      k = j + 1;
      ++*its;
      if (std::fabs(g[j + 1]) <= rtol * bnorm || !(hn > 0.0)) break;
    }
    std::vector<double> y(k);
    for (int i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int c = i + 1; c < k; c++) s -= H[i * m + c] * y[c];
      y[i] = s / H[i * m + i];
    }
    // x += P (V y)
    HIPCHK(hipMemsetAsync(w, 0, bytes, st));
    TRY(bdf_axpy(h, k, V, ld, y.data(), w));
    TRY(precond(w, r));
    HIPCHK(vec_lincomb(st, n3, 1.0, x, 1.0, r, 0, nullptr, 0, nullptr, x));
    // true residual for the restart
    TRY(apply(x, w));
    HIPCHK(vec_lincomb(st, n3, 1.0, b, -1.0, w, 0, nullptr, 0, nullptr, r));
    TRY(bdf_norm(h, r, &beta));
    if (beta <= rtol * bnorm) return 0;
  }
  return 0;
}

// Y_{k+1} from the recorded step k, the history in the state buffer and the
// current parameters: Newton on Y - gamma f(Y) - sum_j A_j Y_{k-j} = 0 from
// Y = Y_k, with (I - gamma J) by GMRES (J the jacobi_ts1 linearisation,
// block-Jacobi preconditioned), until the residual's RMS is below rtol.
int bdf_replay_step(nsim_handle* h, const double* amps, int k, double rtol) {
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const int n = A->n, n3 = 3 * n;
  const size_t bytes = (size_t)n3 * sizeof(double);
  const BdfStep& S = A->bdf_steps[k];
  double* Y = A->bdf_states + (size_t)(k + 1) * n3;
  auto state = [&](int i) { return A->bdf_states + (size_t)i * n3; };
  // history term
  HIPCHK(vec_lincomb(st, n3, S.A[0], state(k), 0, nullptr, 0, nullptr, 0, nullptr, A->bdf_avec));
  for (int j = 1; j < S.q; j += 3)
    HIPCHK(vec_lincomb(st, n3, 1.0, A->bdf_avec, S.A[j], state(k - j),
                       j + 1 < S.q ? S.A[j + 1] : 0.0, j + 1 < S.q ? state(k - j - 1) : nullptr,
                       j + 2 < S.q ? S.A[j + 2] : 0.0, j + 2 < S.q ? state(k - j - 2) : nullptr,
                       A->bdf_avec));
  HIPCHK(hipMemcpyAsync(Y, state(k), bytes, hipMemcpyDeviceToDevice, st));
  TRY(bdf_set_window(h, amps, S.window));
  const double g = S.gamma;
  for (int it = 0; it < 60; it++) {
    TRY(device_rhs(h, Y, A->bdf_mu));
    // r = avec + gamma f - Y
    HIPCHK(vec_lincomb(st, n3, 1.0, A->bdf_avec, g, A->bdf_mu, -1.0, Y, 0, nullptr, A->tmp));
    double rn = 0.0;
    TRY(bdf_norm(h, A->tmp, &rn));
    if (rn <= rtol * std::sqrt((double)n3)) return 0;
    HIPCHK(bjac_setup(st, n, h->d.coefs, h->d.inst, g, Y, h->Htotal, h->kappa, A->bdf_Minv));
    auto apply = [&](const double* v, double* out) -> int {
      TRY(device_jtimes(h, Y, v, out));
      HIPCHK(vec_lincomb(st, n3, 1.0, v, -g, out, 0, nullptr, 0, nullptr, out));
      return 0;
    };
    auto pc = [&](const double* v, double* out) -> int {
      HIPCHK(bjac_solve(st, n, A->bdf_Minv, v, out));
      return 0;
    };
    int its = 0;
    TRY(bdf_gmres(h, apply, pc, A->tmp, A->ybar, 1e-4, 400, &its));
    HIPCHK(vec_lincomb(st, n3, 1.0, Y, 1.0, A->ybar, 0, nullptr, 0, nullptr, Y));
  }
  return fail("nsim_bdf_replay: Newton did not reach the tolerance at step " + std::to_string(k));
}

// The recorded forward run: a host-driven integrator (so that every
// controller step passes the hook), window by window, that appends every
// accepted step's coefficients and state.
int bdf_record(nsim_handle* h, const TsOptions& o, const nsim_bdf_traj* tr,
               const double* amps) {
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const size_t n3 = 3 * (size_t)A->n, bytes = n3 * sizeof(double);
  Integrator I(h, o);
  TRY(I.init(0.0));
  int window = -1;
  int rc_hook = 0;
  I.set_cycle_hook([&](const TsState& at, const TsState& next) -> int {
    if (next.steps <= at.steps) return 0;  // rejected or another Newton iteration
    const int k = (int)A->bdf_steps.size();
    if (k + 1 >= A->bdf_cap) {
      rc_hook = fail("nsim_bdf_forward: more than max_steps steps");
      return rc_hook;
    }
    BdfStep S{};
    S.t1 = next.th[0];
    S.h = at.h;
    S.gamma = at.gamma;
    S.q = at.q;
    S.window = window;
    for (int j = 0; j < kMaxOrder; j++) S.A[j] = j < at.q ? at.A[j] : 0.0;
    A->bdf_steps.push_back(S);
    HIPCHK(hipMemcpyAsync(A->bdf_states + (size_t)(k + 1) * n3, I.iterate(), bytes,
                          hipMemcpyDeviceToDevice, st));
    return 0;
  });
  // window by window, never stepping across a pulse edge
  double t = 0.0;
  const int nw = tr->n_pulses;
  for (int w = 0; w <= nw && t < tr->t_final; w++) {
    const double edge = w < nw ? std::min(tr->t_final, (w + 1) * tr->pulse_len) : tr->t_final;
    if (edge <= t) continue;
    window = w < nw ? w : -1;
    TRY(bdf_set_window(h, amps, window));
    double reached = t;
    const int rc = I.advance(edge, 0, true, &reached);
    if (rc) return rc_hook ? rc_hook : rc;
    t = reached;
  }
  return 0;
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
      TRY(rk4_step(h, dt, A->y, A->stages + (size_t)(s - s0) * 4 * n3,
                   A->sfields + (size_t)(s - s0) * 16 * n3));
    }
    // backward through the segment
    for (int s = s1 - 1; s >= s0; s--) {
      TRY(set_window(h, tr, amps, s));
      const int w = window_of(tr, s);
      AdjAccum acc{A->k1bar, A->msbar, w >= 0 ? A->hbar + (size_t)w * n : nullptr,
                   {tr->pulse_dir[0], tr->pulse_dir[1], tr->pulse_dir[2]}};
      const double* Y = A->stages + (size_t)(s - s0) * 4 * n3;
      // the fields of each stage, kept by the recompute pass above
      const double* F = A->sfields + (size_t)(s - s0) * 16 * n3;
      HIPCHK(vec_scale4(st, n3, A->ybar, dt / 6.0, kb[0], dt / 3.0, kb[1], dt / 3.0, kb[2],
                        dt / 6.0, kb[3]));
      // stage 4: Y4 = y + dt k3
      TRY(vjp(h, Y + 3 * (size_t)n3, kb[3], A->out, acc, true, F + 12 * (size_t)n3));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, dt, kb[2]));
      // stage 3: Y3 = y + dt/2 k2
      TRY(vjp(h, Y + 2 * (size_t)n3, kb[2], A->out, acc, true, F + 8 * (size_t)n3));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, 0.5 * dt, kb[1]));
      // stage 2: Y2 = y + dt/2 k1
      TRY(vjp(h, Y + (size_t)n3, kb[1], A->out, acc, true, F + 4 * (size_t)n3));
      HIPCHK(vec_update2(st, n3, A->out, 1.0, A->ybar, 0.5 * dt, kb[0]));
      // stage 1: Y1 = y
      TRY(vjp(h, Y, kb[0], A->out, acc, true, F));
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

int nsim_bdf_forward(nsim_handle* h, const nsim_bdf_traj* tr, const double* amps, int* n_steps) {
  if (!h || !tr) return fail("nsim_bdf_forward: null argument");
  if (!(tr->t_final > 0.0) || !(tr->rtol >= 0.0) || !(tr->atol > 0.0) || tr->max_order < 1 ||
      tr->max_order > kMaxOrder || tr->max_steps < 1 || tr->n_pulses < 0)
    return fail("nsim_bdf_forward: need t_final > 0, rtol >= 0, atol > 0, max_order 1..5, "
                "max_steps >= 1, n_pulses >= 0");
  if (tr->n_pulses > 0 && (!amps || !(tr->pulse_len > 0.0)))
    return fail("nsim_bdf_forward: pulses need amplitudes and pulse_len > 0");
  TRY(ensure_bdf(h, tr));
  AdjointState* A = h->adj;
  hipStream_t st = h->stream;
  const size_t n3 = 3 * (size_t)A->n, bytes = n3 * sizeof(double);
  A->had_hext = h->have_hext;
  if (h->have_hext)
    HIPCHK(hipMemcpyAsync(A->hstatic, h->Hext, bytes, hipMemcpyDeviceToDevice, st));
  else
    HIPCHK(hipMemsetAsync(A->hstatic, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(A->bdf_states, h->m, bytes, hipMemcpyDeviceToDevice, st));
  A->bdf_ready = false;
  A->bdf_steps.clear();
  A->bdf_tr = *tr;
  A->bdf_amps.assign(amps ? amps : nullptr, amps ? amps + tr->n_pulses : nullptr);
  TsOptions o = h->ts_opt;
  o.rtol = tr->rtol;
  o.atol = tr->atol;
  o.max_order = tr->max_order;
  o.device_gmres = 0;
  const int rc = bdf_record(h, o, tr, amps);
  // the recording integrator took the handle's halt word; give it back
  if (h->ts) h->ts->rebind_halt();
  TRY(restore_hext(h));
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(st));
  if (cg_flags(h) != 0) return fail("nsim_bdf_forward: a Poisson solve reported an error");
  A->bdf_ready = true;
  if (n_steps) *n_steps = (int)A->bdf_steps.size();
  return 0;
}

int nsim_bdf_steps(nsim_handle* h, int max, double* t, double* hs, int* orders) {
  if (!h || !h->adj || !h->adj->bdf_ready) return fail("nsim_bdf_steps: no recorded trajectory");
  const auto& S = h->adj->bdf_steps;
  const int k = std::min<int>(max, (int)S.size());
  for (int i = 0; i < k; i++) {
    if (t) t[i] = S[i].t1;
    if (hs) hs[i] = S[i].h;
    if (orders) orders[i] = S[i].q;
  }
  return (int)S.size();
}

int nsim_bdf_replay(nsim_handle* h, const double* amps, double newton_rtol) {
  if (!h) return fail("nsim_bdf_replay: null handle");
  AdjointState* A = h->adj;
  if (!A || !A->bdf_ready) return fail("nsim_bdf_replay: run nsim_bdf_forward first");
  if (A->bdf_tr.n_pulses > 0 && !amps) return fail("nsim_bdf_replay: null pulse amplitudes");
  if (!(newton_rtol > 0.0)) return fail("nsim_bdf_replay: need newton_rtol > 0");
  hipStream_t st = h->stream;
  const size_t n3 = 3 * (size_t)A->n, bytes = n3 * sizeof(double);
  A->had_hext = h->have_hext;
  if (h->have_hext)
    HIPCHK(hipMemcpyAsync(A->hstatic, h->Hext, bytes, hipMemcpyDeviceToDevice, st));
  else
    HIPCHK(hipMemsetAsync(A->hstatic, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(A->bdf_states, h->m, bytes, hipMemcpyDeviceToDevice, st));
  int rc = 0;
  for (int k = 0; k < (int)A->bdf_steps.size() && rc == 0; k++)
    rc = bdf_replay_step(h, amps, k, newton_rtol);
  const int rr = restore_hext(h);
  if (rc) return rc;
  TRY(rr);
  HIPCHK(hipMemcpyAsync(h->m, A->bdf_states + A->bdf_steps.size() * n3, bytes,
                        hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  if (cg_flags(h) != 0) return fail("nsim_bdf_replay: a Poisson solve reported an error");
  if (A->bdf_tr.n_pulses > 0) A->bdf_amps.assign(amps, amps + A->bdf_tr.n_pulses);
  return 0;
}

int nsim_bdf_adjoint(nsim_handle* h, const double* amps, const double* m_bar_final,
                     double* m_bar_initial, double* K1_bar, double* Ms_bar, double* amps_bar,
                     double lin_rtol) {
  if (!h || !m_bar_final) return fail("nsim_bdf_adjoint: null argument");
  AdjointState* A = h->adj;
  if (!A || !A->bdf_ready) return fail("nsim_bdf_adjoint: run nsim_bdf_forward first");
  const nsim_bdf_traj& tr = A->bdf_tr;
  if (tr.n_pulses > 0 &&
      (!amps || !std::equal(amps, amps + tr.n_pulses, A->bdf_amps.begin())))
    return fail("nsim_bdf_adjoint: pulse amplitudes differ from the last forward or replay");
  if (!(lin_rtol > 0.0)) return fail("nsim_bdf_adjoint: need lin_rtol > 0");
  hipStream_t st = h->stream;
  const int n = A->n, n3 = 3 * n;
  const size_t bytes = (size_t)n3 * sizeof(double);
  const int K = (int)A->bdf_steps.size();
  for (double* r : A->bdf_ring) HIPCHK(hipMemsetAsync(r, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(A->bdf_ring[K % kHist], m_bar_final, bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(A->k1bar, 0, n * sizeof(double), st));
  HIPCHK(hipMemsetAsync(A->msbar, 0, n * sizeof(double), st));
  if (tr.n_pulses > 0)
    HIPCHK(hipMemsetAsync(A->hbar, 0, (size_t)tr.n_pulses * n * sizeof(double), st));
  A->had_hext = h->have_hext;
  if (h->have_hext)
    HIPCHK(hipMemcpyAsync(A->hstatic, h->Hext, bytes, hipMemcpyDeviceToDevice, st));
  else
    HIPCHK(hipMemsetAsync(A->hstatic, 0, bytes, st));
  const AdjAccum none{nullptr, nullptr, nullptr, {0.0, 0.0, 0.0}};
  int rc = 0;
  for (int k = K - 1; k >= 0 && rc == 0; k--) {
    const BdfStep& S = A->bdf_steps[k];
    const double* Y1 = A->bdf_states + (size_t)(k + 1) * n3;
    const double g = S.gamma;
    // mu = (I - gamma J(Y_{k+1})^T)^-1 ybar_{k+1}
    rc = bdf_set_window(h, amps, S.window);
    if (rc == 0) rc = device_rhs(h, Y1, A->bdf_mu);
    if (rc) break;
    HIPCHK(bjac_setup(st, n, h->d.coefs, h->d.inst, g, Y1, h->Htotal, h->kappa, A->bdf_Minv));
    auto apply = [&](const double* v, double* out) -> int {
      TRY(vjp(h, Y1, v, out, none, false));
      HIPCHK(vec_lincomb(st, n3, 1.0, v, -g, out, 0, nullptr, 0, nullptr, out));
      return 0;
    };
    auto pc = [&](const double* v, double* out) -> int {
      HIPCHK(bjac_solve_t(st, n, A->bdf_Minv, v, out));
      return 0;
    };
    int its = 0;
    rc = bdf_gmres(h, apply, pc, A->bdf_ring[(k + 1) % kHist], A->bdf_mu, lin_rtol, 2000, &its);
    if (rc) break;
    // parameter cotangents: (df/dp)^T (gamma mu)
    HIPCHK(vec_lincomb(st, n3, g, A->bdf_mu, 0, nullptr, 0, nullptr, 0, nullptr, A->lam));
    AdjAccum acc{A->k1bar, A->msbar, S.window >= 0 ? A->hbar + (size_t)S.window * n : nullptr,
                 {tr.pulse_dir[0], tr.pulse_dir[1], tr.pulse_dir[2]}};
    rc = vjp(h, Y1, A->lam, A->out, acc, false);
    if (rc) break;
    // the history: ybar_{k-j} += A_j mu
    for (int j = 0; j < S.q; j++)
      HIPCHK(vec_lincomb(st, n3, 1.0, A->bdf_ring[(k - j) % kHist], S.A[j], A->bdf_mu, 0,
                         nullptr, 0, nullptr, A->bdf_ring[(k - j) % kHist]));
    HIPCHK(hipMemsetAsync(A->bdf_ring[(k + 1) % kHist], 0, bytes, st));
  }
  const int rr = restore_hext(h);
  if (rc) return rc;
  TRY(rr);
  if (m_bar_initial)
    HIPCHK(hipMemcpyAsync(m_bar_initial, A->bdf_ring[0], bytes, hipMemcpyDeviceToHost, st));
  if (K1_bar)
    HIPCHK(hipMemcpyAsync(K1_bar, A->k1bar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  std::vector<double> ms(n), hb((size_t)tr.n_pulses * n);
  HIPCHK(hipMemcpyAsync(ms.data(), A->msbar, n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (tr.n_pulses > 0)
    HIPCHK(hipMemcpyAsync(hb.data(), A->hbar, hb.size() * sizeof(double), hipMemcpyDeviceToHost,
                          st));
  HIPCHK(hipStreamSynchronize(st));
  if (Ms_bar) {
    double s = 0.0;
    for (double v : ms) s += v;
    *Ms_bar = s;
  }
  if (amps_bar)
    for (int w = 0; w < tr.n_pulses; w++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += hb[(size_t)w * n + i];
      amps_bar[w] = s;
    }
  if (cg_flags(h) != 0) return fail("nsim_bdf_adjoint: a Poisson solve reported an error");
  return 0;
}

}  // extern "C"
