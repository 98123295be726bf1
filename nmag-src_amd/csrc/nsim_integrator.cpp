// Device-resident BDF / Newton-GMRES time integrator (see nsim_integrator.h).
//
// Step controller (host) -- all vectors stay in HBM, the host sees only the
// reduction partials of norms and inner products:
//  * BDF of order q <= 2 on the actual (variable) step history: the corrector
//    is P'(t_{n+1}) = f(y_{n+1}) for the polynomial P through y_{n+1..n+1-q}.
//  * predictor = extrapolation of the polynomial through y_{n..n-q} (explicit
//    Euler on the first step).  The local error is estimated by Milne's device,
//    err = E_c/(E_c+E_p) ||y - y_pred||_WRMS, with the variable-step error
//    constants of both formulas computed from the node positions.
//  * Newton on y - gamma f(y) - a = 0 with (I - gamma J) solved by scaled GMRES
//    (CGS2 orthogonalisation), J.v from the jacobi_ts1 linearisation and a 3x3
//    block-Jacobi right preconditioner; WRMS weights 1/(rtol|y| + atol).
#include "nsim_integrator.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "nsim_device.h"
#include "nsim_vec.h"

namespace nsim {

namespace {

constexpr int kMaxCor = 3;          // Newton iterations per attempt (CVODE NLS_MAXCOR)
constexpr double kNewtonCoef = 0.1; // Newton convergence coefficient
constexpr double kLinCoef = 0.05;   // linear-solve tolerance factor (CVODE eplifac)
constexpr int kMaxAttempts = 20;

// l_j'(T[0]) of the Lagrange basis polynomial of node j among T[0..m].
double lagrange_deriv0(const double* T, int m, int j) {
  if (j == 0) {
    double s = 0.0;
    for (int i = 1; i <= m; i++) s += 1.0 / (T[0] - T[i]);
    return s;
  }
  double num = 1.0, den = 1.0;
  for (int i = 0; i <= m; i++) {
    if (i == j) continue;
    den *= T[j] - T[i];
    if (i != 0) num *= T[0] - T[i];
  }
  return num / den;
}

double factorial(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; i++) f *= i;
  return f;
}

}  // namespace

int Integrator::check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  set_error(std::string("integrator: ") + what + ": " + hipGetErrorString(e));
  return -1;
}

#define TRY(expr)                     \
  do {                                \
    int _rc = (expr);                 \
    if (_rc) return _rc;              \
  } while (0)
#define HTRY(expr) TRY(check((expr), #expr))

Integrator::Integrator(nsim_handle* h, const TsOptions& o) : h_(h), opt_(o) {
  s_ = handle_stream(h);
  nn_ = handle_n(h);
  n3_ = 3 * nn_;
  opt_.krylov_max = std::max(1, std::min(opt_.krylov_max, 500));
  opt_.max_order = std::max(1, std::min(opt_.max_order, 2));
  auto alloc = [&](double** p, size_t count) {
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(double)) != hipSuccess) return false;
    allocs_.push_back(*p);
    return hipMemset(*p, 0, count * sizeof(double)) == hipSuccess;
  };
  const size_t n = n3_;
  np_ = vec_reduce_grid(n3_);
  ok_ = alloc(&yn_, n) && alloc(&ynm1_, n) && alloc(&ynm2_, n) && alloc(&fn_, n) &&
        alloc(&yp_, n) && alloc(&ycur_, n) && alloc(&avec_, n) && alloc(&f_, n) &&
        alloc(&res_, n) && alloc(&del_, n) && alloc(&tmp_, n) && alloc(&tmp2_, n) &&
        alloc(&ewt_, n) && alloc(&ystart_, n) &&
        alloc(&V_, n * (size_t)(opt_.krylov_max + 1)) && alloc(&Minv_, 9 * (size_t)nn_) &&
        alloc(&coef_, (size_t)opt_.krylov_max + 8) &&
        alloc(&hdev_, 2 * (size_t)opt_.krylov_max + 8) &&
        alloc(&part_, ((size_t)opt_.krylov_max + 8) * kReduceBlocks);
  host_part_.resize(8 * kReduceBlocks);
  if (ok_ && hipHostMalloc(reinterpret_cast<void**>(&hhost_),
                           (2 * (size_t)opt_.krylov_max + 8) * sizeof(double)) != hipSuccess)
    ok_ = false;
  // persistent GMRES: one co-resident grid over the SELL slices
  const DeviceOps& d = handle_ops(h);
  int grid = 0, spw = 0;
  if (ok_ && nsim::gmres_grid(d.nslices, opt_.krylov_max, handle_device(h), 1 << 20,
                        std::max(1, opt_.gmres_spw), &grid, &spw)) {
    gw_.V = V_;
    gw_.gmax = grid;
    gw_.pstride = (opt_.krylov_max + 2 + 7) & ~7;
    const size_t maxl = opt_.krylov_max;
    bool g = alloc(&gw_.Z, n * maxl) && alloc(&gw_.part, 2 * (size_t)grid * gw_.pstride) &&
             alloc(&gw_.R, maxl * maxl) && alloc(&gw_.y, maxl) && alloc(&gw_.status, 8);
    const size_t tag_bytes = 2 * (size_t)grid * kGmresTagStride * sizeof(unsigned long long);
    if (g && hipMalloc(reinterpret_cast<void**>(&gw_.tag), tag_bytes) == hipSuccess) {
      allocs_.push_back(gw_.tag);
      g = hipMemset(gw_.tag, 0, tag_bytes) == hipSuccess;
    } else {
      g = false;
    }
    if (g && hipMalloc(reinterpret_cast<void**>(&gw_.err), sizeof(unsigned)) == hipSuccess) {
      allocs_.push_back(gw_.err);
      g = hipMemset(gw_.err, 0, sizeof(unsigned)) == hipSuccess;
    } else {
      g = false;
    }
    if (g) {
      ggrid_ = grid;
      gspw_ = spw;
    }
  }
}

Integrator::~Integrator() {
  (void)hipStreamSynchronize(s_);
  for (void* p : allocs_) (void)hipFree(p);
  if (hhost_) (void)hipHostFree(hhost_);
}

int Integrator::reduce(int nparts, int k, std::vector<double>* out) {
  HTRY(hipMemcpyAsync(host_part_.data(), part_, sizeof(double) * nparts * k,
                      hipMemcpyDeviceToHost, s_));
  HTRY(hipStreamSynchronize(s_));
  out->assign(k, 0.0);
  for (int j = 0; j < k; j++)
    for (int b = 0; b < nparts; b++) (*out)[j] += host_part_[j * nparts + b];
  return 0;
}

double Integrator::wrms(const double* x) {
  int np = 0;
  if (vec_wsq_partials(s_, n3_, x, ewt_, part_, &np) != hipSuccess) return NAN;
  std::vector<double> s;
  if (reduce(np, 1, &s)) return NAN;
  return std::sqrt(s[0] / n3_);
}

int Integrator::init(double t0) {
  if (!ok_) {
    set_error("integrator: device allocation failed");
    return -1;
  }
  HTRY(hipMemcpyAsync(yn_, handle_m(h_), sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
  HTRY(hipStreamSynchronize(s_));
  tn_ = t0;
  nhist_ = 0;
  hnext_ = 0.0;
  qnext_ = 1;
  gamma_pc_ = -1.0;
  crate_ = 1.0;
  gamma_crate_ = 0.0;
  st_ = TsStats();
  st_.t = t0;
  return 0;
}

int Integrator::apply_A(const double* z, double* w) {
  // w = z - gamma J z, J at the current Newton iterate (H_total of its rhs)
  TRY(device_jtimes(h_, ycur_, z, tmp2_));
  HTRY(vec_lincomb(s_, n3_, 1.0, z, -gamma_, tmp2_, 0.0, nullptr, 0.0, nullptr, w));
  return 0;
}

// Right-preconditioned GMRES in the space scaled by the error weights, so the
// Euclidean residual is sqrt(n) times its WRMS norm.  x = (I - gamma J)^-1 r.
int Integrator::gmres(const double* r, double* x, int* iters) {
  const int maxl = opt_.krylov_max;
  const size_t ld = n3_;
  double* V0 = V_;
  HTRY(vec_mul(s_, n3_, r, ewt_, false, V0));
  int np = 0;
  std::vector<double> red;
  HTRY(vec_wsq_partials(s_, n3_, V0, nullptr, part_, &np));
  TRY(reduce(np, 1, &red));
  const double beta = std::sqrt(red[0]);
  const double tol = kLinCoef * kNewtonCoef * std::sqrt((double)n3_);
  *iters = 0;
  if (!(beta > tol)) {
    HTRY(hipMemsetAsync(x, 0, sizeof(double) * n3_, s_));
    return 0;
  }
  HTRY(vec_lincomb(s_, n3_, 1.0 / beta, V0, 0, nullptr, 0, nullptr, 0, nullptr, V0));
  std::vector<double> H((size_t)(maxl + 1) * maxl, 0.0), g(maxl + 1, 0.0), cs(maxl), sn(maxl);
  g[0] = beta;
  int k = 0;
  for (int j = 0; j < maxl; j++) {
    double* Vj = V_ + j * ld;
    double* Vn = V_ + (j + 1) * ld;
    HTRY(vec_mul(s_, n3_, Vj, ewt_, true, tmp_));
    if (opt_.precond) HTRY(bjac_solve(s_, nn_, Minv_, tmp_, del_));
    else HTRY(hipMemcpyAsync(del_, tmp_, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
    TRY(apply_A(del_, tmp_));
    HTRY(vec_mul(s_, n3_, tmp_, ewt_, false, Vn));
    // classical Gram-Schmidt, twice; the projections are summed and applied on
    // the device, and the host reads the two coefficient columns and the new
    // norm with one synchronisation per Arnoldi step
    for (int pass = 0; pass < 2; pass++) {
      for (int c0 = 0; c0 <= j; c0 += 8) {
        const int kc = std::min(8, j + 1 - c0);
        int npc = 0;  // always vec_reduce_grid(n3_)
        HTRY(vec_multidot_partials(s_, n3_, kc, V_ + c0 * ld, ld, Vn, part_ + c0 * np_, &npc));
      }
      HTRY(vec_reduce_parts(s_, np_, j + 1, part_, hdev_ + pass * (j + 1), coef_));
      for (int c0 = 0; c0 <= j; c0 += 64) {
        const int kc = std::min(64, j + 1 - c0);
        HTRY(vec_axpy_multi(s_, n3_, kc, V_ + c0 * ld, ld, coef_ + c0, Vn));
      }
    }
    HTRY(vec_wsq_partials(s_, n3_, Vn, nullptr, part_, &np));
    HTRY(vec_reduce_parts(s_, np, 1, part_, hdev_ + 2 * (j + 1), nullptr));
    HTRY(hipMemcpyAsync(hhost_, hdev_, sizeof(double) * (2 * (j + 1) + 1),
                        hipMemcpyDeviceToHost, s_));
    HTRY(hipStreamSynchronize(s_));
    for (int i = 0; i <= j; i++) H[i * maxl + j] += hhost_[i] + hhost_[j + 1 + i];
    const double hn = std::sqrt(hhost_[2 * (j + 1)]);
    H[(j + 1) * maxl + j] = hn;
    if (hn > 0.0)
      HTRY(vec_lincomb(s_, n3_, 1.0 / hn, Vn, 0, nullptr, 0, nullptr, 0, nullptr, Vn));
    for (int i = 0; i < j; i++) {  // previous Givens rotations
      const double a = H[i * maxl + j], b = H[(i + 1) * maxl + j];
      H[i * maxl + j] = cs[i] * a + sn[i] * b;
      H[(i + 1) * maxl + j] = -sn[i] * a + cs[i] * b;
    }
    const double a = H[j * maxl + j], b = H[(j + 1) * maxl + j];
    const double rr = std::hypot(a, b);
    cs[j] = rr > 0 ? a / rr : 1.0;
    sn[j] = rr > 0 ? b / rr : 0.0;
    H[j * maxl + j] = rr;
    H[(j + 1) * maxl + j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    k = j + 1;
    st_.linear++;
    if (std::fabs(g[j + 1]) <= tol || hn == 0.0) break;
  }
  // back substitution, then x = M^-1 (sum y_i V_i / ewt)
  std::vector<double> y(k);
  for (int i = k - 1; i >= 0; i--) {
    double s = g[i];
    for (int c = i + 1; c < k; c++) s -= H[i * maxl + c] * y[c];
    y[i] = s / H[i * maxl + i];
  }
  HTRY(hipMemsetAsync(tmp_, 0, sizeof(double) * n3_, s_));
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int kc = std::min(64, k - c0);
    HTRY(hipMemcpyAsync(coef_, y.data() + c0, sizeof(double) * kc, hipMemcpyHostToDevice, s_));
    HTRY(vec_axpy_multi(s_, n3_, kc, V_ + c0 * ld, ld, coef_, tmp_));
  }
  HTRY(vec_mul(s_, n3_, tmp_, ewt_, true, tmp_));
  if (opt_.precond) HTRY(bjac_solve(s_, nn_, Minv_, tmp_, x));
  else HTRY(hipMemcpyAsync(x, tmp_, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
  *iters = k;
  return 0;
}

// One Newton iteration after the rhs at ycur: residual, block-Jacobi set-up,
// GMRES, ycur += del and the weighted norms in one persistent launch
// (nsim_gmres.hip); the host reads the iteration count, the correction's
// WRMS norm *d and the sum of squares of the local error with one
// synchronisation.
int Integrator::newton_device(double* d, double* errsq) {
  NewtonSystem sys;
  sys.ycur = ycur_;
  sys.Ht = handle_Htotal(h_);
  sys.f = f_;
  sys.avec = avec_;
  sys.yp = yp_;
  sys.ewt = ewt_;
  sys.kappa = handle_kappa(h_);
  sys.Minv = opt_.precond ? Minv_ : nullptr;
  sys.x = del_;
  sys.gamma = gamma_;
  sys.tol = kLinCoef * kNewtonCoef * std::sqrt((double)n3_);
  sys.sweeps = opt_.gmres_pc_sweeps;
  HTRY(launch_newton_gmres(handle_ops(h_), handle_mixed(h_), gw_, ggrid_, gspw_,
                           opt_.krylov_max, sys, s_));
  HTRY(hipMemcpyAsync(hhost_, gw_.status, 4 * sizeof(double), hipMemcpyDeviceToHost, s_));
  HTRY(hipStreamSynchronize(s_));
  if (hhost_[2] != 0.0) {
    (void)hipMemset(gw_.err, 0, sizeof(unsigned));
    set_error("integrator: GMRES exchange timed out");
    return -1;
  }
  st_.linear += (long)hhost_[0];
  *d = std::sqrt(hhost_[1] / n3_);
  *errsq = hhost_[3];
  return 0;
}

// Predictor into yp; E_p is the predictor's error constant (coefficient of
// y^(q+1), with y_pred - y_exact = -E_p y^(q+1)).
// yp = L0 y_n + L1 p1 + L2 p2 (null p drop their terms).
void Integrator::predictor(double h, int q, Predictor* p) const {
  const double t1 = tn_ + h;
  if (nhist_ == 0) {  // explicit Euler from f(y_n)
    *p = Predictor{{1.0, h, 0.0}, fn_, nullptr, 0.5 * h * h};
    return;
  }
  const double P[3] = {tn_, tnm1_, tnm2_};
  double L[3] = {0, 0, 0}, ep = 1.0;
  for (int j = 0; j <= q; j++) {
    double w = 1.0;
    for (int i = 0; i <= q; i++)
      if (i != j) w *= (t1 - P[i]) / (P[j] - P[i]);
    L[j] = w;
    ep *= t1 - P[j];
  }
  *p = Predictor{{L[0], L[1], q >= 2 ? L[2] : 0.0}, ynm1_, q >= 2 ? ynm2_ : nullptr,
                 ep / factorial(q + 1)};
}

int Integrator::step_attempt(double h, int q, bool* accepted, double* err_out) {
  *accepted = false;
  const double T[3] = {tn_ + h, tn_, tnm1_};
  const double* Y[3] = {nullptr, yn_, ynm1_};
  double dl[3];
  for (int j = 0; j <= q; j++) dl[j] = lagrange_deriv0(T, q, j);
  gamma_ = 1.0 / dl[0];
  double R = 0.0;
  for (int j = 1; j <= q; j++) R += dl[j] * std::pow(T[j] - T[0], q + 1) / factorial(q + 1);
  const double E_c = -R / dl[0];
  Predictor pr;
  predictor(h, q, &pr);
  const double C = E_c / (E_c + pr.E_p);
  // one pass: error weights of y_n, the predictor yp (= ycur), and
  // a = -gamma sum_{j>=1} l_j' y_j
  HTRY(vec_step_begin(s_, n3_, yn_, pr.L[0], pr.p1, pr.L[1], pr.p2, pr.L[2], -gamma_ * dl[1],
                      q >= 2 ? Y[2] : nullptr, q >= 2 ? -gamma_ * dl[2] : 0.0, opt_.rtol,
                      opt_.atol, ewt_, yp_, ycur_, avec_));
  const bool dev = opt_.device_gmres && ggrid_ > 0;
  double errsq = NAN;  // sum ((ycur - yp) ewt)^2 after the last Newton update
  // The convergence-rate estimate carries over between steps (as CVODE's
  // crate does) unless gamma changed a lot, so a small first correction can
  // be accepted without a second rhs evaluation.
  if (gamma_crate_ <= 0.0 || std::fabs(gamma_ / gamma_crate_ - 1.0) > 0.3) crate_ = 1.0;
  gamma_crate_ = gamma_;
  double del_old = 0.0, crate = crate_;
  bool converged = false;
  for (int k = 0; k < kMaxCor; k++) {
    TRY(device_rhs(h_, ycur_, f_));
    st_.rhs++;
    st_.newton++;
    double d = 0.0;
    if (dev) {
      // residual, preconditioner, solve, update and both norms in one launch
      TRY(newton_device(&d, &errsq));
    } else {
      // res = -G = a + gamma f - y
      HTRY(vec_lincomb(s_, n3_, 1.0, avec_, gamma_, f_, -1.0, ycur_, 0, nullptr, res_));
      if (opt_.precond)
        HTRY(bjac_setup(s_, nn_, handle_coef_table(h_), handle_inst(h_), gamma_, ycur_,
                        handle_Htotal(h_), handle_kappa(h_), Minv_));
      int its = 0;
      TRY(gmres(res_, del_, &its));
      HTRY(vec_lincomb(s_, n3_, 1.0, ycur_, 1.0, del_, 0, nullptr, 0, nullptr, ycur_));
      d = wrms(del_);
    }
    if (!std::isfinite(d)) break;
    if (k > 0) crate = std::max(0.3 * crate, d / del_old);
    if (d * std::min(1.0, crate) * C <= kNewtonCoef) {
      converged = true;
      break;
    }
    if (k > 0 && d > 2.0 * del_old) break;
    del_old = d;
  }
  crate_ = crate;
  if (!converged) {
    st_.conv_fails++;
    crate_ = 1.0;
    hnext_ = 0.25 * h;
    return 0;
  }
  double err = 0.0;
  if (dev) {
    err = C * std::sqrt(errsq / n3_);
  } else {
    HTRY(vec_lincomb(s_, n3_, 1.0, ycur_, -1.0, yp_, 0, nullptr, 0, nullptr, tmp_));
    err = C * wrms(tmp_);
  }
  *err_out = err;
  if (!std::isfinite(err)) {
    st_.conv_fails++;
    hnext_ = 0.25 * h;
    return 0;
  }
  if (err > 1.0) {
    st_.err_fails++;
    hnext_ = h * std::max(0.2, 0.9 * std::pow(err, -1.0 / (q + 1)));
    return 0;
  }
  // accept: rotate the history
  std::swap(ynm2_, ynm1_);
  std::swap(ynm1_, yn_);
  std::swap(yn_, ycur_);
  tnm2_ = tnm1_;
  tnm1_ = tn_;
  tn_ = tn_ + h;
  nhist_ = std::min(nhist_ + 1, 2);
  st_.steps++;
  st_.h_last = h;
  st_.q_last = q;
  st_.t = tn_;
  const double etamax = st_.steps <= 10 ? 100.0 : 10.0;
  double eta = err > 0.0 ? 0.9 * std::pow(err, -1.0 / (q + 1)) : etamax;
  eta = std::min(etamax, std::max(0.2, eta));
  if (eta > 1.0 && eta < 1.5) eta = 1.0;
  hnext_ = h * eta;
  qnext_ = std::min(opt_.max_order, std::max(1, nhist_));
  *accepted = true;
  return 0;
}

int Integrator::advance(double t_final, int max_steps, bool exact_tstop, double* t_reached) {
  if (!ok_) {
    set_error("integrator: not initialised");
    return -1;
  }
  double* m = handle_m(h_);
  HTRY(hipMemcpyAsync(ystart_, m, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
  const double eps_t = 1e-12 * std::max(1.0, std::fabs(t_final));
  int steps = 0;
  while (true) {
    if (max_steps > 0 && steps >= max_steps) break;
    if (exact_tstop ? (t_final - tn_ <= eps_t) : (tn_ >= t_final)) break;
    if (hnext_ <= 0.0) {  // first step: h0 from ||f(y0)|| (each attempt sets ewt again)
      HTRY(vec_ewt(s_, n3_, yn_, opt_.rtol, opt_.atol, ewt_));
      TRY(device_rhs(h_, yn_, fn_));
      st_.rhs++;
      const double fnorm = wrms(fn_);
      double h0 = fnorm > 0 ? 0.1 / fnorm : 1.0;
      h0 = std::min(h0, 0.1 * std::fabs(t_final - tn_));
      hnext_ = h0 > 0 ? h0 : 1e-6;
    }
    double h = hnext_;
    int q = qnext_;
    int fails = 0;
    bool accepted = false;
    double err = 0.0;
    while (!accepted) {
      if (exact_tstop && tn_ + h > t_final - 0.01 * h) h = t_final - tn_;
      if (fails >= 2) q = 1;
      TRY(step_attempt(h, q, &accepted, &err));
      if (accepted) break;
      if (++fails >= kMaxAttempts || hnext_ < 1e-14 * std::max(1.0, std::fabs(tn_))) {
        set_error("integrator: step size underflow or repeated failures at t=" +
                  std::to_string(tn_));
        return 1;
      }
      h = hnext_;
    }
    steps++;
  }
  if (!exact_tstop && tn_ > t_final && nhist_ >= 1) {
    // interpolate the step polynomial back to t_final
    const int q = nhist_;
    const double P[3] = {tn_, tnm1_, tnm2_};
    const double* Y[3] = {yn_, ynm1_, ynm2_};
    double L[3] = {0, 0, 0};
    for (int j = 0; j <= q; j++) {
      double w = 1.0;
      for (int i = 0; i <= q; i++)
        if (i != j) w *= (t_final - P[i]) / (P[j] - P[i]);
      L[j] = w;
    }
    HTRY(vec_lincomb(s_, n3_, L[0], Y[0], L[1], Y[1], q >= 2 ? L[2] : 0.0,
                     q >= 2 ? Y[2] : nullptr, 0, nullptr, m));
    *t_reached = t_final;
  } else {
    HTRY(hipMemcpyAsync(m, yn_, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
    *t_reached = tn_;
  }
  int np = 0;
  HTRY(vec_maxdist_partials(s_, nn_, m, ystart_, part_, &np));
  HTRY(hipMemcpyAsync(host_part_.data(), part_, sizeof(double) * np, hipMemcpyDeviceToHost, s_));
  HTRY(hipStreamSynchronize(s_));
  double mx = 0.0;
  for (int b = 0; b < np; b++) mx = std::max(mx, host_part_[b]);
  st_.max_dm = mx;
  return 0;
}

}  // namespace nsim
