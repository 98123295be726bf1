// Step controller of the BDF integrator (order <= 2, Newton with GMRES), as
// __host__ __device__ code on one state record: the device runs it in a
// one-thread kernel after every Newton iteration, so that whole steps are
// queued without host round trips.  It stands in for the parts of CVODE the
// reference drives (src/nsim.ml:2485-2580, src/sundials_sp_stubs.c:700-1308):
// variable-step BDF coefficients, predictor and error constant, Newton
// convergence test with a carried rate estimate, error test and step-size
// selection, and the stop conditions of an advance call.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace nsim {

constexpr int kMaxCor = 3;           // Newton iterations per attempt (CVODE NLS_MAXCOR)
constexpr double kNewtonCoef = 0.1;  // Newton convergence coefficient
constexpr double kLinCoef = 0.05;    // linear-solve tolerance factor (CVODE eplifac)
constexpr int kMaxAttempts = 20;     // failed attempts of one step before giving up

enum TsMode : int { kTsBegin = 0, kTsNewton = 1, kTsDone = 2, kTsError = 3 };
enum TsSlot : int { kSlotNone = -1, kSlotF = -2 };  // predictor sources besides the ring

// Everything the controller knows.  The history y_n, y_{n-1}, y_{n-2} lives in
// a ring of three device vectors addressed by slot index.
struct TsState {
  // history
  double tn, tnm1, tnm2;
  int nhist;           // valid previous points (0..2)
  int sn, snm1, snm2;  // ring slots of y_n, y_{n-1}, y_{n-2}
  double hnext;        // <= 0: the first step size is still to be chosen
  int qnext, max_order;
  double crate, gamma_crate;  // Newton rate estimate, carried while gamma is similar
  double rtol, atol;
  // the current attempt
  double h, gamma, C;  // step, Newton matrix coefficient, error-estimate factor
  int q, k, fails;     // order, Newton iteration, failed attempts of this step
  double del_old, crate_cur;
  double L[3];         // predictor yp = L0 y_n + L1 src(p1) + L2 src(p2)
  int p1, p2;
  double a1, a2;       // history term avec = a1 y_n + a2 src(a2s)
  int a2s;
  int copy;            // ring slot that receives ycur at the next begin (-1: none)
  int mode;            // TsMode of the next cycle
  int halt;            // nonzero: the queued kernels return at once
  int err_code;        // 1: a GMRES exchange failed, 2: step size underflow
  // the advance call
  double t_final;
  int exact_tstop, max_steps, steps_call;
  // statistics
  long long steps, rhs, newton, linear, err_fails, conv_fails;
  double h_last;
  int q_last;
};

// l_j'(T[0]) of the Lagrange basis polynomial of node j among T[0..m]
// (m <= 2).  The loops have fixed bounds so that, unrolled at a call with a
// constant j, the small arrays stay in registers.
__host__ __device__ inline double ts_lagrange_deriv0(const double* T, int m, int j) {
  if (j == 0) {
    double s = 0.0;
#pragma unroll
    for (int i = 1; i < 3; i++)
      if (i <= m) s += 1.0 / (T[0] - T[i]);
    return s;
  }
  double num = 1.0, den = 1.0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (i > m || i == j) continue;
    den *= T[j] - T[i];
    if (i != 0) num *= T[0] - T[i];
  }
  return num / den;
}

__host__ __device__ inline double ts_factorial(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; i++) f *= i;
  return f;
}

// Finite (neither NaN nor +-inf), usable from host and device code alike.
__host__ __device__ inline bool ts_finite(double v) { return fabs(v) <= DBL_MAX; }

// d^(q+1) and err^(-1/(q+1)) for the orders built (q = 1, 2): products,
// sqrt and cbrt instead of the general power function.
__host__ __device__ inline double ts_pow_q1(double d, int q) {
  return q >= 2 ? d * d * d : d * d;
}
__host__ __device__ inline double ts_root_q1(double err, int q) {
  return q >= 2 ? 1.0 / cbrt(err) : 1.0 / sqrt(err);
}

__host__ __device__ inline bool ts_finished(const TsState& s) {
  if (s.max_steps > 0 && s.steps_call >= s.max_steps) return true;
  const double eps_t = 1e-12 * fmax(1.0, fabs(s.t_final));
  return s.exact_tstop ? (s.t_final - s.tn <= eps_t) : (s.tn >= s.t_final);
}

// Coefficients of an attempt of step h at order q from the history: the BDF
// corrector y - gamma f(y) - avec = 0 with the exact variable-step
// coefficients, the extrapolation predictor and the Milne error factor C.
__host__ __device__ inline void ts_prepare(TsState& s, double h, int q) {
  if (s.exact_tstop && s.tn + h > s.t_final - 0.01 * h) h = s.t_final - s.tn;
  s.h = h;
  s.q = q;
  s.k = 0;
  s.del_old = 0.0;
  const double T[3] = {s.tn + h, s.tn, s.tnm1};
  double dl[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 3; j++)
    if (j <= q) dl[j] = ts_lagrange_deriv0(T, q, j);
  s.gamma = 1.0 / dl[0];
  double R = 0.0;
#pragma unroll
  for (int j = 1; j < 3; j++)
    if (j <= q) R += dl[j] * ts_pow_q1(T[j] - T[0], q) / ts_factorial(q + 1);
  const double E_c = -R / dl[0];
  double E_p;
  if (s.nhist == 0) {  // explicit Euler from f(y_n)
    s.L[0] = 1.0;
    s.L[1] = h;
    s.L[2] = 0.0;
    s.p1 = kSlotF;
    s.p2 = kSlotNone;
    E_p = 0.5 * h * h;
  } else {
    const double t1 = s.tn + h;
    const double P[3] = {s.tn, s.tnm1, s.tnm2};
    double L[3] = {0.0, 0.0, 0.0}, ep = 1.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j > q) continue;
      double w = 1.0;
#pragma unroll
      for (int i = 0; i < 3; i++)
        if (i <= q && i != j) w *= (t1 - P[i]) / (P[j] - P[i]);
      L[j] = w;
      ep *= t1 - P[j];
    }
    s.L[0] = L[0];
    s.L[1] = L[1];
    s.L[2] = q >= 2 ? L[2] : 0.0;
    s.p1 = s.snm1;
    s.p2 = q >= 2 ? s.snm2 : kSlotNone;
    E_p = ep / ts_factorial(q + 1);
  }
  s.C = E_c / (E_c + E_p);
  s.a1 = -s.gamma * dl[1];
  s.a2 = q >= 2 ? -s.gamma * dl[2] : 0.0;
  s.a2s = q >= 2 ? s.snm1 : kSlotNone;
  // the rate estimate carries over between steps (as CVODE's crate does)
  // unless gamma changed a lot, so a small first correction can be accepted
  // without a second rhs evaluation
  if (s.gamma_crate <= 0.0 || fabs(s.gamma / s.gamma_crate - 1.0) > 0.3) s.crate = 1.0;
  s.gamma_crate = s.gamma;
  s.crate_cur = s.crate;
  s.mode = kTsBegin;
}

__host__ __device__ inline void ts_reject(TsState& s) {
  s.fails++;
  if (s.fails >= kMaxAttempts || s.hnext < 1e-14 * fmax(1.0, fabs(s.tn))) {
    s.mode = kTsError;
    s.err_code = 2;
    s.halt = 1;
    return;
  }
  ts_prepare(s, s.hnext, s.fails >= 2 ? 1 : s.q);
}

__host__ __device__ inline void ts_accept(TsState& s, double err) {
  const int oldest = s.snm2;  // receives the accepted iterate at the next begin
  s.snm2 = s.snm1;
  s.snm1 = s.sn;
  s.sn = oldest;
  s.copy = oldest;
  s.tnm2 = s.tnm1;
  s.tnm1 = s.tn;
  s.tn = s.tn + s.h;
  s.nhist = s.nhist + 1 < 2 ? s.nhist + 1 : 2;
  s.steps++;
  s.steps_call++;
  s.h_last = s.h;
  s.q_last = s.q;
  const double etamax = s.steps <= 10 ? 100.0 : 10.0;
  double eta = err > 0.0 ? 0.9 * ts_root_q1(err, s.q) : etamax;
  eta = fmin(etamax, fmax(0.2, eta));
  if (eta > 1.0 && eta < 1.5) eta = 1.0;
  s.hnext = s.h * eta;
  const int qn = s.nhist > 1 ? s.nhist : 1;
  s.qnext = qn < s.max_order ? qn : s.max_order;
  s.fails = 0;
  if (ts_finished(s)) {
    s.mode = kTsDone;
    s.halt = 1;
    return;
  }
  ts_prepare(s, s.hnext, s.qnext);
}

// After a Newton iteration of the current attempt: d is the correction's WRMS
// norm, errsq the weighted sum of squares of ycur - yp (n3 entries), its the
// Krylov iterations.  Decides: another Newton iteration, accept, or reject.
__host__ __device__ inline void ts_after_newton(TsState& s, double d, double errsq, int its,
                                                int n3) {
  s.copy = kSlotNone;  // the begin of this attempt has run
  s.rhs++;
  s.newton++;
  s.linear += its;
  bool converged = false, failed = false;
  if (!ts_finite(d)) {
    failed = true;
  } else {
    if (s.k > 0) s.crate_cur = fmax(0.3 * s.crate_cur, d / s.del_old);
    if (d * fmin(1.0, s.crate_cur) * s.C <= kNewtonCoef)
      converged = true;
    else if ((s.k > 0 && d > 2.0 * s.del_old) || s.k + 1 >= kMaxCor)
      failed = true;
  }
  if (!converged && !failed) {
    s.del_old = d;
    s.k++;
    s.mode = kTsNewton;
    return;
  }
  s.crate = s.crate_cur;
  if (failed) {
    s.conv_fails++;
    s.crate = 1.0;
    s.hnext = 0.25 * s.h;
    ts_reject(s);
    return;
  }
  const double err = s.C * sqrt(errsq / n3);
  if (!ts_finite(err)) {
    s.conv_fails++;
    s.hnext = 0.25 * s.h;
    ts_reject(s);
    return;
  }
  if (err > 1.0) {
    s.err_fails++;
    s.hnext = s.h * fmax(0.2, 0.9 * ts_root_q1(err, s.q));
    ts_reject(s);
    return;
  }
  ts_accept(s, err);
}

// The controller's step after a Newton iteration whose results are its
// Krylov iterations, the correction's weighted sum of squares xw2, an error
// flag and errsq (as ts_after_newton).  One thread, on a register copy of the
// state written back whole.  Nothing to do once the run has stopped.
__host__ __device__ inline void ts_control(TsState* s, double its, double xw2, double bad,
                                           double errsq, int n3) {
  TsState st = *s;
  if (st.mode == kTsDone || st.mode == kTsError || st.halt) return;
  if (bad != 0.0) {
    st.mode = kTsError;
    st.err_code = 1;
    st.halt = 1;
  } else {
    ts_after_newton(st, sqrt(xw2 / n3), errsq, (int)its, n3);
  }
  *s = st;
}

// Launchers (nsim_tsctl.hip).
// Start of an attempt when s->mode == kTsBegin: the accepted iterate into
// its ring slot, error weights of y_n, predictor yp = ycur, history term avec.
hipError_t launch_ts_begin(hipStream_t st, int n3, const TsState* s, double* const* slots,
                           const double* fn, double* ycur, double* yp, double* avec,
                           double* ewt);
// The controller after a Newton iteration whose results are in status
// ([0] iterations, [1] sum (x ewt)^2, [2] error flag, [3] sum ((y - yp) ewt)^2).
hipError_t launch_ts_ctrl(hipStream_t st, TsState* s, const double* status, int n3);

}  // namespace nsim
