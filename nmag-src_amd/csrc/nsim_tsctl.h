// Step controller of the BDF integrator (orders 1 to 5, Newton with GMRES),
// as __host__ __device__ code on one state record: the device runs it at the
// end of every Newton iteration, so that whole steps are queued without host
// round trips.  It stands in for the parts of CVODE the reference drives
// (src/nsim.ml:2485-2580, src/sundials_sp_stubs.c:700-1308): variable-step
// BDF coefficients, predictor and error constant, Newton convergence test with
// a carried rate estimate, error test, step-size and order selection, and the
// stop conditions of an advance call.
//
// Order: with a maximum order of 1 or 2 the order rises to the maximum as the
// history allows (the reference's default maximum is 2, main.py:838).  With a
// maximum of 3 to 5 the order is chosen as CVODE chooses it: after q + 1 steps
// at order q, the local errors that orders q - 1 and q + 1 would have made on
// the step just taken are estimated from divided differences of the history,
// and the order whose step-size factor is largest is taken (biases 6, 6 and
// 10 against q - 1, q and q + 1).  Step-size growth at orders 3 to 5 is
// capped where variable-step BDF stays zero-stable.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace nsim {

constexpr int kMaxCor = 3;           // Newton iterations per attempt (CVODE NLS_MAXCOR)
constexpr double kNewtonCoef = 0.1;  // Newton convergence coefficient
constexpr double kLinCoef = 0.05;    // linear-solve tolerance factor (CVODE eplifac)
constexpr int kMaxAttempts = 20;     // failed attempts of one step before giving up
constexpr int kMaxOrder = 5;         // CVODE's largest BDF order
constexpr int kHist = kMaxOrder + 1; // history vectors y_n ... y_{n-5} (a ring)

enum TsMode : int { kTsBegin = 0, kTsNewton = 1, kTsDone = 2, kTsError = 3 };
enum TsSlot : int { kSlotNone = -1, kSlotF = -2 };  // predictor sources besides the ring

// Everything the controller knows.  The history y_n, y_{n-1}, ... lives in a
// ring of kHist device vectors addressed by slot index.
struct TsState {
  // history: t_{n-i} and the ring slot of y_{n-i} for i = 0 .. nhist
  double th[kHist];
  int sh[kHist];
  int nhist;           // valid previous points besides y_n (0 .. kHist - 1)
  double hnext;        // <= 0: the first step size is still to be chosen
  int qnext, max_order;
  int select;          // order selection (max_order >= 3)
  int qwait;           // accepted steps before an order change is considered
  double crate, gamma_crate;  // Newton rate estimate, carried while gamma is similar
  double rtol, atol;
  // the current attempt
  double h, gamma, C;  // step, Newton matrix coefficient, error-estimate factor
  int q, k, fails;     // order, Newton iteration, failed attempts of this step
  double del_old, crate_cur;
  double L[kHist];     // predictor yp = sum_j L[j] src(p[j]), j <= q
  int p[kHist];
  double A[kMaxOrder]; // history term avec = sum_j A[j] y_{n-j}, j < q
  int as[kMaxOrder];
  // order-selection estimates of this attempt: the local errors of orders
  // q - 1 and q + 1 are the WRMS norms of Em[0] ycur + sum_{i=1..em} Em[i]
  // y_{n+1-i} and of the same with Ep (em, ep: history vectors used; 0: none)
  double Em[kHist + 1], Ep[kHist + 1];
  int em, ep;
  int copy;            // ring slot that receives ycur at the next begin (-1: none)
  int mode;            // TsMode of the next cycle
  int halt;            // nonzero: the queued kernels return at once
  int err_code;        // 1: a GMRES exchange failed, 2: step size underflow
  // the advance call
  double t_final;
  int exact_tstop, max_steps, steps_call;
  // statistics
  long long steps, rhs, newton, linear, err_fails, conv_fails;
  long long linear_max;  // most Krylov iterations of one Newton iteration
  long long order_steps[kMaxOrder + 1];  // accepted steps per order
  double h_last;
  int q_last;
};

// l_j'(T[0]) of the Lagrange basis polynomial of node j among T[0..m]
// (m <= kMaxOrder).  Fixed loop bounds, so that the small arrays stay in
// registers.
__host__ __device__ inline double ts_lagrange_deriv0(const double* T, int m, int j) {
  if (j == 0) {
    double s = 0.0;
#pragma unroll
    for (int i = 1; i <= kMaxOrder; i++)
      if (i <= m) s += 1.0 / (T[0] - T[i]);
    return s;
  }
  double num = 1.0, den = 1.0;
#pragma unroll
  for (int i = 0; i <= kMaxOrder; i++) {
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

// d^(q+1) and err^(-1/(q+1)): products, sqrt and cbrt for orders 1 and 2,
// the power function above.
__host__ __device__ inline double ts_pow_q1(double d, int q) {
  if (q == 1) return d * d;
  if (q == 2) return d * d * d;
  return pow(d, q + 1.0);
}
__host__ __device__ inline double ts_root_q1(double err, int q) {
  if (q == 1) return 1.0 / sqrt(err);
  if (q == 2) return 1.0 / cbrt(err);
  return pow(err, -1.0 / (q + 1.0));
}

__host__ __device__ inline bool ts_finished(const TsState& s) {
  if (s.max_steps > 0 && s.steps_call >= s.max_steps) return true;
  const double eps_t = 1e-12 * fmax(1.0, fabs(s.t_final));
  return s.exact_tstop ? (s.t_final - s.th[0] <= eps_t) : (s.th[0] >= s.t_final);
}

// Weights of the scaled divided difference over T[0..m] (m + 1 nodes):
// c k! h^(k+1) f[T_0 .. T_m] = sum_i W[i] y(T_i), m = k + 1, with c the
// error constant of BDF-k, beta_k / (k + 1), beta_k = 1 / (1 + 1/2 + ... + 1/k):
// the local error that a step of order k would make.
__host__ __device__ inline void ts_error_weights(const double* T, int m, double h, double* W) {
  const int k = m - 1;
  double hk = 0.0;
  for (int i = 1; i <= k; i++) hk += 1.0 / i;
  const double sc = ts_factorial(k) / hk / (k + 1) * ts_pow_q1(h, k) * (k + 1);
#pragma unroll
  for (int i = 0; i <= kHist; i++) {
    if (i > m) {
      W[i] = 0.0;
      continue;
    }
    double den = 1.0;
#pragma unroll
    for (int l = 0; l <= kHist; l++)
      if (l <= m && l != i) den *= T[i] - T[l];
    W[i] = sc / den;
  }
}

// Coefficients of an attempt of step h at order q from the history: the BDF
// corrector y - gamma f(y) - avec = 0 with the exact variable-step
// coefficients, the extrapolation predictor and the Milne error factor C.
// The record is read once into locals and written once at the end (on the
// device the controller works on it in global memory, where interleaved reads
// and writes would each wait for the last).
__host__ __device__ inline void ts_prepare(TsState& s, double h, int q) {
  const double tn = s.th[0], t_final = s.t_final;
  const int nhist = s.nhist, select = s.select, max_order = s.max_order;
  if (s.exact_tstop && tn + h > t_final - 0.01 * h) h = t_final - tn;
  double T[kHist + 1];  // t_{n+1}, t_n, t_{n-1}, ...
  int sh[kHist];
  T[0] = tn + h;
#pragma unroll
  for (int i = 0; i < kHist; i++) {
    T[i + 1] = s.th[i];
    sh[i] = s.sh[i];
  }
  double dl[kMaxOrder + 1];
#pragma unroll
  for (int j = 0; j <= kMaxOrder; j++) dl[j] = j <= q ? ts_lagrange_deriv0(T, q, j) : 0.0;
  const double gamma = 1.0 / dl[0];
  double R = 0.0;
#pragma unroll
  for (int j = 1; j <= kMaxOrder; j++)
    if (j <= q) R += dl[j] * ts_pow_q1(T[j] - T[0], q) / ts_factorial(q + 1);
  const double E_c = -R / dl[0];
  double E_p;
  double L[kHist];
  int P[kHist];
#pragma unroll
  for (int j = 0; j < kHist; j++) {
    L[j] = 0.0;
    P[j] = kSlotNone;
  }
  if (nhist == 0) {  // explicit Euler from f(y_n)
    L[0] = 1.0;
    P[0] = sh[0];
    L[1] = h;
    P[1] = kSlotF;
    E_p = 0.5 * h * h;
  } else {
    const double t1 = T[0];
    double ep = 1.0;
#pragma unroll
    for (int j = 0; j < kHist; j++) {
      if (j > q) continue;
      double w = 1.0;
#pragma unroll
      for (int i = 0; i < kHist; i++)
        if (i <= q && i != j) w *= (t1 - T[i + 1]) / (T[j + 1] - T[i + 1]);
      L[j] = w;
      P[j] = sh[j];
      ep *= t1 - T[j + 1];
    }
    E_p = ep / ts_factorial(q + 1);
  }
  // order-selection estimates (orders q - 1 and q + 1)
  double Em[kHist + 1], Ep[kHist + 1];
  int em = 0, epn = 0;
  if (select && q >= 2 && nhist >= q - 1) {
    ts_error_weights(T, q, h, Em);
    em = q;
  }
  if (select && q < max_order && nhist >= q + 1) {
    ts_error_weights(T, q + 2, h, Ep);
    epn = q + 2;
  }
  // the rate estimate carries over between steps (as CVODE's crate does)
  // unless gamma changed a lot, so a small first correction can be accepted
  // without a second rhs evaluation
  const double gamma_crate = s.gamma_crate;
  double crate = s.crate;
  if (gamma_crate <= 0.0 || fabs(gamma / gamma_crate - 1.0) > 0.3) crate = 1.0;
  s.h = h;
  s.q = q;
  s.k = 0;
  s.del_old = 0.0;
  s.gamma = gamma;
  s.C = E_c / (E_c + E_p);
#pragma unroll
  for (int j = 0; j < kHist; j++) {
    s.L[j] = L[j];
    s.p[j] = P[j];
  }
#pragma unroll
  for (int j = 0; j < kMaxOrder; j++) {
    s.A[j] = j < q ? -gamma * dl[j + 1] : 0.0;
    s.as[j] = j < q ? sh[j] : kSlotNone;
  }
  s.em = em;
  s.ep = epn;
#pragma unroll
  for (int j = 0; j <= kHist; j++) {
    s.Em[j] = em ? Em[j] : 0.0;
    s.Ep[j] = epn ? Ep[j] : 0.0;
  }
  s.crate = crate;
  s.gamma_crate = gamma;
  s.crate_cur = crate;
  s.mode = kTsBegin;
}

__host__ __device__ inline void ts_reject(TsState& s) {
  s.fails++;
  if (s.fails >= kMaxAttempts || s.hnext < 1e-14 * fmax(1.0, fabs(s.th[0]))) {
    s.mode = kTsError;
    s.err_code = 2;
    s.halt = 1;
    return;
  }
  const int q = s.fails >= 2 ? 1 : s.q;
  if (q != s.q) s.qwait = q + 1;
  ts_prepare(s, s.hnext, q);
}

// Largest step-size growth factor at order q: variable-step BDF of order 3, 4
// and 5 stays zero-stable for step ratios below about 1.6, 1.3 and 1.15.
__host__ __device__ inline double ts_eta_cap(int q, double etamax) {
  if (q == 3) return fmin(etamax, 1.5);
  if (q == 4) return fmin(etamax, 1.25);
  if (q >= 5) return fmin(etamax, 1.1);
  return etamax;
}

// errm, errp: the estimated local errors of orders q - 1 and q + 1 on this
// step (negative: not estimated).
__host__ __device__ inline void ts_accept(TsState& s, double err, double errm, double errp) {
  const int oldest = s.sh[kHist - 1];  // receives the accepted iterate at the next begin
#pragma unroll
  for (int i = kHist - 1; i > 0; i--) {
    s.sh[i] = s.sh[i - 1];
    s.th[i] = s.th[i - 1];
  }
  s.sh[0] = oldest;
  s.copy = oldest;
  s.th[0] = s.th[1] + s.h;
  s.nhist = s.nhist + 1 < kHist - 1 ? s.nhist + 1 : kHist - 1;
  s.steps++;
  s.steps_call++;
  s.h_last = s.h;
  s.q_last = s.q;
  if (s.q >= 1 && s.q <= kMaxOrder) s.order_steps[s.q]++;
  const double etamax = s.steps <= 10 ? 100.0 : 10.0;
  double eta = err > 0.0 ? 0.9 * ts_root_q1(err, s.q) : etamax;
  int qn = s.q;
  if (!s.select || s.q < 2) {
    // ramp: the order rises to the maximum as the history allows
    qn = s.nhist > 1 ? s.nhist : 1;
    if (qn > s.max_order) qn = s.max_order;
    if (s.select && qn > 2) qn = 2;  // above 2 the order is chosen below
    if (qn != s.q) s.qwait = qn + 1;
  } else if (s.qwait > 0 && --s.qwait == 0) {
    const double bias = 1e-6;
    const double etaq = 1.0 / (pow(6.0 * fmax(err, 1e-300), 1.0 / (s.q + 1)) + bias);
    const double etam = errm >= 0.0 ? 1.0 / (pow(6.0 * fmax(errm, 1e-300), 1.0 / s.q) + bias)
                                    : 0.0;
    const double etap = errp >= 0.0 && s.q < s.max_order
                            ? 1.0 / (pow(10.0 * fmax(errp, 1e-300), 1.0 / (s.q + 2)) + bias)
                            : 0.0;
    // compared as the growth each order may actually take: a higher order
    // whose step is held back by its zero-stability cap loses to a lower one
    const double cq = fmin(etaq, ts_eta_cap(s.q, etamax));
    const double cm = fmin(etam, ts_eta_cap(s.q - 1, etamax));
    const double cp = fmin(etap, ts_eta_cap(s.q + 1, etamax));
    if (cm > cq && cm >= cp) {
      qn = s.q - 1;
      eta = etam;
    } else if (cp > cq) {
      qn = s.q + 1;
      eta = etap;
    }
    s.qwait = qn != s.q ? qn + 1 : 2;
  }
  eta = fmin(ts_eta_cap(qn, etamax), fmax(0.2, eta));
  if (eta > 1.0 && eta < (qn >= 3 ? 1.05 : 1.5)) eta = 1.0;
  s.hnext = s.h * eta;
  s.qnext = qn;
  s.fails = 0;
  if (ts_finished(s)) {
    s.mode = kTsDone;
    s.halt = 1;
    return;
  }
  ts_prepare(s, s.hnext, s.qnext);
}

// After a Newton iteration of the current attempt: d is the correction's WRMS
// norm; errsq the weighted sum of squares of ycur - yp and errsq_m, errsq_p
// those of the order q - 1 and q + 1 estimates (n3 entries); its the Krylov
// iterations.  Decides: another Newton iteration, accept, or reject.
__host__ __device__ inline void ts_after_newton(TsState& s, double d, double errsq,
                                                double errsq_m, double errsq_p, int its,
                                                int n3) {
  s.copy = kSlotNone;  // the begin of this attempt has run
  s.rhs++;
  s.newton++;
  s.linear += its;
  if (its > s.linear_max) s.linear_max = its;
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
  const double errm = s.em > 0 && ts_finite(errsq_m) ? sqrt(errsq_m / n3) : -1.0;
  const double errp = s.ep > 0 && ts_finite(errsq_p) ? sqrt(errsq_p / n3) : -1.0;
  ts_accept(s, err, errm, errp);
}

// The controller's step after a Newton iteration whose results are its
// Krylov iterations, the correction's weighted sum of squares xw2, an error
// flag and errsq, errsq_m, errsq_p (as ts_after_newton), on the record st in
// place.  One thread.  Nothing to do once the run has stopped.
__host__ __device__ inline void ts_control_at(TsState& st, double its, double xw2, double bad,
                                              double errsq, double errsq_m, double errsq_p,
                                              int n3) {
  if (st.mode == kTsDone || st.mode == kTsError || st.halt) return;
  if (bad != 0.0) {
    st.mode = kTsError;
    st.err_code = 1;
    st.halt = 1;
  } else {
    ts_after_newton(st, sqrt(xw2 / n3), errsq, errsq_m, errsq_p, (int)its, n3);
  }
}

// The same on a register copy of *s, written back whole (the loads in flight
// together rather than one field at a time).
__host__ __device__ inline void ts_control(TsState* s, double its, double xw2, double bad,
                                           double errsq, double errsq_m, double errsq_p,
                                           int n3) {
  TsState st = *s;
  ts_control_at(st, its, xw2, bad, errsq, errsq_m, errsq_p, n3);
  *s = st;
}

// Launchers (nsim_tsctl.hip).
// Start of an attempt when s->mode == kTsBegin: the accepted iterate into
// its ring slot, error weights of y_n, predictor yp = ycur, history term avec.
hipError_t launch_ts_begin(hipStream_t st, int n3, const TsState* s, double* const* slots,
                           const double* fn, double* ycur, double* yp, double* avec,
                           double* ewt);
// The controller after a Newton iteration whose results are in status
// ([0] iterations, [1] sum (x ewt)^2, [2] error flag, [3] sum ((y - yp) ewt)^2,
// [4], [5] the sums of squares of the order q - 1 and q + 1 estimates).
hipError_t launch_ts_ctrl(hipStream_t st, TsState* s, const double* status, int n3);

}  // namespace nsim
