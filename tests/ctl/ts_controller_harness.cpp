// Host driver of the BDF step controller (nmag-src_amd/csrc/nsim_tsctl.h):
// runs the controller's own host functions through forced scenarios, with
// the Newton results made up on the host, and through whole integrations of
// a linear ODE y' = A y whose Newton systems are solved exactly on the host.
// Prints one JSON object per scenario.  No HIP runtime call is made, so it
// runs without a GPU.
#include <cmath>
#include <cstdio>
#include <limits>

#include "nsim_tsctl.h"

using namespace nsim;

namespace {

TsState fresh(double t_final, int max_steps, bool exact, int max_order = 2,
              bool select = false) {
  TsState s{};
  for (int i = 0; i < kHist; i++) {
    s.th[i] = 0.0;
    s.sh[i] = i;
  }
  s.nhist = 0;
  s.copy = kSlotNone;
  s.max_order = max_order;
  s.select = select ? 1 : 0;
  s.qwait = 2;
  s.rtol = s.atol = 1e-6;
  s.t_final = t_final;
  s.exact_tstop = exact ? 1 : 0;
  s.max_steps = max_steps;
  s.steps_call = 0;
  s.gamma_crate = 0.0;
  ts_prepare(s, 0.01, 1);
  return s;
}

// Drives cycles until the controller stops or `limit` cycles pass.  Each
// cycle reports the Newton correction norm d and the error sum errsq that
// the scenario supplies.
template <typename F>
void run(const char* name, TsState s, int limit, F newton) {
  int cycles = 0;
  const int n3 = 30;
  while (cycles < limit && s.mode != kTsDone && s.mode != kTsError) {
    double d, errsq;
    newton(s, &d, &errsq);
    ts_after_newton(s, d, errsq, 0.0, 0.0, 2, n3);
    cycles++;
  }
  std::printf("{\"name\": \"%s\", \"cycles\": %d, \"mode\": %d, \"err_code\": %d, "
              "\"halt\": %d, \"steps\": %lld, \"steps_call\": %d, \"tn\": %.17g, "
              "\"t_final\": %.17g, \"fails\": %d, \"err_fails\": %lld, "
              "\"conv_fails\": %lld, \"h\": %.17g, \"q_last\": %d}\n",
              name, cycles, (int)s.mode, s.err_code, s.halt, s.steps, s.steps_call, s.th[0],
              s.t_final, s.fails, s.err_fails, s.conv_fails, s.h, s.q_last);
}

// y' = A y in three components: a damped rotation and a decay,
//   A = [[-a, -w, 0], [w, -a, 0], [0, 0, -b]],
// exact solution e^{-a t} R(w t) y0 and e^{-b t} y0_z.  The controller
// integrates it as the device integrator does (k_ts_begin, Newton on
// y - gamma A y - avec = 0 solved exactly, the error sums of the GMRES
// kernel's last pass), on a ring of kHist vectors.
struct Linear {
  double a = 0.3, w = 2.0, b = 1.5;
  void f(const double* y, double* out) const {
    out[0] = -a * y[0] - w * y[1];
    out[1] = w * y[0] - a * y[1];
    out[2] = -b * y[2];
  }
  // (I - g A) y = r
  void solve(double g, const double* r, double* y) const {
    const double p = 1.0 + g * a, q = g * w;  // [[p, q], [-q, p]]
    const double det = p * p + q * q;
    y[0] = (p * r[0] - q * r[1]) / det;
    y[1] = (q * r[0] + p * r[1]) / det;
    y[2] = r[2] / (1.0 + g * b);
  }
  void exact(double t, const double* y0, double* y) const {
    const double e = std::exp(-a * t), c = std::cos(w * t), s = std::sin(w * t);
    y[0] = e * (c * y0[0] - s * y0[1]);
    y[1] = e * (s * y0[0] + c * y0[1]);
    y[2] = std::exp(-b * t) * y0[2];
  }
};

void integrate(const char* name, int max_order, bool select, double tol, double T) {
  const Linear L;
  const double y0[3] = {1.0, 0.5, 2.0};
  double ring[kHist][3] = {};
  double fn[3], ycur[3], yp[3], avec[3], ewt[3];
  for (int c = 0; c < 3; c++) ring[0][c] = y0[c];
  TsState s = fresh(T, -1, true, max_order, select);
  s.rtol = s.atol = tol;
  // first step from |f(y0)|, as the integrator's advance does
  L.f(y0, fn);
  double fw = 0.0;
  for (int c = 0; c < 3; c++) {
    const double wc = 1.0 / (tol * std::fabs(y0[c]) + tol);
    fw += fn[c] * wc * fn[c] * wc;
  }
  fw = std::sqrt(fw / 3);
  s.rhs = 1;
  ts_prepare(s, std::min(0.1 / fw, 0.1 * T), 1);
  int cycles = 0;
  auto src = [&](int slot) -> const double* { return slot == kSlotF ? fn : ring[slot]; };
  while (s.mode != kTsDone && s.mode != kTsError && cycles < 2000000) {
    if (s.mode == kTsBegin) {  // k_ts_begin
      const double* y = s.copy >= 0 ? ycur : ring[s.sh[0]];
      double yn[3] = {y[0], y[1], y[2]};
      if (s.copy >= 0)
        for (int c = 0; c < 3; c++) ring[s.copy][c] = yn[c];
      for (int c = 0; c < 3; c++) {
        ewt[c] = 1.0 / (s.rtol * std::fabs(yn[c]) + s.atol);
        double v = s.L[0] * yn[c], av = s.A[0] * yn[c];
        for (int j = 1; j < kHist; j++)
          if (s.p[j] != kSlotNone) v += s.L[j] * src(s.p[j])[c];
        for (int j = 1; j < kMaxOrder; j++)
          if (s.as[j] != kSlotNone) av += s.A[j] * ring[s.as[j]][c];
        yp[c] = v;
        ycur[c] = v;
        avec[c] = av;
      }
    }
    // Newton iteration, exact for a linear f: y = (I - gamma A)^-1 avec
    double y[3], d2 = 0.0, e2 = 0.0, m2 = 0.0, p2 = 0.0;
    L.solve(s.gamma, avec, y);
    for (int c = 0; c < 3; c++) {
      const double dx = (y[c] - ycur[c]) * ewt[c];
      d2 += dx * dx;
      ycur[c] = y[c];
      const double de = (ycur[c] - yp[c]) * ewt[c];
      e2 += de * de;
      if (s.em > 0) {
        double t = s.Em[0] * ycur[c];
        for (int i = 1; i <= s.em; i++) t += s.Em[i] * ring[s.sh[i - 1]][c];
        m2 += t * ewt[c] * t * ewt[c];
      }
      if (s.ep > 0) {
        double t = s.Ep[0] * ycur[c];
        for (int i = 1; i <= s.ep; i++) t += s.Ep[i] * ring[s.sh[i - 1]][c];
        p2 += t * ewt[c] * t * ewt[c];
      }
    }
    ts_control_at(s, 1.0, d2, 0.0, e2, m2, p2, 3);
    cycles++;
  }
  // the last accepted iterate
  const double* yT = s.copy >= 0 ? ycur : ring[s.sh[0]];
  double ex[3], err = 0.0, ref = 0.0;
  L.exact(s.th[0], y0, ex);
  for (int c = 0; c < 3; c++) {
    err = std::fmax(err, std::fabs(yT[c] - ex[c]));
    ref = std::fmax(ref, std::fabs(ex[c]));
  }
  std::printf("{\"name\": \"%s\", \"max_order\": %d, \"select\": %d, \"tol\": %.3g, "
              "\"mode\": %d, \"tn\": %.17g, \"steps\": %lld, \"rhs\": %lld, "
              "\"err_fails\": %lld, \"conv_fails\": %lld, \"err\": %.6g, \"q_last\": %d, "
              "\"order_steps\": [%lld, %lld, %lld, %lld, %lld]}\n",
              name, max_order, select ? 1 : 0, tol, (int)s.mode, s.th[0], s.steps, s.rhs,
              s.err_fails, s.conv_fails, err / ref, s.q_last, s.order_steps[1],
              s.order_steps[2], s.order_steps[3], s.order_steps[4], s.order_steps[5]);
}

}  // namespace

int main() {
  // every attempt converges at once but fails the error test
  run("error_test_always_fails", fresh(1.0, -1, true), 1000, [](const TsState&, double* d,
                                                                 double* e) {
    *d = 1e-9;
    *e = 1e30;
  });
  // Newton corrections grow: convergence failures, h / 4 each time
  run("newton_diverges", fresh(1.0, -1, true), 1000, [](const TsState& s, double* d, double* e) {
    *d = 10.0 * (s.k + 1);
    *e = 0.0;
  });
  // a NaN correction norm
  run("nan_correction", fresh(1.0, -1, true), 1000, [](const TsState&, double* d, double* e) {
    *d = std::numeric_limits<double>::quiet_NaN();
    *e = 0.0;
  });
  // well-behaved: small corrections and errors, exact stop at t_final
  run("smooth_exact_stop", fresh(1.0, -1, true), 100000, [](const TsState& s, double* d,
                                                            double* e) {
    *d = 1e-6;
    *e = 30.0 * 1e-4 * s.h;  // error grows with the step
  });
  // the same with a step budget that ends the call before t_final
  run("smooth_max_steps", fresh(1e9, 7, true), 100000, [](const TsState& s, double* d,
                                                          double* e) {
    *d = 1e-6;
    *e = 30.0 * 1e-4 * s.h;
  });
  // two rejections, then acceptance: the order drops to 1 after two failures
  {
    int calls = 0;
    TsState s = fresh(1e9, 1, true);
    // start from a history of two points at order 2
    s.nhist = 2;
    s.th[1] = -0.01;
    s.th[2] = -0.02;
    ts_prepare(s, 0.01, 2);
    run("two_rejections_then_accept", s, 1000,
        [&calls](const TsState&, double* d, double* e) {
          *d = 1e-9;
          *e = calls++ < 2 ? 1e30 : 0.0;
        });
  }
  // whole integrations of the linear ODE to t = 10
  for (int q = 1; q <= kMaxOrder; q++)
    for (double tol : {1e-6, 1e-8, 1e-10}) integrate("fixed", q, false, tol, 10.0);
  for (int q = 3; q <= kMaxOrder; q++)
    for (double tol : {1e-6, 1e-8, 1e-10}) integrate("select", q, true, tol, 10.0);
  return 0;
}
