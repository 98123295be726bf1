// Host driver of the BDF step controller (nmag-src_amd/csrc/nsim_tsctl.h):
// runs the controller's own host functions through forced scenarios, with
// the Newton results made up on the host, and prints one JSON object per
// scenario.  No HIP runtime call is made, so it runs without a GPU.
#include <cmath>
#include <cstdio>
#include <limits>

#include "nsim_tsctl.h"

using namespace nsim;

namespace {

TsState fresh(double t_final, int max_steps, bool exact) {
  TsState s{};
  s.tn = 0.0;
  s.nhist = 0;
  s.sn = 0;
  s.snm1 = 1;
  s.snm2 = 2;
  s.max_order = 2;
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
    ts_after_newton(s, d, errsq, 2, n3);
    cycles++;
  }
  std::printf("{\"name\": \"%s\", \"cycles\": %d, \"mode\": %d, \"err_code\": %d, "
              "\"halt\": %d, \"steps\": %lld, \"steps_call\": %d, \"tn\": %.17g, "
              "\"t_final\": %.17g, \"fails\": %d, \"err_fails\": %lld, "
              "\"conv_fails\": %lld, \"h\": %.17g, \"q_last\": %d}\n",
              name, cycles, (int)s.mode, s.err_code, s.halt, s.steps, s.steps_call, s.tn,
              s.t_final, s.fails, s.err_fails, s.conv_fails, s.h, s.q_last);
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
    s.tnm1 = -0.01;
    s.tnm2 = -0.02;
    ts_prepare(s, 0.01, 2);
    run("two_rejections_then_accept", s, 1000,
        [&calls](const TsState&, double* d, double* e) {
          *d = 1e-9;
          *e = calls++ < 2 ? 1e30 : 0.0;
        });
  }
  return 0;
}
