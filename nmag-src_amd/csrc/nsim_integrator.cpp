// Device-resident BDF / Newton-GMRES time integrator (see nsim_integrator.h).
//
// All vectors stay in HBM.  The step controller (nsim_tsctl.h) is the same
// code on host and device:
//  * BDF of order q <= 5 on the actual (variable) step history: the corrector
//    is P'(t_{n+1}) = f(y_{n+1}) for the polynomial P through y_{n+1..n+1-q}.
//  * predictor = extrapolation of the polynomial through y_{n..n-q} (explicit
//    Euler on the first step).  The local error is estimated by Milne's device,
//    err = E_c/(E_c+E_p) ||y - y_pred||_WRMS, with the variable-step error
//    constants of both formulas computed from the node positions.
//  * Newton on y - gamma f(y) - a = 0 with (I - gamma J) solved by scaled GMRES
//    (CGS2 orthogonalisation), J.v from the jacobi_ts1 linearisation and a
//    block-Jacobi right preconditioner; WRMS weights 1/(rtol|y| + atol).
// With the persistent GMRES an advance call queues Newton cycles in batches
// (k_ts_begin, the field evaluation, k_gmres, k_ts_ctrl); a finished run
// sets a halt word that turns the cycles still queued into empty launches.
// The host-driven GMRES runs one cycle at a time and reads the state back.
#include "nsim_integrator.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "nsim_device.h"
#include "nsim_vec.h"

namespace nsim {

#define TRY(expr)                     \
  do {                                \
    int _rc = (expr);                 \
    if (_rc) return _rc;              \
  } while (0)
#define HTRY(expr) TRY(check((expr), #expr))

int Integrator::check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  set_error(std::string("integrator: ") + what + ": " + hipGetErrorString(e));
  return -1;
}

Integrator::Integrator(nsim_handle* h, const TsOptions& o) : h_(h), opt_(o) {
  s_ = handle_stream(h);
  nn_ = handle_n(h);
  n3_ = 3 * nn_;
  opt_.krylov_max = std::max(1, std::min(opt_.krylov_max, 500));
  opt_.max_order = std::max(1, std::min(opt_.max_order, kMaxOrder));
  auto alloc = [&](double** p, size_t count) {
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(double)) != hipSuccess) return false;
    allocs_.push_back(*p);
    return hipMemset(*p, 0, count * sizeof(double)) == hipSuccess;
  };
  const size_t n = n3_;
  np_ = vec_reduce_grid(n3_);
  ok_ = true;
  for (int i = 0; i < kHist && ok_; i++) ok_ = alloc(&slot_[i], n);
  ok_ = ok_ && alloc(&fn_, n) &&
        alloc(&yp_, n) && alloc(&ycur_, n) && alloc(&avec_, n) && alloc(&f_, n) &&
        alloc(&res_, n) && alloc(&del_, n) && alloc(&tmp_, n) && alloc(&tmp2_, n) &&
        alloc(&ewt_, n) && alloc(&ystart_, n) && alloc(&status_, 8) &&
        alloc(&V_, n * (size_t)(opt_.krylov_max + 1)) && alloc(&Minv_, 9 * (size_t)nn_) &&
        alloc(&coef_, (size_t)opt_.krylov_max + 8) &&
        alloc(&hdev_, 2 * (size_t)opt_.krylov_max + 8) &&
        alloc(&part_, ((size_t)opt_.krylov_max + 8) * kReduceBlocks);
  host_part_.resize(8 * kReduceBlocks);
  if (ok_ && hipHostMalloc(reinterpret_cast<void**>(&hhost_),
                           (2 * (size_t)opt_.krylov_max + 8) * sizeof(double)) != hipSuccess)
    ok_ = false;
  // controller state: device copy, pinned mirrors, the ring's pointer table
  if (ok_) {
    ok_ = hipMalloc(reinterpret_cast<void**>(&ds_), sizeof(TsState)) == hipSuccess &&
          hipMemset(ds_, 0, sizeof(TsState)) == hipSuccess &&
          hipHostMalloc(reinterpret_cast<void**>(&pin_), 3 * sizeof(TsState)) == hipSuccess &&
          hipMalloc(reinterpret_cast<void**>(&slots_dev_), kHist * sizeof(double*)) ==
              hipSuccess &&
          hipMemcpy(slots_dev_, slot_, kHist * sizeof(double*), hipMemcpyHostToDevice) ==
              hipSuccess &&
          hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming) == hipSuccess &&
          hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming) == hipSuccess;
    if (ds_) allocs_.push_back(ds_);
    if (slots_dev_) allocs_.push_back(slots_dev_);
  }
  // persistent GMRES: one co-resident grid over the SELL slices
  const DeviceOps& d = handle_ops(h);
  int grid = 0, spw = 0, resident = 0;
  if (ok_ && nsim::gmres_grid(d.nslices, opt_.krylov_max, handle_device(h), 1 << 20,
                              std::max(1, opt_.gmres_spw), &grid, &spw, &resident)) {
    gw_.V = V_;
    gw_.gmax = grid;
    gw_.resident = resident;
    gw_.pstride = gmres_pstride(opt_.krylov_max);
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
  if (ok_) handle_set_halt(h_, &ds_->halt);
}

Integrator::~Integrator() {
  (void)hipStreamSynchronize(s_);
  if (ok_) handle_set_halt(h_, nullptr);
  for (void* p : allocs_) (void)hipFree(p);
  if (hhost_) (void)hipHostFree(hhost_);
  if (pin_) (void)hipHostFree(pin_);
  for (hipEvent_t e : ev_)
    if (e) (void)hipEventDestroy(e);
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
  HTRY(hipMemcpyAsync(slot_[0], handle_m(h_), sizeof(double) * n3_, hipMemcpyDeviceToDevice,
                      s_));
  HTRY(hipStreamSynchronize(s_));
  hs_ = TsState{};
  for (int i = 0; i < kHist; i++) {
    hs_.th[i] = t0;
    hs_.sh[i] = i;
  }
  hs_.hnext = 0.0;
  hs_.qnext = 1;
  hs_.qwait = 2;
  hs_.max_order = opt_.max_order;
  hs_.select = opt_.max_order >= 3 && !opt_.fixed_order;
  hs_.crate = 1.0;
  hs_.gamma_crate = 0.0;
  hs_.copy = kSlotNone;
  hs_.mode = kTsDone;
  st_ = TsStats();
  st_.t = t0;
  return 0;
}

int Integrator::upload_state() {
  pin_[2] = hs_;
  HTRY(hipMemcpyAsync(ds_, &pin_[2], sizeof(TsState), hipMemcpyHostToDevice, s_));
  return 0;
}

int Integrator::download_state() {
  HTRY(hipMemcpyAsync(&pin_[2], ds_, sizeof(TsState), hipMemcpyDeviceToHost, s_));
  HTRY(hipStreamSynchronize(s_));
  hs_ = pin_[2];
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

// One Newton cycle of the device-driven run: start of attempt (when the
// controller asks for one), field evaluation at ycur, and the persistent GMRES
// (residual, preconditioner, solve, ycur += delta, norms), whose last act is
// the controller's step.
int Integrator::enqueue_cycle() {
  HTRY(launch_ts_begin(s_, n3_, ds_, slots_dev_, fn_, ycur_, yp_, avec_, ewt_));
  TRY(device_rhs(h_, ycur_, f_));
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
  sys.gamma = 0.0;
  sys.gamma_p = &ds_->gamma;
  sys.halt = &ds_->halt;
  sys.tol = kLinCoef * kNewtonCoef * std::sqrt((double)n3_);
  sys.sweeps = opt_.gmres_pc_sweeps;
  sys.onex_steps = opt_.gmres_onex_steps;
  sys.ts = ds_;
  sys.slots = slots_dev_;
  gw_.cooperative = opt_.cooperative;
  HTRY(launch_newton_gmres(handle_ops(h_), handle_mixed(h_), gw_, ggrid_, gspw_,
                           opt_.krylov_max, sys, s_));
  return 0;
}

// Queues batches of cycles, two in flight: while one runs the host waits for
// the other's copy of the state, stops when it reads "done", and sizes the
// next batch from the steps left at the current step size.
int Integrator::run_device_cycles() {
  const int maxb = std::max(1, std::min(opt_.max_batch, 64));
  int batches = 0, waited = 0;
  // cycles still needed: about one per step left at the current step size
  auto cycles_left = [&](const TsState& m) {
    double left = 1.0;
    if (m.max_steps > 0) left = m.max_steps - m.steps_call;
    if (m.h > 0.0) {
      const double steps = std::ceil((m.t_final - m.th[0]) / m.h);
      left = m.max_steps > 0 ? std::min(left, steps) : steps;
    }
    return (int)std::min<double>(maxb, std::max(1.0, left));
  };
  auto enqueue_batch = [&](int cycles) -> int {
    for (int c = 0; c < cycles; c++) TRY(enqueue_cycle());
    TsState* mirror = &pin_[batches & 1];
    HTRY(hipMemcpyAsync(mirror, ds_, sizeof(TsState), hipMemcpyDeviceToHost, s_));
    HTRY(hipEventRecord(ev_[batches & 1], s_));
    batches++;
    return 0;
  };
  // the first batch covers the estimate, the second a couple of retries
  TRY(enqueue_batch(cycles_left(hs_)));
  TRY(enqueue_batch(2));
  while (true) {
    HTRY(hipEventSynchronize(ev_[waited & 1]));
    const TsState m = pin_[waited & 1];
    waited++;
    if (m.mode == kTsDone || m.mode == kTsError || m.halt) {
      // the state does not change after the end: this copy is final
      hs_ = m;
      return 0;
    }
    TRY(enqueue_batch(cycles_left(m)));
  }
}

// WRMS norm of W[0] ycur + sum_{i=1..cnt} W[i] y_{n+1-i} (0 when cnt is 0).
int Integrator::estimate(int cnt, const double* W, double* out) {
  *out = 0.0;
  if (cnt <= 0) return 0;
  HTRY(vec_lincomb(s_, n3_, W[0], ycur_, 0, nullptr, 0, nullptr, 0, nullptr, tmp_));
  for (int i = 1; i <= cnt; i += 3) {
    const double* y1 = slot_[hs_.sh[i - 1]];
    const double* y2 = i + 1 <= cnt ? slot_[hs_.sh[i]] : nullptr;
    const double* y3 = i + 2 <= cnt ? slot_[hs_.sh[i + 1]] : nullptr;
    HTRY(vec_lincomb(s_, n3_, 1.0, tmp_, W[i], y1, y2 ? W[i + 1] : 0.0, y2,
                     y3 ? W[i + 2] : 0.0, y3, tmp_));
  }
  *out = wrms(tmp_);
  return 0;
}

// Host-driven GMRES: one cycle at a time, the controller's state read back
// after each.
int Integrator::run_host_cycles() {
  for (;;) {
    HTRY(launch_ts_begin(s_, n3_, ds_, slots_dev_, fn_, ycur_, yp_, avec_, ewt_));
    TRY(device_rhs(h_, ycur_, f_));
    gamma_ = hs_.gamma;
    // res = -G = a + gamma f - y
    HTRY(vec_lincomb(s_, n3_, 1.0, avec_, gamma_, f_, -1.0, ycur_, 0, nullptr, res_));
    if (opt_.precond)
      HTRY(bjac_setup(s_, nn_, handle_coef_table(h_), handle_inst(h_), gamma_, ycur_,
                      handle_Htotal(h_), handle_kappa(h_), Minv_));
    int its = 0;
    TRY(gmres(res_, del_, &its));
    HTRY(vec_lincomb(s_, n3_, 1.0, ycur_, 1.0, del_, 0, nullptr, 0, nullptr, ycur_));
    const double d = wrms(del_);
    HTRY(vec_lincomb(s_, n3_, 1.0, ycur_, -1.0, yp_, 0, nullptr, 0, nullptr, tmp_));
    const double e = wrms(tmp_);
    // the order q - 1 and q + 1 estimates the controller asked for
    double em = 0.0, ep = 0.0;
    TRY(estimate(hs_.em, hs_.Em, &em));
    TRY(estimate(hs_.ep, hs_.Ep, &ep));
    hhost_[0] = its;
    hhost_[1] = d * d * n3_;
    hhost_[2] = 0.0;
    hhost_[3] = e * e * n3_;
    hhost_[4] = em * em * n3_;
    hhost_[5] = ep * ep * n3_;
    HTRY(hipMemcpyAsync(status_, hhost_, 6 * sizeof(double), hipMemcpyHostToDevice, s_));
    const TsState attempt = hs_;
    HTRY(launch_ts_ctrl(s_, ds_, status_, n3_));
    TRY(download_state());
    if (hook_) TRY(hook_(attempt, hs_));
    if (hs_.mode == kTsDone || hs_.mode == kTsError) return 0;
  }
}

int Integrator::advance(double t_final, int max_steps, bool exact_tstop, double* t_reached) {
  if (!ok_) {
    set_error("integrator: not initialised");
    return -1;
  }
  double* m = handle_m(h_);
  HTRY(hipMemcpyAsync(ystart_, m, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
  hs_.t_final = t_final;
  hs_.exact_tstop = exact_tstop ? 1 : 0;
  hs_.max_steps = max_steps;
  hs_.steps_call = 0;
  hs_.halt = 0;
  hs_.err_code = 0;
  hs_.rtol = opt_.rtol;
  hs_.atol = opt_.atol;
  hs_.max_order = opt_.max_order;
  hs_.select = opt_.max_order >= 3 && !opt_.fixed_order;
  if (!ts_finished(hs_)) {
    double* yn = slot_[hs_.sh[0]];
    if (hs_.hnext <= 0.0) {  // first step: h0 from ||f(y0)||
      HTRY(vec_ewt(s_, n3_, yn, opt_.rtol, opt_.atol, ewt_));
      TRY(device_rhs(h_, yn, fn_));
      hs_.rhs++;
      const double fnorm = wrms(fn_);
      double h0 = fnorm > 0 ? 0.1 / fnorm : 1.0;
      h0 = std::min(h0, 0.1 * std::fabs(t_final - hs_.th[0]));
      hs_.hnext = h0 > 0 ? h0 : 1e-6;
      hs_.qnext = 1;
    }
    ts_prepare(hs_, hs_.hnext, hs_.qnext);
    TRY(upload_state());
    const bool dev = opt_.device_gmres && ggrid_ > 0 && !hook_;
    const int rc = dev ? run_device_cycles() : run_host_cycles();
    // cycles queued after the end run as empty launches; behind them on the
    // stream, lift the halt word so that later field evaluations run
    const hipError_t me = hipMemsetAsync(&ds_->halt, 0, sizeof(int), s_);
    if (rc) {
      (void)hipStreamSynchronize(s_);
      return rc;
    }
    HTRY(me);
    if (hs_.copy >= 0) {  // the accepted iterate of the last step becomes y_n
      HTRY(hipMemcpyAsync(slot_[hs_.copy], ycur_, sizeof(double) * n3_,
                          hipMemcpyDeviceToDevice, s_));
      hs_.copy = kSlotNone;
    }
    hs_.halt = 0;
    st_.steps = hs_.steps;
    st_.rhs = hs_.rhs;
    st_.newton = hs_.newton;
    st_.linear = hs_.linear;
    st_.linear_max = hs_.linear_max;
    st_.err_fails = hs_.err_fails;
    st_.conv_fails = hs_.conv_fails;
    st_.h_last = hs_.h_last;
    st_.q_last = hs_.q_last;
    st_.t = hs_.th[0];
    for (int q = 0; q <= kMaxOrder; q++) st_.order_steps[q] = hs_.order_steps[q];
    if (hs_.mode == kTsError) {
      if (hs_.err_code == 1) {
        (void)hipMemset(gw_.err, 0, sizeof(unsigned));
        set_error("integrator: GMRES exchange timed out");
        return -1;
      }
      set_error("integrator: step size underflow or repeated failures at t=" +
                std::to_string(hs_.th[0]));
      return 1;
    }
  }
  const double* yn = slot_[hs_.sh[0]];
  if (!exact_tstop && hs_.th[0] > t_final && hs_.nhist >= 1) {
    // interpolate the polynomial of the last step's order back to t_final
    const int q = std::min(hs_.nhist, std::max(1, hs_.q_last));
    double L[kHist] = {};
    for (int j = 0; j <= q; j++) {
      double w = 1.0;
      for (int i = 0; i <= q; i++)
        if (i != j) w *= (t_final - hs_.th[i]) / (hs_.th[j] - hs_.th[i]);
      L[j] = w;
    }
    HTRY(vec_lincomb(s_, n3_, L[0], yn, L[1], slot_[hs_.sh[1]], q >= 2 ? L[2] : 0.0,
                     q >= 2 ? slot_[hs_.sh[2]] : nullptr, q >= 3 ? L[3] : 0.0,
                     q >= 3 ? slot_[hs_.sh[3]] : nullptr, m));
    if (q >= 4)
      HTRY(vec_lincomb(s_, n3_, 1.0, m, L[4], slot_[hs_.sh[4]], q >= 5 ? L[5] : 0.0,
                       q >= 5 ? slot_[hs_.sh[5]] : nullptr, 0, nullptr, m));
    *t_reached = t_final;
  } else {
    HTRY(hipMemcpyAsync(m, yn, sizeof(double) * n3_, hipMemcpyDeviceToDevice, s_));
    *t_reached = hs_.th[0];
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
