// The CVODE preconditioner and Jacobian-times-vector callbacks of the C ABI
// (nsim_psetup, nsim_psolve, nsim_jtimes), with the reference's semantics
// (src/nsim.ml:2210-2263, 2400-2484; src/sundials_sp_stubs.c:352-678):
//
// * psetup rebuilds the jacobi_ts1 Jacobian at y when CVODE says it is not
//   current (jok == 0) or none exists yet: m = y and the fields of y, as the
//   reference's re-Jacobi first runs the field part of the rhs (the
//   linearisation point is frozen in the handle).  With jok != 0 the last
//   Jacobian is kept.  Either way it forms M = I - gamma J.
// * psolve solves M z = r with the M of the last psetup (its own gamma and
//   delta are not used, as in the reference) by restarted GMRES(30) with zero
//   initial guess and left preconditioning, stopped when the preconditioned
//   residual falls below max(pc_rtol * its initial value, pc_atol), after
//   pc_maxit iterations, or when it grows past pc_dtol times its initial
//   value: PETSc's default KSP settings with the reference's pc_rtol 1e-2 and
//   pc_atol 1e-7 (interface/nmag/main.py:728-729).  The preconditioner is the
//   3x3 node-block Jacobi splitting of M, where the reference's KSP uses
//   bjacobi with ILU(0) sub-blocks (a sequential factorisation); the solve
//   tolerance, not the preconditioner, fixes what psolve returns.
// * jtimes multiplies by the Jacobian of the last re-Jacobi; it reads neither
//   y nor fy (the reference does the same).  Before any psetup it uses J at y
//   with H_total of the last rhs.
//
// The Arnoldi loop runs on the host over device vector kernels: one
// synchronisation per iteration to read the Hessenberg column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "nsim_device.h"
#include "nsim_handle.h"
#include "nsim_integrator.h"
#include "nsim_vec.h"

using namespace nsim;

namespace {

constexpr int kRestart = 30;  // PETSc's default GMRES restart

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
    if (_rc) return _rc;   \
  } while (0)

// Device workspace of one solve, carved from one block of doubles.
struct PcWork {
  double* V = nullptr;     // (kRestart + 1) * n3 Krylov basis
  double* Ju = nullptr;    // n3
  double* u = nullptr;     // n3
  double* x = nullptr;     // n3 solution being accumulated
  double* part = nullptr;  // (kRestart + 1) * np reduction partials
  double* hdev = nullptr;  // 2 (kRestart + 1) + 8 Hessenberg entries and norms
  double* coef = nullptr;  // kRestart + 8 projection / update coefficients
};

size_t pc_work_doubles(int n3) {
  const size_t np = (size_t)vec_reduce_grid(n3);
  // every piece starts on a 32-double (256-byte) boundary
  auto up = [](size_t v) { return (v + 31) & ~(size_t)31; };
  return up((size_t)(kRestart + 1) * n3) + 3 * up((size_t)n3) + up((kRestart + 1) * np) +
         up(2 * (kRestart + 1) + 8) + up(kRestart + 8);
}

PcWork pc_work_layout(double* base, int n3) {
  const size_t np = (size_t)vec_reduce_grid(n3);
  auto up = [](size_t v) { return (v + 31) & ~(size_t)31; };
  PcWork w;
  double* p = base;
  w.V = p;
  p += up((size_t)(kRestart + 1) * n3);
  w.Ju = p;
  p += up((size_t)n3);
  w.u = p;
  p += up((size_t)n3);
  w.x = p;
  p += up((size_t)n3);
  w.part = p;
  p += up((kRestart + 1) * np);
  w.hdev = p;
  p += up(2 * (kRestart + 1) + 8);
  w.coef = p;
  return w;
}

// The linearisation point of the Jacobian the callbacks use: the frozen one
// of the last re-Jacobi, else y with H_total of the last rhs.
void jac_point(nsim_handle* h, const double* y, const double** m, const double** H) {
  if (h->pc_have_J) {
    *m = h->pc_m;
    *H = h->pc_H;
  } else {
    *m = y;
    *H = h->Htotal;
  }
}

// u = M v = v - gamma J v with the frozen Jacobian.
int apply_M(nsim_handle* h, const PcWork& w, const double* v, double* out) {
  const int n3 = 3 * h->d.n;
  HIPCHK(launch_jtimes(h->d, h->mixed, h->pc_m, h->pc_H, v, w.Ju, h->stream));
  HIPCHK(vec_lincomb(h->stream, n3, 1.0, v, -h->pc_gamma, w.Ju, 0.0, nullptr, 0.0, nullptr,
                     out));
  return 0;
}

// out = P^-1 in, P the 3x3 node blocks of M.
int apply_P(nsim_handle* h, const double* in, double* out) {
  HIPCHK(bjac_solve(h->stream, h->d.n, h->pc_Minv, in, out));
  return 0;
}

// |v|^2 into hdev[slot] (device), via the reduction partials.
int norm2_to(nsim_handle* h, const PcWork& w, const double* v, int slot) {
  int np = 0;
  HIPCHK(vec_wsq_partials(h->stream, 3 * h->d.n, v, nullptr, w.part, &np));
  HIPCHK(vec_reduce_parts(h->stream, np, 1, w.part, w.hdev + slot, nullptr));
  return 0;
}

int read_host(nsim_handle* h, const double* dev, double* host, int count) {
  HIPCHK(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// Restarted, left-preconditioned GMRES on M z = b (device vectors).
int pc_gmres(nsim_handle* h, const PcWork& w, const double* b, double* z) {
  hipStream_t st = h->stream;
  const int n3 = 3 * h->d.n;
  const size_t ld = (size_t)n3;
  const int np = vec_reduce_grid(n3);
  double H[(kRestart + 1) * kRestart], g[kRestart + 1], cs[kRestart], sn[kRestart],
      y[kRestart], hh[2 * (kRestart + 1) + 1];
  HIPCHK(hipMemsetAsync(w.x, 0, sizeof(double) * n3, st));
  // initial residual b (zero guess), preconditioned
  TRY(apply_P(h, b, w.V));
  TRY(norm2_to(h, w, w.V, 0));
  TRY(read_host(h, w.hdev, hh, 1));
  const double beta0 = std::sqrt(hh[0]);
  const double tol = std::fmax(h->pc_rtol * beta0, h->pc_atol);
  int its = 0;
  bool converged = !(beta0 > tol), diverged = false;
  double beta = beta0;
  while (!converged && !diverged && its < h->pc_maxit) {
    if (its > 0) {  // restart: r = b - M x, preconditioned
      TRY(apply_M(h, w, w.x, w.u));
      HIPCHK(vec_lincomb(st, n3, 1.0, b, -1.0, w.u, 0.0, nullptr, 0.0, nullptr, w.u));
      TRY(apply_P(h, w.u, w.V));
      TRY(norm2_to(h, w, w.V, 0));
      TRY(read_host(h, w.hdev, hh, 1));
      beta = std::sqrt(hh[0]);
      if (!(beta > tol)) {
        converged = true;
        break;
      }
    }
    HIPCHK(vec_lincomb(st, n3, 1.0 / beta, w.V, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr, w.V));
    std::memset(H, 0, sizeof(H));
    std::memset(g, 0, sizeof(g));
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < kRestart && its < h->pc_maxit; j++) {
      double* Vj = w.V + j * ld;
      double* Vn = w.V + (j + 1) * ld;
      TRY(apply_M(h, w, Vj, w.u));
      TRY(apply_P(h, w.u, Vn));
      // classical Gram-Schmidt applied twice, projections applied on the device
      for (int pass = 0; pass < 2; pass++) {
        for (int c0 = 0; c0 <= j; c0 += 8) {
          const int kc = std::min(8, j + 1 - c0);
          int npc = 0;
          HIPCHK(vec_multidot_partials(st, n3, kc, w.V + c0 * ld, ld, Vn, w.part + c0 * np,
                                       &npc));
        }
        HIPCHK(vec_reduce_parts(st, np, j + 1, w.part, w.hdev + pass * (j + 1), w.coef));
        HIPCHK(vec_axpy_multi(st, n3, j + 1, w.V, ld, w.coef, Vn));
      }
      TRY(norm2_to(h, w, Vn, 2 * (j + 1)));
      TRY(read_host(h, w.hdev, hh, 2 * (j + 1) + 1));
      for (int i = 0; i <= j; i++) H[i * kRestart + j] = hh[i] + hh[j + 1 + i];
      const double hn = std::sqrt(hh[2 * (j + 1)]);
      H[(j + 1) * kRestart + j] = hn;
      if (hn > 0.0)
        HIPCHK(vec_lincomb(st, n3, 1.0 / hn, Vn, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr, Vn));
      for (int i = 0; i < j; i++) {  // earlier Givens rotations
        const double a = H[i * kRestart + j], c = H[(i + 1) * kRestart + j];
        H[i * kRestart + j] = cs[i] * a + sn[i] * c;
        H[(i + 1) * kRestart + j] = -sn[i] * a + cs[i] * c;
      }
      const double a = H[j * kRestart + j], c = H[(j + 1) * kRestart + j];
      const double rr = std::hypot(a, c);
      cs[j] = rr > 0.0 ? a / rr : 1.0;
      sn[j] = rr > 0.0 ? c / rr : 0.0;
      H[j * kRestart + j] = rr;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      k = j + 1;
      its++;
      const double res = std::fabs(g[j + 1]);
      if (res <= tol) {
        converged = true;
        break;
      }
      if (res > h->pc_dtol * beta0 || !std::isfinite(res)) {
        diverged = true;
        break;
      }
      if (!(hn > 0.0)) break;  // invariant subspace: the update solves exactly
    }
    // x += V y, y = R^-1 g
    for (int i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int c = i + 1; c < k; c++) s -= H[i * kRestart + c] * y[c];
      y[i] = s / H[i * kRestart + i];
    }
    if (k > 0) {
      HIPCHK(hipMemcpy(w.coef, y, sizeof(double) * k, hipMemcpyHostToDevice));
      HIPCHK(vec_axpy_multi(st, n3, k, w.V, ld, w.coef, w.x));
    }
    if (k == 0) break;
  }
  HIPCHK(hipMemcpyAsync(z, w.x, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
  h->pc_solves++;
  h->pc_iters_last = its;
  h->pc_iters_total += its;
  if (!converged) h->pc_unconverged++;
  return 0;
}

// Workspace of nsim_psolve_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  16 MiB: the GMRES(30) workspace of a
// magnet of up to about 20,400 nodes (816 bytes per node).  In host code the
// array's name is only a stand-in; its device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kScratchDoubles = 2u * 1024u * 1024u;
constexpr int kScratchMaxNodes = 20400;
__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of its copy

int scratch_ptr(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return 0;
  }
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != device) HIPCHK(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != device) HIPCHK(hipSetDevice(cur));
  if (e != hipSuccess || !p)
    return fail(std::string("nsim_psolve_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(p);
  *out = static_cast<double*>(p);
  return 0;
}

static_assert(kScratchMaxNodes * 816ull + 64 * 1024 <= kScratchDoubles * sizeof(double),
              "the stated node limit must fit the scratch");

}  // namespace

namespace nsim {

// Device buffers of the frozen Jacobian and the block preconditioner.
int pc_ensure_state(nsim_handle* h) {
  const size_t n = h->d.n;
  if (!h->pc_Minv) {
    HIPCHK(hipMalloc(&h->pc_Minv, 9 * n * sizeof(double)));
    h->track(h->pc_Minv);
  }
  if (!h->pc_m) {
    HIPCHK(hipMalloc(&h->pc_m, 3 * n * sizeof(double)));
    h->track(h->pc_m);
  }
  if (!h->pc_H) {
    HIPCHK(hipMalloc(&h->pc_H, 3 * n * sizeof(double)));
    h->track(h->pc_H);
  }
  return 0;
}

}  // namespace nsim

extern "C" {

int nsim_jtimes(nsim_handle* h, double t, const double* v, double* Jv, const double* y,
                const double* fy, int device_ptrs) {
  (void)t;
  (void)fy;
  if (!h || !v || !Jv || (!y && !h->pc_have_J)) return fail("nsim_jtimes: null argument");
  const size_t bytes = 3 * (size_t)h->d.n * sizeof(double);
  const double *m = nullptr, *H = nullptr;
  if (device_ptrs) {
    jac_point(h, y, &m, &H);
    HIPCHK(launch_jtimes(h->d, h->mixed, m, H, v, Jv, h->stream));
    return 0;
  }
  if (!h->pc_have_J) HIPCHK(hipMemcpyAsync(h->io_a, y, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->io_b, v, bytes, hipMemcpyHostToDevice, h->stream));
  jac_point(h, h->io_a, &m, &H);
  HIPCHK(launch_jtimes(h->d, h->mixed, m, H, h->io_b, h->io_c, h->stream));
  HIPCHK(hipMemcpyAsync(Jv, h->io_c, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int nsim_psetup(nsim_handle* h, double t, const double* y, const double* fy, int jok,
                int* jcur, double gamma, int device_ptrs) {
  (void)t;
  (void)fy;
  if (!h || !y) return fail("nsim_psetup: null argument");
  if (!std::isfinite(gamma)) return fail("nsim_psetup: gamma is not finite");
  TRY(pc_ensure_state(h));
  const size_t bytes = 3 * (size_t)h->d.n * sizeof(double);
  const bool rebuild = !jok || !h->pc_have_J;
  if (rebuild) {
    // re-Jacobi at y: the linearisation point and the fields of y
    HIPCHK(hipMemcpyAsync(h->pc_m, y, bytes,
                          device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          h->stream));
    TRY(device_rhs(h, h->pc_m, h->io_c));
    HIPCHK(hipMemcpyAsync(h->pc_H, h->Htotal, bytes, hipMemcpyDeviceToDevice, h->stream));
    h->pc_have_J = true;
    h->pc_jac_builds++;
  }
  h->pc_gamma = gamma;
  HIPCHK(bjac_setup(h->stream, h->d.n, h->d.coefs, h->d.inst, gamma, h->pc_m, h->pc_H,
                    h->kappa, h->pc_Minv));
  if (jcur) *jcur = rebuild ? 1 : 0;
  if (!device_ptrs) HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int nsim_psolve(nsim_handle* h, double t, const double* y, const double* fy, const double* r,
                double* z, double gamma, double delta, int lr, int device_ptrs) {
  (void)t;
  (void)y;
  (void)fy;
  (void)gamma;
  (void)delta;
  if (!h || !r || !z) return fail("nsim_psolve: null argument");
  if (lr != 1) return fail("nsim_psolve: only left preconditioning (lr == 1) is supported");
  if (!h->pc_have_J) return fail("nsim_psolve: call nsim_psetup first");
  const int n3 = 3 * h->d.n;
  const size_t bytes = (size_t)n3 * sizeof(double);
  const double* rd = r;
  double* zd = z;
  if (!device_ptrs) {
    HIPCHK(hipMemcpyAsync(h->io_a, r, bytes, hipMemcpyHostToDevice, h->stream));
    rd = h->io_a;
    zd = h->io_b;
  }
  if (h->pc_full) {
    if (!h->pc_ws) {
      HIPCHK(hipMalloc(&h->pc_ws, pc_work_doubles(n3) * sizeof(double)));
      h->track(h->pc_ws);
    }
    TRY(pc_gmres(h, pc_work_layout(h->pc_ws, n3), rd, zd));
  } else {
    TRY(apply_P(h, rd, zd));
  }
  if (!device_ptrs) {
    HIPCHK(hipMemcpyAsync(z, zd, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return 0;
}

int nsim_psolve_noalloc(nsim_handle* h, double t, const double* y, const double* fy,
                        const double* r, double* z, double gamma, double delta, int lr,
                        int device_ptrs) {
  (void)t;
  (void)y;
  (void)fy;
  (void)gamma;
  (void)delta;
  if (!h || !r || !z) return fail("nsim_psolve_noalloc: null argument");
  if (lr != 1)
    return fail("nsim_psolve_noalloc: only left preconditioning (lr == 1) is supported");
  if (!h->pc_have_J) return fail("nsim_psolve_noalloc: call nsim_psetup first");
  const int n3 = 3 * h->d.n;
  const size_t need = pc_work_doubles(n3);
  if (h->pc_full && need > kScratchDoubles)
    return fail("nsim_psolve_noalloc: " + std::to_string(h->d.n) + " nodes need " +
                std::to_string(need * sizeof(double)) + " bytes of workspace, above the " +
                std::to_string(kScratchDoubles * sizeof(double)) +
                " bytes of the static scratch (at most about " +
                std::to_string(kScratchMaxNodes) + " nodes)");
  double* ws = nullptr;
  TRY(scratch_ptr(h->device, &ws));
  const size_t bytes = (size_t)n3 * sizeof(double);
  const double* rd = r;
  double* zd = z;
  if (!device_ptrs) {
    HIPCHK(hipMemcpyAsync(h->io_a, r, bytes, hipMemcpyHostToDevice, h->stream));
    rd = h->io_a;
    zd = h->io_b;
  }
  if (h->pc_full) TRY(pc_gmres(h, pc_work_layout(ws, n3), rd, zd));
  else TRY(apply_P(h, rd, zd));
  if (!device_ptrs) {
    HIPCHK(hipMemcpyAsync(z, zd, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return 0;
}

int nsim_ts_set_tols(nsim_handle* h, double ksp_rtol, double ksp_atol, double ksp_dtol,
                     int ksp_maxit, double cvode_rtol, double cvode_atol) {
  if (!h) return fail("nsim_ts_set_tols: null handle");
  if (ksp_rtol >= 0.0) h->pc_rtol = ksp_rtol;
  if (ksp_atol >= 0.0) h->pc_atol = ksp_atol;
  if (ksp_dtol > 0.0) h->pc_dtol = ksp_dtol;
  if (ksp_maxit > 0) h->pc_maxit = ksp_maxit;
  if (cvode_rtol >= 0.0) h->ts_opt.rtol = cvode_rtol;
  if (cvode_atol > 0.0) h->ts_opt.atol = cvode_atol;
  if (h->ts && (cvode_rtol >= 0.0 || cvode_atol > 0.0)) {
    h->ts->options().rtol = h->ts_opt.rtol;
    h->ts->options().atol = h->ts_opt.atol;
  }
  return 0;
}

}  // extern "C"
