// Dense demag potential operator.
//
// The Fredkin-Koehler potential of the reference (nmag_lam.py:742-759) is
// linear in the magnetic charge rho:
//     phi1 = K^+ rho                  (Neumann solve, solve_neg_laplace_phi)
//     phi2 = E B R phi1               (BEM on the boundary, then the Dirichlet
//                                      solve solve_laplace_DBC: E extends
//                                      boundary values harmonically)
//     phi  = phi1 + phi2 = M rho,     M = (I + E B R) K^+.
// With 288 GB of HBM per GPU, M fits for magnets up to ~20k nodes (3.2 GB in
// FP64 at 19,926 nodes), and one GEMV streaming M replaces two latency-bound
// Krylov solves per field evaluation.  M is built once per magnet here with
// dense Cholesky factorisations (rocSOLVER) and GEMMs (rocBLAS, FP64 MFMA):
//   1. K^+ = (K + (tau/N) 1 1^T)^-1 - 1 1^T / (tau N)   (K 1 = 0; the shifted
//      matrix is SPD and this is the exact pseudo-inverse for any tau > 0)
//   2. W = B K^+(b, :)                 (boundary rows of K^+, then the BEM)
//   3. Z = K_II^-1 K_IB W              (harmonic extension of W's columns)
//   4. M = K^+ with M(b, :) += W and M(I, :) -= Z.
// Result identical to the two solves up to rounding: the gauge constant of
// K^+ cancels exactly as in the reference (B 1 = -1, E 1 = 1).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "nsim_dense.h"
#include "nsim_handle.h"
#include "nsim_hmatrix_build.h"

namespace nsim {

namespace {

int fail(const std::string& msg) {
  set_error(msg);
  return -1;
}

struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  bool alloc(T** out, size_t count) {
    *out = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(out), std::max<size_t>(count, 1) * sizeof(T)) !=
        hipSuccess)
      return false;
    p.push_back(*out);
    return true;
  }
  template <typename T>
  bool upload(T** out, const std::vector<T>& v) {
    return alloc(out, v.size()) &&
           (v.empty() || hipMemcpy(*out, v.data(), v.size() * sizeof(T),
                                   hipMemcpyHostToDevice) == hipSuccess);
  }
  void release(void* q) {
    auto it = std::find(p.begin(), p.end(), q);
    if (it != p.end()) {
      (void)hipFree(q);
      p.erase(it);
    }
  }
};

struct Blas {
  rocblas_handle h = nullptr;
  ~Blas() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

#define HCHK(expr, what)                                                     \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(_e)); \
  } while (0)
#define RCHK(expr, what)                                                     \
  do {                                                                       \
    rocblas_status _s = (expr);                                              \
    if (_s != rocblas_status_success)                                        \
      return fail(std::string(what) + ": rocBLAS/rocSOLVER status " + std::to_string((int)_s)); \
  } while (0)

int check_info(const rocblas_int* dinfo, hipStream_t st, const std::string& what) {
  rocblas_int info = 0;
  HCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st), what);
  HCHK(hipStreamSynchronize(st), what);
  if (info != 0) return fail(what + " failed (info " + std::to_string(info) + ")");
  return 0;
}

}  // namespace

int build_potential(nsim_handle* h) {
  DeviceOps& d = h->d;
  const HostOperators& o = h->host;
  hipStream_t st = h->stream;
  DenseOp& p = d.pot;
  p.type = h->opt.precision == NSIM_MIXED_F24 ? kBemF24 : d.bem.type;
  p.kind = 1;
  const int ninst = d.ninst;
  p.h_n.clear();
  p.h_ld.clear();
  p.h_off.clear();
  p.h_aoff.clear();
  long long elems = 0;
  for (int k = 0; k < ninst; k++) {
    const int nk = h->node_off[k + 1] - h->node_off[k];
    const int ld = ((nk + 63) / 64) * 64;
    p.h_n.push_back(nk);
    p.h_ld.push_back(ld);
    p.h_off.push_back(h->node_off[k]);
    p.h_aoff.push_back(elems);
    elems += (long long)nk * ld;
  }
  p.elems = elems;
  // H-matrix route: each magnet's M is built dense in FP64 as below, then
  // compressed on the host; FP64 entries for an FP64 handle, FP32 otherwise
  const bool hm = h->use_hmatrix;
  const int hm_type = h->opt.precision == NSIM_FP64 ? kBemF64 : kBemF32;
  const double hm_eps = h->opt.compress_eps > 0.0 ? h->opt.compress_eps
                        : hm_type == kBemF64      ? 1e-12
                                                  : 1e-7;
  const double hm_eta = h->opt.compress_eta > 0.0 ? h->opt.compress_eta : 3.0;
  HMatrixHost H;
  H.group = hm_type == kBemF32 ? 4 : 2;
  std::vector<double> hostP;
  double hm_seconds = 0.0;
  const int hm_threads =
      std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  if (hm) {
    elems = 0;
    p.elems = 0;
  } else {
    HCHK(hipMalloc(&p.A, dense_bytes(p.type, elems)), "potential operator alloc");
    h->track(p.A);
  }
  for (auto* v : {&p.h_n, &p.h_ld, &p.h_off}) {
    int* dv = nullptr;
    HCHK(hipMalloc(&dv, v->size() * sizeof(int)), "potential tables");
    h->track(dv);
    HCHK(hipMemcpy(dv, v->data(), v->size() * sizeof(int), hipMemcpyHostToDevice),
         "potential tables");
    if (v == &p.h_n) p.n = dv;
    else if (v == &p.h_ld) p.ld = dv;
    else p.off = dv;
  }
  HCHK(hipMalloc(&p.aoff, p.h_aoff.size() * sizeof(long long)), "potential tables");
  h->track(p.aoff);
  HCHK(hipMemcpy(p.aoff, p.h_aoff.data(), p.h_aoff.size() * sizeof(long long),
                 hipMemcpyHostToDevice),
       "potential tables");
  p.map = d.iota;

  Scratch s;
  Blas blas;
  RCHK(rocblas_create_handle(&blas.h), "rocblas_create_handle");
  RCHK(rocblas_set_stream(blas.h, st), "rocblas_set_stream");
  int *rp = nullptr, *ci = nullptr;
  double* kv = nullptr;
  rocblas_int* info = nullptr;
  if (!s.upload(&rp, o.rp) || !s.upload(&ci, o.ci) || !s.upload(&kv, o.K) ||
      !s.alloc(&info, 1))
    return fail("potential set-up: upload of the stiffness failed");
  const double one = 1.0, zero = 0.0;

  for (int k = 0; k < ninst; k++) {
    const int N = p.h_n[k], o0 = h->node_off[k];
    // local boundary / interior node lists of this magnet
    std::vector<int> bl, il, posb(N, -1), posi(N, -1);
    for (int l = h->bnd_off[k]; l < h->bnd_off[k + 1]; l++) bl.push_back(o.bnode[l] - o0);
    for (int i = 0, b = 0; i < N; i++) {
      if (b < (int)bl.size() && bl[b] == i) {
        posb[i] = b++;
      } else {
        posi[i] = (int)il.size();
        il.push_back(i);
      }
    }
    const int Nb = (int)bl.size(), Ni = (int)il.size();
    if (Nb == 0) return fail("potential set-up: a magnet without boundary");
    double kd = 0.0;
    for (int i = 0; i < N; i++) kd += o.Kdiag[o0 + i];
    const double tau = kd / N;
    int *dbl = nullptr, *dil = nullptr, *dposb = nullptr, *dposi = nullptr;
    if (!s.upload(&dbl, bl) || !s.upload(&dil, il) || !s.upload(&dposb, posb) ||
        !s.upload(&dposi, posi))
      return fail("potential set-up: upload of the node lists failed");

    // 1. K^+ through the Cholesky factor of K + (tau/N) 1 1^T
    double* P = nullptr;
    if (!s.alloc(&P, (size_t)N * N)) return fail("potential set-up: out of device memory");
    HCHK(dense_fill((long long)N * N, tau / N, P, st), "fill");
    HCHK(dense_scatter_csr(N, o0, o0, rp, ci, kv, nullptr, nullptr, P, N, st), "scatter K");
    RCHK(rocsolver_dpotrf(blas.h, rocblas_fill_lower, N, P, N, info), "dpotrf");
    if (check_info(info, st, "Cholesky of the shifted stiffness") != 0) return -1;
    RCHK(rocsolver_dpotri(blas.h, rocblas_fill_lower, N, P, N, info), "dpotri");
    if (check_info(info, st, "inverse of the shifted stiffness") != 0) return -1;
    HCHK(dense_sym_shift(N, P, N, 1.0 / (tau * N), st), "symmetrise");

    // 2. W = B K^+(b, :)   (B row-major = B^T column-major)
    const int ldb = d.bem.h_ld[k];
    double *Kb = nullptr, *B64 = nullptr, *W = nullptr;
    if (!s.alloc(&Kb, (size_t)Nb * N) || !s.alloc(&B64, (size_t)Nb * ldb) ||
        !s.alloc(&W, (size_t)Nb * N))
      return fail("potential set-up: out of device memory");
    HCHK(dense_gather_rows(Nb, N, dbl, P, N, Kb, Nb, st), "gather boundary rows");
    HCHK(launch_bem_assemble_f64(d, k, B64, st), "FP64 BEM segment");
    RCHK(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, Nb, N, Nb,
                       &one, B64, ldb, Kb, Nb, &zero, W, Nb),
         "dgemm B K+");
    HCHK(hipStreamSynchronize(st), "sync");
    s.release(Kb);
    s.release(B64);

    // 3. Z = K_II^-1 K_IB W, 4. M = K^+ + [W on b; -Z on I]
    HCHK(dense_add_rows(Nb, N, dbl, 1.0, W, Nb, P, N, st), "add boundary rows");
    if (Ni > 0) {
      double *KIB = nullptr, *T = nullptr, *KII = nullptr;
      if (!s.alloc(&KIB, (size_t)Ni * Nb) || !s.alloc(&T, (size_t)Ni * N) ||
          !s.alloc(&KII, (size_t)Ni * Ni))
        return fail("potential set-up: out of device memory");
      HCHK(dense_fill((long long)Ni * Nb, 0.0, KIB, st), "fill");
      HCHK(dense_scatter_csr(N, o0, o0, rp, ci, kv, dposi, dposb, KIB, Ni, st), "scatter K_IB");
      RCHK(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, Ni, N, Nb, &one,
                         KIB, Ni, W, Nb, &zero, T, Ni),
           "dgemm K_IB W");
      HCHK(dense_fill((long long)Ni * Ni, 0.0, KII, st), "fill");
      HCHK(dense_scatter_csr(N, o0, o0, rp, ci, kv, dposi, dposi, KII, Ni, st), "scatter K_II");
      RCHK(rocsolver_dpotrf(blas.h, rocblas_fill_lower, Ni, KII, Ni, info), "dpotrf K_II");
      if (check_info(info, st, "Cholesky of the interior stiffness") != 0) return -1;
      RCHK(rocsolver_dpotrs(blas.h, rocblas_fill_lower, Ni, N, KII, Ni, T, Ni), "dpotrs");
      HCHK(dense_add_rows(Ni, N, dil, -1.0, T, Ni, P, N, st), "add interior rows");
      HCHK(hipStreamSynchronize(st), "sync");
      s.release(KIB);
      s.release(T);
      s.release(KII);
    }
    s.release(W);

    if (hm) {
      // to the host, then the H-matrix compression of this magnet
      hostP.resize((size_t)N * N);
      HCHK(hipMemcpyAsync(hostP.data(), P, sizeof(double) * (size_t)N * N,
                          hipMemcpyDeviceToHost, st),
           "potential operator to host");
      HCHK(hipStreamSynchronize(st), "sync");
      const auto t0 = std::chrono::steady_clock::now();
      const std::string err = hmatrix_compress(N, hostP.data(), N, &o.x[3 * (size_t)o0], o0,
                                               hm_eps, hm_eta, hm_threads, H);
      hm_seconds +=
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (!err.empty()) return fail("potential set-up: " + err);
    } else {
      // store row-major in the storage type
      HCHK(dense_store_rowmajor(p.type, N, p.h_ld[k], P, N,
                                (char*)p.A + dense_bytes(p.type, p.h_aoff[k]), st),
           "store");
      HCHK(hipStreamSynchronize(st), "sync");
    }
    s.release(P);
    for (int* q : {dbl, dil, dposb, dposi}) s.release(q);
  }
  if (hm) {
    hostP.clear();
    hostP.shrink_to_fit();
    hmatrix_finish(H);
    HMatrixOp& op = d.hm;
    HCHK(upload_hmatrix(H, hm_type, h->num_cus, op, &h->allocations), "H-matrix upload");
    op.eps = hm_eps;
    op.eta = hm_eta;
    op.full_entries = H.full_entries;
    op.dense_entries = H.dense_entries;
    op.lowrank_entries = H.lowrank_entries;
    op.lowrank_blocks = H.lowrank_blocks;
    op.dense_blocks = H.dense_blocks;
    op.max_rank = H.max_rank;
    op.fallback_blocks = H.fallback_blocks;
    op.build_s = hm_seconds;
    op.halt = d.halt;
  }
  return 0;
}

}  // namespace nsim
