// Device side of the H-matrix demag potential operator: upload of the panels
// built by hmatrix_compress (nsim_hmatrix_build.cpp) and their two-phase
// product phi = M rho.
//
// Each panel is R = 16 rows by C columns, stored as groups of G columns
// (G = 16 bytes / entry size) with a group's R x G entries contiguous.  A
// workgroup of 256 threads takes one work item (a panel, or a chunk of at
// most kItemGroups groups of one): thread (r, l), r = tid % 16, reads row r
// of the groups l, l + 16, ... with one 16-byte non-temporal load each, so a
// wavefront reads 1 KiB of contiguous memory per load.  It multiplies them by
// the group's G consecutive sources (entries of rho in the cluster order,
// which k_exch_div writes beside rho, or of the phase-1 vector t) and
// accumulates in FP64.  The 16 partial sums of a row meet in LDS in a fixed
// order (deterministic), and the row's destination receives the total.
// Items are many (about 4,000 and 6,000 at 20k nodes) and of uneven length:
// one workgroup each, largest first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "nsim_dev_common.h"
#include "nsim_device.h"
#include "nsim_hmatrix_build.h"

namespace nsim {

namespace {

constexpr int kThreads = 256;

constexpr int kRows = 16;
static_assert(sizeof(PanelDesc) == 40, "work item layout");

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// Broadcast of lane e of each quad of lanes (DPP quad_perm [e, e, e, e]) of
// a double, as its two 32-bit halves.
template <int E>
__device__ inline double quad_bcast(double x) {
  constexpr int ctrl = E * 0x55;
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), ctrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Row r of work item d, over the groups cl, cl + 16, ...; RUN: the item's
// columns are one run of sources, else gs (in LDS) holds each group's first
// source.  Every group's G sources are G consecutive doubles; lane r of a row
// loads only the (r mod G)-th of them, and the quad's DPP broadcasts hand each
// lane the others: a quarter of the source bytes through the vector memory
// path.  Groups go U = 4 at a time, then 2, then 1 (a short item still has
// all of a lane's loads in flight at once); the loads of a step are issued
// before any of its products.
template <typename T, bool RUN>
__device__ inline double item_row(const T* __restrict__ data, const PanelDesc& d,
                                  const int* gs, const double* __restrict__ xp,
                                  const double* __restrict__ t, int r, int cl) {
  constexpr int G = 16 / sizeof(T);
  constexpr int R = kRows;
  constexpr int CL = kThreads / R;
  typedef typename Vec16<T>::type V;
  const V* base = reinterpret_cast<const V*>(data + d.off);
  const double* run = d.src0 >= 0 ? xp + d.src0 : t + (-d.src0 - 1);
  auto source = [&](int g) -> const double* {
    if (RUN) return run + g * G;
    const int s = gs[g];
    return s >= 0 ? xp + s : t + (-s - 1);
  };
  const int me = r & (G - 1);
  const int ng = d.ng;
  double acc = 0.0;
  auto step = [&](int g, auto UU) {
    constexpr int U = decltype(UU)::value;
    V v[U];
    double sm[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(base + (size_t)(g + u * CL) * R + r);
#pragma unroll
    for (int u = 0; u < U; u++) sm[u] = source(g + u * CL)[me];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      acc += (double)v[u][0] * quad_bcast<0>(sm[u]);
      acc += (double)v[u][1] * quad_bcast<1>(sm[u]);
      if constexpr (G == 4) {
        acc += (double)v[u][2] * quad_bcast<2>(sm[u]);
        acc += (double)v[u][3] * quad_bcast<3>(sm[u]);
      }
    }
  };
  int g = cl;
  for (; g + 3 * CL < ng; g += 4 * CL) step(g, std::integral_constant<int, 4>());
  if (g + CL < ng) {
    step(g, std::integral_constant<int, 2>());
    g += 2 * CL;
  }
  if (g < ng) step(g, std::integral_constant<int, 1>());
  return acc;
}

// One workgroup per work item (a panel or a chunk of one).  Items are ordered
// by decreasing size and the dispatcher starts workgroups in index order.
// The 16 partial sums of a row meet in LDS in a fixed order.  A whole panel
// writes its rows; a chunk publishes its row sums with device-coherent
// stores, waits for them, and counts itself on the panel's counter; the chunk
// that counts last reads every chunk's sums (device-coherent loads), adds
// them in chunk order, writes the rows and resets the counter.  So the result
// does not depend on which chunk finishes last.  (A shared work queue was
// slower: its atomics on one address serialise.)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_panel_gemv(
    const T* __restrict__ data, const PanelDesc* __restrict__ desc,
    const int* __restrict__ gsrc, const int* __restrict__ rowdst,
    const double* __restrict__ xp, const double* __restrict__ t, double* __restrict__ out,
    double* __restrict__ partial, unsigned* __restrict__ cnt, const int* halt) {
  if (halt && *halt) return;  // queued by a finished integrator run: nothing counted
  constexpr int R = kRows;
  constexpr int CL = kThreads / R;
  __shared__ double red[CL][R];
  __shared__ int last;
  __shared__ int gs[kItemGroups];
  const PanelDesc d = desc[blockIdx.x];
  const int r = threadIdx.x % R, cl = threadIdx.x / R;
  if (d.g0 >= 0) {
    // the item's group sources, read once into LDS (every row reads them)
    for (int i = threadIdx.x; i < d.ng; i += kThreads) gs[i] = gsrc[d.g0 + i];
    __syncthreads();
    red[cl][r] = item_row<T, false>(data, d, gs, xp, t, r, cl);
  } else {
    red[cl][r] = item_row<T, true>(data, d, nullptr, xp, t, r, cl);
  }
  __syncthreads();
  double sum = 0.0;
  if (threadIdx.x < R) {
#pragma unroll
    for (int q = 0; q < CL; q++) sum += red[q][threadIdx.x];
  }
  if (d.pslot < 0) {
    if (threadIdx.x < R) {
      const int dst = rowdst[d.r0 + threadIdx.x];
      if (dst >= 0) out[dst] = sum;
    }
    return;
  }
  if (threadIdx.x < R) {
    dev::st_coh(partial + d.pslot + d.chunk * R + threadIdx.x, sum);
    dev::wait_all_memory();
  }
  __syncthreads();
  if (threadIdx.x == 0)
    last = __hip_atomic_fetch_add(cnt + d.cslot, 1u, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) == (unsigned)(d.nchunk - 1);
  __syncthreads();
  if (!last) return;
  if (threadIdx.x < R) {
    double tot = 0.0;
    for (int c = 0; c < d.nchunk; c++) tot += dev::ld_coh(partial + d.pslot + c * R + threadIdx.x);
    const int dst = rowdst[d.r0 + threadIdx.x];
    if (dst >= 0) out[dst] = tot;
  }
  if (threadIdx.x == 0) dev::st_coh(cnt + d.cslot, 0u);
}

template <typename T>
hipError_t upload_vec(T** dst, const std::vector<T>& v, std::vector<void*>* track) {
  *dst = nullptr;
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  track->push_back(*dst);
  return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

hipError_t upload_panels(const HPanels& P, int type, int num_cus, PanelOp& op,
                         std::vector<void*>* track) {
  // the kernel stages an item's group sources in LDS: at most kItemGroups
  for (const PanelDesc& d : P.desc)
    if (d.ng < 0 || d.ng > kItemGroups) return hipErrorInvalidValue;
  op.type = type;
  op.panels = P.panels();
  op.entries = P.data.size();
  (void)num_cus;
  hipError_t e = upload_vec(&op.desc, P.desc, track);
  if (e == hipSuccess) e = upload_vec(&op.gsrc, P.gsrc, track);
  if (e == hipSuccess) e = upload_vec(&op.rowdst, P.rowdst, track);
  if (e != hipSuccess) return e;
  // partial sums and counters of the chunked panels (counters start at zero
  // and every launch leaves them at zero)
  e = hipMalloc(&op.partial, sizeof(double) * std::max(P.npartial, 1));
  if (e != hipSuccess) return e;
  track->push_back(op.partial);
  e = hipMalloc(&op.cnt, sizeof(unsigned) * std::max(P.ncounters, 1));
  if (e != hipSuccess) return e;
  track->push_back(op.cnt);
  e = hipMemset(op.cnt, 0, sizeof(unsigned) * std::max(P.ncounters, 1));
  if (e != hipSuccess) return e;
  if (type == kBemF32) {
    std::vector<float> f(P.data.begin(), P.data.end());
    float* d = nullptr;
    e = upload_vec(&d, f, track);
    op.data = d;
  } else {
    double* d = nullptr;
    e = upload_vec(&d, P.data, track);
    op.data = d;
  }
  op.bytes = dense_bytes(type, (long long)P.data.size()) + sizeof(PanelDesc) * P.desc.size() +
             sizeof(int) * (P.gsrc.size() + P.rowdst.size());
  return e;
}

template <typename T>
hipError_t launch_panels(const PanelOp& op, const double* xp, const double* t, double* out,
                         const int* halt, hipStream_t st) {
  if (op.panels <= 0) return hipSuccess;
  k_panel_gemv<T><<<op.panels, kThreads, 0, st>>>(static_cast<const T*>(op.data), op.desc,
                                                   op.gsrc, op.rowdst, xp, t, out, op.partial,
                                                   op.cnt, halt);
  return hipGetLastError();
}

}  // namespace

hipError_t upload_hmatrix(const HMatrixHost& H, int type, int num_cus, HMatrixOp& op,
                          std::vector<void*>* track) {
  if (H.group != (type == kBemF32 ? 4 : 2)) return hipErrorInvalidValue;
  hipError_t e = upload_panels(H.p1, type, num_cus, op.p1, track);
  if (e == hipSuccess) e = upload_panels(H.p2, type, num_cus, op.p2, track);
  if (e != hipSuccess) return e;
  op.nt = H.nt;
  // t and rho in cluster order carry a group of zero padding at their ends:
  // a padded group reads past its run of sources
  const size_t tn = (size_t)H.nt + 8, xn = H.pos.size() + 8;
  e = hipMalloc(&op.t, sizeof(double) * tn);
  if (e != hipSuccess) return e;
  track->push_back(op.t);
  e = hipMemset(op.t, 0, sizeof(double) * tn);
  if (e == hipSuccess) e = hipMalloc(&op.xp, sizeof(double) * xn);
  if (e != hipSuccess) return e;
  track->push_back(op.xp);
  e = hipMemset(op.xp, 0, sizeof(double) * xn);
  if (e == hipSuccess) e = upload_vec(&op.pos, H.pos, track);
  return e;
}

hipError_t launch_hmatrix(const HMatrixOp& op, double* y, hipStream_t st) {
  if (op.p2.panels <= 0) return hipSuccess;
  hipError_t e;
  if (op.p1.type == kBemF32) {
    e = launch_panels<float>(op.p1, op.xp, nullptr, op.t, op.halt, st);
    if (e == hipSuccess) e = launch_panels<float>(op.p2, op.xp, op.t, y, op.halt, st);
  } else {
    e = launch_panels<double>(op.p1, op.xp, nullptr, op.t, op.halt, st);
    if (e == hipSuccess) e = launch_panels<double>(op.p2, op.xp, op.t, y, op.halt, st);
  }
  return e;
}

}  // namespace nsim
