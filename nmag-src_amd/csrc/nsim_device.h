// Device-side data of one magnet and the launch interface of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace nsim {

enum BemType { kBemF64 = 0, kBemF32 = 1, kBemBF16 = 2, kBemF24 = 3 };

// Bytes of `elems` entries of a dense operator in storage type `type`.
inline size_t dense_bytes(int type, long long elems) {
  const size_t per = type == kBemF64 ? 8 : type == kBemF32 ? 4 : type == kBemF24 ? 3 : 2;
  return (size_t)elems * per;
}

// Material and LLG coefficients of one magnet (simulation units).
struct Coeffs {
  double Ms, negJ, c1, c2, c3, mu0;
};

// A block-diagonal set of dense row-major matrices (one segment per magnet)
// applied by one grouped GEMV launch as
//     y[map[off_s + r]] = sum_j A_s[r][j] x[off_s + j].
// Used for the BEM matrix (x and y on the boundary nodes) and for the dense
// demag potential operator (x and y on all nodes).
struct DenseOp {
  void* A = nullptr;
  int type = kBemF64;                 // storage: FP64, FP32, BF16 or 24-bit
  int kind = 0;                       // 0: BEM matrix, 1: potential operator
  std::vector<int> h_n, h_ld, h_off;  // per segment: order, row stride (x64), x/map offset
  std::vector<long long> h_aoff;      // per segment: first element in A
  long long elems = 0;
  int* n = nullptr;
  int* ld = nullptr;
  int* off = nullptr;
  long long* aoff = nullptr;
  const int* map = nullptr;  // output index map (device)
  int* wg_seg = nullptr;     // plan: workgroup -> segment
  int* wg_row0 = nullptr;    //       workgroup -> first row in the segment
  int grid = 0, rpw = 0, max_ld = 0;
  int rows = 2, wg_per_cu = 2, waves = 4;  // tile shape: rows per pass (0: rows
                                           // claimed dynamically), workgroups per
                                           // CU, wavefronts per workgroup
  unsigned* ctr = nullptr;                 // row counter of the dynamic variant
  const int* halt = nullptr;               // as DeviceOps::halt
};

// One phase of the H-matrix potential operator on the device: panels of 16
// rows (see nsim_hmatrix_build.h for the layout).
struct PanelDesc;
struct PanelOp {
  int type = kBemF64;  // entries in FP64 or FP32
  int panels = 0;  // work items
  long long entries = 0;
  size_t bytes = 0;  // entries and index arrays
  PanelDesc* desc = nullptr;  // per panel: entries, groups, rows
  int* gsrc = nullptr;      // per group: first source (>= 0 rho in cluster
                            // order, < 0 the t entry -s - 1); G consecutive
  int* rowdst = nullptr;
  double* partial = nullptr;  // row sums of the chunks of long panels
  unsigned* cnt = nullptr;    // per chunked panel: chunks finished (reset to 0)
  void* data = nullptr;
};

// The H-matrix potential operator: phase 1 (t = V^T rho per low-rank block)
// and phase 2 (phi = dense near field + U t).
struct HMatrixOp {
  PanelOp p1, p2;
  double* t = nullptr;      // phase-1 results, nt entries (+ padding)
  int nt = 0;
  int* pos = nullptr;       // node -> cluster-order position (null: no H-matrix)
  double* xp = nullptr;     // rho in cluster order, written by k_exch_div
  double eps = 0.0;        // relative tolerance per block
  double eta = 0.0;        // admissibility parameter
  const int* halt = nullptr;
  long long full_entries = 0, dense_entries = 0, lowrank_entries = 0;
  int lowrank_blocks = 0, dense_blocks = 0, max_rank = 0, fallback_blocks = 0;
  double build_s = 0.0;    // host compression time
};

// Operators resident in HBM (one copy per precision actually used).
struct DeviceOps {
  int n = 0, nb = 0, ntri = 0, nslices = 0, sell_total = 0;
  int max_width = 0;  // widest SELL slice (slots per row)
  int* sell_off = nullptr;
  int* sell_col = nullptr;
  double* K = nullptr;  // FP64 values (always present)
  double* D = nullptr;
  double* G = nullptr;
  float* Kf = nullptr;  // FP32 copies (mixed precision only); the divergence
  float* Gf = nullptr;  // D stays FP64, it feeds the singular Neumann problem
  double* Vinv = nullptr;      // 1/V_i
  double* dinv_nbc = nullptr;  // 1/K_ii
  double* dinv_dbc = nullptr;  // 1/K_ii on interior rows, 0 on boundary rows
  int* bnode = nullptr;        // local boundary index -> node
  int* blocal = nullptr;       // node -> local boundary index or -1
  DenseOp bem;                      // BEM matrix, one segment per magnet
  std::vector<double> h_bem_geo_eps;  // per segment geometric tolerance
  DenseOp pot;                      // dense demag potential operator (A null: iterative)
  HMatrixOp hm;                     // compressed potential operator (p2.panels 0: none)
  int* iota = nullptr;              // identity map of the nodes
  const int* halt = nullptr;        // nonzero word: the field kernels return at once
                                    // (set by a device-driven integrator run that
                                    // has finished while kernels are still queued)
  // coefficients: one entry per magnet; inst[i] = magnet of node i (null
  // when the handle holds a single magnet)
  Coeffs* coefs = nullptr;
  int* inst = nullptr;
  int ninst = 1;
  // BEM geometry (set-up only)
  double* coords = nullptr;
  int* tri_node = nullptr;     // 3*ntri global node indices
  double* tnorm = nullptr;
  double* omega = nullptr;
  int* inc_rp = nullptr;
  int* inc_tri = nullptr;
};

// Communication buffers of the persistent CG kernel.
struct CgWork {
  double* pub = nullptr;            // 2*n published vector (by phase parity)
  double* part = nullptr;           // 2*gmax records of kCgRecord partial sums
  unsigned long long* tag = nullptr;  // 2*gmax phase tags, dev::kTagStride words apart
  unsigned* sync = nullptr;         // [2] error flag [3] last iterations
                                    // [4] total iterations [5] solves
                                    // [6] solves stopped at maxit
  int gmax = 0;
  unsigned long long epoch = 0;     // launch counter; makes every phase tag unique
  const int* halt = nullptr;        // as DeviceOps::halt
  // The CG grid never exceeds what the device holds at once and the stream
  // orders it after and before other work, so a plain launch keeps its
  // workgroups co-resident; a cooperative launch guarantees it at about
  // 20 us more dispatch cost.  Every exchange wait is bounded either way.
  bool cooperative = false;
};

// One term c mx^p[0] my^p[1] mz^p[2] of a polynomial anisotropy energy
// density (simulation units).
struct AnisTerm {
  int p[3];
  int pad;
  double c;
};

// Field inputs of one evaluation (device pointers; optional ones may be null).
struct FieldArgs {
  const double* m;
  const double* Hext;
  const double* k1;
  const double* k2;
  const double* axis;
  const double* pin;
  // polynomial anisotropy: magnet k's terms are poly[poly_off[k] ...
  // poly_off[k + 1]) (null: none on any magnet)
  const AnisTerm* poly;
  const int* poly_off;
};

struct FieldBufs {
  double *Hexch, *Hdemag, *Hanis, *Htotal, *rho, *phi1, *phi2, *phi, *phi1b;
};

// Launchers (all asynchronous on `st`).  Return hipSuccess or the launch error.
hipError_t launch_bem_assemble(const DeviceOps& d, hipStream_t st);
hipError_t launch_exch_div(const DeviceOps& d, bool mixed, const double* m, double* Hexch,
                           double* rho, bool want_rho, hipStream_t st);
// The CG's reordered stiffness operator in HBM (see CgOperator).
struct CgDevice {
  int parts = 0, npad = 0, rpt = 1, cap_shift = 8, agg_shift = 0, sub_shift = 0, naggs = 0;
  int maxw = 0, max_width = 0;
  bool fine = false;  // sub-part aggregates, FP32 coarse inverse in LDS (ainvf_*)
  float *ainvf_nbc = nullptr, *ainvf_dbc = nullptr;
  int *need = nullptr, *need_n = nullptr;  // per part: the aggregates its rows touch
  int need_max = 0, need_lg = 2;            // see CgOperator
  int* sell_off = nullptr;
  int* sell_col = nullptr;
  double* K = nullptr;
  int* p2o = nullptr;
  double *dinv_nbc = nullptr, *dinv_dbc = nullptr;
  double *ainv_nbc = nullptr, *ainv_dbc = nullptr;
  int resident = 0;  // workgroups of the kernel the device holds at once
};
// Persistent pipelined two-level PCG on the stiffness matrix; see the kernel.
// b, x, xadd, xsum are in the original node order.
hipError_t launch_cg(const CgDevice& c, CgWork& w, bool dirichlet, bool coarse,
                     const double* b, double* x, double rtol, int maxit, const double* xadd,
                     double* xsum, double* restrict_out, const int* blocal,
                     hipStream_t st);
// Uploads a compressed operator built on the host (entries converted to
// `type`, kBemF64 or kBemF32); allocations are appended to `track`.
struct HMatrixHost;
hipError_t upload_hmatrix(const HMatrixHost& H, int type, int num_cus, HMatrixOp& op,
                          std::vector<void*>* track);
// phi = M rho with the H-matrix potential operator (two launches); rho is
// read in the cluster order from op.xp, which k_exch_div fills.
hipError_t launch_hmatrix(const HMatrixOp& op, double* y, hipStream_t st);
// Grouped GEMV of a DenseOp (see there).
hipError_t launch_dense_gemv(const DenseOp& op, const double* x, double* y, hipStream_t st);
// Host plan of the grouped GEMV for a tile shape (rows_per_pass 1, 2 or 4 rows
// streamed together by a wavefront; wg_per_cu 1..8 workgroups per CU; waves 4
// or 16 wavefronts per workgroup): fills op.grid/rpw/max_ld/rows/wg_per_cu/
// waves and the workgroup -> (segment, first row) table to upload.  False for
// a shape outside those ranges.
bool plan_dense_gemv(DenseOp& op, int num_cus, int rows_per_pass, int wg_per_cu, int waves,
                     std::vector<int>* wg_seg, std::vector<int>* wg_row0);
// Largest segment order the grouped GEMV can stage (its x in LDS).
int dense_gemv_max_order(int type);
// Assembles BEM segment s in FP64 into `out` (row stride op.h_ld[s]), whatever
// the storage type of d.bem.
hipError_t launch_bem_assemble_f64(const DeviceOps& d, int s, double* out, hipStream_t st);
hipError_t launch_demag_llg(const DeviceOps& d, bool mixed, const FieldArgs& f,
                            const FieldBufs& b, double* dmdt, bool do_demag, hipStream_t st);
hipError_t launch_jtimes(const DeviceOps& d, bool mixed, const double* m, const double* Htotal,
                         const double* v, double* Jv, hipStream_t st);

// Register-slot width of the CG kernel for an operator whose widest slice has
// `width` slots at `rpt` rows per thread (0: slots streamed from memory).
int cg_slot_variant(int width, int rpt);
// Cooperative-launch capacity of the CG kernel (workgroups of 256 threads).
int cg_max_workgroups(int rpt, int maxw, int device, bool fine);
// Workgroups of `kernel` (`threads` threads each) that one CU holds at once:
// the runtime's occupancy figure, capped by kLdsPerCu over the kernel's LDS
// rounded up to whole allocation granules (the runtime's figure divides by
// the unrounded size).  0 if the device cannot be queried.
int resident_per_cu(const void* kernel, int threads, int device);

}  // namespace nsim
