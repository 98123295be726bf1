// Device-side data of one magnet and the launch interface of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace nsim {

enum BemType { kBemF64 = 0, kBemF32 = 1, kBemBF16 = 2 };

// Material and LLG coefficients of one magnet (simulation units).
struct Coeffs {
  double Ms, negJ, c1, c2, c3, mu0;
};

// Operators resident in HBM (one copy per precision actually used).
struct DeviceOps {
  int n = 0, nb = 0, ntri = 0, nslices = 0, sell_total = 0;
  int ldb = 0;        // widest BEM segment's row stride (multiple of 64)
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
  void* B = nullptr;           // block-diagonal dense segments, each nb_s x ldb_s row-major
  int btype = kBemF64;
  // BEM segments (one per independent magnet): host copies and device tables
  std::vector<int> h_seg_nb, h_seg_ldb, h_seg_boff;
  std::vector<long long> h_seg_Boff;
  std::vector<double> h_seg_geo_eps;
  int* seg_nb = nullptr;
  int* seg_ldb = nullptr;
  int* seg_boff = nullptr;
  long long* seg_Boff = nullptr;
  int* wg_seg = nullptr;   // GEMV workgroup -> segment
  int* wg_row0 = nullptr;  // GEMV workgroup -> first row in the segment
  int gemv_grid = 0, gemv_rpw = 0, max_ldb = 0;
  int gemv_rows = 2, gemv_wg_per_cu = 2;  // tile shape: rows per pass, workgroups per CU
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
  double* part = nullptr;           // 2*gmax*2 per-workgroup partial sums
  unsigned long long* tag = nullptr;  // 2*gmax phase tags
  unsigned* sync = nullptr;         // [2] error flag [3] last iterations
                                    // [4] total iterations [5] solves
                                    // [6] solves stopped at maxit
  int gmax = 0;
  unsigned long long epoch = 0;     // launch counter; makes every phase tag unique
};

// Field inputs of one evaluation (device pointers; optional ones may be null).
struct FieldArgs {
  const double* m;
  const double* Hext;
  const double* k1;
  const double* k2;
  const double* axis;
  const double* pin;
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
  int parts = 0, npad = 0, rpt = 1, cap_shift = 8, agg_shift = 0, naggs = 0;
  int maxw = 0, max_width = 0;
  int* sell_off = nullptr;
  int* sell_col = nullptr;
  double* K = nullptr;
  int* p2o = nullptr;
  double *dinv_nbc = nullptr, *dinv_dbc = nullptr;
  double *ainv_nbc = nullptr, *ainv_dbc = nullptr;
};
// Persistent pipelined two-level PCG on the stiffness matrix; see the kernel.
// b, x, xadd, xsum are in the original node order.
hipError_t launch_cg(const CgDevice& c, CgWork& w, bool dirichlet, bool coarse,
                     const double* b, double* x, double rtol, int maxit, const double* xadd,
                     double* xsum, double* restrict_out, const int* blocal,
                     hipStream_t st);
hipError_t launch_bem_gemv(const DeviceOps& d, const double* phi1b, double* phi2,
                           hipStream_t st, int num_cus);
// Host plan of the grouped GEMV for a tile shape (rows_per_pass 1, 2 or 4 rows
// streamed together by a wavefront; wg_per_cu 1..8 workgroups per CU): fills
// d.gemv_* and max_ldb and the workgroup -> (segment, first row) table to
// upload.  False for a shape outside those ranges.
bool plan_bem_gemv(DeviceOps& d, int num_cus, int rows_per_pass, int wg_per_cu,
                   std::vector<int>* wg_seg, std::vector<int>* wg_row0);
hipError_t launch_demag_llg(const DeviceOps& d, bool mixed, const FieldArgs& f,
                            const FieldBufs& b, double* dmdt, bool do_demag, hipStream_t st);
hipError_t launch_jtimes(const DeviceOps& d, bool mixed, const double* m, const double* Htotal,
                         const double* v, double* Jv, hipStream_t st);

// Register-slot width of the CG kernel for an operator whose widest slice has
// `width` slots at `rpt` rows per thread (0: slots streamed from memory).
int cg_slot_variant(int width, int rpt);
// Cooperative-launch capacity of the CG kernel (workgroups of 256 threads).
int cg_max_workgroups(int rpt, int maxw, int device);

}  // namespace nsim
