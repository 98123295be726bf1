// Internal structures shared by the host set-up, the kernels and the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nsim.h"

namespace nsim {

constexpr int kSlice = 64;  // SELL slice height = one wavefront

// Host-side mesh preprocessing result (all FP64, host memory).
struct HostOperators {
  int n = 0, T = 0;
  std::vector<double> x;       // 3n coordinates
  std::vector<int> tet;        // 4T
  std::vector<double> V;       // n box volumes (src/fem.ml:1757-1788)
  // node-adjacency CSR (sorted, diagonal included)
  std::vector<int> rp, ci;
  std::vector<double> K;       // stiffness
  std::vector<double> D;       // 3/nnz: rho_i = Ms * sum_j D_ij . m_j
  std::vector<double> G;       // 3/nnz: H_demag_i = sum_j G_ij phi_j
  std::vector<double> Kdiag;
  // SELL-64 copy of the same pattern
  int nslices = 0;
  std::vector<int> sell_off;   // nslices+1, element offsets
  std::vector<int> sell_col;
  std::vector<double> sK, sD, sG;
  // boundary
  int nb = 0, ntri = 0;
  std::vector<int> bnode;      // nb, ascending global index
  std::vector<int> blocal;     // n -> local boundary index or -1
  std::vector<int> tri;        // 3*ntri (local boundary indices)
  std::vector<double> tnorm;   // 3*ntri outward unit normals
  std::vector<double> omega;   // nb interior solid angles
  // triangles incident to each boundary node: (triangle, vertex slot)
  std::vector<int> inc_rp, inc_tri;
  double geo_eps = 0.0;
  // batch of independent magnets: node -> magnet (empty for one magnet)
  int ninst = 1;
  std::vector<int> inst;
};

// Builds every host-side operator; returns "" on success, else a message.
std::string build_host_operators(int n, const double* coords, int T, const int* tets,
                                 HostOperators& ops);

// The stiffness operator in the CG's own row order: the nodes are split by
// recursive coordinate bisection into `parts` spatially compact parts of at
// most `cap` nodes; part g occupies rows [g*cap, g*cap + size_g) and the rest
// of its rows are inactive padding.  Consecutive parts are merged in groups of
// 2^agg_shift into at most 64 coarse aggregates; ainv_* are the inverses of the
// Galerkin coarse matrices P^T K P (Neumann: the pseudo-inverse, whose null
// space holds the constants of every magnet; Dirichlet: free rows and columns
// only).  The magnets of a batch are laid side by side for the bisection, so
// parts rarely straddle two magnets.
// Fine coarse spaces: a part is split into 2^sub_shift sub-parts, each its own
// aggregate, while parts x sub-parts stays within kCgFineAggs and every part's
// rows touch at most kCgNeed aggregates (its workgroup keeps those rows of the
// coarse inverse in LDS in FP32 and evaluates only them; one workgroup per CU,
// so at most kCgFineParts parts); otherwise runs of parts form at most 64
// aggregates.
constexpr int kCgFineAggs = 320;
constexpr int kCgFineParts = 256;
constexpr int kCgNeed = 64;
constexpr int kCgRecord = 12;  // doubles a CG workgroup publishes per exchange
struct CgOperator {
  int cap = 256, parts = 0, npad = 0, nslices = 0, max_width = 0;
  int agg_shift = 0, sub_shift = 0, naggs = 0;
  std::vector<int> p2o;        // npad: original node or -1
  std::vector<int> sell_off, sell_col;
  std::vector<double> sK;
  std::vector<double> dinv_nbc, dinv_dbc;  // npad, 0 on inactive / fixed rows
  std::vector<double> ainv_nbc, ainv_dbc;  // naggs x naggs
  // fine coarse spaces: per part, kCgNeed slots listing the aggregates its
  // rows' columns touch (its own sub-parts first), need_n of them used
  std::vector<int> need, need_n;
  int need_max = 0, need_lg = 2;  // longest list; log2 threads per coarse row (2^need_lg * need_max <= 256)
};
// fine: allow sub-part aggregates (see kCgFineAggs).
std::string build_cg_operator(const HostOperators& o, int cap, CgOperator& cg, bool fine);

}  // namespace nsim
