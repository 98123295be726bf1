// Hierarchical (H-matrix) compression of a magnet's dense demag potential
// operator M = (I + E B R) K^+, host side.  Pure C++: no HIP types, so the
// builder also runs in a CPU test harness.
//
// The magnet's nodes are ordered by a geometric cluster tree (bisection of
// the bounding box's longest side, cluster sizes in whole multiples of 64
// except along the last branch, leaves of at most 128 nodes).  Pairs of
// clusters that are far apart relative to their size,
//     min(diam s, diam t) <= eta dist(s, t),
// hold a smooth part of M (the potential of a distant charge) and are stored
// as U V^T of low rank; the remaining pairs of leaves are stored dense.  The
// low-rank factors come from adaptive cross approximation with partial
// pivoting, recompressed by QR and a Jacobi SVD to the smallest rank whose
// discarded singular values have Frobenius norm at most eps/4 of the block's,
// and checked against the block itself with Gaussian probe vectors (blocks
// that fail are stored dense).  The compression error of a block is thus at
// most about eps of its Frobenius norm.
//
// The compressed operator is applied in two phases, each a list of "panels":
//   phase 1: t = V^T x for every low-rank block (the V^T of all blocks with
//            the same column cluster stacked, one panel row per t entry);
//   phase 2: y = A_dense x + U t, one panel per run of R rows of a leaf, its
//            columns the dense blocks' columns and the low-rank blocks' t.
// A panel of R = 16 rows and C columns (C a multiple of G) is stored as
// groups of G
// consecutive columns, each group R rows x G entries: entry (r, c) at
// ((c / G) R + r) G + c % G, so a wavefront's 16-byte loads read contiguous
// memory.  The inputs are read in the cluster order (x' = x permuted, pos
// below).  Each group of G columns names its first source: >= 0 the entry of
// x', < 0 the t entry -src - 1; the group's other columns take the next
// sources in order (runs of columns are padded with zero entries to whole
// groups, so x' and t need G entries of padding at their ends).  Every row
// names its destination (-1: padding).
#pragma once

#include <string>
#include <vector>

namespace nsim {

// Column groups of one work item at most: longer panels are cut into chunks
// of this many groups (64 KiB of FP32 entries).
constexpr int kItemGroups = 256;

// One work item: a panel, or a run of at most 256 of its column groups (a
// chunk).  Its entries start at data[off]; its ng column groups' sources are
// gsrc[g0 ...] (or, when the panel's columns are one run of sources, src0 +
// g G for its group g, and g0 = -1); its 16 rows' destinations are
// rowdst[r0 ...].  The chunks of a panel add their row sums into
// partial[pslot + chunk 16 + r]; the last chunk to finish (counter cslot)
// sums them in chunk order and writes the rows.  pslot = -1: a whole panel.
// (40 bytes, the same layout on the device.)
struct PanelDesc {
  long long off;
  int g0, ng, r0, src0;
  int pslot, chunk, nchunk, cslot;
};

struct HPanels {
  std::vector<PanelDesc> desc;       // work items, by decreasing size
  std::vector<int> gsrc;             // per column group: its first source
  std::vector<int> rowdst;           // 16 per panel: destination index or -1
  std::vector<double> data;          // the entries, layout above
  int npartial = 0, ncounters = 0;   // partial sums and counters of chunked panels
  int panels() const { return (int)desc.size(); }
};

struct HMatrixHost {
  int group = 4;  // G: columns per group (4 for FP32 storage, 2 for FP64)
  HPanels p1, p2;
  int nt = 0;     // t entries (sum of the ranks)
  std::vector<int> pos;  // input index -> position in the cluster order
  // statistics of the last compression(s), accumulated over magnets
  long long dense_entries = 0, lowrank_entries = 0, full_entries = 0;
  int lowrank_blocks = 0, dense_blocks = 0, max_rank = 0, fallback_blocks = 0;
};

// Compresses one magnet: A is its N x N operator, column-major with leading
// dimension lda, in the magnet's local node order; xyz its 3N node
// coordinates.  Panels are appended to out (input and output indices offset
// by node_off, t indices after out.nt).  eps: relative Frobenius tolerance of
// each block; eta: admissibility parameter; threads: host threads to use.
// Returns "" on success, else a message.
std::string hmatrix_compress(int N, const double* A, long long lda, const double* xyz,
                             int node_off, double eps, double eta, int threads,
                             HMatrixHost& out);

// Cuts long panels into chunks and orders the work items (after the last
// hmatrix_compress of a handle, before the upload or hmatrix_apply_host).
void hmatrix_finish(HMatrixHost& out);

// y = H x applied on the host from the panels (phase 1 then phase 2), for
// tests: x over the input indices, y over the output indices (entries no
// panel row names are left untouched).
void hmatrix_apply_host(const HMatrixHost& H, const double* x, double* y);

}  // namespace nsim
