// Host side of the H-matrix compression of the demag potential operator (see
// nsim_hmatrix_build.h for the format).  Pure C++ with std::thread.
#include "nsim_hmatrix_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

namespace nsim {

namespace {

constexpr int kLeafMax = 128;  // largest leaf cluster
constexpr int kAlign = 64;     // cluster sizes in whole multiples of this
constexpr int kProbes = 6;     // Gaussian probe vectors of the block check
constexpr int kRows = 16;      // rows per panel
constexpr int kChunk = kItemGroups;  // column groups per work item

struct Cluster {
  int b = 0, e = 0;  // range of the cluster in the tree's node order
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int child[2] = {-1, -1};
  bool leaf() const { return child[0] < 0; }
  int size() const { return e - b; }
};

struct Tree {
  std::vector<int> perm;  // tree order -> local node
  std::vector<Cluster> c;
};

int build_cluster(Tree& t, const double* xyz, int b, int e) {
  const int id = (int)t.c.size();
  t.c.emplace_back();
  Cluster cl;
  cl.b = b;
  cl.e = e;
  for (int d = 0; d < 3; d++) {
    cl.lo[d] = 1e300;
    cl.hi[d] = -1e300;
  }
  for (int q = b; q < e; q++)
    for (int d = 0; d < 3; d++) {
      const double v = xyz[3 * t.perm[q] + d];
      cl.lo[d] = std::min(cl.lo[d], v);
      cl.hi[d] = std::max(cl.hi[d], v);
    }
  const int n = e - b;
  if (n > kLeafMax) {
    int c1 = kAlign * std::max(1, (int)std::lround(n / (2.0 * kAlign)));
    if (c1 >= n) c1 = kAlign * ((n - 1) / kAlign);
    int axis = 0;
    for (int d = 1; d < 3; d++)
      if (cl.hi[d] - cl.lo[d] > cl.hi[axis] - cl.lo[axis]) axis = d;
    // the c1 nodes with the smallest coordinate along the axis (ties by node
    // index: a unique set), then each side on its own
    std::nth_element(t.perm.begin() + b, t.perm.begin() + b + c1, t.perm.begin() + e,
                     [&](int p, int q) {
                       const double a = xyz[3 * p + axis], c = xyz[3 * q + axis];
                       return a < c || (a == c && p < q);
                     });
    const int l = build_cluster(t, xyz, b, b + c1);
    const int r = build_cluster(t, xyz, b + c1, e);
    cl.child[0] = l;
    cl.child[1] = r;
  } else {
    std::sort(t.perm.begin() + b, t.perm.begin() + e);
  }
  t.c[id] = cl;
  return id;
}

bool admissible(const Cluster& a, const Cluster& b, double eta) {
  double da = 0.0, db = 0.0, dist = 0.0;
  for (int d = 0; d < 3; d++) {
    da += (a.hi[d] - a.lo[d]) * (a.hi[d] - a.lo[d]);
    db += (b.hi[d] - b.lo[d]) * (b.hi[d] - b.lo[d]);
    const double g = std::max(0.0, std::max(a.lo[d] - b.hi[d], b.lo[d] - a.hi[d]));
    dist += g * g;
  }
  return dist > 0.0 && std::sqrt(std::min(da, db)) <= eta * std::sqrt(dist);
}

struct Block {
  int s, t;
  bool lowrank;
};

void partition(const Tree& T, int s, int t, double eta, std::vector<Block>& out) {
  const Cluster &cs = T.c[s], &ct = T.c[t];
  if (admissible(cs, ct, eta)) {
    out.push_back({s, t, true});
    return;
  }
  if (cs.leaf() && ct.leaf()) {
    out.push_back({s, t, false});
    return;
  }
  if (cs.leaf()) {
    for (int q : ct.child) partition(T, s, q, eta, out);
    return;
  }
  if (ct.leaf()) {
    for (int q : cs.child) partition(T, q, t, eta, out);
    return;
  }
  for (int p : cs.child)
    for (int q : ct.child) partition(T, p, q, eta, out);
}

// Entry (i, j) of the block of clusters (s, t), in tree order.
struct Acc {
  const double* A;
  long long lda;
  const int* rows;  // local node of each block row
  const int* cols;
  double operator()(int i, int j) const { return A[rows[i] + (long long)cols[j] * lda]; }
};

double dot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

struct LowRank {
  int k = -1;              // rank; -1: store dense
  std::vector<double> U;   // m x k, column l at U[l m]
  std::vector<double> V;   // n x k
};

// Adaptive cross approximation with partial pivoting, continued from the
// current (U, V) until the estimated relative error is below eps or the rank
// reaches kmax.  Returns true if it converged.
bool aca(const Acc& A, int m, int n, double eps, int kmax, std::vector<double>& U,
         std::vector<double>& V, int& k) {
  std::vector<char> used(m, 0);
  std::vector<double> r(n), c(m);
  double norm2 = 0.0;
  for (int l = 0; l < k; l++)  // the estimate of |U V^T|_F^2 of what is already there
    for (int q = 0; q < k; q++)
      norm2 += dot(&U[(size_t)l * m], &U[(size_t)q * m], m) *
               dot(&V[(size_t)l * n], &V[(size_t)q * n], n);
  int i = 0, zero_rows = 0;
  while (k < kmax) {
    for (int j = 0; j < n; j++) {
      double v = A(i, j);
      for (int l = 0; l < k; l++) v -= U[(size_t)l * m + i] * V[(size_t)l * n + j];
      r[j] = v;
    }
    used[i] = 1;
    int jp = 0;
    for (int j = 1; j < n; j++)
      if (std::fabs(r[j]) > std::fabs(r[jp])) jp = j;
    const double piv = r[jp];
    if (piv == 0.0) {  // this row is represented already: try another one
      int nxt = -1;
      for (int q = 0; q < m; q++)
        if (!used[q]) {
          nxt = q;
          break;
        }
      if (nxt < 0 || ++zero_rows > 8) return k > 0;
      i = nxt;
      continue;
    }
    zero_rows = 0;
    for (int j = 0; j < n; j++) r[j] /= piv;
    for (int q = 0; q < m; q++) {
      double v = A(q, jp);
      for (int l = 0; l < k; l++) v -= V[(size_t)l * n + jp] * U[(size_t)l * m + q];
      c[q] = v;
    }
    const double nu2 = dot(c.data(), c.data(), m), nv2 = dot(r.data(), r.data(), n);
    double cross = 0.0;
    for (int l = 0; l < k; l++)
      cross += dot(&U[(size_t)l * m], c.data(), m) * dot(&V[(size_t)l * n], r.data(), n);
    norm2 += nu2 * nv2 + 2.0 * cross;
    U.insert(U.end(), c.begin(), c.end());
    V.insert(V.end(), r.begin(), r.end());
    k++;
    if (std::sqrt(nu2 * nv2) <= eps * std::sqrt(std::fabs(norm2))) return true;
    int nxt = -1;
    for (int q = 0; q < m; q++)
      if (!used[q] && (nxt < 0 || std::fabs(c[q]) > std::fabs(c[nxt]))) nxt = q;
    if (nxt < 0) return true;  // every row used: the approximation is exact
    i = nxt;
  }
  return false;
}

// Q R of the m x k matrix Q (in place, modified Gram-Schmidt applied twice);
// R is k x k upper triangular, column-major.
void qr(std::vector<double>& Q, int m, int k, std::vector<double>& R) {
  R.assign((size_t)k * k, 0.0);
  for (int j = 0; j < k; j++) {
    double* qj = &Q[(size_t)j * m];
    for (int pass = 0; pass < 2; pass++)
      for (int i = 0; i < j; i++) {
        const double* qi = &Q[(size_t)i * m];
        const double h = dot(qi, qj, m);
        R[i + (size_t)j * k] += h;
        for (int q = 0; q < m; q++) qj[q] -= h * qi[q];
      }
    const double nrm = std::sqrt(dot(qj, qj, m));
    R[j + (size_t)j * k] = nrm;
    if (nrm > 0.0)
      for (int q = 0; q < m; q++) qj[q] /= nrm;
  }
}

// One-sided Jacobi SVD of the k x k matrix X (column-major): on return the
// columns of X are sigma_i w_i and those of J the right singular vectors,
// both sorted by decreasing sigma; sig holds the sigma_i.
void jacobi_svd(std::vector<double>& X, int k, std::vector<double>& J,
                std::vector<double>& sig) {
  J.assign((size_t)k * k, 0.0);
  for (int i = 0; i < k; i++) J[i + (size_t)i * k] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    bool rotated = false;
    for (int p = 0; p < k - 1; p++)
      for (int q = p + 1; q < k; q++) {
        double* xp = &X[(size_t)p * k];
        double* xq = &X[(size_t)q * k];
        const double a = dot(xp, xp, k), b = dot(xq, xq, k), g = dot(xp, xq, k);
        if (!(std::fabs(g) > 1e-15 * std::sqrt(a * b)) || g == 0.0) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        double* jp = &J[(size_t)p * k];
        double* jq = &J[(size_t)q * k];
        for (int i = 0; i < k; i++) {
          const double u = xp[i], v = xq[i];
          xp[i] = cs * u - sn * v;
          xq[i] = sn * u + cs * v;
          const double s = jp[i], w = jq[i];
          jp[i] = cs * s - sn * w;
          jq[i] = sn * s + cs * w;
        }
      }
    if (!rotated) break;
  }
  std::vector<int> order(k);
  sig.resize(k);
  for (int i = 0; i < k; i++) {
    order[i] = i;
    sig[i] = std::sqrt(dot(&X[(size_t)i * k], &X[(size_t)i * k], k));
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
  std::vector<double> X2(X.size()), J2(J.size()), s2(k);
  for (int i = 0; i < k; i++) {
    std::copy_n(&X[(size_t)order[i] * k], k, &X2[(size_t)i * k]);
    std::copy_n(&J[(size_t)order[i] * k], k, &J2[(size_t)i * k]);
    s2[i] = sig[order[i]];
  }
  X.swap(X2);
  J.swap(J2);
  sig.swap(s2);
}

// U V^T (rank k) recompressed to the smallest rank whose discarded singular
// values have Frobenius norm at most tol times that of all of them.
void recompress(std::vector<double>& U, std::vector<double>& V, int m, int n, int& k,
                double tol) {
  if (k <= 0) return;
  std::vector<double> Ru, Rv;
  qr(U, m, k, Ru);
  qr(V, n, k, Rv);
  std::vector<double> C((size_t)k * k, 0.0);  // Ru Rv^T
  for (int j = 0; j < k; j++)
    for (int l = 0; l < k; l++) {
      const double rv = Rv[j + (size_t)l * k];
      if (rv == 0.0) continue;
      for (int i = 0; i < k; i++) C[i + (size_t)j * k] += Ru[i + (size_t)l * k] * rv;
    }
  std::vector<double> J, sig;
  jacobi_svd(C, k, J, sig);
  double total = 0.0;
  for (double s : sig) total += s * s;
  int kk = k;
  double tail = 0.0;
  while (kk > 0 && tail + sig[kk - 1] * sig[kk - 1] <= tol * tol * total) {
    tail += sig[kk - 1] * sig[kk - 1];
    kk--;
  }
  std::vector<double> U2((size_t)m * kk, 0.0), V2((size_t)n * kk, 0.0);
  for (int c = 0; c < kk; c++) {
    for (int l = 0; l < k; l++) {
      const double a = C[l + (size_t)c * k], b = J[l + (size_t)c * k];
      for (int i = 0; i < m; i++) U2[(size_t)c * m + i] += U[(size_t)l * m + i] * a;
      for (int j = 0; j < n; j++) V2[(size_t)c * n + j] += V[(size_t)l * n + j] * b;
    }
  }
  U.swap(U2);
  V.swap(V2);
  k = kk;
}

// |A w - U V^T w|^2 and |A w|^2 summed over Gaussian probes w (seeded).
void probe(const Acc& A, int m, int n, const LowRank& lr, unsigned long long seed,
           double* err2, double* ref2) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<double> w((size_t)n * kProbes), y((size_t)m * kProbes, 0.0);
  for (auto& v : w) v = nd(rng);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      const double a = A(i, j);
      for (int p = 0; p < kProbes; p++) y[(size_t)p * m + i] += a * w[(size_t)p * n + j];
    }
  double e = 0.0, r = 0.0;
  std::vector<double> z(lr.k > 0 ? lr.k : 1);
  for (int p = 0; p < kProbes; p++) {
    for (int l = 0; l < lr.k; l++) z[l] = dot(&lr.V[(size_t)l * n], &w[(size_t)p * n], n);
    for (int i = 0; i < m; i++) {
      double v = 0.0;
      for (int l = 0; l < lr.k; l++) v += lr.U[(size_t)l * m + i] * z[l];
      const double yi = y[(size_t)p * m + i];
      e += (yi - v) * (yi - v);
      r += yi * yi;
    }
  }
  *err2 = e;
  *ref2 = r;
}

// Low-rank factors of one admissible block, or k = -1 (store it dense).
LowRank compress_block(const Acc& A, int m, int n, double eps, unsigned long long seed,
                       int* fallback) {
  LowRank lr;
  const int kmax = std::max(1, (int)(((long long)m * n) / (m + n)) - 1);
  double e_aca = 0.25 * eps;
  std::vector<double> U, V;
  int k = 0;
  for (int attempt = 0; attempt < 3; attempt++) {
    const bool ok = aca(A, m, n, e_aca, kmax, U, V, k);
    if (!ok) break;  // no rank below kmax reaches the tolerance
    LowRank t;
    t.U = U;
    t.V = V;
    t.k = k;
    recompress(t.U, t.V, m, n, t.k, 0.25 * eps);
    double err2 = 0.0, ref2 = 0.0;
    probe(A, m, n, t, seed + attempt, &err2, &ref2);
    if (err2 <= eps * eps * ref2) return t;
    e_aca *= 0.125;  // continue the cross approximation further
  }
  (*fallback)++;
  return lr;
}

}  // namespace

std::string hmatrix_compress(int N, const double* A, long long lda, const double* xyz,
                             int node_off, double eps, double eta, int threads,
                             HMatrixHost& out) {
  if (N <= 0 || !A || !xyz || !(eps > 0.0) || !(eta > 0.0))
    return "hmatrix_compress: invalid arguments";
  Tree T;
  T.perm.resize(N);
  for (int i = 0; i < N; i++) T.perm[i] = i;
  build_cluster(T, xyz, 0, N);
  std::vector<Block> blocks;
  partition(T, 0, 0, eta, blocks);
  const int nb = (int)blocks.size();

  // low-rank factors, block by block on the host threads
  std::vector<LowRank> lr(nb);
  std::vector<int> fallback(nb, 0);
  std::atomic<int> next(0);
  auto work = [&]() {
    for (int b = next++; b < nb; b = next++) {
      if (!blocks[b].lowrank) continue;
      const Cluster &s = T.c[blocks[b].s], &t = T.c[blocks[b].t];
      Acc acc{A, lda, &T.perm[s.b], &T.perm[t.b]};
      lr[b] = compress_block(acc, s.size(), t.size(), eps, 0x9e3779b97f4a7c15ull * (b + 1),
                             &fallback[b]);
    }
  };
  const int nthr = std::max(1, std::min(threads, 64));
  std::vector<std::thread> pool;
  for (int q = 1; q < nthr; q++) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();

  // t offsets (this magnet's, from 0; out_nt0 = entries of earlier magnets);
  // statistics
  const int out_nt0 = out.nt;
  std::vector<int> toff(nb, -1);
  for (int b = 0; b < nb; b++) {
    const Cluster &s = T.c[blocks[b].s], &t = T.c[blocks[b].t];
    const long long mn = (long long)s.size() * t.size();
    out.full_entries += mn;
    if (blocks[b].lowrank && lr[b].k >= 0) {
      toff[b] = out.nt - out_nt0;
      out.nt += lr[b].k;
      out.lowrank_entries += (long long)lr[b].k * (s.size() + t.size());
      out.lowrank_blocks++;
      out.max_rank = std::max(out.max_rank, lr[b].k);
    } else {
      out.dense_entries += mn;
      out.dense_blocks++;
    }
    out.fallback_blocks += fallback[b];
  }
  // node -> position in the cluster order (global, as the inputs)
  if ((int)out.pos.size() < node_off + N) out.pos.resize(node_off + N, -1);
  for (int q = 0; q < N; q++) out.pos[node_off + T.perm[q]] = node_off + q;

  // A panel's columns are runs ("segments") of consecutive sources: a
  // cluster's rho entries in cluster order, or a block's t entries.  Each
  // segment is padded to whole groups of G columns, the padding continuing the
  // source numbering with zero entries, so a group's G sources are always
  // consecutive and one index per group names them.
  const int G = out.group;
  struct Seg {
    int src0;   // first source (encoded as in gsrc)
    int width;  // real columns
    int b;      // block (dense or low-rank), or -1 for a phase-1 stack
  };
  auto add_panel = [&](HPanels& P, int R, const std::vector<Seg>& segs, auto&& dst_of,
                       auto&& entry) {
    int C = 0;
    for (const Seg& g : segs) C += (g.width + G - 1) / G * G;
    const long long base = (long long)P.data.size();
    P.data.resize(base + (long long)R * C, 0.0);
    // one run of sources: no per-group index needed
    const bool run = segs.size() == 1;
    P.desc.push_back({base, run ? -1 : (int)P.gsrc.size(), C / G, (int)P.rowdst.size(),
                      segs[0].src0, -1, 0, 1, -1});
    for (int r = 0; r < R; r++) P.rowdst.push_back(dst_of(r));
    int c = 0;
    for (size_t q = 0; q < segs.size(); q++) {
      const Seg& g = segs[q];
      const int wpad = (g.width + G - 1) / G * G;
      if (!run)
        for (int j = 0; j < wpad; j += G) P.gsrc.push_back(g.src0 >= 0 ? g.src0 + j : g.src0 - j);
      for (int j = 0; j < g.width; j++)
        for (int r = 0; r < R; r++)
          if (dst_of(r) >= 0)
            P.data[base + ((long long)((c + j) / G) * R + r) * G + (c + j) % G] =
                entry((int)q, r, j);
      c += wpad;
    }
  };

  // phase 1: per column cluster, the stacked V^T of its low-rank blocks
  {
    std::vector<std::vector<int>> by_col(T.c.size());
    for (int b = 0; b < nb; b++)
      if (toff[b] >= 0) by_col[blocks[b].t].push_back(b);
    for (size_t t = 0; t < T.c.size(); t++) {
      if (by_col[t].empty()) continue;
      const Cluster& ct = T.c[t];
      std::vector<std::pair<int, int>> stack;  // (block, rank index)
      for (int b : by_col[t])
        for (int l = 0; l < lr[b].k; l++) stack.emplace_back(b, l);
      const std::vector<Seg> segs{{node_off + ct.b, ct.size(), -1}};
      for (size_t r0 = 0, R = kRows; r0 < stack.size(); r0 += R) {
        auto row = [&](int r) -> const std::pair<int, int>* {
          return r0 + r < stack.size() ? &stack[r0 + r] : nullptr;
        };
        add_panel(
            out.p1, (int)R, segs,
            [&](int r) { auto p = row(r); return p ? out_nt0 + toff[p->first] + p->second : -1; },
            [&](int, int r, int j) {
              auto p = row(r);
              return lr[p->first].V[(size_t)p->second * ct.size() + j];
            });
      }
    }
  }
  // phase 2: per leaf, runs of R rows over its dense blocks and the U of the
  // low-rank blocks whose row cluster holds it
  {
    std::vector<std::vector<int>> anc(T.c.size());  // leaf -> blocks covering its rows
    std::vector<int> leaf_of_pos(N, -1);
    for (size_t c = 0; c < T.c.size(); c++)
      if (T.c[c].leaf())
        for (int q = T.c[c].b; q < T.c[c].e; q++) leaf_of_pos[q] = (int)c;
    for (int b = 0; b < nb; b++) {
      const Cluster& s = T.c[blocks[b].s];
      int prev = -1;
      for (int q = s.b; q < s.e; q++)
        if (leaf_of_pos[q] != prev) {
          prev = leaf_of_pos[q];
          anc[prev].push_back(b);
        }
    }
    for (size_t c = 0; c < T.c.size(); c++) {
      const Cluster& leaf = T.c[c];
      if (!leaf.leaf()) continue;
      std::vector<Seg> segs;
      for (int b : anc[c]) {
        if (toff[b] >= 0) segs.push_back({-(out_nt0 + toff[b]) - 1, lr[b].k, b});
        else segs.push_back({node_off + T.c[blocks[b].t].b, T.c[blocks[b].t].size(), b});
      }
      for (int r0 = 0, R = kRows; r0 < leaf.size(); r0 += R) {
        add_panel(
            out.p2, R, segs,
            [&](int r) { return r0 + r < leaf.size() ? node_off + T.perm[leaf.b + r0 + r] : -1; },
            [&](int q, int r, int j) {
              const int b = segs[q].b;
              const int pos = leaf.b + r0 + r;  // tree position of the row
              const Cluster& s = T.c[blocks[b].s];
              if (toff[b] >= 0) return lr[b].U[(size_t)j * s.size() + (pos - s.b)];
              return A[T.perm[pos] + (long long)T.perm[T.c[blocks[b].t].b + j] * lda];
            });
      }
    }
  }
  return "";
}

void hmatrix_finish(HMatrixHost& out) {
  // Work items: panels longer than kChunk groups are cut into chunks (the
  // time of one workgroup over the longest panel would otherwise bound the
  // launch), then the items are ordered by decreasing size, so the longest
  // start first (the dispatcher starts workgroups in index order).
  for (HPanels* P : {&out.p1, &out.p2}) {
    std::vector<PanelDesc> items;
    for (const PanelDesc& p : P->desc) {
      const int nchunk = (p.ng + kChunk - 1) / kChunk;
      if (nchunk <= 1) {
        items.push_back(p);
        continue;
      }
      const int pslot = P->npartial, cslot = P->ncounters++;
      P->npartial += 16 * nchunk;
      for (int c = 0; c < nchunk; c++) {
        const int ga = c * kChunk, gb = std::min(p.ng, ga + kChunk);
        PanelDesc q = p;
        q.off = p.off + (long long)ga * kRows * out.group;
        q.g0 = p.g0 < 0 ? -1 : p.g0 + ga;
        q.src0 = p.g0 >= 0 ? p.src0 : p.src0 >= 0 ? p.src0 + ga * out.group
                                                  : p.src0 - ga * out.group;
        q.ng = gb - ga;
        q.pslot = pslot;
        q.chunk = c;
        q.nchunk = nchunk;
        q.cslot = cslot;
        items.push_back(q);
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const PanelDesc& a, const PanelDesc& b) { return a.ng > b.ng; });
    P->desc.swap(items);
  }
}

void hmatrix_apply_host(const HMatrixHost& H, const double* x, double* y) {
  const int G = H.group;
  const int n = (int)H.pos.size();
  std::vector<double> xp(n + G, 0.0), t(H.nt + G, 0.0);
  for (int i = 0; i < n; i++)
    if (H.pos[i] >= 0) xp[H.pos[i]] = x[i];
  for (int phase = 0; phase < 2; phase++) {
    const HPanels& P = phase == 0 ? H.p1 : H.p2;
    std::vector<double> partial(P.npartial, 0.0);
    auto put = [&](int dst, double v) {
      if (dst < 0) return;
      if (phase == 0) t[dst] = v;
      else y[dst] = v;
    };
    for (const PanelDesc& d : P.desc) {
      for (int r = 0; r < kRows; r++) {
        double acc = 0.0;
        for (int g = 0; g < d.ng; g++) {
          const int s = d.g0 < 0 ? (d.src0 >= 0 ? d.src0 + g * G : d.src0 - g * G)
                                 : P.gsrc[d.g0 + g];
          const double* src = s >= 0 ? &xp[s] : &t[-s - 1];
          for (int e = 0; e < G; e++)
            acc += P.data[d.off + ((long long)g * kRows + r) * G + e] * src[e];
        }
        if (d.pslot < 0) put(P.rowdst[d.r0 + r], acc);
        else partial[d.pslot + d.chunk * kRows + r] = acc;
      }
    }
    for (const PanelDesc& d : P.desc)  // the chunked panels, once each
      if (d.pslot >= 0 && d.chunk == 0)
        for (int r = 0; r < kRows; r++) {
          double s = 0.0;
          for (int c = 0; c < d.nchunk; c++) s += partial[d.pslot + c * kRows + r];
          put(P.rowdst[d.r0 + r], s);
        }
  }
}

}  // namespace nsim
