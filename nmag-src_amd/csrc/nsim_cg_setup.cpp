// Host set-up of the CG's reordered stiffness operator and its two-level
// preconditioner (see CgOperator in nsim_internal.h).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "nsim_internal.h"

namespace nsim {

namespace {

// Splits nodes idx[lo, hi) into k parts of near-equal size by recursive
// coordinate bisection along the longest bounding-box axis; emits the parts
// in tree order, so consecutive parts are spatial neighbours.
void rcb(const std::vector<double>& x, std::vector<int>& idx, int lo, int hi, int k,
         std::vector<std::pair<int, int>>& parts) {
  if (k <= 1) {
    parts.push_back({lo, hi});
    return;
  }
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = lo; i < hi; i++)
    for (int d = 0; d < 3; d++) {
      mn[d] = std::min(mn[d], x[3 * idx[i] + d]);
      mx[d] = std::max(mx[d], x[3 * idx[i] + d]);
    }
  int ax = 0;
  for (int d = 1; d < 3; d++)
    if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
  const int kl = k / 2;
  const int mid = lo + (int)((long long)(hi - lo) * kl / k);
  std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
    const double xa = x[3 * a + ax], xb = x[3 * b + ax];
    return xa < xb || (xa == xb && a < b);
  });
  rcb(x, idx, lo, mid, kl, parts);
  rcb(x, idx, mid, hi, k - kl, parts);
}

// In-place inverse of an SPD matrix (Cholesky); false if not positive definite.
bool spd_inverse(std::vector<double>& A, int n) {
  std::vector<double> L(n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double s = A[j * n + j];
    for (int k = 0; k < j; k++) s -= L[j * n + k] * L[j * n + k];
    if (!(s > 0.0)) return false;
    L[j * n + j] = std::sqrt(s);
    for (int i = j + 1; i < n; i++) {
      double t = A[i * n + j];
      for (int k = 0; k < j; k++) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / L[j * n + j];
    }
  }
  // A^-1 = L^-T L^-1, column by column
  std::vector<double> inv(n * n, 0.0), y(n);
  for (int c = 0; c < n; c++) {
    for (int i = 0; i < n; i++) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) t -= L[i * n + k] * y[k];
      y[i] = t / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double t = y[i];
      for (int k = i + 1; k < n; k++) t -= L[k * n + i] * inv[k * n + c];
      inv[i * n + c] = t / L[i * n + i];
    }
  }
  A.swap(inv);
  return true;
}

// In-place pseudo-inverse of a symmetric positive semi-definite matrix from its
// eigen-decomposition by cyclic Jacobi rotations; eigenvalues at or below
// rel * max(lambda_max, scale) are taken as zero (the null space gets no
// correction).  `scale` is the operator's own magnitude: a matrix whose every
// eigenvalue is roundoff (one aggregate) must come out as zero, not 1/roundoff.
void sym_pinv(std::vector<double>& A, int n, double rel, double scale) {
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, dg = 0.0;
    for (int p = 0; p < n; p++) {
      dg += A[p * n + p] * A[p * n + p];
      for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
    }
    if (off <= 1e-32 * dg) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (apq == 0.0) continue;
        const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p, q
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p, q
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  double lmax = scale;
  for (int k = 0; k < n; k++) lmax = std::max(lmax, A[k * n + k]);
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int k = 0; k < n; k++) {
    const double lam = A[k * n + k];
    if (!(lam > rel * lmax)) continue;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) inv[i * n + j] += V[i * n + k] * V[j * n + k] / lam;
  }
  A.swap(inv);
}

// Pseudo-inverse of the Neumann coarse matrix A by deflation.  Its null space
// is spanned by the vectors constant on each connected set of aggregates (the
// aggregates that one mesh component covers, linked through the edges of the
// stiffness pattern; an aggregate without nodes is a set of its own).  With Z
// their normalised indicators, A + c Z Z^T is SPD and
//     A^+ = (A + c Z Z^T)^-1 - Z Z^T / c.
// False (A unchanged) if the shifted matrix is not positive definite.
bool neumann_pinv(const HostOperators& o, const std::vector<int>& agg_of, std::vector<double>& A,
                  int na, double c) {
  std::vector<int> up(na);
  std::iota(up.begin(), up.end(), 0);
  auto find = [&](int a) {
    while (up[a] != a) a = up[a] = up[up[a]];
    return a;
  };
  for (int i = 0; i < o.n; i++)
    for (int k = o.rp[i]; k < o.rp[i + 1]; k++) {
      const int a = find(agg_of[i]), b = find(agg_of[o.ci[k]]);
      if (a != b) up[std::max(a, b)] = std::min(a, b);
    }
  std::vector<int> root(na), size(na, 0);
  for (int a = 0; a < na; a++) size[root[a] = find(a)]++;
  std::vector<double> B(A);
  for (int a = 0; a < na; a++)
    for (int b = 0; b < na; b++)
      if (root[a] == root[b]) B[a * na + b] += c / size[root[a]];
  if (!spd_inverse(B, na)) return false;
  for (int a = 0; a < na; a++)
    for (int b = 0; b < na; b++)
      if (root[a] == root[b]) B[a * na + b] -= 1.0 / (c * size[root[a]]);
  A.swap(B);
  return true;
}

// Coordinates for the bisection: a batch's magnets are shifted to lie side by
// side along x with gaps, so that bisection separates magnets before it cuts
// one; a single magnet keeps its own coordinates.
std::vector<double> bisection_coords(const HostOperators& o) {
  if (o.ninst <= 1 || o.inst.empty()) return o.x;
  const int ni = o.ninst;
  std::vector<double> lo(3 * ni, 1e300), hi(3 * ni, -1e300);
  for (int i = 0; i < o.n; i++)
    for (int d = 0; d < 3; d++) {
      const int k = o.inst[i];
      lo[3 * k + d] = std::min(lo[3 * k + d], o.x[3 * i + d]);
      hi[3 * k + d] = std::max(hi[3 * k + d], o.x[3 * i + d]);
    }
  double gap = 0.0;
  for (int k = 0; k < ni; k++)
    for (int d = 0; d < 3; d++) gap = std::max(gap, hi[3 * k + d] - lo[3 * k + d]);
  std::vector<double> shift(ni, 0.0);
  double at = 0.0;
  for (int k = 0; k < ni; k++) {
    if (hi[3 * k] < lo[3 * k]) continue;  // magnet without nodes
    shift[k] = at - lo[3 * k];
    at += hi[3 * k] - lo[3 * k] + gap + 1.0;
  }
  std::vector<double> x(o.x);
  for (int i = 0; i < o.n; i++) {
    const int k = o.inst[i];
    x[3 * i] += shift[k];
    x[3 * i + 1] -= lo[3 * k + 1];
    x[3 * i + 2] -= lo[3 * k + 2];
  }
  return x;
}

}  // namespace

std::string build_cg_operator(const HostOperators& o, int cap, CgOperator& cg, bool fine) {
  const int n = o.n;
  cg.cap = cap;
  cg.parts = (n + cap - 1) / cap;
  cg.npad = cg.parts * cap;
  // sub-parts per part: the most (at most 4, at least 64 rows each) that keep
  // the aggregates within kCgFineAggs
  cg.sub_shift = 0;
  if (fine && cg.parts <= kCgFineParts)
    while (cg.sub_shift < 2 && (cap >> (cg.sub_shift + 1)) >= 64 &&
           (cg.parts << (cg.sub_shift + 1)) <= kCgFineAggs)
      cg.sub_shift++;
  const int S = 1 << cg.sub_shift, sub = cap >> cg.sub_shift;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  // parts x S leaves in tree order: leaf s of part g fills rows
  // g cap + s sub ... (each leaf holds at most n / (parts S) rounded up <= sub)
  std::vector<std::pair<int, int>> parts;
  rcb(bisection_coords(o), idx, 0, n, cg.parts * S, parts);
  if ((int)parts.size() != cg.parts * S) return "partitioning failed";
  cg.p2o.assign(cg.npad, -1);
  std::vector<int> o2p(n, -1), part_of(n, -1), agg_of(n, -1);
  for (int l = 0; l < cg.parts * S; l++) {
    const int sz = parts[l].second - parts[l].first;
    if (sz > sub) return "partition part exceeds its capacity";
    const int g = l / S, s = l % S;
    for (int i = 0; i < sz; i++) {
      const int node = idx[parts[l].first + i];
      const int row = g * cap + s * sub + i;
      cg.p2o[row] = node;
      o2p[node] = row;
      part_of[node] = g;
      agg_of[node] = l;
    }
  }
  // coarse aggregates: the sub-parts, or runs of 2^agg_shift consecutive
  // parts, at most 64
  cg.agg_shift = 0;
  if (S > 1) {
    cg.naggs = cg.parts * S;
  } else {
    while (((cg.parts + (1 << cg.agg_shift) - 1) >> cg.agg_shift) > 64) cg.agg_shift++;
    cg.naggs = (cg.parts + (1 << cg.agg_shift) - 1) >> cg.agg_shift;
    for (int i = 0; i < n; i++) agg_of[i] = part_of[i] >> cg.agg_shift;
  }

  // SELL-64 operator in the permuted order
  cg.nslices = cg.npad / kSlice;
  cg.sell_off.assign(cg.nslices + 1, 0);
  cg.max_width = 0;
  for (int s = 0; s < cg.nslices; s++) {
    int w = 0;
    for (int l = 0; l < kSlice; l++) {
      const int node = cg.p2o[s * kSlice + l];
      if (node >= 0) w = std::max(w, o.rp[node + 1] - o.rp[node]);
    }
    cg.max_width = std::max(cg.max_width, w);
    cg.sell_off[s + 1] = cg.sell_off[s] + w * kSlice;
  }
  const int tot = cg.sell_off[cg.nslices];
  cg.sell_col.assign(tot, 0);
  cg.sK.assign(tot, 0.0);
  for (int s = 0; s < cg.nslices; s++) {
    const int w = (cg.sell_off[s + 1] - cg.sell_off[s]) / kSlice;
    for (int l = 0; l < kSlice; l++) {
      const int row = s * kSlice + l, node = cg.p2o[row];
      const int len = node >= 0 ? o.rp[node + 1] - o.rp[node] : 0;
      for (int k = 0; k < w; k++) {
        const size_t e = (size_t)cg.sell_off[s] + (size_t)k * kSlice + l;
        if (k < len) {
          cg.sell_col[e] = o2p[o.ci[o.rp[node] + k]];
          cg.sK[e] = o.K[o.rp[node] + k];
        } else {
          cg.sell_col[e] = row;  // padding: the row itself with value 0
        }
      }
    }
  }
  cg.dinv_nbc.assign(cg.npad, 0.0);
  cg.dinv_dbc.assign(cg.npad, 0.0);
  for (int r = 0; r < cg.npad; r++) {
    const int node = cg.p2o[r];
    if (node < 0) continue;
    cg.dinv_nbc[r] = 1.0 / o.Kdiag[node];
    cg.dinv_dbc[r] = o.blocal[node] >= 0 ? 0.0 : cg.dinv_nbc[r];
  }

  // Galerkin coarse matrices over the aggregates
  const int na = cg.naggs;
  std::vector<double> An(na * na, 0.0), Ad(na * na, 0.0);
  std::vector<char> has_free(na, 0);
  for (int i = 0; i < n; i++) {
    const int a = agg_of[i];
    const bool fi = o.blocal[i] < 0;
    if (fi) has_free[a] = 1;
    for (int k = o.rp[i]; k < o.rp[i + 1]; k++) {
      const int j = o.ci[k], b = agg_of[j];
      An[a * na + b] += o.K[k];
      if (fi && o.blocal[j] < 0) Ad[a * na + b] += o.K[k];
    }
  }
  // The Neumann coarse matrix is singular along every vector that is constant
  // on each magnet (one such direction per magnet whose aggregates it does not
  // share), and along each aggregate without nodes; its pseudo-inverse gives
  // no coarse correction there.  It is formed by deflating those directions
  // with the operator scale c (the mean of sum K_ii over an aggregate), in
  // O(naggs^3); the eigen-decomposition is the fallback.  Roundoff leaves the
  // null eigenvalues near 1e-16 of c; the smallest genuine one is above 1e-4
  // of the largest at 64 aggregates and falls roughly as naggs^(-2/3) beyond.
  double kd = 0.0;
  for (int i = 0; i < n; i++) kd += o.Kdiag[i];
  if (!neumann_pinv(o, agg_of, An, na, kd / na)) sym_pinv(An, na, 1e-9, kd / na);
  for (int a = 0; a < na; a++)
    if (!has_free[a]) {
      for (int b = 0; b < na; b++) Ad[a * na + b] = Ad[b * na + a] = 0.0;
      Ad[a * na + a] = 1.0;
    }
  if (!spd_inverse(Ad, na)) return "Dirichlet coarse matrix is not positive definite";
  for (int a = 0; a < na; a++)
    if (!has_free[a]) Ad[a * na + a] = 0.0;  // no coarse correction there
  cg.ainv_nbc.swap(An);
  cg.ainv_dbc.swap(Ad);

  // fine coarse spaces: the aggregates each part's rows touch, its own first
  if (S > 1) {
    cg.need.assign((size_t)cg.parts * kCgNeed, 0);
    cg.need_n.assign(cg.parts, 0);
    std::vector<int> mark(na, -1), other;
    for (int g = 0; g < cg.parts; g++) {
      int* nd = &cg.need[(size_t)g * kCgNeed];
      int c = 0;
      for (int s = 0; s < S; s++) {
        nd[c++] = g * S + s;
        mark[g * S + s] = g;
      }
      other.clear();
      for (int r = g * cap; r < (g + 1) * cap; r++) {
        const int node = cg.p2o[r];
        if (node < 0) continue;
        for (int k = o.rp[node]; k < o.rp[node + 1]; k++) {
          const int b = agg_of[o.ci[k]];
          if (mark[b] != g) {
            mark[b] = g;
            other.push_back(b);
          }
        }
      }
      if (c + (int)other.size() > kCgNeed)
        return "a CG part touches more coarse aggregates than a workgroup holds";
      std::sort(other.begin(), other.end());
      for (int b : other) nd[c++] = b;
      cg.need_n[g] = c;
      cg.need_max = std::max(cg.need_max, c);
    }
    // as many threads per coarse row (at most a wavefront) as the longest
    // list leaves
    cg.need_lg = 6;
    while ((cg.need_max << cg.need_lg) > 256) cg.need_lg--;
  }
  return "";
}

}  // namespace nsim
