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

}  // namespace

std::string build_cg_operator(const HostOperators& o, int cap, CgOperator& cg) {
  const int n = o.n;
  cg.cap = cap;
  cg.parts = (n + cap - 1) / cap;
  cg.npad = cg.parts * cap;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<std::pair<int, int>> parts;
  rcb(o.x, idx, 0, n, cg.parts, parts);
  if ((int)parts.size() != cg.parts) return "partitioning failed";
  cg.p2o.assign(cg.npad, -1);
  std::vector<int> o2p(n, -1), part_of(n, -1);
  for (int g = 0; g < cg.parts; g++) {
    const int sz = parts[g].second - parts[g].first;
    if (sz > cap) return "partition part exceeds its capacity";
    for (int i = 0; i < sz; i++) {
      const int node = idx[parts[g].first + i];
      cg.p2o[g * cap + i] = node;
      o2p[node] = g * cap + i;
      part_of[node] = g;
    }
  }
  // coarse aggregates: runs of 2^agg_shift consecutive parts, at most 64
  cg.agg_shift = 0;
  while (((cg.parts + (1 << cg.agg_shift) - 1) >> cg.agg_shift) > 64) cg.agg_shift++;
  cg.naggs = (cg.parts + (1 << cg.agg_shift) - 1) >> cg.agg_shift;

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
    const int a = part_of[i] >> cg.agg_shift;
    const bool fi = o.blocal[i] < 0;
    if (fi) has_free[a] = 1;
    for (int k = o.rp[i]; k < o.rp[i + 1]; k++) {
      const int j = o.ci[k], b = part_of[j] >> cg.agg_shift;
      An[a * na + b] += o.K[k];
      if (fi && o.blocal[j] < 0) Ad[a * na + b] += o.K[k];
    }
  }
  // The Neumann coarse matrix is singular along the constants e.  With
  // tau > 0, (A + (tau/na) e e^T)^-1 = A^+ + e e^T / (tau na), so subtracting
  // 1/(tau na) from every entry gives the pseudo-inverse: no coarse
  // correction along the constants, whatever tau.  tau is kept at the scale of
  // the operator (never the roundoff-sized trace of a single aggregate).
  double tr = 0.0, kd = 0.0;
  for (int a = 0; a < na; a++) tr += An[a * na + a];
  for (int i = 0; i < n; i++) kd += o.Kdiag[i];
  const double tau = std::max(tr / na, kd / n);
  for (int a = 0; a < na; a++)
    for (int b = 0; b < na; b++) An[a * na + b] += tau / na;
  for (int a = 0; a < na; a++)
    if (!has_free[a]) {
      for (int b = 0; b < na; b++) Ad[a * na + b] = Ad[b * na + a] = 0.0;
      Ad[a * na + a] = 1.0;
    }
  if (!spd_inverse(An, na)) return "Neumann coarse matrix is not positive definite";
  for (double& v : An) v -= 1.0 / (tau * na);
  if (!spd_inverse(Ad, na)) return "Dirichlet coarse matrix is not positive definite";
  for (int a = 0; a < na; a++)
    if (!has_free[a]) Ad[a * na + a] = 0.0;  // no coarse correction there
  cg.ainv_nbc.swap(An);
  cg.ainv_dbc.swap(Ad);
  return "";
}

}  // namespace nsim
