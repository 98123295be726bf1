// Host-side mesh preprocessing: the set-up half of raw_make_lam.
//
// What the reference assembles through its operator vivificators
// (src/fem.ml:4890-5347) for the strings of interface/nmag/nmag_lam.py:445-451,
// 587-605 is computed here directly from the P1 element formulas:
//   K_ij = sum_t vol_t grad(l_i).grad(l_j)                (op_H_exch, op_neg_laplace_phi)
//   D_ij = sum_t vol_t/4 grad(l_i)|_t                      (op_div_m then SCALE -1)
//   G_ij = -(1/V_i) sum_t vol_t/4 grad(l_j)|_t             (op_grad_phi then CFBOX)
//   V_i  = sum_t vol_t/4                                   (src/fem.ml:1757-1788)
// plus the boundary geometry of the BEM (src/bem3d.ml:32-153, src/mesh.ml:1792).
// Set-up is O(n) and runs once per magnet; the operators are then uploaded
// and every evaluation runs on the device.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nsim_internal.h"

namespace nsim {

namespace {

struct V3 {
  double x, y, z;
};
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 at(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

// Unsigned solid angle of triangle (p0,p1,p2) seen from o (src/snippets.ml:1229).
double solid_angle(const V3& o, const V3& p0, const V3& p1, const V3& p2) {
  V3 r0 = sub(p0, o), r1 = sub(p1, o), r2 = sub(p2, o);
  double l0 = std::sqrt(dot(r0, r0)), l1 = std::sqrt(dot(r1, r1)), l2 = std::sqrt(dot(r2, r2));
  double d01 = dot(r0, r1), d12 = dot(r1, r2), d20 = dot(r2, r0);
  double num = l0 * l1 * l2 + l0 * d12 + l1 * d20 + l2 * d01;
  double den = std::sqrt(2.0 * (l1 * l2 + d12) * (l2 * l0 + d20) * (l0 * l1 + d01));
  if (den == 0.0) return 0.0;
  double q = std::min(1.0, std::max(-1.0, num / den));
  return std::fabs(2.0 * std::acos(q));
}

}  // namespace

std::string build_host_operators(int n, const double* coords, int T, const int* tets,
                                 HostOperators& o) {
  if (n <= 0 || T <= 0) return "empty mesh";
  if (n >= (1 << 21)) return "mesh too large for 21-bit face keys";
  for (int k = 0; k < 4 * T; k++)
    if (tets[k] < 0 || tets[k] >= n) return "tetrahedron references a node out of range";
  o.n = n;
  o.T = T;
  o.x.assign(coords, coords + 3 * n);
  o.tet.assign(tets, tets + 4 * T);

  // per-tet barycentric gradients and volumes
  std::vector<double> grad(12 * (size_t)T), vol(T);
  o.V.assign(n, 0.0);
  double gmax = 0.0;
  for (int i = 0; i < 3 * n; i++) gmax = std::max(gmax, std::fabs(coords[i]));
  o.geo_eps = gmax * 1e-13;
  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    V3 x0 = at(coords, v[0]);
    V3 e1 = sub(at(coords, v[1]), x0), e2 = sub(at(coords, v[2]), x0),
       e3 = sub(at(coords, v[3]), x0);
    V3 c23 = cross(e2, e3), c31 = cross(e3, e1), c12 = cross(e1, e2);
    double det = dot(e1, c23);
    if (det == 0.0) return "degenerate tetrahedron";
    double* g = grad.data() + 12 * (size_t)t;
    const V3 rows[3] = {c23, c31, c12};
    for (int a = 0; a < 3; a++) {
      g[3 * (a + 1) + 0] = rows[a].x / det;
      g[3 * (a + 1) + 1] = rows[a].y / det;
      g[3 * (a + 1) + 2] = rows[a].z / det;
    }
    for (int d = 0; d < 3; d++) g[d] = -(g[3 + d] + g[6 + d] + g[9 + d]);
    vol[t] = std::fabs(det) / 6.0;
    for (int a = 0; a < 4; a++) o.V[v[a]] += 0.25 * vol[t];
  }

  // node adjacency: rows sorted, unique, diagonal included
  std::vector<std::vector<int>> adj(n);
  for (int t = 0; t < T; t++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) adj[tets[4 * t + a]].push_back(tets[4 * t + b]);
  o.rp.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    auto& r = adj[i];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    o.rp[i + 1] = o.rp[i] + (int)r.size();
  }
  const int nnz = o.rp[n];
  o.ci.resize(nnz);
  for (int i = 0; i < n; i++) std::copy(adj[i].begin(), adj[i].end(), o.ci.begin() + o.rp[i]);
  adj.clear();
  auto find = [&](int i, int j) {
    auto b = o.ci.begin() + o.rp[i], e = o.ci.begin() + o.rp[i + 1];
    return (int)(std::lower_bound(b, e, j) - o.ci.begin());
  };

  o.K.assign(nnz, 0.0);
  o.D.assign(3 * (size_t)nnz, 0.0);
  o.G.assign(3 * (size_t)nnz, 0.0);
  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    const double* g = grad.data() + 12 * (size_t)t;
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        int k = find(v[a], v[b]);
        o.K[k] += vol[t] * (g[3 * a] * g[3 * b] + g[3 * a + 1] * g[3 * b + 1] +
                            g[3 * a + 2] * g[3 * b + 2]);
        for (int d = 0; d < 3; d++) {
          o.D[3 * (size_t)k + d] += 0.25 * vol[t] * g[3 * a + d];
          o.G[3 * (size_t)k + d] -= 0.25 * vol[t] * g[3 * b + d];
        }
      }
  }
  o.Kdiag.resize(n);
  for (int i = 0; i < n; i++) {
    o.Kdiag[i] = o.K[find(i, i)];
    if (!(o.Kdiag[i] > 0.0)) return "non-positive stiffness diagonal";
    for (int k = o.rp[i]; k < o.rp[i + 1]; k++)
      for (int d = 0; d < 3; d++) o.G[3 * (size_t)k + d] /= o.V[i];
  }

  // SELL-64: slice s holds rows 64s..64s+63, column-major inside the slice
  o.nslices = (n + kSlice - 1) / kSlice;
  o.sell_off.assign(o.nslices + 1, 0);
  for (int s = 0; s < o.nslices; s++) {
    int w = 0;
    for (int l = 0; l < kSlice; l++) {
      int r = s * kSlice + l;
      if (r < n) w = std::max(w, o.rp[r + 1] - o.rp[r]);
    }
    o.sell_off[s + 1] = o.sell_off[s] + w * kSlice;
  }
  const int tot = o.sell_off[o.nslices];
  o.sell_col.assign(tot, 0);
  o.sK.assign(tot, 0.0);
  o.sD.assign(3 * (size_t)tot, 0.0);
  o.sG.assign(3 * (size_t)tot, 0.0);
  for (int s = 0; s < o.nslices; s++) {
    int w = (o.sell_off[s + 1] - o.sell_off[s]) / kSlice;
    for (int l = 0; l < kSlice; l++) {
      int r = s * kSlice + l;
      for (int k = 0; k < w; k++) {
        size_t e = (size_t)o.sell_off[s] + (size_t)k * kSlice + l;
        if (r < n && k < o.rp[r + 1] - o.rp[r]) {
          int src = o.rp[r] + k;
          o.sell_col[e] = o.ci[src];
          o.sK[e] = o.K[src];
          for (int d = 0; d < 3; d++) {
            o.sD[3 * e + d] = o.D[3 * (size_t)src + d];
            o.sG[3 * e + d] = o.G[3 * (size_t)src + d];
          }
        } else {
          o.sell_col[e] = r < n ? r : 0;  // padding: a valid index with value 0
        }
      }
    }
  }

  // boundary faces: faces owned by exactly one tetrahedron
  std::vector<std::pair<uint64_t, int>> faces;
  faces.reserve(4 * (size_t)T);
  for (int t = 0; t < T; t++)
    for (int opp = 0; opp < 4; opp++) {
      int f[3], q = 0;
      for (int a = 0; a < 4; a++)
        if (a != opp) f[q++] = tets[4 * t + a];
      std::sort(f, f + 3);
      uint64_t key = ((uint64_t)f[0] << 42) | ((uint64_t)f[1] << 21) | (uint64_t)f[2];
      faces.push_back({key, 4 * t + opp});
    }
  std::sort(faces.begin(), faces.end());
  std::vector<int> bface;  // 4*tet + opposite vertex slot
  for (size_t k = 0; k < faces.size();) {
    size_t e = k + 1;
    while (e < faces.size() && faces[e].first == faces[k].first) e++;
    if (e - k == 1) bface.push_back(faces[k].second);
    k = e;
  }
  o.blocal.assign(n, -1);
  for (int f : bface) {
    int t = f / 4, opp = f % 4;
    for (int a = 0; a < 4; a++)
      if (a != opp) o.blocal[tets[4 * t + a]] = 0;
  }
  o.bnode.clear();
  for (int i = 0; i < n; i++)
    if (o.blocal[i] == 0) {
      o.blocal[i] = (int)o.bnode.size();
      o.bnode.push_back(i);
    }
  o.nb = (int)o.bnode.size();
  o.ntri = (int)bface.size();
  o.tri.resize(3 * (size_t)o.ntri);
  o.tnorm.resize(3 * (size_t)o.ntri);
  for (int k = 0; k < o.ntri; k++) {
    int t = bface[k] / 4, opp = bface[k] % 4;
    int gv[3], q = 0;
    for (int a = 0; a < 4; a++)
      if (a != opp) gv[q++] = tets[4 * t + a];  // tet order, opposite vertex dropped
    for (int a = 0; a < 3; a++) o.tri[3 * k + a] = o.blocal[gv[a]];
    V3 p0 = at(coords, gv[0]);
    V3 nn = cross(sub(at(coords, gv[1]), p0), sub(at(coords, gv[2]), p0));
    V3 inward = sub(at(coords, tets[4 * t + opp]), p0);
    double s = (dot(nn, inward) > 0.0 ? -1.0 : 1.0) / std::sqrt(dot(nn, nn));
    o.tnorm[3 * k] = s * nn.x;
    o.tnorm[3 * k + 1] = s * nn.y;
    o.tnorm[3 * k + 2] = s * nn.z;
  }

  // interior solid angle at each boundary node
  o.omega.assign(o.nb, 0.0);
  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    for (int a = 0; a < 4; a++) {
      int lb = o.blocal[v[a]];
      if (lb < 0) continue;
      int w[3], q = 0;
      for (int b = 0; b < 4; b++)
        if (b != a) w[q++] = v[b];
      o.omega[lb] += solid_angle(at(coords, v[a]), at(coords, w[0]), at(coords, w[1]),
                                 at(coords, w[2]));
    }
  }

  // boundary node -> incident (triangle, slot), in ascending triangle order
  o.inc_rp.assign(o.nb + 1, 0);
  for (int k = 0; k < 3 * o.ntri; k++) o.inc_rp[o.tri[k] + 1]++;
  for (int j = 0; j < o.nb; j++) o.inc_rp[j + 1] += o.inc_rp[j];
  o.inc_tri.assign(3 * (size_t)o.ntri, 0);
  std::vector<int> fill(o.inc_rp.begin(), o.inc_rp.end() - 1);
  for (int k = 0; k < o.ntri; k++)
    for (int a = 0; a < 3; a++) o.inc_tri[fill[o.tri[3 * k + a]]++] = 3 * k + a;
  return "";
}

}  // namespace nsim
