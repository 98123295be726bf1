/*
 * CPU oracle for the nsim LLG right-hand-side path (TEST INFRASTRUCTURE ONLY).
 * FP64 restatement of the reference algorithm; see nsim_oracle.h for the
 * reference anchors of every function.  Loaded only by tests/, the smoke check
 * and bench.py's CPU-baseline leg.  OpenMP parallel where it is free to be.
 */
#include "nsim_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793238462643383279502884

struct orc_mesh {
  int n, T;
  double* x;      /* 3n coords                                   */
  int* tet;       /* 4T                                          */
  double* tvol;   /* T   |volume|                                */
  double* tgrad;  /* 12T gradients of the barycentric functions  */
  double* V;      /* n   box volumes                             */
  int* rp;        /* n+1 CSR of node adjacency (incl. diagonal)  */
  int* ci;
  double* K;      /* stiffness values on that pattern            */
  double* Kdiag;
  int nb;         /* boundary nodes (ascending global index)     */
  int* bnode;     /* nb                                          */
  int* blocal;    /* n -> local boundary index or -1             */
  int ntri;
  int* tri;       /* 3*ntri local boundary indices               */
  double* tnorm;  /* 3*ntri outward unit normals                 */
  double* omega;  /* nb interior solid angle                     */
  double* B;      /* nb*nb dense row-major, or NULL              */
  double geo_eps;
};

/* ------------------------------------------------------------------ mesh */

int orc_kuhn_cuboid(int nx, int ny, int nz, double h, double* coords, int* tets) {
  /* node index = i + (nx+1)*(j + (ny+1)*k); each cube (i,j,k) is split into the
   * 6 Kuhn simplices {v0, v0+e_a, v0+e_a+e_b, v0+(1,1,1)} for (a,b,c) running
   * over the permutations of the axes. */
  int sx = nx + 1, sy = ny + 1;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        int id = i + sx * (j + sy * k);
        coords[3 * id + 0] = i * h;
        coords[3 * id + 1] = j * h;
        coords[3 * id + 2] = k * h;
      }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                 {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int t = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int p = 0; p < 6; p++) {
          int c[3] = {i, j, k};
          int v = 0;
          tets[4 * t + v++] = c[0] + sx * (c[1] + sy * c[2]);
          for (int s = 0; s < 3; s++) {
            c[perm[p][s]] += 1;
            tets[4 * t + v++] = c[0] + sx * (c[1] + sy * c[2]);
          }
          t++;
        }
  return t;
}

static void sub3(double* r, const double* a, const double* b) {
  r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void cross3(double* r, const double* a, const double* b) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static double len3(const double* a) { return sqrt(dot3(a, a)); }

/* Gradients of the 4 barycentric coordinates and |volume| of one tetrahedron. */
static double tet_geometry(const double* x0, const double* x1, const double* x2,
                           const double* x3, double* g /* 12 */) {
  double e1[3], e2[3], e3[3], c23[3], c31[3], c12[3];
  sub3(e1, x1, x0); sub3(e2, x2, x0); sub3(e3, x3, x0);
  cross3(c23, e2, e3); cross3(c31, e3, e1); cross3(c12, e1, e2);
  double det = dot3(e1, c23);
  /* rows of E^{-1} (E has columns e1,e2,e3) are the gradients of lambda_1..3 */
  for (int d = 0; d < 3; d++) {
    g[3 + d] = c23[d] / det;
    g[6 + d] = c31[d] / det;
    g[9 + d] = c12[d] / det;
    g[d] = -(g[3 + d] + g[6 + d] + g[9 + d]);
  }
  return fabs(det) / 6.0;
}

/* Solid angle subtended at x by the triangle (p0,p1,p2): src/snippets.ml:1229. */
double orc_vertex_solid_angle(const double* x, const double* p0, const double* p1,
                              const double* p2) {
  double r0[3], r1[3], r2[3];
  sub3(r0, p0, x); sub3(r1, p1, x); sub3(r2, p2, x);
  double r0l = len3(r0), r1l = len3(r1), r2l = len3(r2);
  double sp01 = dot3(r0, r1), sp12 = dot3(r1, r2), sp20 = dot3(r2, r0);
  double num = r0l * r1l * r2l + r0l * sp12 + r1l * sp20 + r2l * sp01;
  double den = sqrt(2.0 * (r1l * r2l + sp12) * (r2l * r0l + sp20) * (r0l * r1l + sp01));
  if (den == 0.0) return 0.0;
  double q = num / den;
  if (q < -1.0) q = -1.0;
  if (q > 1.0) q = 1.0;
  return fabs(2.0 * acos(q));
}

/* Lindholm linear-shape double-layer weights: src/bem3d_stubs.c:413-610
 * (identity lattice transform, grey factor 1, min_dist 0).
 * The triangle's solid angle is 2*acos(num/den)*sign(zeta) in the reference.
 * Since den^2 = num^2 + (r0.(r1 x r2))^2 and zeta has the sign of r0.(r1 x r2),
 * that equals 2*atan2(r0.(r1 x r2), num), which keeps full precision for the
 * small angles of distant triangles (acos near 1 loses about half the digits).
 * acos_form != 0 selects the reference's literal acos form. */
static void lindholm_impl(const double* obs, const double* p0, const double* p1,
                          const double* p2, const double* n_out, double geo_eps,
                          double* lh, int acos_form) {
  lh[0] = lh[1] = lh[2] = 0.0;
  double s0[3], s1[3], s2[3], xi0[3], xi1[3], xi2[3], zv[3];
  sub3(s0, p2, p1); sub3(s1, p0, p2); sub3(s2, p1, p0);
  double s0l = len3(s0), s1l = len3(s1), s2l = len3(s2);
  if (s0l == 0.0 || s1l == 0.0 || s2l == 0.0) return;
  for (int d = 0; d < 3; d++) {
    xi0[d] = s0[d] / s0l; xi1[d] = s1[d] / s1l; xi2[d] = s2[d] / s2l;
  }
  cross3(zv, s1, s2);
  double zl = len3(zv);
  if (zl == 0.0) return;
  double area = 0.5 * zl;
  for (int d = 0; d < 3; d++) zv[d] /= zl;
  double c01 = dot3(xi0, xi1), c12 = dot3(xi1, xi2), c20 = dot3(xi2, xi0);
  double eta0[3], eta1[3], eta2[3];
  cross3(eta0, zv, xi0); cross3(eta1, zv, xi1); cross3(eta2, zv, xi2);

  double r0[3], r1[3], r2[3];
  sub3(r0, p0, obs); sub3(r1, p1, obs); sub3(r2, p2, obs);
  double r0l = len3(r0), r1l = len3(r1), r2l = len3(r2);
  if (r0l <= 0.0 || r1l <= 0.0 || r2l <= 0.0) return; /* observer on a vertex */
  double lg[3];
  lg[0] = log((r1l + r2l + s0l) / (r1l + r2l - s0l));
  lg[1] = log((r2l + r0l + s1l) / (r2l + r0l - s1l));
  lg[2] = log((r0l + r1l + s2l) / (r0l + r1l - s2l));
  double sp01 = dot3(r0, r1), sp12 = dot3(r1, r2), sp20 = dot3(r2, r0);
  double num = r0l * r1l * r2l + r0l * sp12 + r1l * sp20 + r2l * sp01;
  double den = sqrt(2.0 * (r1l * r2l + sp12) * (r2l * r0l + sp20) * (r0l * r1l + sp01));
  if (den == 0.0) return;
  double zeta = dot3(zv, r0);
  double angle;
  if (acos_form) {
    double q = num / den;
    if (q < -1.0) q = -1.0;
    if (q > 1.0) q = 1.0;
    angle = 2.0 * acos(q);
    if (zeta < 0.0) angle = -angle;
  } else {
    double r1xr2[3];
    cross3(r1xr2, r1, r2);
    angle = 2.0 * atan2(dot3(r0, r1xr2), num);
  }
  double f = 1.0 / (8.0 * PI * area);
  if (dot3(zv, n_out) < 0.0) f = -f;
  if (fabs(zeta) < geo_eps) f = 0.0;
  double g0 = lg[0] + c01 * lg[1] + c20 * lg[2];
  double g1 = c01 * lg[0] + lg[1] + c12 * lg[2];
  double g2 = c20 * lg[0] + c12 * lg[1] + lg[2];
  lh[0] = f * s0l * (angle * dot3(eta0, r1) - zeta * g0);
  lh[1] = f * s1l * (angle * dot3(eta1, r2) - zeta * g1);
  lh[2] = f * s2l * (angle * dot3(eta2, r0) - zeta * g2);
}

void orc_lindholm(const double* obs, const double* p0, const double* p1,
                  const double* p2, const double* n_out, double geo_eps, double* lh) {
  lindholm_impl(obs, p0, p1, p2, n_out, geo_eps, lh, 0);
}

void orc_lindholm_acos(const double* obs, const double* p0, const double* p1,
                       const double* p2, const double* n_out, double geo_eps, double* lh) {
  lindholm_impl(obs, p0, p1, p2, n_out, geo_eps, lh, 1);
}

/* ------------------------------------------------------------- set-up */

typedef struct { int a, b, c, tet, opp; } face_rec;

static int face_cmp(const void* pa, const void* pb) {
  const face_rec* x = (const face_rec*)pa;
  const face_rec* y = (const face_rec*)pb;
  if (x->a != y->a) return x->a < y->a ? -1 : 1;
  if (x->b != y->b) return x->b < y->b ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}
static int int_cmp(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}
static void sort3(int* v) {
  int t;
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
  if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

static int csr_find(const orc_mesh* m, int i, int j) {
  int lo = m->rp[i], hi = m->rp[i + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (m->ci[mid] == j) return mid;
    if (m->ci[mid] < j) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static void build_bem(orc_mesh* m, int acos_form) {
  int nb = m->nb;
  m->B = (double*)calloc((size_t)nb * nb, sizeof(double));
  /* Row fill over every surface triangle (src/bem3d.ml:938-1046); rows are
   * independent, so observer rows are split across threads. */
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < nb; i++) {
    const double* xo = m->x + 3 * m->bnode[i];
    double* row = m->B + (size_t)i * nb;
    for (int t = 0; t < m->ntri; t++) {
      const int* tv = m->tri + 3 * t;
      double lh[3];
      lindholm_impl(xo, m->x + 3 * m->bnode[tv[0]], m->x + 3 * m->bnode[tv[1]],
                    m->x + 3 * m->bnode[tv[2]], m->tnorm + 3 * t, m->geo_eps, lh, acos_form);
      row[tv[0]] += lh[0];
      row[tv[1]] += lh[1];
      row[tv[2]] += lh[2];
    }
    row[i] += m->omega[i] / (4.0 * PI) - 1.0;
  }
}

orc_mesh* orc_setup(int n, const double* coords, int T, const int* tets, int build_b) {
  orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
  m->n = n; m->T = T;
  m->x = (double*)malloc(sizeof(double) * 3 * n);
  memcpy(m->x, coords, sizeof(double) * 3 * n);
  m->tet = (int*)malloc(sizeof(int) * 4 * T);
  memcpy(m->tet, tets, sizeof(int) * 4 * T);
  m->tvol = (double*)malloc(sizeof(double) * T);
  m->tgrad = (double*)malloc(sizeof(double) * 12 * T);
  m->V = (double*)calloc(n, sizeof(double));
  double gmax = 0.0;
  for (int i = 0; i < 3 * n; i++) if (fabs(coords[i]) > gmax) gmax = fabs(coords[i]);
  m->geo_eps = gmax * 1e-13;

  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    m->tvol[t] = tet_geometry(coords + 3 * v[0], coords + 3 * v[1], coords + 3 * v[2],
                              coords + 3 * v[3], m->tgrad + 12 * t);
    for (int a = 0; a < 4; a++) m->V[v[a]] += 0.25 * m->tvol[t];
  }

  /* node adjacency pattern: each row sorted, duplicates removed */
  int* deg = (int*)calloc(n, sizeof(int));
  for (int t = 0; t < T; t++)
    for (int a = 0; a < 4; a++) deg[tets[4 * t + a]] += 4;
  int* tmp_rp = (int*)malloc(sizeof(int) * (n + 1));
  tmp_rp[0] = 0;
  for (int i = 0; i < n; i++) tmp_rp[i + 1] = tmp_rp[i] + deg[i];
  int* tmp = (int*)malloc(sizeof(int) * tmp_rp[n]);
  memset(deg, 0, sizeof(int) * n);
  for (int t = 0; t < T; t++)
    for (int a = 0; a < 4; a++) {
      int i = tets[4 * t + a];
      for (int b = 0; b < 4; b++) tmp[tmp_rp[i] + deg[i]++] = tets[4 * t + b];
    }
  m->rp = (int*)malloc(sizeof(int) * (n + 1));
  m->rp[0] = 0;
  for (int i = 0; i < n; i++) {
    int* row = tmp + tmp_rp[i];
    qsort(row, deg[i], sizeof(int), int_cmp);
    int u = 0;
    for (int k = 0; k < deg[i]; k++)
      if (u == 0 || row[k] != row[u - 1]) row[u++] = row[k];
    deg[i] = u;
    m->rp[i + 1] = m->rp[i] + u;
  }
  m->ci = (int*)malloc(sizeof(int) * m->rp[n]);
  for (int i = 0; i < n; i++) memcpy(m->ci + m->rp[i], tmp + tmp_rp[i], sizeof(int) * deg[i]);
  free(tmp); free(tmp_rp); free(deg);

  /* stiffness K_ij = sum_t vol_t grad(l_a).grad(l_b) */
  m->K = (double*)calloc(m->rp[n], sizeof(double));
  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    const double* g = m->tgrad + 12 * t;
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        int k = csr_find(m, v[a], v[b]);
        m->K[k] += m->tvol[t] * dot3(g + 3 * a, g + 3 * b);
      }
  }
  m->Kdiag = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) m->Kdiag[i] = m->K[csr_find(m, i, i)];

  /* boundary faces: faces owned by exactly one tetrahedron */
  face_rec* f = (face_rec*)malloc(sizeof(face_rec) * 4 * (size_t)T);
  for (int t = 0; t < T; t++)
    for (int o = 0; o < 4; o++) {
      int v[3], q = 0;
      for (int a = 0; a < 4; a++) if (a != o) v[q++] = tets[4 * t + a];
      sort3(v);
      face_rec* r = f + 4 * t + o;
      r->a = v[0]; r->b = v[1]; r->c = v[2]; r->tet = t; r->opp = o;
    }
  qsort(f, 4 * (size_t)T, sizeof(face_rec), face_cmp);
  int nf = 0;
  face_rec* bf = (face_rec*)malloc(sizeof(face_rec) * 4 * (size_t)T);
  for (size_t k = 0; k < 4 * (size_t)T;) {
    size_t e = k + 1;
    while (e < 4 * (size_t)T && face_cmp(f + k, f + e) == 0) e++;
    if (e - k == 1) bf[nf++] = f[k];
    k = e;
  }
  free(f);
  m->blocal = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) m->blocal[i] = -1;
  for (int k = 0; k < nf; k++) {
    m->blocal[bf[k].a] = 1; m->blocal[bf[k].b] = 1; m->blocal[bf[k].c] = 1;
  }
  m->nb = 0;
  for (int i = 0; i < n; i++) if (m->blocal[i] > 0) m->blocal[i] = m->nb++;
  m->bnode = (int*)malloc(sizeof(int) * (m->nb > 0 ? m->nb : 1));
  for (int i = 0; i < n; i++) if (m->blocal[i] >= 0) m->bnode[m->blocal[i]] = i;
  m->ntri = nf;
  m->tri = (int*)malloc(sizeof(int) * 3 * (nf > 0 ? nf : 1));
  m->tnorm = (double*)malloc(sizeof(double) * 3 * (nf > 0 ? nf : 1));
  for (int k = 0; k < nf; k++) {
    /* vertex order: the tet's own order with the opposite vertex removed */
    const int* v = tets + 4 * bf[k].tet;
    int q = 0, gv[3];
    for (int a = 0; a < 4; a++) if (a != bf[k].opp) gv[q++] = v[a];
    for (int a = 0; a < 3; a++) m->tri[3 * k + a] = m->blocal[gv[a]];
    double e1[3], e2[3], nn[3], to[3];
    sub3(e1, coords + 3 * gv[1], coords + 3 * gv[0]);
    sub3(e2, coords + 3 * gv[2], coords + 3 * gv[0]);
    cross3(nn, e1, e2);
    sub3(to, coords + 3 * v[bf[k].opp], coords + 3 * gv[0]);
    double s = (dot3(nn, to) > 0.0 ? -1.0 : 1.0) / len3(nn);
    for (int d = 0; d < 3; d++) m->tnorm[3 * k + d] = s * nn[d];
  }
  free(bf);

  /* interior solid angle at every boundary node: src/mesh.ml:1792 */
  m->omega = (double*)calloc(m->nb > 0 ? m->nb : 1, sizeof(double));
  for (int t = 0; t < T; t++) {
    const int* v = tets + 4 * t;
    for (int a = 0; a < 4; a++) {
      int lb = m->blocal[v[a]];
      if (lb < 0) continue;
      int o[3], q = 0;
      for (int b = 0; b < 4; b++) if (b != a) o[q++] = v[b];
      m->omega[lb] += orc_vertex_solid_angle(coords + 3 * v[a], coords + 3 * o[0],
                                             coords + 3 * o[1], coords + 3 * o[2]);
    }
  }
  if (build_b) build_bem(m, build_b == 2);
  return m;
}

void orc_free(orc_mesh* m) {
  if (!m) return;
  free(m->x); free(m->tet); free(m->tvol); free(m->tgrad); free(m->V);
  free(m->rp); free(m->ci); free(m->K); free(m->Kdiag); free(m->bnode);
  free(m->blocal); free(m->tri); free(m->tnorm); free(m->omega); free(m->B);
  free(m);
}

/* OpenMP threads of the parallel loops (n <= 0: all processors); returns
 * the count in effect, so a CPU-baseline timing can state its cores. */
int orc_set_threads(int n) {
  omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
  return omp_get_max_threads();
}

void orc_sizes(const orc_mesh* m, int* s) {
  s[0] = m->n; s[1] = m->rp[m->n]; s[2] = m->nb; s[3] = m->ntri; s[4] = m->T;
}
void orc_get_volumes(const orc_mesh* m, double* V) { memcpy(V, m->V, sizeof(double) * m->n); }
void orc_get_stiffness(const orc_mesh* m, int* rp, int* ci, double* val) {
  memcpy(rp, m->rp, sizeof(int) * (m->n + 1));
  memcpy(ci, m->ci, sizeof(int) * m->rp[m->n]);
  memcpy(val, m->K, sizeof(double) * m->rp[m->n]);
}
void orc_get_boundary(const orc_mesh* m, int* bnodes, int* tris, double* normals,
                      double* sa) {
  if (bnodes) memcpy(bnodes, m->bnode, sizeof(int) * m->nb);
  if (tris) memcpy(tris, m->tri, sizeof(int) * 3 * m->ntri);
  if (normals) memcpy(normals, m->tnorm, sizeof(double) * 3 * m->ntri);
  if (sa) memcpy(sa, m->omega, sizeof(double) * m->nb);
}
void orc_get_bem(const orc_mesh* m, double* B) {
  if (m->B) memcpy(B, m->B, sizeof(double) * (size_t)m->nb * m->nb);
}

/* ----------------------------------------------------------- operators */

void orc_apply_K(const orc_mesh* m, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m->n; i++) {
    double s = 0.0;
    for (int k = m->rp[i]; k < m->rp[i + 1]; k++) s += m->K[k] * x[m->ci[k]];
    y[i] = s;
  }
}

/* rho_i = Ms * int grad(phi_i) . m dV  (the reference's volume + surface charge
 * operator followed by SCALE -1; identical after integration by parts). */
void orc_div(const orc_mesh* m, double Ms, const double* mag, double* rho) {
  memset(rho, 0, sizeof(double) * m->n);
  for (int t = 0; t < m->T; t++) {
    const int* v = m->tet + 4 * t;
    const double* g = m->tgrad + 12 * t;
    double mb[3] = {0, 0, 0};
    for (int a = 0; a < 4; a++)
      for (int d = 0; d < 3; d++) mb[d] += 0.25 * mag[3 * v[a] + d];
    for (int a = 0; a < 4; a++) rho[v[a]] += Ms * m->tvol[t] * dot3(g + 3 * a, mb);
  }
}

/* H_i = -(1/V_i) int phi_i grad(phi) dV  (op_grad_phi + CFBOX). */
void orc_grad(const orc_mesh* m, const double* phi, double* H) {
  memset(H, 0, sizeof(double) * 3 * m->n);
  for (int t = 0; t < m->T; t++) {
    const int* v = m->tet + 4 * t;
    const double* g = m->tgrad + 12 * t;
    double gp[3] = {0, 0, 0};
    for (int a = 0; a < 4; a++)
      for (int d = 0; d < 3; d++) gp[d] += phi[v[a]] * g[3 * a + d];
    for (int a = 0; a < 4; a++)
      for (int d = 0; d < 3; d++) H[3 * v[a] + d] -= 0.25 * m->tvol[t] * gp[d];
  }
  for (int i = 0; i < m->n; i++)
    for (int d = 0; d < 3; d++) H[3 * i + d] /= m->V[i];
}

/* Jacobi-preconditioned CG on K x = b.  Rows with fixed[i] != 0 are Dirichlet
 * rows: x_i keeps its value and the row is dropped from the system.  Stops when
 * sqrt(r.Dinv.r) <= rtol * sqrt(b_eff.Dinv.b_eff), b_eff = b - K x_fixed. */
int orc_cg(const orc_mesh* m, const double* b, double* x, const unsigned char* fixed,
           double rtol, int maxit, double* res_out) {
  int n = m->n;
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* xf = (double*)calloc(n, sizeof(double));
  for (int i = 0; i < n; i++) if (fixed && fixed[i]) xf[i] = x[i];
  orc_apply_K(m, xf, q);
  double bn = 0.0;
  for (int i = 0; i < n; i++) {
    if (fixed && fixed[i]) continue;
    double be = (b ? b[i] : 0.0) - q[i];
    bn += be * be / m->Kdiag[i];
  }
  orc_apply_K(m, x, q);
  double rz = 0.0;
  for (int i = 0; i < n; i++) {
    if (fixed && fixed[i]) { r[i] = z[i] = p[i] = 0.0; continue; }
    r[i] = (b ? b[i] : 0.0) - q[i];
    z[i] = r[i] / m->Kdiag[i];
    p[i] = z[i];
    rz += r[i] * z[i];
  }
  double stop = rtol * rtol * bn;
  int it = 0;
  while (rz > stop && it < maxit) {
    orc_apply_K(m, p, q);
    double pq = 0.0;
    for (int i = 0; i < n; i++) if (!(fixed && fixed[i])) pq += p[i] * q[i];
    double alpha = rz / pq;
    double rz_new = 0.0;
    for (int i = 0; i < n; i++) {
      if (fixed && fixed[i]) continue;
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = r[i] / m->Kdiag[i];
      rz_new += r[i] * z[i];
    }
    double beta = rz_new / rz;
    rz = rz_new;
    for (int i = 0; i < n; i++) if (!(fixed && fixed[i])) p[i] = z[i] + beta * p[i];
    it++;
  }
  if (res_out) *res_out = bn > 0 ? sqrt(rz / bn) : sqrt(rz);
  free(r); free(z); free(p); free(q); free(xf);
  return it;
}

static void llg(const orc_params* p, const double* m, const double* H, double* f) {
  double mxH[3], mxmxH[3];
  cross3(mxH, m, H);
  cross3(mxmxH, m, mxH);
  double mm = dot3(m, m);
  for (int d = 0; d < 3; d++)
    f[d] = p->c1 * mxH[d] + p->c2 * mxmxH[d] + p->c3 * (1.0 - mm) * m[d];
}

int orc_fields(const orc_mesh* M, const orc_params* p, const double* mag,
               const double* H_ext, const double* k1, const double* k2,
               const double* axis, double cg_rtol, int cg_maxit, int do_demag,
               double* phi1_io, double* phi2_io, double* rho_o, double* Hex_o,
               double* Hd_o, double* Ha_o, double* Ht_o, double* dmdt_o, int* iters) {
  int n = M->n;
  double* Hex = (double*)malloc(sizeof(double) * 3 * n);
  double* Hd = (double*)calloc(3 * n, sizeof(double));
  double* Ha = (double*)calloc(3 * n, sizeof(double));
  double* Ht = (double*)malloc(sizeof(double) * 3 * n);
  double* comp = (double*)malloc(sizeof(double) * n);
  double* Kc = (double*)malloc(sizeof(double) * n);
  int it1 = 0, it2 = 0;

  /* exchange: H_exch = negJ K m / V (per component) */
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n; i++) comp[i] = mag[3 * i + d];
    orc_apply_K(M, comp, Kc);
    for (int i = 0; i < n; i++) Hex[3 * i + d] = p->negJ * Kc[i] / M->V[i];
  }

  if (do_demag) {
    double* rho = (double*)malloc(sizeof(double) * n);
    double* phi1 = (double*)calloc(n, sizeof(double));
    double* phi2 = (double*)calloc(n, sizeof(double));
    double* phi = (double*)malloc(sizeof(double) * n);
    unsigned char* fixed = (unsigned char*)calloc(n, 1);
    if (phi1_io) memcpy(phi1, phi1_io, sizeof(double) * n);
    if (phi2_io) memcpy(phi2, phi2_io, sizeof(double) * n);
    orc_div(M, p->Ms, mag, rho);
    /* Neumann problem: the constant null space is removed from the load */
    double mean = 0.0;
    for (int i = 0; i < n; i++) mean += rho[i];
    mean /= n;
    double* b = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) b[i] = rho[i] - mean;
    it1 = orc_cg(M, b, phi1, NULL, cg_rtol, cg_maxit, NULL);
    free(b);
    /* phi2 on the boundary: dense BEM matrix times phi1 restricted */
    for (int i = 0; i < M->nb; i++) {
      const double* row = M->B + (size_t)i * M->nb;
      double s = 0.0;
      for (int j = 0; j < M->nb; j++) s += row[j] * phi1[M->bnode[j]];
      phi2[M->bnode[i]] = s;
      fixed[M->bnode[i]] = 1;
    }
    /* Dirichlet Laplace problem for the interior of phi2 */
    it2 = orc_cg(M, NULL, phi2, fixed, cg_rtol, cg_maxit, NULL);
    for (int i = 0; i < n; i++) phi[i] = phi1[i] + phi2[i];
    orc_grad(M, phi, Hd);
    if (rho_o) memcpy(rho_o, rho, sizeof(double) * n);
    if (phi1_io) memcpy(phi1_io, phi1, sizeof(double) * n);
    if (phi2_io) memcpy(phi2_io, phi2, sizeof(double) * n);
    free(rho); free(phi1); free(phi2); free(phi); free(fixed);
  }

  /* uniaxial anisotropy, E = -K1 (a.m)^2 - K2 (a.m)^4, tangential gradient */
  if (k1 && axis) {
    for (int i = 0; i < n; i++) {
      const double* a = axis + 3 * i;
      const double* mi = mag + 3 * i;
      double u = dot3(a, mi);
      double dE = 2.0 * k1[i] * u + (k2 ? 4.0 * k2[i] * u * u * u : 0.0);
      double s = dE / (p->mu0 * p->Ms);
      for (int d = 0; d < 3; d++) Ha[3 * i + d] = s * (a[d] - u * mi[d]);
    }
  }

  for (int i = 0; i < 3 * n; i++)
    Ht[i] = (H_ext ? H_ext[i] : 0.0) + Hex[i] + Ha[i] + Hd[i];
  if (dmdt_o)
    for (int i = 0; i < n; i++) llg(p, mag + 3 * i, Ht + 3 * i, dmdt_o + 3 * i);
  if (Hex_o) memcpy(Hex_o, Hex, sizeof(double) * 3 * n);
  if (Hd_o) memcpy(Hd_o, Hd, sizeof(double) * 3 * n);
  if (Ha_o) memcpy(Ha_o, Ha, sizeof(double) * 3 * n);
  if (Ht_o) memcpy(Ht_o, Ht, sizeof(double) * 3 * n);
  if (iters) { iters[0] = it1; iters[1] = it2; }
  free(Hex); free(Hd); free(Ha); free(Ht); free(comp); free(Kc);
  return 0;
}

/* Polynomial anisotropy (src/nsim_anisotropy.ml:151-234): the energy density
 * E(m) = sum_j c_j mx^px_j my^py_j mz^pz_j, homogeneous of degree `order`
 * (the fit of taylor_coefficients_of_spherical_function, :119-133); the field
 * is minus the unit-sphere tangential gradient,
 *   H = -s (g - m (m.g)),  g = grad E(m),  s = 1/(mu0 Ms)
 * (unit_sphere_tangential_polypotential_gradient, :167-197, with the
 * correction factor 1/(mu0 Ms) of interface/nmag/nmag_lam.py:490-497).
 * Adds H to Ha (n x 3). */
static double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; i++) r *= x;
  return r;
}

void orc_poly_anis(int n, const double* mag, int nterms, const int* powers,
                   const double* coeffs, double s, double* Ha) {
  for (int i = 0; i < n; i++) {
    const double* m = mag + 3 * i;
    double g[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < nterms; j++) {
      const int* p = powers + 3 * j;
      for (int d = 0; d < 3; d++) {
        if (p[d] == 0) continue;
        double t = coeffs[j] * p[d];
        for (int e = 0; e < 3; e++) t *= ipow(m[e], e == d ? p[e] - 1 : p[e]);
        g[d] += t;
      }
    }
    const double mg = dot3(m, g);
    for (int d = 0; d < 3; d++) Ha[3 * i + d] += -s * (g[d] - m[d] * mg);
  }
}

void orc_jtimes(const orc_mesh* M, const orc_params* p, const double* mag,
                const double* H, const double* v, double* Jv) {
  int n = M->n;
  double* comp = (double*)malloc(sizeof(double) * n);
  double* Kc = (double*)malloc(sizeof(double) * n);
  double* h = (double*)malloc(sizeof(double) * 3 * n);
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n; i++) comp[i] = v[3 * i + d];
    orc_apply_K(M, comp, Kc);
    for (int i = 0; i < n; i++) h[3 * i + d] = p->negJ * Kc[i] / M->V[i];
  }
  for (int i = 0; i < n; i++) {
    const double* m = mag + 3 * i;
    const double* vi = v + 3 * i;
    const double* Hi = H + 3 * i;
    const double* hi = h + 3 * i;
    double a[3], b[3], c[3], d1[3], e[3], f[3], g[3], k[3];
    cross3(a, vi, Hi);       /* v x H          */
    cross3(b, m, hi);        /* m x h          */
    cross3(c, m, Hi);        /* m x H          */
    cross3(d1, vi, c);       /* v x (m x H)    */
    cross3(e, m, a);         /* m x (v x H)    */
    cross3(f, m, hi);
    cross3(g, m, f);         /* m x (m x h)    */
    for (int q = 0; q < 3; q++) {
      k[q] = p->c1 * (a[q] + b[q]) + p->c2 * (d1[q] + e[q] + g[q]);
      Jv[3 * i + q] = k[q];
    }
  }
  free(comp); free(Kc); free(h);
}
