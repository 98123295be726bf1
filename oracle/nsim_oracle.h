/*
 * CPU oracle for the nsim LLG right-hand-side path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C, FP64 restatement of the reference algorithm.  It is the
 * checker that the HIP path is compared against, and the "port" CPU baseline
 * that bench.py times on the host.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Reference anchors (paths relative to the nmag-src tree):
 *   box volumes              src/fem.ml:1757-1788
 *   CFBOX (divide by V_i)    src/fem.ml:2052-2092
 *   exchange operator        interface/nmag/nmag_lam.py:445,587-588,734-739
 *   divergence / charges     interface/nmag/nmag_lam.py:446-448,589-590,742-743
 *   Neumann / Dirichlet ops  interface/nmag/nmag_lam.py:591-605,614-633,744-754
 *   gradient -> H_demag      interface/nmag/nmag_lam.py:597-598,755-759
 *   BEM surface geometry     src/bem3d.ml:32-153, src/mesh.ml:1792, src/snippets.ml:1229
 *   Lindholm weights         src/bem3d_stubs.c:413-610
 *   BEM row fill + diagonal  src/bem3d.ml:938-1046
 *   uniaxial anisotropy      interface/nmag/anisotropy.py:92-136,
 *                            src/nsim_anisotropy.ml:167-234
 *   H_total                  interface/nmag/nmag_lam.py:377-411
 *   LLG right-hand side      interface/nmag/nmag_lam.py:417-430
 *   Jacobian (frozen H)      interface/nmag/nmag_lam.py:432-436,684-702,
 *                            src/nsim.ml:1004-1042
 */
#ifndef NSIM_ORACLE_H
#define NSIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;

/* Material / LLG coefficients, all in nmag simulation units. */
typedef struct {
  double Ms;       /* saturation magnetisation                               */
  double negJ;     /* -2A/(mu0 Ms): exchange prefactor (nmag_lam.py:256)     */
  double c1;       /* -gamma/(1+alpha^2)                                      */
  double c2;       /* c1*alpha                                                */
  double c3;       /* llg_normalisationfactor                                 */
  double mu0;      /* 0.4*pi in su                                            */
} orc_params;

/* Structured cuboid, every cube split into 6 Kuhn tetrahedra. */
int orc_kuhn_cuboid(int nx, int ny, int nz, double h, double* coords, int* tets);

/* build_bem: 0 none, 1 dense BEM (atan2 solid angle), 2 dense BEM with the
 * reference's literal acos solid-angle form. */
orc_mesh* orc_setup(int n_nodes, const double* coords, int n_tets, const int* tets,
                    int build_bem);
void orc_free(orc_mesh* m);
/* OpenMP threads used from now on (n <= 0: all processors); returns the
 * count in effect. */
int orc_set_threads(int n);

/* sizes[0..4] = n_nodes, nnz(K), n_boundary_nodes, n_boundary_triangles, n_tets */
void orc_sizes(const orc_mesh* m, int* sizes);
void orc_get_volumes(const orc_mesh* m, double* V);
void orc_get_stiffness(const orc_mesh* m, int* rowptr, int* col, double* val);
void orc_get_boundary(const orc_mesh* m, int* bnodes, int* tris, double* normals,
                      double* solid_angles);
void orc_get_bem(const orc_mesh* m, double* B);

/* Lindholm weights of one triangle seen from one observer (reference formula;
 * _acos keeps the reference's literal acos solid-angle form). */
void orc_lindholm(const double* observer, const double* p0, const double* p1,
                  const double* p2, const double* outward_normal, double geo_eps,
                  double* lh3);
void orc_lindholm_acos(const double* observer, const double* p0, const double* p1,
                       const double* p2, const double* outward_normal, double geo_eps,
                       double* lh3);
double orc_vertex_solid_angle(const double* x, const double* p0, const double* p1,
                              const double* p2);

/* Field evaluation.  m: 3N site-major.  H_ext: 3N or NULL.  k1,k2: N or NULL,
 * axis: 3N (unit) or NULL.  phi1/phi2 are in/out warm starts (may be NULL ->
 * zero start).  Outputs may be NULL.  iters[0..1] = CG iterations used. */
int orc_fields(const orc_mesh* m, const orc_params* p, const double* mag,
               const double* H_ext, const double* k1, const double* k2,
               const double* axis, double cg_rtol, int cg_maxit, int do_demag,
               double* phi1, double* phi2, double* rho, double* H_exch,
               double* H_demag, double* H_anis, double* H_total, double* dmdt,
               int* iters);

/* Individual operators, exposed so tests can pin each one. */
void orc_apply_K(const orc_mesh* m, const double* x, double* y);
void orc_div(const orc_mesh* m, double Ms, const double* mag, double* rho);
void orc_grad(const orc_mesh* m, const double* phi, double* H);
int orc_cg(const orc_mesh* m, const double* b, double* x, const unsigned char* fixed,
           double rtol, int maxit, double* res_out);

/* Jacobian-times-vector of the LLG rhs with H_total frozen except for the
 * exchange part (dH/dm = negJ K / V), as in the reference's jacobi_ts1. */
/* Polynomial anisotropy field, added to Ha (src/nsim_anisotropy.ml:151-234). */
void orc_poly_anis(int n, const double* mag, int nterms, const int* powers,
                   const double* coeffs, double s, double* Ha);
void orc_jtimes(const orc_mesh* m, const orc_params* p, const double* mag,
                const double* H_total, const double* v, double* Jv);

#ifdef __cplusplus
}
#endif
#endif
