/*
 * nsim-mi355x: C ABI of the MI355X-native LLG right-hand-side path.
 *
 * Plain C types only (no torch, no HIP types).  Every entry point returns an
 * int status: 0 = ok, > 0 = recoverable (the caller may retry with a smaller
 * step, as CVODE does), < 0 = fatal.  nsim_last_error() returns the message of
 * the last failure on the calling thread.  A handle owns one HIP stream and all
 * device memory of one magnet; calls on one handle are not re-entrant.
 *
 * The reference exposes this path through two stacked boundaries, and each
 * entry point below names the interface it replaces:
 *   B1  Python -> "ocaml" module (src/nsimexec.ml:4467-4676 registry)
 *   B2  CVODE callbacks (src/sundials_sp_stubs.c)
 */
#ifndef NSIM_H
#define NSIM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsim_handle nsim_handle;

/* Tetrahedral mesh: coordinates in simulation length units (nm). */
typedef struct {
  int n_nodes;
  int n_tets;
  const double* coords; /* 3*n_nodes, site-major                      */
  const int* tets;      /* 4*n_tets node indices                      */
} nsim_mesh;

/* LLG and material coefficients in simulation units; the same numbers that
 * MagMaterial computes (interface/nmag/main.py:423-471). */
typedef struct {
  double Ms;   /* su_Ms                                                   */
  double negJ; /* -su_exch_prefactor = -2A/(mu0 Ms) (nmag_lam.py:256)     */
  double c1;   /* su_llg_coeff1 = -gamma/(1+alpha^2)                       */
  double c2;   /* su_llg_coeff2 = c1*alpha                                 */
  double c3;   /* su_llg_normalisationfactor                               */
  double mu0;  /* su_mu0                                                  */
} nsim_material;

enum {
  NSIM_FP64 = 0,       /* everything in FP64                                */
  NSIM_MIXED_FP32 = 1, /* FP32 SpMV and BEM, FP64 dots/state/norms          */
  NSIM_MIXED_BF16 = 2, /* as MIXED_FP32 but the BEM matrix stored in BF16   */
  NSIM_MIXED_F24 = 3   /* as MIXED_FP32 but the dense potential operator
                        * stored in 24 bits: FP32 rounded to 15 stored
                        * significand bits (unit roundoff 2^-16), 3/4 of
                        * FP32's bytes.  No counterpart in the reference
                        * (whose path is FP64 throughout).                  */
};

/* Solver options; the KSP tolerances mirror Simulation(ksp_tolerances=...)
 * (interface/nmag/main.py:819-843, nmag_lam.py:614-633). */
typedef struct {
  int precision;   /* NSIM_FP64 / _MIXED_FP32 / _MIXED_BF16 / _MIXED_F24    */
  int do_demag;    /* 0: exchange/anisotropy/applied field only            */
  double nbc_rtol; /* Neumann solve relative tolerance                     */
  double dbc_rtol; /* Dirichlet solve relative tolerance                   */
  int nbc_maxits;
  int dbc_maxits;
  /* Demag route.  -1: the reference's route, two Poisson solves (persistent
   * CG, warm-started, to nbc_rtol / dbc_rtol) around the BEM GEMV.  1: a
   * dense potential operator M = (I + E B R) K^+ per magnet, built once at
   * create time (Cholesky factorisations), so that phi = M rho is one GEMV
   * streamed from HBM; exact up to rounding, the tolerances are unused.
   * 0: 1 when every magnet has at most 20480 nodes (40960 with BF16
   * storage) and all operators fit 32 GiB, else -1.  2: the same operator M
   * compressed as an H-matrix (dense near-field blocks, low-rank far-field
   * blocks to a relative tolerance of compress_eps per block), applied in two
   * launches; magnets above 20480 nodes allowed.  No counterpart in the
   * reference (its hlib H-matrices compress the BEM matrix B only,
   * src/bem3d.ml:555-621).  nsim_get_stat "potential" tells which route a
   * handle took (1 dense, 2 H-matrix, 0 iterative). */
  int potential;
  /* Relative Frobenius tolerance of each low-rank block of the H-matrix
   * route; 0 selects the default, 1e-12 with FP64 storage and 1e-7 with FP32
   * storage (the other precisions), an error of the order of their rounding. */
  double compress_eps;
  /* Admissibility of the H-matrix route: a pair of clusters is stored low
   * rank when min(diam) <= compress_eta dist; 0 selects 3, which at the
   * 20k-node slab stores 8 % fewer bytes than HLib's default 2 at the same
   * error (fewer blocks fail their probe check and fall back to dense). */
  double compress_eta;
} nsim_options;

/* Replaces raw_make_lam (src/nsimexec.ml:3605 -> src/nsim.ml:2828): builds
 * every operator (stiffness, divergence, gradient, dense BEM) on the device. */
int nsim_create(const nsim_mesh* mesh, const nsim_material* mat,
                const nsim_options* opt, nsim_handle** out);
void nsim_destroy(nsim_handle* h);
const char* nsim_last_error(void);

/* Batched form of nsim_create: n_magnets independent magnets (meshes[k],
 * mats[k]) in one handle.  Replaces one raw_make_lam per magnet
 * (src/nsimexec.ml:3605), the "batched variants take an instance-offset
 * table" form of the boundary.  The magnets' nodes are concatenated in order:
 * every field vector of the handle is the concatenation of the magnets'
 * vectors, and one rhs evaluation evaluates every magnet (grouped SpMV, one
 * persistent CG over all magnets, one grouped BEM GEMV with a dense matrix per
 * magnet).  The CG stopping test is on the residual of the whole batch. */
int nsim_create_batch(int n_magnets, const nsim_mesh* meshes, const nsim_material* mats,
                      const nsim_options* opt, nsim_handle** out);
/* ---- Differentiation form (no counterpart in the reference, which has no
 * adjoint).  A handle with a single magnet only. ---- */

/* Vector-Jacobian product of the right-hand side at y: y_bar = J(y)^T lam
 * with the full Jacobian (exchange, anisotropy, demag through both Poisson
 * solves and B^T, the LLG torque); optionally K1_bar = (df/dK1)^T lam per
 * node, *Ms_bar = (df/dMs)^T lam, Hext_bar = (df/dH_ext)^T lam (3N).  Host
 * pointers.  The handle's fields are left at y. */
int nsim_rhs_vjp(nsim_handle* h, double t, const double* y, const double* lam, double* y_bar,
                 double* K1_bar, double* Ms_bar, double* Hext_bar);

/* A fixed-step classical Runge-Kutta trajectory from the handle's m, with the
 * handle's H_ext plus piecewise-constant pulses: during window w = [w L,
 * (w+1) L) (L = pulse_len, a multiple of dt) the applied field gains
 * amps[w] * pulse_dir.  States are kept every checkpoint_every steps for the
 * adjoint, which recomputes the steps in between. */
typedef struct {
  double dt;            /* step, simulation time units (ps)                 */
  int n_steps;
  int checkpoint_every;
  int n_pulses;
  double pulse_len;
  double pulse_dir[3];
} nsim_traj;

/* Runs the trajectory; the handle's m becomes m(n_steps dt). */
int nsim_traj_forward(nsim_handle* h, const nsim_traj* tr, const double* amps);
/* Discrete adjoint of the last nsim_traj_forward (same tr and amps): given
 * m_bar_final = dLoss/dm(T) (3N), returns dLoss/dm(0) (3N), dLoss/dK1 per node
 * (N), dLoss/dMs and dLoss/damps (n_pulses).  Outputs may be null. */
int nsim_traj_adjoint(nsim_handle* h, const nsim_traj* tr, const double* amps,
                      const double* m_bar_final, double* m_bar_initial, double* K1_bar,
                      double* Ms_bar, double* amps_bar);

/* ---- The BDF trajectory and its discrete adjoint (no counterpart in the
 * reference, which has no adjoint; the forward is the reference's stepper,
 * src/nsim.ml:2537-2580).  A single-magnet handle.
 *
 * nsim_bdf_forward integrates from the handle's m at t = 0 to t_final with
 * the BDF integrator (tolerances, max_order as nsim_ts_init), never stepping
 * across a pulse edge (the applied field is the handle's H_ext plus
 * amps[w] * pulse_dir during [w pulse_len, (w+1) pulse_len)), and records
 * every accepted step: its state Y_{k+1}, step size, order, gamma and
 * history coefficients, so that Y_{k+1} - gamma f(Y_{k+1}) - sum_j A_j Y_{k-j}
 * = 0.  All states stay in device memory (max_steps + 1 of 3N doubles).
 * The handle's m becomes Y_K.
 *
 * nsim_bdf_replay re-integrates the recorded step sequence (the same step
 * sizes, orders and coefficients) from the handle's m with the current
 * material, anisotropy, H_ext and amps, solving every step's equation by
 * Newton to an RMS residual below newton_rtol; its states replace the
 * recorded ones.  With the step sequence fixed the map from parameters to
 * Y_K is smooth, and finite differences of it check the adjoint.
 *
 * nsim_bdf_adjoint: the discrete adjoint of the recorded (or replayed)
 * trajectory, each step's equation taken as solved exactly: from
 * m_bar_final = dLoss/dY_K it returns dLoss/dY_0 (3N), dLoss/dK1 per node,
 * dLoss/dMs and dLoss/damps, solving (I - gamma J(Y_{k+1})^T) mu = Ybar_{k+1}
 * per step by GMRES to lin_rtol with the full transposed Jacobian (demag
 * included).  Outputs may be null. */
typedef struct {
  double t_final;       /* end of the trajectory, from t = 0 (ps)            */
  double rtol, atol;    /* the integrator's tolerances                       */
  int max_order;        /* 1..5                                              */
  int max_steps;        /* steps the state buffer holds                      */
  int n_pulses;         /* as nsim_traj                                      */
  double pulse_len;
  double pulse_dir[3];
} nsim_bdf_traj;

int nsim_bdf_forward(nsim_handle* h, const nsim_bdf_traj* tr, const double* amps, int* n_steps);
/* The recorded steps: end times t_{k+1}, step sizes and orders (up to max
 * entries each; arrays may be null).  Returns the number of steps. */
int nsim_bdf_steps(nsim_handle* h, int max, double* t, double* hs, int* orders);
int nsim_bdf_replay(nsim_handle* h, const double* amps, double newton_rtol);
int nsim_bdf_adjoint(nsim_handle* h, const double* amps, const double* m_bar_final,
                     double* m_bar_initial, double* K1_bar, double* Ms_bar, double* amps_bar,
                     double lin_rtol);

/* Changes magnet k's material coefficients in place (the operators do not
 * depend on them).  Replaces what the reference does when a simulation is
 * rebuilt with new MagMaterial parameters (interface/nmag/main.py:423-471);
 * used to vary Ms in gradient checks. */
int nsim_set_material(nsim_handle* h, int magnet, const nsim_material* mat);

/* Polynomial anisotropy of magnet k (k = -1: every magnet): the energy
 * density E(m) = sum_j coeffs[j] mx^p0 my^p1 mz^p2, powers[3j..3j+2] = p,
 * in simulation units (0..256 terms, powers 0..16; nterms 0 clears it).  The
 * field is H_anis = -(g - m (m.g)) / (mu0 Ms), g = grad E, added to the
 * uniaxial field of the K1/K2/axis fields when both are set.  Replaces the
 * LAM equations that nsim_anisotropy_equations builds from the material's
 * fitted polynomial (src/nsim_anisotropy.ml:216-234, called from
 * interface/nmag/nmag_lam.py:482-501); the fit itself is the caller's
 * (interface/nmag/anisotropy.py: uniaxial, cubic or any function with an
 * order).  Like the reference's default (do_anisotropy_jacobian = False,
 * nmag_lam.py:133), nsim_jtimes leaves it out of J. */
int nsim_set_anisotropy(nsim_handle* h, int magnet, int nterms, const int* powers,
                        const double* coeffs);

/* Autotuner (no counterpart in the reference, whose PETSc kernels have no
 * tile shapes): times the candidates on the device and keeps the fastest.
 * what bit 0: the BEM GEMV tile shape (1/2/4 rows streamed together per
 * wavefront x 1/2/4/8 workgroups per CU); bit 1: the CG layout (1/2/4 rows per
 * thread, i.e. parts of 256/512/1024 nodes), timed on a cold Neumann solve of
 * the current m's load (so set m first).  The choice shows in nsim_get_stat
 * ("bem_rows", "bem_wg_per_cu", "cg_rows_per_thread", "tuned_bem_us",
 * "tuned_cg_us"); nsim_set_tuning sets the same knobs by hand. */
int nsim_autotune(nsim_handle* h, int what);

/* Instance-offset table of a handle: magnet k owns nodes
 * [node_off[k], node_off[k+1]) and boundary nodes [boundary_off[k],
 * boundary_off[k+1]); both arrays have n_magnets + 1 entries (either may be
 * null).  A handle from nsim_create is a batch of one. */
int nsim_get_batch_layout(nsim_handle* h, int* node_off, int* boundary_off);

/* sizes[0..5] = n_nodes, n_boundary_nodes, n_boundary_triangles, n_tets,
 * stored operator entries (padded SELL slots), stiffness non-zeros. */
int nsim_get_sizes(nsim_handle* h, int* sizes);

/* Replaces lam_set_field / lam_get_field (src/nsimexec.ml:4286-4305).
 * Settable: "m", "H_ext" (3N), "K1", "K2" (N), "axis" (3N), "pin" (N).
 * Readable additionally: "H_exch","H_demag","H_anis","H_total","dmdt" (3N),
 * "rho","phi","phi1","phi2","volumes" (N), "solid_angles" (N_b),
 * "boundary_nodes" (N_b, as doubles). */
int nsim_set_field(nsim_handle* h, const char* name, const double* host_data);
int nsim_get_field(nsim_handle* h, const char* name, double* host_out);

/* Replaces lam_execute (src/nsimexec.ml:4077-4095) for the path's programs:
 * "update_dmdt", "set_H_demag", "set_H_exch", "set_H_anis", "set_H_total". */
int nsim_execute(nsim_handle* h, const char* program);

/* Dense BEM matrix (N_b x N_b, row-major, FP64 copy) for inspection; for a
 * batch, the magnets' matrices one after another (each N_b,k x N_b,k). */
int nsim_get_bem(nsim_handle* h, double* host_out);

/* Dense potential operator M = (I + E B R) K^+ of each magnet (N_k x N_k,
 * row-major, FP64 copy, one after another) for inspection: phi = M rho.  Fails
 * on a handle that took the iterative route.  No counterpart in the
 * reference, which only ever applies K^-1 through KSP solves. */
int nsim_get_potential(nsim_handle* h, double* host_out);

/* CVODE rhs callback (src/sundials_sp_stubs.c:264, src/nsim.ml:2273-2313):
 * ydot = f(t, y).  y and ydot are 3N arrays; device_ptrs != 0 means they are
 * device pointers, else host pointers. */
int nsim_rhs(nsim_handle* h, double t, const double* y, double* ydot, int device_ptrs);

/* CVODE jtimes callback (src/sundials_sp_stubs.c:579, src/nsim.ml:2400-2417):
 * Jv = J v with the reference's jacobi_ts1 linearisation (nmag_lam.py:432-436:
 * H frozen except for the exchange term).  J is the Jacobian built by the
 * last nsim_psetup that rebuilt it, whatever y is, as in the reference, whose
 * jtimes multiplies by the matrix of its last re-Jacobi and reads neither y
 * nor fy.  Before the first psetup, J is taken at y with H_total of the last
 * rhs.  fy (CVODE's f(t, y)) is accepted for the callback's shape and not
 * read; y may be null once a psetup has run. */
int nsim_jtimes(nsim_handle* h, double t, const double* v, double* Jv,
                const double* y, const double* fy, int device_ptrs);

/* CVODE psetup callback (src/sundials_sp_stubs.c:352, src/nsim.ml:2419-2461):
 * when jok == 0 or no Jacobian exists yet, rebuilds the jacobi_ts1 Jacobian
 * at y (the fields of y are evaluated, as the reference's re-Jacobi runs the
 * field part of the rhs) and sets *jcur = 1; when jok != 0 keeps the last
 * Jacobian and sets *jcur = 0.  Then forms M = I - gamma J and its 3x3
 * node-block Jacobi splitting, the preconditioner of nsim_psolve's solve.
 * fy is not read (nor is it in the reference). */
int nsim_psetup(nsim_handle* h, double t, const double* y, const double* fy, int jok,
                int* jcur, double gamma, int device_ptrs);
/* CVODE psolve callback (src/sundials_sp_stubs.c:473, src/nsim.ml:2465-2484):
 * z solves M z = r, M = I - gamma J from the last nsim_psetup (psolve's gamma
 * and delta are not used, as in the reference).  Restarted GMRES(30) from a
 * zero guess, left-preconditioned by M's 3x3 node blocks (the reference's KSP
 * uses bjacobi/ILU), until the preconditioned residual is at most
 * max(pc_rtol x its initial value, pc_atol), for at most pc_maxit iterations,
 * or until it exceeds pc_dtol x its initial value: pc_rtol 1e-2, pc_atol 1e-7
 * (interface/nmag/main.py:728-729), dtol 1e5 and maxit 10000 (PETSc's
 * defaults); set them with nsim_ts_set_tols.  nsim_set_tuning
 * "pc_full_solve" 0 makes z = M_block^-1 r instead (one block solve).  Only
 * left preconditioning (lr == 1) is accepted, as in the reference
 * (src/nsim.ml:2469-2472).  Statistics: "pc_solves", "pc_iters_last",
 * "pc_iters_total", "pc_unconverged", "pc_jac_builds". */
int nsim_psolve(nsim_handle* h, double t, const double* y, const double* fy, const double* r,
                double* z, double gamma, double delta, int lr, int device_ptrs);

/* nsim_psolve with nothing allocated: the same solve, its GMRES workspace
 * taken from a static device array of the library (16 MiB per GPU, held until
 * the process ends) instead of memory nsim_psolve allocates at its first
 * call.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight solver loops, harnesses that forbid
 * hipMalloc).  Magnets of up to about 20,400 nodes (3N-vector workspace of
 * 816 bytes per node); larger ones fail with an error naming the limit.
 * nsim_psetup must have run (it allocates the frozen Jacobian).  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (two handles on
 * two host threads) race; calls in sequence are independent.  The call
 * synchronises its stream once per Arnoldi step, so it cannot be captured in a
 * HIP graph. */
int nsim_psolve_noalloc(nsim_handle* h, double t, const double* y, const double* fy,
                        const double* r, double* z, double gamma, double delta, int lr,
                        int device_ptrs);

/* Replaces lam_ts_set_tols (src/nsimexec.ml:4148-4171 -> src/nsim.ml:2581-2600):
 * psolve's KSP tolerances (rtol, atol, dtol, maxit) and the integrator's
 * rtol and atol.  A negative value (maxit: <= 0) keeps the current setting,
 * as the reference's optional arguments do.  Integrator tolerances take
 * effect at the next nsim_advance. */
int nsim_ts_set_tols(nsim_handle* h, double ksp_rtol, double ksp_atol, double ksp_dtol,
                     int ksp_maxit, double cvode_rtol, double cvode_atol);

/* Replaces lam_ts_init (src/nsimexec.ml:4113-4128 -> src/nsim.ml:2485-2536):
 * (re)initialises the device BDF integrator from the handle's m at t0 with
 * scalar tolerances (CVodeReInit semantics: history is discarded).
 * max_order <= 0 keeps 2, krylov_max <= 0 keeps 300 (main.py:838-839).
 * BDF orders 1 to 5 (CVODE's maxord): with max_order 1 or 2 the order rises
 * to it as the history allows; with 3 to 5 the order is chosen step by step
 * among q - 1, q, q + 1 from local-error estimates, as CVODE does
 * (nsim_set_tuning "ts_fixed_order" 1 makes it rise to max_order instead).
 * max_order above 5 and krylov_max above 500 are errors (status < 0). */
int nsim_ts_init(nsim_handle* h, double t0, double rtol, double atol, int max_order,
                 int krylov_max);
/* Replaces lam_ts_advance (src/nsimexec.ml:4130-4146 -> src/nsim.ml:2537-2580):
 * integrates towards t_final, at most max_steps steps when max_steps > 0.
 * exact_tstop != 0 never steps past t_final (CV_NORMAL_TSTOP); otherwise the
 * solution is interpolated back to t_final.  The handle's m receives y at
 * *t_reached.  Returns > 0 on step-size underflow / repeated failures. */
int nsim_advance(nsim_handle* h, double t_final, int max_steps, int exact_tstop,
                 double* t_reached);

/* Statistics: "cg_iters_nbc", "cg_iters_dbc" (last solve), "rhs_evals",
 * "cg_iters_nbc_total", "cg_iters_dbc_total", "cg_error"; and of the
 * integrator (cvode_get_stats, src/sundials_sp_stubs.c:1409-1528):
 * "ts_steps", "ts_rhs", "ts_newton", "ts_linear", "ts_err_fails",
 * "ts_conv_fails", "ts_h", "ts_order", "ts_t", "ts_max_dm" (max_i |m_i(t1) -
 * m_i(t0)| over the last nsim_advance, as mwe_norm_dist_fun),
 * "ts_order_steps_<q>" (accepted steps at order q = 1..5 since ts_init) and
 * "ts_linear_max" (most Krylov iterations of one Newton iteration). */
int nsim_get_stat(nsim_handle* h, const char* name, double* out);

/* Per-kernel device timers, the analogue of the reference's TSTART/TSTOP
 * (src/nsim.ml:3414-3453): when enabled, each launch on the handle's stream is
 * bracketed by HIP events.  name: "exch_div", "cg_nbc", "bem", "cg_dbc",
 * "demag_llg"; out[0] = total ms, out[1] = launches. */
int nsim_timers_enable(nsim_handle* h, int on);
/* Launch-shape knobs (used by the tile autotuner); key "cg_rows_per_thread"
 * takes 1, 2, 4 or 8.  No counterpart in the reference (PETSc picks its own).
 * "cg_coarse" (default 1) turns the CG's coarse correction on; "cg_fine"
 * (default 1) lets it use two or four sub-part aggregates per part (at most
 * 256 parts and 320 aggregates), else aggregates of whole parts (setting it
 * rebuilds the CG operator).
 * Integrator keys: "ts_fixed_order" (see nsim_ts_init), "gmres_device",
 * "gmres_pc_sweeps", "gmres_slices_per_wg", "cooperative_launch",
 * "gmres_onex_steps" (0 .. 32, default 32: the persistent GMRES's Arnoldi
 * steps that take one grid exchange each; the later ones take two), "gmres_precond"
 * (1, the default: block-Jacobi preconditioning of the Newton systems; 0:
 * none). */
int nsim_set_tuning(nsim_handle* h, const char* key, int value);
int nsim_get_timer(nsim_handle* h, const char* name, double* out);
int nsim_synchronize(nsim_handle* h);

#ifdef __cplusplus
}
#endif
#endif
