// C ABI of the MI355X-native LLG right-hand-side path (declared in include/nsim.h).
//
// A handle is one magnet: its operators in HBM, its field vectors, one HIP
// stream and the persistent-CG communication buffers.  The rhs program follows
// update_dmdt (interface/nmag/nmag_lam.py:784-793): set_H_demag, set_H_exch,
// set_H_anis, set_H_total, then the site-wise LLG, all on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "nsim_device.h"
#include "nsim_handle.h"
#include "nsim_integrator.h"
#include "nsim_internal.h"
#include "nsim_vec.h"

using namespace nsim;

namespace {

thread_local std::string g_last_error;

int fail(const std::string& msg, int code = -1) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      return fail(std::string(#expr) + ": " + hipGetErrorString(_e));             \
  } while (0)

template <typename T>
hipError_t upload(T** dst, const T* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), count * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

template <typename T>
hipError_t zeros(T** dst, size_t count) {
  *dst = nullptr;
  if (count == 0) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), count * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemset(*dst, 0, count * sizeof(T));
}

// Records a named device interval around a launch when timers are enabled.
struct Span {
  nsim_handle* h;
  Timer* t = nullptr;
  hipEvent_t a{}, b{};
  Span(nsim_handle* hh, const char* name) : h(hh) {
    if (!h->timers_on) return;
    t = &h->timers[name];
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, h->stream);
  }
  ~Span() {
    if (!t) return;
    (void)hipEventRecord(b, h->stream);
    t->pending.push_back({a, b});
  }
};

void flush_timers(nsim_handle* h) {
  (void)hipStreamSynchronize(h->stream);
  for (auto& kv : h->timers) {
    for (auto& ev : kv.second.pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) kv.second.ms += ms;
      kv.second.count++;
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    kv.second.pending.clear();
  }
}

// (Re)plans a grouped dense GEMV (the BEM or the potential operator) for a
// tile shape and uploads its workgroup table.
int setup_gemv(nsim_handle* h, DenseOp& op, int rows_per_pass, int wg_per_cu, int waves) {
  if (op.h_n.empty()) return 0;
  if (!op.ctr) {  // row counter of the dynamic variant; it resets itself after use
    HIPCHK(zeros(&op.ctr, 2));
  }
  DenseOp trial = op;
  std::vector<int> wg_seg, wg_row0;
  if (!plan_dense_gemv(trial, h->num_cus, rows_per_pass, wg_per_cu, waves, &wg_seg, &wg_row0))
    return fail("GEMV tile shape out of range (rows 0/1/2/4 with 0 = dynamic rows on one "
                "segment, workgroups per CU 1..8, wavefronts per workgroup 4/16)");
  int *ds = nullptr, *dr = nullptr;
  hipError_t e = upload(&ds, wg_seg.data(), wg_seg.size());
  if (e == hipSuccess) e = upload(&dr, wg_row0.data(), wg_row0.size());
  if (e != hipSuccess) {
    (void)hipFree(ds);
    (void)hipFree(dr);
    return fail(std::string("GEMV plan upload: ") + hipGetErrorString(e));
  }
  drop_rhs_graphs(h);  // they hold the old plan's tables and launch shape
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(op.wg_seg);
  (void)hipFree(op.wg_row0);
  trial.wg_seg = ds;
  trial.wg_row0 = dr;
  op = trial;
  return 0;
}

// (Re)builds the CG operator with `rpt` rows per thread (parts of 256*rpt
// nodes) and uploads it; fails if the parts do not fit one co-resident grid.
int setup_cg(nsim_handle* h, int rpt) {
  // the fine coarse space (sub-part aggregates) when the parts allow it and
  // it fits a co-resident grid, else aggregates of whole parts
  CgOperator op;
  std::string err;
  int maxw = 0, resident = 0;
  bool fine = false;
  for (int pass = h->cg_fine ? 0 : 1; pass < 2; pass++) {
    op = CgOperator();
    err = build_cg_operator(h->host, 256 * rpt, op, pass == 0);
    if (!err.empty()) continue;
    fine = op.sub_shift > 0;
    maxw = cg_slot_variant(op.max_width, rpt);
    resident = cg_max_workgroups(rpt, maxw, h->device, fine);
    if (op.parts <= 1024 && op.parts <= resident) break;
    err = "CG parts do not fit one co-resident grid at " + std::to_string(rpt) +
          " rows per thread";
  }
  if (!err.empty()) return fail(err);
  for (void* p : h->cg_allocs) (void)hipFree(p);
  h->cg_allocs.clear();
  CgDevice c;
  c.parts = op.parts;
  c.npad = op.npad;
  c.rpt = rpt;
  c.cap_shift = rpt == 1 ? 8 : rpt == 2 ? 9 : 10;
  c.agg_shift = op.agg_shift;
  c.sub_shift = op.sub_shift;
  c.need_max = op.need_max;
  c.need_lg = op.need_lg;
  c.fine = fine;
  c.naggs = op.naggs;
  c.maxw = maxw;
  c.max_width = op.max_width;
  c.resident = resident;
  auto up = [&](auto** dst, const auto& vec) {
    hipError_t e = upload(dst, vec.data(), vec.size());
    if (*dst) h->cg_allocs.push_back(*dst);
    return e == hipSuccess;
  };
  bool ok = up(&c.sell_off, op.sell_off) && up(&c.sell_col, op.sell_col) && up(&c.K, op.sK) &&
            up(&c.p2o, op.p2o) && up(&c.dinv_nbc, op.dinv_nbc) &&
            up(&c.dinv_dbc, op.dinv_dbc) && up(&c.ainv_nbc, op.ainv_nbc) &&
            up(&c.ainv_dbc, op.ainv_dbc);
  if (ok && fine) {
    // the fine coarse inverses in FP32 (the kernel keeps them in LDS)
    const std::vector<float> fn(op.ainv_nbc.begin(), op.ainv_nbc.end());
    const std::vector<float> fd(op.ainv_dbc.begin(), op.ainv_dbc.end());
    ok = up(&c.ainvf_nbc, fn) && up(&c.ainvf_dbc, fd) && up(&c.need, op.need) &&
         up(&c.need_n, op.need_n);
  }
  if (!ok) return fail("CG operator upload failed");
  h->cg = c;
  return 0;
}

}  // namespace

namespace nsim {

FieldArgs field_args(nsim_handle* h, const double* m) {
  FieldArgs f;
  f.m = m;
  f.Hext = h->have_hext ? h->Hext : nullptr;
  f.k1 = h->have_anis ? h->k1 : nullptr;
  f.k2 = (h->have_anis && h->have_k2) ? h->k2 : nullptr;
  f.axis = h->have_anis ? h->axis : nullptr;
  f.pin = h->have_pin ? h->pin : nullptr;
  f.poly = h->poly;
  f.poly_off = h->poly ? h->poly_off : nullptr;
  return f;
}

// Full field evaluation for magnetisation m (device) into dmdt (device).
namespace {

// phi = M rho on the potential route: one dense GEMV or the two H-matrix
// phases.
hipError_t launch_potential(nsim_handle* h, const double* rho, double* phi) {
  // the H-matrix phases read rho in their cluster order, which k_exch_div
  // wrote beside rho
  if (h->use_hmatrix) return launch_hmatrix(h->d.hm, phi, h->stream);
  return launch_dense_gemv(h->d.pot, rho, phi, h->stream);
}

// The dense route's three launches, as recorded into a graph.
hipError_t dense_rhs_launches(nsim_handle* h, const double* m, double* dmdt, bool demag) {
  hipError_t e = launch_exch_div(h->d, h->mixed, m, h->Hexch, h->rho, demag, h->stream);
  if (e == hipSuccess && demag) e = launch_potential(h, h->rho, h->phi);
  if (e != hipSuccess) return e;
  FieldBufs b{h->Hexch, h->Hdemag, h->Hanis, h->Htotal, h->rho,
              h->phi1,  h->phi2,   h->phi,   h->phi1b};
  return launch_demag_llg(h->d, h->mixed, field_args(h, m), b, dmdt, demag, h->stream);
}

bool same_args(const FieldArgs& a, const FieldArgs& b) {
  return a.m == b.m && a.Hext == b.Hext && a.k1 == b.k1 && a.k2 == b.k2 && a.axis == b.axis &&
         a.pin == b.pin && a.poly == b.poly && a.poly_off == b.poly_off;
}

// Replays the captured evaluation for (m, dmdt) or captures it first.  Graph
// launches carry the same kernels and arguments as the eager path, so results
// are bit-identical; what they save is the per-launch host work.
int graph_rhs(nsim_handle* h, const double* m, double* dmdt, bool demag) {
  const FieldArgs f = field_args(h, m);
  auto same = [&](const RhsGraph& g) {
    return g.m == m && g.dmdt == dmdt && g.demag == demag && same_args(g.f, f);
  };
  for (RhsGraph& g : h->graphs)
    if (same(g)) {
      HIPCHK(hipGraphLaunch(g.exec, h->stream));
      h->graph_launches++;
      return 0;
    }
  const auto it = std::find_if(h->seen.begin(), h->seen.end(), same);
  if (it == h->seen.end()) {
    // first sight of this key: evaluate eagerly and remember it
    if (h->seen.size() >= kMaxRhsSeen) h->seen.erase(h->seen.begin());
    RhsGraph k;
    k.m = m;
    k.dmdt = dmdt;
    k.f = f;
    k.demag = demag;
    h->seen.push_back(k);
    HIPCHK(dense_rhs_launches(h, m, dmdt, demag));
    return 0;
  }
  h->seen.erase(it);
  if (h->graphs.size() >= kMaxRhsGraphs) {
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipGraphExecDestroy(h->graphs.front().exec);
    h->graphs.erase(h->graphs.begin());
  }
  // This evaluation runs eagerly (capture only records); the eager launches
  // also set the GEMV's LDS attribute, so nothing but kernels is recorded.
  HIPCHK(dense_rhs_launches(h, m, dmdt, demag));
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t le = dense_rhs_launches(h, m, dmdt, demag);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(h->stream, &graph);
  if (le != hipSuccess || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(std::string("graph capture of the field evaluation failed: ") +
                hipGetErrorString(le != hipSuccess ? le : ce));
  }
  RhsGraph g;
  g.m = m;
  g.dmdt = dmdt;
  g.f = f;
  g.demag = demag;
  const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ie);
  h->graphs.push_back(g);
  h->graph_captures++;
  return 0;
}

}  // namespace

void drop_rhs_graphs(nsim_handle* h) {
  h->seen.clear();
  if (h->graphs.empty()) return;
  (void)hipStreamSynchronize(h->stream);
  for (RhsGraph& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

int device_rhs(nsim_handle* h, const double* m, double* dmdt) {
  const bool demag = h->opt.do_demag != 0 && h->d.nb > 0;
  if (h->rhs_graph && !h->timers_on && (!demag || h->use_potential)) {
    // the eager first evaluation inside graph_rhs counts as the same one
    const int rc = graph_rhs(h, m, dmdt, demag);
    if (rc == 0) h->rhs_evals++;
    return rc;
  }
  {
    Span s(h, "exch_div");
    HIPCHK(launch_exch_div(h->d, h->mixed, m, h->Hexch, h->rho, demag, h->stream));
  }
  if (demag && h->use_potential) {
    // phi = M rho with the precomputed potential operator M = (I + E B R) K^+
    Span s(h, "potential");
    HIPCHK(launch_potential(h, h->rho, h->phi));
  } else if (demag) {
    CgWork& w = h->cgw;
    {
      Span s(h, "cg_nbc");
      w.sync = h->sync_nbc;
      HIPCHK(launch_cg(h->cg, w, false, h->cg_coarse, h->rho, h->phi1,
                       h->opt.nbc_rtol, h->opt.nbc_maxits, nullptr, nullptr, h->phi1b,
                       h->d.blocal, h->stream));
    }
    {
      Span s(h, "bem");
      HIPCHK(launch_dense_gemv(h->d.bem, h->phi1b, h->phi2, h->stream));
    }
    {
      Span s(h, "cg_dbc");
      w.sync = h->sync_dbc;
      HIPCHK(launch_cg(h->cg, w, true, h->cg_coarse, nullptr, h->phi2,
                       h->opt.dbc_rtol, h->opt.dbc_maxits, h->phi1, h->phi, nullptr,
                       h->d.blocal, h->stream));
    }
  }
  {
    Span s(h, "demag_llg");
    FieldBufs b{h->Hexch, h->Hdemag, h->Hanis, h->Htotal, h->rho,
                h->phi1,  h->phi2,   h->phi,   h->phi1b};
    HIPCHK(launch_demag_llg(h->d, h->mixed, field_args(h, m), b, dmdt, demag, h->stream));
  }
  h->rhs_evals++;
  return 0;
}

int device_jtimes(nsim_handle* h, const double* m, const double* v, double* Jv) {
  Span s(h, "jtimes");
  HIPCHK(launch_jtimes(h->d, h->mixed, m, h->Htotal, v, Jv, h->stream));
  return 0;
}

int refresh_coefficients(nsim_handle* h) {
  const HostOperators& o = h->host;
  std::vector<double> kap(o.n);
  for (int i = 0; i < o.n; i++)
    kap[i] = h->coefs[o.inst.empty() ? 0 : o.inst[i]].negJ * o.Kdiag[i] / o.V[i];
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(h->kappa, kap.data(), kap.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d.coefs, h->coefs.data(), h->coefs.size() * sizeof(Coeffs),
                   hipMemcpyHostToDevice));
  h->coef = h->coefs[0];
  return 0;
}

hipStream_t handle_stream(nsim_handle* h) { return h->stream; }
int handle_n(nsim_handle* h) { return h->d.n; }
double* handle_m(nsim_handle* h) { return h->m; }
const double* handle_Htotal(nsim_handle* h) { return h->Htotal; }
const double* handle_kappa(nsim_handle* h) { return h->kappa; }
const Coeffs* handle_coef_table(nsim_handle* h) { return h->d.coefs; }
const int* handle_inst(nsim_handle* h) { return h->d.inst; }
const DeviceOps& handle_ops(nsim_handle* h) { return h->d; }
bool handle_mixed(nsim_handle* h) { return h->mixed; }
int handle_device(nsim_handle* h) { return h->device; }
void handle_set_halt(nsim_handle* h, const int* halt) {
  drop_rhs_graphs(h);  // their kernel arguments hold the old word
  h->d.halt = halt;
  h->d.pot.halt = halt;
  h->d.hm.halt = halt;
  h->d.bem.halt = halt;
  h->cgw.halt = halt;
}
int handle_error_flags(nsim_handle* h, unsigned* out) {
  unsigned a[6], b[6];
  HIPCHK(hipMemcpyAsync(a, h->sync_nbc, sizeof(a), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(b, h->sync_dbc, sizeof(b), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = a[2] | b[2];
  return 0;
}
void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace nsim

namespace {

// Builds a handle for `ninst` independent magnets (one for nsim_create).  The
// magnets' meshes are concatenated, node indices offset, into one operator
// set: the sparse operators are block-diagonal by construction and the BEM
// matrix is stored as one dense segment per magnet.
int create_impl(const char* who, int ninst, const nsim_mesh* meshes,
                const nsim_material* mats, const nsim_options* opt, nsim_handle** out) {
  const std::string me(who);
  if (!meshes || !mats || !out) return fail(me + ": null argument");
  *out = nullptr;
  if (ninst < 1) return fail(me + ": need at least one magnet");
  nsim_handle* h = new nsim_handle();
  auto cleanup = [&](int code) {
    nsim_destroy(h);
    return code;
  };
  std::vector<int> node_off(ninst + 1, 0);
  std::vector<double> inst_eps(ninst, 0.0);
  for (int k = 0; k < ninst; k++) {
    const nsim_mesh& mk = meshes[k];
    if (mk.n_nodes <= 0 || mk.n_tets <= 0 || !mk.coords || !mk.tets)
      return cleanup(fail(me + ": magnet " + std::to_string(k) + " has an empty mesh"));
    for (long long e = 0; e < 4LL * mk.n_tets; e++)
      if (mk.tets[e] < 0 || mk.tets[e] >= mk.n_nodes)
        return cleanup(fail(me + ": magnet " + std::to_string(k) +
                            ": tetrahedron references a node out of range"));
    double gmax = 0.0;
    for (long long e = 0; e < 3LL * mk.n_nodes; e++) gmax = std::max(gmax, std::fabs(mk.coords[e]));
    inst_eps[k] = gmax * 1e-13;  // the single-magnet geometric tolerance
    if ((long long)node_off[k] + mk.n_nodes >= (1LL << 21))
      return cleanup(fail(me + ": more than 2^21 nodes in one handle"));
    node_off[k + 1] = node_off[k] + mk.n_nodes;
  }
  std::string err;
  if (ninst == 1) {
    err = build_host_operators(meshes[0].n_nodes, meshes[0].coords, meshes[0].n_tets,
                               meshes[0].tets, h->host);
  } else {
    std::vector<double> xs;
    std::vector<int> ts;
    xs.reserve(3 * (size_t)node_off[ninst]);
    for (int k = 0; k < ninst; k++) {
      xs.insert(xs.end(), meshes[k].coords, meshes[k].coords + 3 * (size_t)meshes[k].n_nodes);
      for (long long e = 0; e < 4LL * meshes[k].n_tets; e++)
        ts.push_back(meshes[k].tets[e] + node_off[k]);
    }
    err = build_host_operators(node_off[ninst], xs.data(), (int)(ts.size() / 4), ts.data(),
                               h->host);
    h->host.ninst = ninst;
    h->host.inst.resize(node_off[ninst]);
    for (int k = 0; k < ninst; k++)
      std::fill(h->host.inst.begin() + node_off[k], h->host.inst.begin() + node_off[k + 1], k);
  }
  if (!err.empty()) return cleanup(fail(me + ": " + err));
  h->node_off = node_off;
  if (opt) h->opt = *opt;
  else {
    h->opt.precision = NSIM_FP64;
    h->opt.do_demag = 1;
    h->opt.nbc_rtol = h->opt.dbc_rtol = 1e-5;
    h->opt.nbc_maxits = h->opt.dbc_maxits = 10000;
    h->opt.potential = 0;
    h->opt.compress_eps = 0.0;
    h->opt.compress_eta = 0.0;
  }
  if (h->opt.potential > 2)
    return cleanup(fail(me + ": unknown potential route " + std::to_string(h->opt.potential)));
  if (!(h->opt.compress_eps >= 0.0) || h->opt.compress_eps >= 1.0)
    return cleanup(fail(me + ": compress_eps must be in [0, 1)"));
  if (!(h->opt.compress_eta >= 0.0) || h->opt.compress_eta > 100.0)
    return cleanup(fail(me + ": compress_eta must be in [0, 100]"));
  if (h->opt.precision < NSIM_FP64 || h->opt.precision > NSIM_MIXED_F24)
    return cleanup(fail(me + ": unknown precision"));
  if (h->opt.nbc_maxits <= 0) h->opt.nbc_maxits = 10000;
  if (h->opt.dbc_maxits <= 0) h->opt.dbc_maxits = 10000;
  h->mixed = h->opt.precision != NSIM_FP64;
  std::vector<Coeffs> coefs(ninst);
  for (int k = 0; k < ninst; k++) {
    const nsim_material& mk = mats[k];
    if (!(mk.Ms > 0.0) || !(mk.mu0 > 0.0))
      return cleanup(fail(me + ": magnet " + std::to_string(k) + " needs Ms > 0 and mu0 > 0"));
    coefs[k] = Coeffs{mk.Ms, mk.negJ, mk.c1, mk.c2, mk.c3, mk.mu0};
  }
  h->coef = coefs[0];
  h->coefs = coefs;

  const HostOperators& o = h->host;
  DeviceOps& d = h->d;
  d.n = o.n;
  d.nb = o.nb;
  d.ntri = o.ntri;
  d.nslices = o.nslices;
  d.sell_total = o.sell_off[o.nslices];
  d.ninst = ninst;
  for (int sl = 0; sl < o.nslices; sl++)
    d.max_width = std::max(d.max_width, (o.sell_off[sl + 1] - o.sell_off[sl]) / kSlice);
  // storage of the BEM matrix, and of the dense potential operator (24-bit
  // storage is offered for the potential operator only)
  const int stype = h->opt.precision == NSIM_FP64        ? kBemF64
                    : h->opt.precision == NSIM_MIXED_BF16 ? kBemBF16
                                                          : kBemF32;
  const int ptype = h->opt.precision == NSIM_MIXED_F24 ? kBemF24 : stype;
  d.bem.type = stype;
  // BEM segments: the boundary nodes are in ascending node order, so each
  // magnet's boundary nodes form one contiguous run
  long long boff_elems = 0;
  int max_ldb = 0;
  for (int k = 0; k < ninst; k++) {
    const int b0 = (int)(std::lower_bound(o.bnode.begin(), o.bnode.end(), node_off[k]) -
                         o.bnode.begin());
    const int b1 = (int)(std::lower_bound(o.bnode.begin(), o.bnode.end(), node_off[k + 1]) -
                         o.bnode.begin());
    const int nbk = b1 - b0, ldk = ((nbk + 63) / 64) * 64;
    d.bem.h_n.push_back(nbk);
    d.bem.h_ld.push_back(ldk);
    d.bem.h_off.push_back(b0);
    d.bem.h_aoff.push_back(boff_elems);
    d.h_bem_geo_eps.push_back(ninst == 1 ? o.geo_eps : inst_eps[k]);
    boff_elems += (long long)nbk * ldk;
    max_ldb = std::max(max_ldb, ldk);
    h->bnd_off.push_back(b0);
  }
  h->bnd_off.push_back(o.nb);
  d.bem.elems = boff_elems;
  if (h->opt.do_demag && max_ldb > dense_gemv_max_order(stype))
    return cleanup(fail(me + ": " + std::to_string(max_ldb) +
                        " boundary nodes in one magnet exceed the BEM GEMV's LDS staging (" +
                        std::to_string(dense_gemv_max_order(stype)) + ")"));
  // dense potential operator: per magnet N_k x N_k, when every magnet fits
  // the GEMV's staging and all of them fit the budget
  {
    long long pot_elems = 0;
    int max_n = 0;
    for (int k = 0; k < ninst; k++) {
      const int nk = node_off[k + 1] - node_off[k];
      pot_elems += (long long)nk * (((nk + 63) / 64) * 64);
      max_n = std::max(max_n, nk);
    }
    const bool fits = max_n <= dense_gemv_max_order(ptype) &&
                      (double)dense_bytes(ptype, pot_elems) <= kPotentialBudgetBytes;
    if (h->opt.potential == 1 && !fits)
      return cleanup(fail(me + ": the dense potential operator needs every magnet to have at "
                               "most " + std::to_string(dense_gemv_max_order(ptype)) +
                          " nodes and at most 32 GiB in all"));
    h->use_hmatrix = h->opt.do_demag && o.nb > 0 && h->opt.potential == 2;
    h->use_potential =
        h->opt.do_demag && o.nb > 0 && (h->opt.potential == 2 || (h->opt.potential >= 0 && fits));
  }

  hipError_t e;
#define UP(dst, vec)                                                        \
  do {                                                                      \
    e = upload(&(dst), (vec).data(), (vec).size());                         \
    if (e != hipSuccess) return cleanup(fail(std::string("upload: ") + hipGetErrorString(e))); \
    h->track(dst);                                                          \
  } while (0)
#define ZEROS(dst, count)                                                   \
  do {                                                                      \
    e = zeros(&(dst), (count));                                             \
    if (e != hipSuccess) return cleanup(fail(std::string("alloc: ") + hipGetErrorString(e))); \
    h->track(dst);                                                          \
  } while (0)

  if (hipGetDevice(&h->device) != hipSuccess) return cleanup(fail("no HIP device"));
  (void)hipDeviceGetAttribute(&h->num_cus, hipDeviceAttributeMultiprocessorCount, h->device);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail("hipStreamCreate failed"));

  UP(d.sell_off, o.sell_off);
  UP(d.sell_col, o.sell_col);
  UP(d.K, o.sK);
  UP(d.D, o.sD);
  UP(d.G, o.sG);
  if (h->mixed) {
    std::vector<float> kf(o.sK.begin(), o.sK.end()), gf(o.sG.begin(), o.sG.end());
    UP(d.Kf, kf);
    UP(d.Gf, gf);
  }
  std::vector<double> vinv(o.n), dn(o.n), dd(o.n);
  for (int i = 0; i < o.n; i++) {
    vinv[i] = 1.0 / o.V[i];
    dn[i] = 1.0 / o.Kdiag[i];
    dd[i] = o.blocal[i] >= 0 ? 0.0 : dn[i];
  }
  UP(d.Vinv, vinv);
  UP(d.dinv_nbc, dn);
  UP(d.dinv_dbc, dd);
  UP(d.blocal, o.blocal);
  std::vector<double> kap(o.n);
  for (int i = 0; i < o.n; i++)
    kap[i] = coefs[o.inst.empty() ? 0 : o.inst[i]].negJ * o.Kdiag[i] / o.V[i];
  UP(h->kappa, kap);
  UP(d.coefs, coefs);
  if (ninst > 1) UP(d.inst, o.inst);

  // BEM matrix, assembled on the device from the boundary geometry
  if (o.nb > 0) {
    std::vector<int> tri_node(3 * (size_t)o.ntri);
    for (size_t k = 0; k < tri_node.size(); k++) tri_node[k] = o.bnode[o.tri[k]];
    UP(d.bnode, o.bnode);
    UP(d.coords, o.x);
    UP(d.tri_node, tri_node);
    UP(d.tnorm, o.tnorm);
    UP(d.omega, o.omega);
    UP(d.inc_rp, o.inc_rp);
    UP(d.inc_tri, o.inc_tri);
    UP(d.bem.n, d.bem.h_n);
    UP(d.bem.ld, d.bem.h_ld);
    UP(d.bem.off, d.bem.h_off);
    UP(d.bem.aoff, d.bem.h_aoff);
    d.bem.map = d.bnode;
    if (setup_gemv(h, d.bem, 2, 2, 4) != 0) return cleanup(-1);
    e = hipMalloc(&d.bem.A, dense_bytes(stype, d.bem.elems));
    if (e != hipSuccess) return cleanup(fail(std::string("BEM alloc: ") + hipGetErrorString(e)));
    h->track(d.bem.A);
    e = launch_bem_assemble(d, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess)
      return cleanup(fail(std::string("BEM assembly: ") + hipGetErrorString(e)));
    if (h->use_potential) {
      std::vector<int> iota(o.n);
      for (int i = 0; i < o.n; i++) iota[i] = i;
      UP(d.iota, iota);
      if (build_potential(h) != 0) return cleanup(-1);
      if (!h->use_hmatrix && setup_gemv(h, d.pot, 2, 1, 16) != 0) return cleanup(-1);
    }
  }

  // field vectors
  const size_t n = o.n;
  ZEROS(h->m, 3 * n);
  ZEROS(h->Hext, 3 * n);
  ZEROS(h->k1, n);
  ZEROS(h->k2, n);
  ZEROS(h->axis, 3 * n);
  ZEROS(h->pin, n);
  ZEROS(h->Hexch, 3 * n);
  ZEROS(h->Hdemag, 3 * n);
  ZEROS(h->Hanis, 3 * n);
  ZEROS(h->Htotal, 3 * n);
  ZEROS(h->dmdt, 3 * n);
  ZEROS(h->rho, n);
  ZEROS(h->phi1, n);
  ZEROS(h->phi2, n);
  ZEROS(h->phi, n);
  ZEROS(h->phi1b, std::max<size_t>(1, o.nb));
  ZEROS(h->io_a, 3 * n);
  ZEROS(h->io_b, 3 * n);
  ZEROS(h->io_c, 3 * n);

  // persistent CG: smallest rows-per-thread whose parts fit one co-resident grid
  int rc = -1;
  for (int rpt = 1; rpt <= 4 && rc != 0; rpt *= 2) rc = setup_cg(h, rpt);
  if (rc != 0) return cleanup(fail("CG set-up failed: " + g_last_error));
  h->cgw.gmax = 1024;
  // every layout pads by less than one part of at most 1024 rows
  ZEROS(h->cgw.pub, 2 * (n + 1024));
  ZEROS(h->cgw.part, 2 * (size_t)kCgRecord * h->cgw.gmax);
  ZEROS(h->cgw.tag, 2 * (size_t)h->cgw.gmax * dev::kTagStride);
  ZEROS(h->sync_nbc, 8);
  ZEROS(h->sync_dbc, 8);
  h->cgw.sync = h->sync_nbc;
#undef UP
#undef ZEROS
  if (hipDeviceSynchronize() != hipSuccess) return cleanup(fail("device sync failed"));
  *out = h;
  return 0;
}

}  // namespace

extern "C" {

const char* nsim_last_error(void) { return g_last_error.c_str(); }

int nsim_create(const nsim_mesh* mesh, const nsim_material* mat, const nsim_options* opt,
                nsim_handle** out) {
  return create_impl("nsim_create", 1, mesh, mat, opt, out);
}

int nsim_create_batch(int n_magnets, const nsim_mesh* meshes, const nsim_material* mats,
                      const nsim_options* opt, nsim_handle** out) {
  return create_impl("nsim_create_batch", n_magnets, meshes, mats, opt, out);
}

int nsim_set_material(nsim_handle* h, int magnet, const nsim_material* mat) {
  if (!h || !mat) return fail("nsim_set_material: null argument");
  if (magnet < 0 || magnet >= h->d.ninst) return fail("nsim_set_material: no such magnet");
  if (!(mat->Ms > 0.0) || !(mat->mu0 > 0.0))
    return fail("nsim_set_material: need Ms > 0 and mu0 > 0");
  h->coefs[magnet] = Coeffs{mat->Ms, mat->negJ, mat->c1, mat->c2, mat->c3, mat->mu0};
  return refresh_coefficients(h);
}

int nsim_set_anisotropy(nsim_handle* h, int magnet, int nterms, const int* powers,
                        const double* coeffs) {
  if (!h) return fail("nsim_set_anisotropy: null handle");
  if (magnet < -1 || magnet >= h->d.ninst) return fail("nsim_set_anisotropy: no such magnet");
  if (nterms < 0 || nterms > kMaxAnisTerms)
    return fail("nsim_set_anisotropy: 0 to " + std::to_string(kMaxAnisTerms) + " terms");
  if (nterms > 0 && (!powers || !coeffs)) return fail("nsim_set_anisotropy: null argument");
  std::vector<AnisTerm> terms(nterms);
  for (int j = 0; j < nterms; j++) {
    for (int d = 0; d < 3; d++) {
      if (powers[3 * j + d] < 0 || powers[3 * j + d] > kMaxAnisPower)
        return fail("nsim_set_anisotropy: powers must be 0 to " +
                    std::to_string(kMaxAnisPower));
      terms[j].p[d] = powers[3 * j + d];
    }
    terms[j].pad = 0;
    if (!std::isfinite(coeffs[j])) return fail("nsim_set_anisotropy: coefficient not finite");
    terms[j].c = coeffs[j];
  }
  h->poly_terms.resize(h->d.ninst);
  for (int k = 0; k < h->d.ninst; k++)
    if (magnet < 0 || k == magnet) h->poly_terms[k] = terms;
  // one table of every magnet's terms, offsets per magnet
  std::vector<AnisTerm> all;
  std::vector<int> off(1, 0);
  for (const auto& t : h->poly_terms) {
    all.insert(all.end(), t.begin(), t.end());
    off.push_back((int)all.size());
  }
  // the old table may still be read by queued launches
  if (hipStreamSynchronize(h->stream) != hipSuccess) return fail("nsim_set_anisotropy: sync");
  drop_rhs_graphs(h);
  if (h->poly) (void)hipFree(h->poly);
  if (h->poly_off) (void)hipFree(h->poly_off);
  h->poly = nullptr;
  h->poly_off = nullptr;
  if (all.empty()) return 0;
  if (hipMalloc(&h->poly, sizeof(AnisTerm) * all.size()) != hipSuccess ||
      hipMalloc(&h->poly_off, sizeof(int) * off.size()) != hipSuccess ||
      hipMemcpy(h->poly, all.data(), sizeof(AnisTerm) * all.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(h->poly_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice) !=
          hipSuccess) {
    if (h->poly) (void)hipFree(h->poly);
    if (h->poly_off) (void)hipFree(h->poly_off);
    h->poly = nullptr;
    h->poly_off = nullptr;
    h->poly_terms.clear();
    return fail("nsim_set_anisotropy: device upload failed");
  }
  return 0;
}

int nsim_get_batch_layout(nsim_handle* h, int* node_off, int* boundary_off) {
  if (!h) return fail("nsim_get_batch_layout: null handle");
  if (node_off) std::copy(h->node_off.begin(), h->node_off.end(), node_off);
  if (boundary_off) std::copy(h->bnd_off.begin(), h->bnd_off.end(), boundary_off);
  return 0;
}

void nsim_destroy(nsim_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h->ts;
  destroy_adjoint(h->adj);
  drop_rhs_graphs(h);
  for (auto& kv : h->timers)
    for (auto& ev : kv.second.pending) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  for (void* p : h->allocations) (void)hipFree(p);
  for (void* p : h->cg_allocs) (void)hipFree(p);
  if (h->poly) (void)hipFree(h->poly);
  if (h->poly_off) (void)hipFree(h->poly_off);
  for (DenseOp* op : {&h->d.bem, &h->d.pot}) {
    if (op->wg_seg) (void)hipFree(op->wg_seg);
    if (op->wg_row0) (void)hipFree(op->wg_row0);
    if (op->ctr) (void)hipFree(op->ctr);
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int nsim_get_sizes(nsim_handle* h, int* s) {
  if (!h || !s) return fail("nsim_get_sizes: null argument");
  s[0] = h->d.n;
  s[1] = h->d.nb;
  s[2] = h->d.ntri;
  s[3] = h->host.T;
  s[4] = h->d.sell_total;
  s[5] = h->host.rp[h->d.n];
  return 0;
}

int nsim_set_tuning(nsim_handle* h, const char* key, int value) {
  if (!h || !key) return fail("nsim_set_tuning: null argument");
  std::string k(key);
  if (k == "cg_rows_per_thread") {
    if (value != 1 && value != 2 && value != 4)
      return fail("cg_rows_per_thread must be 1, 2 or 4");
    HIPCHK(hipStreamSynchronize(h->stream));
    return setup_cg(h, value);
  }
  if (k == "cg_coarse") {
    h->cg_coarse = value != 0;
    return 0;
  }
  if (k == "cg_fine") {  // sub-part aggregates of the coarse space (rebuilds the CG)
    h->cg_fine = value != 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    return setup_cg(h, h->cg.rpt);
  }
  if (k == "cooperative_launch") {  // persistent CG / GMRES launches
    h->cgw.cooperative = value != 0;
    h->ts_opt.cooperative = value != 0;
    if (h->ts) h->ts->options().cooperative = value != 0;
    return 0;
  }
  if (k == "pc_full_solve") {
    h->pc_full = value != 0;
    return 0;
  }
  if (k == "gmres_device") {
    h->ts_opt.device_gmres = value != 0;
    if (h->ts) h->ts->options().device_gmres = h->ts_opt.device_gmres;
    return 0;
  }
  if (k == "ts_fixed_order") {
    h->ts_opt.fixed_order = value != 0;
    if (h->ts) h->ts->options().fixed_order = h->ts_opt.fixed_order;
    return 0;
  }
  if (k == "gmres_pc_sweeps") {
    if (value < 0 || value > 8) return fail("gmres_pc_sweeps must be 0..8");
    h->ts_opt.gmres_pc_sweeps = value;
    if (h->ts) h->ts->options().gmres_pc_sweeps = value;
    return 0;
  }
  if (k == "gmres_precond") {  // 0: no preconditioner (tests of long Krylov solves)
    h->ts_opt.precond = value != 0 ? 1 : 0;
    if (h->ts) h->ts->options().precond = h->ts_opt.precond;
    return 0;
  }
  if (k == "gmres_onex_steps") {
    if (value < 0 || value > kGmresOneExMax)
      return fail("gmres_onex_steps must be 0.." + std::to_string(kGmresOneExMax));
    h->ts_opt.gmres_onex_steps = value;
    if (h->ts) h->ts->options().gmres_onex_steps = value;
    return 0;
  }
  if (k == "gmres_slices_per_wg") {  // takes effect at the next nsim_ts_init
    if (value < 1 || value > 64) return fail("gmres_slices_per_wg must be 1..64");
    h->ts_opt.gmres_spw = value;
    return 0;
  }
  if (k == "rhs_graph") {
    h->rhs_graph = value != 0;
    if (!h->rhs_graph) drop_rhs_graphs(h);
    return 0;
  }
  DenseOp& b = h->d.bem;
  DenseOp& p = h->d.pot;
  if (k == "bem_rows") return setup_gemv(h, b, value, b.wg_per_cu, b.waves);
  if (k == "bem_wg_per_cu") return setup_gemv(h, b, b.rows, value, b.waves);
  if (k == "bem_waves") return setup_gemv(h, b, b.rows, b.wg_per_cu, value);
  if (k == "pot_rows") return setup_gemv(h, p, value, p.wg_per_cu, p.waves);
  if (k == "pot_wg_per_cu") return setup_gemv(h, p, p.rows, value, p.waves);
  if (k == "pot_waves") return setup_gemv(h, p, p.rows, p.wg_per_cu, value);
  return fail("unknown tuning key: " + k);
}

}  // extern "C"

namespace {

// Times every tile shape of a grouped dense GEMV and keeps the fastest;
// returns its time per launch in microseconds (0 if the operator is empty).
double tune_dense(nsim_handle* h, DenseOp& op, const double* x, double* y, hipEvent_t a,
                  hipEvent_t b) {
  if (!op.A) return 0.0;
  double best = 1e30;
  int br = op.rows, bw = op.wg_per_cu, bv = op.waves;
  for (int v : {4, 16})
    for (int r : {0, 1, 2, 4})
      for (int w : {1, 2, 4, 8}) {
        if (setup_gemv(h, op, r, w, v) != 0) continue;
        bool ok = true;
        for (int k = 0; k < 2 && ok; k++)
          ok = launch_dense_gemv(op, x, y, h->stream) == hipSuccess;
        if (!ok) continue;
        (void)hipEventRecord(a, h->stream);
        for (int k = 0; k < 10; k++) (void)launch_dense_gemv(op, x, y, h->stream);
        (void)hipEventRecord(b, h->stream);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
          continue;
        const double us = ms * 100.0;  // per launch
        if (us < best) best = us, br = r, bw = w, bv = v;
      }
  if (setup_gemv(h, op, br, bw, bv) != 0) return -1.0;
  return best;
}

}  // namespace

extern "C" {

int nsim_autotune(nsim_handle* h, int what) {
  if (!h) return fail("nsim_autotune: null handle");
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  auto done = [&](int rc) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
  };
  auto elapsed_ms = [&]() {
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
      return 1e30;
    return (double)ms;
  };
  const bool demag = h->opt.do_demag != 0 && h->d.nb > 0;
  if ((what & 1) && h->d.nb > 0) {
    // BEM GEMV tile shape (scratch output: the GEMV result is not kept)
    h->tuned_bem_us = tune_dense(h, h->d.bem, h->phi1b, h->phi2, a, b);
    if (h->tuned_bem_us < 0.0) return done(-1);
  }
  if ((what & 4) && h->d.pot.A) {
    h->tuned_pot_us = tune_dense(h, h->d.pot, h->rho, h->phi2, a, b);
    if (h->tuned_pot_us < 0.0) return done(-1);
  }
  if ((what & 2) && demag && !h->use_potential) {
    // CG layout: rows per thread (part size), timed on a cold Neumann solve
    // of the load of the current m
    if (launch_exch_div(h->d, h->mixed, h->m, h->Hexch, h->rho, true, h->stream) != hipSuccess)
      return done(fail("nsim_autotune: load launch failed"));
    double best = 1e30;
    int brpt = h->cg.rpt;
    for (int rpt : {1, 2, 4}) {
      HIPCHK(hipStreamSynchronize(h->stream));
      if (setup_cg(h, rpt) != 0) continue;
      CgWork& w = h->cgw;
      w.sync = h->sync_nbc;
      double tot = 0.0;
      for (int k = 0; k < 3; k++) {
        HIPCHK(hipMemsetAsync(h->phi1, 0, sizeof(double) * h->d.n, h->stream));
        (void)hipEventRecord(a, h->stream);
        HIPCHK(launch_cg(h->cg, w, false, h->cg_coarse, h->rho, h->phi1, h->opt.nbc_rtol,
                         h->opt.nbc_maxits, nullptr, nullptr, h->phi1b, h->d.blocal,
                         h->stream));
        (void)hipEventRecord(b, h->stream);
        tot += elapsed_ms();
      }
      if (tot < best) best = tot, brpt = rpt;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (setup_cg(h, brpt) != 0) return done(-1);
    HIPCHK(hipMemsetAsync(h->phi1, 0, sizeof(double) * h->d.n, h->stream));
    HIPCHK(hipMemsetAsync(h->phi2, 0, sizeof(double) * h->d.n, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->tuned_cg_us = best * 1e3 / 3.0;
  }
  g_last_error.clear();
  return done(0);
}

int nsim_set_field(nsim_handle* h, const char* name, const double* data) {
  if (!h || !name || !data) return fail("nsim_set_field: null argument");
  std::string f(name);
  const size_t n = h->d.n;
  double* dst = nullptr;
  size_t cnt = 0;
  if (f == "m") dst = h->m, cnt = 3 * n;
  else if (f == "H_ext") dst = h->Hext, cnt = 3 * n, h->have_hext = true;
  else if (f == "K1") dst = h->k1, cnt = n, h->have_anis = true;
  else if (f == "K2") dst = h->k2, cnt = n, h->have_k2 = true;
  else if (f == "axis") dst = h->axis, cnt = 3 * n;
  else if (f == "pin") dst = h->pin, cnt = n, h->have_pin = true;
  else if (f == "phi1") dst = h->phi1, cnt = n;
  else if (f == "phi2") dst = h->phi2, cnt = n;
  else return fail("nsim_set_field: unknown or read-only field '" + f + "'");
  HIPCHK(hipMemcpyAsync(dst, data, cnt * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int nsim_get_field(nsim_handle* h, const char* name, double* out) {
  if (!h || !name || !out) return fail("nsim_get_field: null argument");
  std::string f(name);
  const size_t n = h->d.n;
  const double* src = nullptr;
  size_t cnt = 0;
  if (f == "volumes") {
    std::copy(h->host.V.begin(), h->host.V.end(), out);
    return 0;
  }
  if (f == "solid_angles") {
    std::copy(h->host.omega.begin(), h->host.omega.end(), out);
    return 0;
  }
  if (f == "boundary_nodes") {
    for (int i = 0; i < h->d.nb; i++) out[i] = h->host.bnode[i];
    return 0;
  }
  if ((f == "phi1" || f == "phi2") && h->use_potential)
    return fail("nsim_get_field: '" + f + "' is not formed on the dense potential route "
                "(phi = phi1 + phi2 is)");
  if (f == "m") src = h->m, cnt = 3 * n;
  else if (f == "H_ext") src = h->Hext, cnt = 3 * n;
  else if (f == "H_exch") src = h->Hexch, cnt = 3 * n;
  else if (f == "H_demag") src = h->Hdemag, cnt = 3 * n;
  else if (f == "H_anis") src = h->Hanis, cnt = 3 * n;
  else if (f == "H_total") src = h->Htotal, cnt = 3 * n;
  else if (f == "dmdt") src = h->dmdt, cnt = 3 * n;
  else if (f == "rho") src = h->rho, cnt = n;
  else if (f == "phi") src = h->phi, cnt = n;
  else if (f == "phi1") src = h->phi1, cnt = n;
  else if (f == "phi2") src = h->phi2, cnt = n;
  else if (f == "K1") src = h->k1, cnt = n;
  else if (f == "pin") src = h->pin, cnt = n;
  else return fail("nsim_get_field: unknown field '" + f + "'");
  HIPCHK(hipMemcpyAsync(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"

namespace {

// Copies a DenseOp's segments to the host as FP64, one after another, each
// n_k x n_k row-major without padding.
int copy_dense(nsim_handle* h, const DenseOp& op, double* out) {
  if (!op.A) return 0;
  std::vector<unsigned char> buf(dense_bytes(op.type, op.elems));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(buf.data(), op.A, buf.size(), hipMemcpyDeviceToHost));
  size_t o = 0;
  for (size_t s = 0; s < op.h_n.size(); s++) {
    const size_t nb = op.h_n[s], ld = op.h_ld[s], base = op.h_aoff[s];
    for (size_t i = 0; i < nb; i++)
      for (size_t j = 0; j < nb; j++) {
        const size_t k = base + i * ld + j;
        double v;
        if (op.type == kBemF64) v = reinterpret_cast<double*>(buf.data())[k];
        else if (op.type == kBemF32) v = reinterpret_cast<float*>(buf.data())[k];
        else if (op.type == kBemF24) {
          // row i of the segment starts at byte 3 (base + i ld); within it
          // entry j sits in a block of up to 512: high 16 bits, then low 8
          const unsigned char* row = buf.data() + 3 * (base + i * ld);
          const size_t j0 = j & ~(size_t)511, nbk = std::min<size_t>(512, ld - j0);
          uint16_t hi;
          std::memcpy(&hi, row + 3 * j0 + 2 * (j - j0), 2);
          const uint32_t u = ((uint32_t)hi << 16) | ((uint32_t)row[3 * j0 + 2 * nbk + (j - j0)] << 8);
          float f;
          std::memcpy(&f, &u, 4);
          v = f;
        } else {
          uint32_t u = (uint32_t)reinterpret_cast<uint16_t*>(buf.data())[k] << 16;
          float f;
          std::memcpy(&f, &u, 4);
          v = f;
        }
        out[o + i * nb + j] = v;
      }
    o += nb * nb;
  }
  return 0;
}

}  // namespace

extern "C" {

int nsim_get_bem(nsim_handle* h, double* out) {
  if (!h || !out) return fail("nsim_get_bem: null argument");
  return copy_dense(h, h->d.bem, out);
}

int nsim_get_potential(nsim_handle* h, double* out) {
  if (!h || !out) return fail("nsim_get_potential: null argument");
  if (!h->d.pot.A)
    return fail(h->use_hmatrix
                    ? "nsim_get_potential: this handle holds the operator compressed (H-matrix)"
                    : "nsim_get_potential: this handle uses the iterative route");
  return copy_dense(h, h->d.pot, out);
}

int nsim_execute(nsim_handle* h, const char* program) {
  if (!h || !program) return fail("nsim_execute: null argument");
  std::string p(program);
  // Every field program of the path recomputes the whole dependency set of
  // update_dmdt: the H_* vectors it leaves behind are those of the current m.
  if (p == "update_dmdt" || p == "set_H_demag" || p == "set_H_exch" || p == "set_H_anis" ||
      p == "set_H_total")
    return device_rhs(h, h->m, h->dmdt);
  return fail("nsim_execute: unknown program '" + p + "'");
}

int nsim_rhs(nsim_handle* h, double t, const double* y, double* ydot, int device_ptrs) {
  (void)t;
  if (!h || !y || !ydot) return fail("nsim_rhs: null argument");
  const size_t bytes = 3 * (size_t)h->d.n * sizeof(double);
  if (device_ptrs) return device_rhs(h, y, ydot);
  HIPCHK(hipMemcpyAsync(h->io_a, y, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = device_rhs(h, h->io_a, h->io_b);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ydot, h->io_b, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int nsim_get_stat(nsim_handle* h, const char* name, double* out) {
  if (!h || !name || !out) return fail("nsim_get_stat: null argument");
  std::string s(name);
  if (s == "rhs_evals") {
    *out = (double)h->rhs_evals;
    return 0;
  }
  if (s == "graph_launches" || s == "graph_captures") {
    *out = (double)(s == "graph_launches" ? h->graph_launches : h->graph_captures);
    return 0;
  }
  if (s == "cg_grid") {
    *out = h->cg.parts;
    return 0;
  }
  if (s == "cg_rows_per_thread") {
    *out = h->cg.rpt;
    return 0;
  }
  if (s == "cg_aggregates") {
    *out = h->cg_coarse ? h->cg.naggs : 0;
    return 0;
  }
  if (s.rfind("ts_", 0) == 0) {
    if (!h->ts) return fail("nsim_get_stat: timestepper not initialised");
    const TsStats& t = h->ts->stats();
    if (s == "ts_steps") *out = (double)t.steps;
    else if (s == "ts_rhs") *out = (double)t.rhs;
    else if (s == "ts_newton") *out = (double)t.newton;
    else if (s == "ts_linear") *out = (double)t.linear;
    else if (s == "ts_linear_max") *out = (double)t.linear_max;
    else if (s == "ts_err_fails") *out = (double)t.err_fails;
    else if (s == "ts_conv_fails") *out = (double)t.conv_fails;
    else if (s == "ts_h") *out = t.h_last;
    else if (s == "ts_order") *out = t.q_last;
    else if (s.rfind("ts_order_steps_", 0) == 0 && s.size() == 16 && s[15] >= '1' &&
             s[15] <= '0' + kMaxOrder)
      *out = (double)t.order_steps[s[15] - '0'];
    else if (s == "ts_t") *out = t.t;
    else if (s == "ts_max_dm") *out = t.max_dm;
    else if (s == "ts_gmres_grid") *out = h->ts->gmres_grid();
    else return fail("nsim_get_stat: unknown statistic '" + s + "'");
    return 0;
  }
  if (s.rfind("pc_", 0) == 0) {
    if (s == "pc_solves") *out = (double)h->pc_solves;
    else if (s == "pc_iters_last") *out = h->pc_iters_last;
    else if (s == "pc_iters_total") *out = (double)h->pc_iters_total;
    else if (s == "pc_unconverged") *out = (double)h->pc_unconverged;
    else if (s == "pc_jac_builds") *out = (double)h->pc_jac_builds;
    else return fail("nsim_get_stat: unknown statistic '" + s + "'");
    return 0;
  }
  if (s == "cg_register_slots") {
    *out = h->cg.maxw;
    return 0;
  }
  {
    const DenseOp& b = h->d.bem;
    const DenseOp& p = h->d.pot;
    const HMatrixOp& hm = h->d.hm;
    const double pot_bytes = h->use_hmatrix ? (double)(hm.p1.bytes + hm.p2.bytes)
                                            : (double)dense_bytes(p.type, p.elems);
    const std::map<std::string, double> simple = {
        {"hm_eps", hm.eps}, {"hm_eta", hm.eta}, {"hm_rank_entries", (double)hm.nt},
        {"hm_full_entries", (double)hm.full_entries},
        {"hm_dense_entries", (double)hm.dense_entries},
        {"hm_lowrank_entries", (double)hm.lowrank_entries},
        {"hm_lowrank_blocks", (double)hm.lowrank_blocks},
        {"hm_dense_blocks", (double)hm.dense_blocks}, {"hm_max_rank", (double)hm.max_rank},
        {"hm_fallback_blocks", (double)hm.fallback_blocks}, {"hm_build_s", hm.build_s},
        {"hm_p1_panels", (double)hm.p1.panels}, {"hm_p2_panels", (double)hm.p2.panels},
        {"hm_p1_entries", (double)hm.p1.entries}, {"hm_p2_entries", (double)hm.p2.entries},
        {"bem_rows", b.rows}, {"bem_wg_per_cu", b.wg_per_cu}, {"bem_waves", b.waves},
        {"bem_grid", b.grid}, {"pot_rows", p.rows}, {"pot_wg_per_cu", p.wg_per_cu},
        {"pot_waves", p.waves}, {"pot_grid", p.grid}, {"n_magnets", h->d.ninst},
        {"potential", h->use_hmatrix ? 2.0 : p.A ? 1.0 : 0.0},
        {"tuned_bem_us", h->tuned_bem_us},
        {"tuned_cg_us", h->tuned_cg_us}, {"tuned_pot_us", h->tuned_pot_us},
        {"potential_bytes", pot_bytes},
        {"bem_bytes", (double)dense_bytes(b.type, b.elems)}};
    const auto it = simple.find(s);
    if (it != simple.end()) {
      *out = it->second;
      return 0;
    }
  }
  unsigned a[8], b[8];
  HIPCHK(hipMemcpyAsync(a, h->sync_nbc, sizeof(a), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(b, h->sync_dbc, sizeof(b), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (s == "cg_iters_nbc") *out = a[3];
  else if (s == "cg_iters_dbc") *out = b[3];
  else if (s == "cg_iters_nbc_total") *out = a[4];
  else if (s == "cg_iters_dbc_total") *out = b[4];
  else if (s == "cg_solves") *out = a[5] + b[5];
  else if (s == "cg_error") *out = (double)(a[2] | b[2]);
  else if (s == "cg_unconverged") *out = (double)(a[6] + b[6]);
  else return fail("nsim_get_stat: unknown statistic '" + s + "'");
  return 0;
}

int nsim_timers_enable(nsim_handle* h, int on) {
  if (!h) return fail("nsim_timers_enable: null handle");
  flush_timers(h);
  if (on) h->timers.clear();
  h->timers_on = on != 0;
  return 0;
}

int nsim_get_timer(nsim_handle* h, const char* name, double* out) {
  if (!h || !name || !out) return fail("nsim_get_timer: null argument");
  flush_timers(h);
  auto it = h->timers.find(name);
  out[0] = it == h->timers.end() ? 0.0 : it->second.ms;
  out[1] = it == h->timers.end() ? 0.0 : (double)it->second.count;
  return 0;
}

int nsim_ts_init(nsim_handle* h, double t0, double rtol, double atol, int max_order,
                 int krylov_max) {
  if (!h) return fail("nsim_ts_init: null handle");
  if (!(rtol >= 0.0) || !(atol > 0.0)) return fail("nsim_ts_init: need rtol >= 0, atol > 0");
  // BDF orders 1 to 5 (CVODE's maxord range; the reference's default is 2,
  // main.py:838); Krylov bases up to 500 vectors
  if (max_order > kMaxOrder)
    return fail("nsim_ts_init: max_order " + std::to_string(max_order) +
                " not available (BDF orders 1 to 5 are)");
  if (krylov_max > 500)
    return fail("nsim_ts_init: krylov_max " + std::to_string(krylov_max) +
                " above the largest Krylov basis (500)");
  TsOptions o = h->ts_opt;
  o.rtol = rtol;
  o.atol = atol;
  if (max_order > 0) o.max_order = max_order;
  if (krylov_max > 0) o.krylov_max = krylov_max;
  if (!h->ts || o.krylov_max != h->ts->options().krylov_max ||
      o.gmres_spw != h->ts->options().gmres_spw) {
    delete h->ts;
    h->ts = new Integrator(h, o);
  } else {
    h->ts->options() = o;
  }
  h->ts_opt = o;
  return h->ts->init(t0);
}

int nsim_advance(nsim_handle* h, double t_final, int max_steps, int exact_tstop,
                 double* t_reached) {
  if (!h || !t_reached) return fail("nsim_advance: null argument");
  if (!h->ts) return fail("nsim_advance: call nsim_ts_init first");
  const int rc = h->ts->advance(t_final, max_steps, exact_tstop != 0, t_reached);
  if (rc) return rc;
  // the Poisson CG runs only on the iterative route: with the potential
  // operator there is no flag to read back (and no extra synchronisation)
  unsigned flags = 0;
  if (!h->use_potential && handle_error_flags(h, &flags) == 0 && flags != 0)
    return fail("nsim_advance: a Poisson solve reported an error (flag " +
                std::to_string(flags) + ")");
  return 0;
}

int nsim_synchronize(nsim_handle* h) {
  if (!h) return fail("nsim_synchronize: null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"
