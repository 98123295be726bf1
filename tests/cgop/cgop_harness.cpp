// Host harness of the CG operator set-up (nsim_cg_setup.cpp) for the CPU
// tests: builds the reordered operator of a mesh and returns the pieces the
// tests check (row order, coarse inverses, per-part aggregate lists).
#include <algorithm>
#include <cstring>
#include <string>

#include "nsim_internal.h"

namespace {
nsim::CgOperator g_op;
}

// Builds the operator; info = {parts, npad, naggs, sub_shift, agg_shift,
// need_max, need_lg}.  Returns 0, or -1 with the message in msg.
extern "C" int cg_build(int n, const double* coords, int T, const int* tets, int cap, int fine,
                        int* info, char* msg, int msg_len) {
  nsim::HostOperators o;
  std::string err = nsim::build_host_operators(n, coords, T, tets, o);
  g_op = nsim::CgOperator();
  if (err.empty()) err = nsim::build_cg_operator(o, cap, g_op, fine != 0);
  if (!err.empty()) {
    std::strncpy(msg, err.c_str(), msg_len - 1);
    msg[msg_len - 1] = 0;
    return -1;
  }
  const int v[] = {g_op.parts, g_op.npad, g_op.naggs, g_op.sub_shift, g_op.agg_shift,
                   g_op.need_max, g_op.need_lg};
  std::memcpy(info, v, sizeof(v));
  return 0;
}

// Copies out the last build: p2o (npad), the two coarse inverses (naggs^2
// each) and, for a fine coarse space, need (parts * kCgNeed) and need_n.
extern "C" int cg_fetch(int* p2o, double* ainv_nbc, double* ainv_dbc, int* need, int* need_n) {
  std::copy(g_op.p2o.begin(), g_op.p2o.end(), p2o);
  std::copy(g_op.ainv_nbc.begin(), g_op.ainv_nbc.end(), ainv_nbc);
  std::copy(g_op.ainv_dbc.begin(), g_op.ainv_dbc.end(), ainv_dbc);
  if (need) std::copy(g_op.need.begin(), g_op.need.end(), need);
  if (need_n) std::copy(g_op.need_n.begin(), g_op.need_n.end(), need_n);
  return nsim::kCgNeed;
}
