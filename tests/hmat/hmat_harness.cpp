// Host harness of the H-matrix builder (nsim_hmatrix_build.cpp) for the CPU
// tests: compresses a dense operator and applies the result from its panels.
#include <cstring>
#include <string>

#include "nsim_hmatrix_build.h"

extern "C" int hm_apply(int N, const double* A, const double* xyz, double eps, double eta,
                        int group, int nvec, const double* x, double* y, long long* stats,
                        char* msg, int msg_len) {
  nsim::HMatrixHost H;
  H.group = group;
  const std::string err = nsim::hmatrix_compress(N, A, N, xyz, 0, eps, eta, 4, H);
  if (!err.empty()) {
    std::strncpy(msg, err.c_str(), msg_len - 1);
    return -1;
  }
  nsim::hmatrix_finish(H);
  for (int v = 0; v < nvec; v++) nsim::hmatrix_apply_host(H, x + (size_t)v * N, y + (size_t)v * N);
  const long long s[] = {H.dense_entries, H.lowrank_entries, H.full_entries, H.lowrank_blocks,
                         H.dense_blocks,  H.max_rank,        H.fallback_blocks, H.nt,
                         H.p1.panels(),   H.p2.panels(),     (long long)H.p1.data.size(),
                         (long long)H.p2.data.size()};
  std::memcpy(stats, s, sizeof(s));
  return 0;
}
