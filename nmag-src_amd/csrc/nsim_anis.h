// Polynomial anisotropy on the device: E(m) = sum_j c_j mx^px my^py mz^pz,
// the homogeneous polynomial that represents a material's anisotropy energy
// density on the unit sphere (src/nsim_anisotropy.ml:119-133), and its field
//   H_anis = -(g - m (m.g)) / (mu0 Ms),  g = grad E(m),
// the unit-sphere tangential gradient of src/nsim_anisotropy.ml:167-197 with
// the factor 1/(mu0 Ms) of interface/nmag/nmag_lam.py:490-497.
#pragma once

#include <hip/hip_runtime.h>

#include "nsim_device.h"

namespace nsim {
namespace dev {

__device__ inline double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; i++) r *= x;
  return r;
}

// g = grad E(m) and, when hs is not null, the Hessian of E at m (row-major
// 3 x 3, symmetric).
__device__ inline void poly_grad(const AnisTerm* __restrict__ t, int nt, const double m[3],
                                 double g[3], double* hs) {
  g[0] = g[1] = g[2] = 0.0;
  if (hs)
    for (int q = 0; q < 9; q++) hs[q] = 0.0;
  for (int j = 0; j < nt; j++) {
    const int p[3] = {t[j].p[0], t[j].p[1], t[j].p[2]};
    const double c = t[j].c;
    for (int d = 0; d < 3; d++) {
      if (p[d] == 0) continue;
      int q[3] = {p[0], p[1], p[2]};
      q[d] -= 1;
      g[d] += c * p[d] * ipow(m[0], q[0]) * ipow(m[1], q[1]) * ipow(m[2], q[2]);
      if (!hs) continue;
      for (int e = 0; e < 3; e++) {
        if (q[e] == 0) continue;
        int r[3] = {q[0], q[1], q[2]};
        r[e] -= 1;
        hs[3 * d + e] += c * p[d] * q[e] * ipow(m[0], r[0]) * ipow(m[1], r[1]) * ipow(m[2], r[2]);
      }
    }
  }
}

// Adds the polynomial anisotropy field of magnet k at m to ha (s = 1/(mu0 Ms)).
__device__ inline void poly_field(const FieldArgs& f, int k, const double m[3], double s,
                                  double ha[3]) {
  const int a = f.poly_off[k], b = f.poly_off[k + 1];
  if (b <= a) return;
  double g[3];
  poly_grad(f.poly + a, b - a, m, g, nullptr);
  const double mg = m[0] * g[0] + m[1] * g[1] + m[2] * g[2];
  for (int d = 0; d < 3; d++) ha[d] -= s * (g[d] - m[d] * mg);
}

// Adds (dH_anis/dm)^T w of the polynomial anisotropy of magnet k at m to av:
// with H = -s (g - m (m.g)),
//   (dH/dm)^T w = -s (Hs w - (m.g) w - (m.w) (g + Hs m)).
__device__ inline void poly_field_vjp(const FieldArgs& f, int k, const double m[3], double s,
                                      const double w[3], double av[3]) {
  const int a = f.poly_off[k], b = f.poly_off[k + 1];
  if (b <= a) return;
  double g[3], hs[9];
  poly_grad(f.poly + a, b - a, m, g, hs);
  const double mg = m[0] * g[0] + m[1] * g[1] + m[2] * g[2];
  const double mw = m[0] * w[0] + m[1] * w[1] + m[2] * w[2];
  for (int d = 0; d < 3; d++) {
    const double hw = hs[3 * d] * w[0] + hs[3 * d + 1] * w[1] + hs[3 * d + 2] * w[2];
    const double hm = hs[3 * d] * m[0] + hs[3 * d + 1] * m[1] + hs[3 * d + 2] * m[2];
    av[d] -= s * (hw - mg * w[d] - mw * (g[d] + hm));
  }
}

}  // namespace dev
}  // namespace nsim
