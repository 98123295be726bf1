// Device helpers shared by the kernels of the rhs path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsim {
namespace dev {

constexpr double kPi = 3.141592653589793238462643383279502884;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a 256-thread block; every thread gets the same (deterministic) value.
__device__ inline void block_sum2(double& a, double& b, double* sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[2 * w] = a;
    sh[2 * w + 1] = b;
  }
  __syncthreads();
  a = (sh[0] + sh[2]) + (sh[4] + sh[6]);
  b = (sh[1] + sh[3]) + (sh[5] + sh[7]);
}

__device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ inline void cross(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// Agent-coherent accesses for data exchanged between workgroups of one launch
// (gfx950 emits global_load/store ... sc1: they bypass the non-coherent part of
// the per-XCD L2, so no cache write-back or invalidation is needed).
__device__ inline void st_coh(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_coh(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coh(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_coh(const unsigned long long* p) {
  return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coh(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_coh(const unsigned* p) {
  return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One row of the jacobi_ts1 Jacobian-times-vector (nmag_lam.py:432-436,
// src/nsim.ml:1004-1042): with H_total frozen except for its exchange part h,
//   Jv = c1 (v x H + m x h) + c2 (v x (m x H) + m x (v x H) + m x (m x h)).
__device__ inline void llg_jv_row(double c1, double c2, const double* h, const double* m,
                                  const double* v, const double* H, double* out) {
  double vxH[3], mxh[3], mxH[3], vxmxH[3], mxvxH[3], mxmxh[3];
  cross(v, H, vxH);
  cross(m, h, mxh);
  cross(m, H, mxH);
  cross(v, mxH, vxmxH);
  cross(m, vxH, mxvxH);
  cross(m, mxh, mxmxh);
  for (int d = 0; d < 3; d++)
    out[d] = c1 * (vxH[d] + mxh[d]) + c2 * (vxmxH[d] + mxvxH[d] + mxmxh[d]);
}

// FP32 -> BF16 with round-to-nearest-even (NaN kept a quiet NaN).
__device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Stores an FP64 value in a dense operator of storage type TB (FP64, FP32
// or BF16 as uint16_t).
template <typename TB>
__device__ inline void store_b(TB* p, double v);
template <>
__device__ inline void store_b<double>(double* p, double v) {
  *p = v;
}
template <>
__device__ inline void store_b<float>(float* p, double v) {
  *p = (float)v;
}
template <>
__device__ inline void store_b<uint16_t>(uint16_t* p, double v) {
  *p = f32_to_bf16((float)v);
}

// Wait until every memory operation this wave issued has completed.  A
// workgroup barrier alone does not wait for stores on gfx950.
__device__ inline void wait_all_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Row `row` of a SELL-64 operator applied to vector v, accumulated in FP64
// whatever the storage type of the values (FP64 or FP32).
template <typename TV>
__device__ inline double sell_dot(const int* __restrict__ off, const int* __restrict__ col,
                                  const TV* __restrict__ val, const double* __restrict__ v,
                                  int row) {
  const int s = row >> 6, lane = row & 63;
  const int b = off[s], w = (off[s + 1] - b) >> 6;
  double acc = 0.0;
  for (int k = 0; k < w; k++) {
    const int e = b + (k << 6) + lane;
    acc += (double)val[e] * v[col[e]];
  }
  return acc;
}

}  // namespace dev
}  // namespace nsim
