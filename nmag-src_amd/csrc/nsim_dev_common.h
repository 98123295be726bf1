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

// LDS is given to workgroups in whole allocation granules.  gfx950 has
// 160 KiB of LDS per CU; its granule is taken as 2 KiB, the largest of the
// granules in use on the CDNA parts, so a residency count made with it never
// exceeds what the CU holds.  The persistent kernels size their grids from
// this count: a workgroup that cannot start would leave the others waiting.
constexpr int kLdsPerCu = 160 * 1024;
constexpr int kLdsGranule = 2048;
__host__ __device__ constexpr int lds_granules(int bytes) {
  return (bytes + kLdsGranule - 1) / kLdsGranule;
}

// Phase tags of the grid-wide exchanges sit one per 128-byte line: every
// workgroup polls all of them, and packed tags put all pollers on a few
// lines (measured on MI355X: about 1 us more per exchange at 312 workgroups).
constexpr int kTagStride = 16;  // unsigned long long words per tag
__host__ __device__ inline size_t tag_slot(int par, int gmax, int b) {
  return ((size_t)par * gmax + b) * kTagStride;
}

// Raw buffer resources for device-coherent loads that the compiler may batch
// (unlike atomic loads, which it issues one at a time).
constexpr int kBufferConfig = 0x00020000;  // raw buffer, gfx9 family
constexpr int kSc1 = 16;                   // cache policy: device-coherent

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           kBufferConfig);
}

// Device-coherent load of element `idx` (out-of-range indices read 0).
__device__ inline double load_coh(__amdgpu_buffer_rsrc_t r, int idx) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, kSc1));
}

// Device-coherent load of the 64-bit word `idx` (a phase tag).
__device__ inline unsigned long long load_tag(__amdgpu_buffer_rsrc_t r, size_t idx) {
  return __builtin_bit_cast(unsigned long long,
                            __builtin_amdgcn_raw_buffer_load_b64(r, (int)(idx * 8), 0, kSc1));
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

// Block-Jacobi preconditioner of one node: M = I - gamma J_ii, J_ii the 3x3
// diagonal block of the linearisation above (H frozen, dH_i/dm_i = kappa I,
// kappa = negJ K_ii / V_i); writes M^-1 row-major to o[9].
__device__ inline void bjac_block(double c1, double c2, double gamma, double kappa,
                                  const double* m, const double* H, double* o) {
  double M[3][3];
  for (int col = 0; col < 3; col++) {
    double v[3] = {0.0, 0.0, 0.0};
    v[col] = 1.0;
    const double h[3] = {kappa * v[0], kappa * v[1], kappa * v[2]};
    double J[3];
    llg_jv_row(c1, c2, h, m, v, H, J);
    for (int r = 0; r < 3; r++) M[r][col] = (r == col ? 1.0 : 0.0) - gamma * J[r];
  }
  const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                     M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  const double id = 1.0 / det;
  o[0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * id;
  o[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
  o[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
  o[3] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * id;
  o[4] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
  o[5] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
  o[6] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * id;
  o[7] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
  o[8] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
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

// 24-bit storage of a row-major dense operator: each entry is the FP32 value
// with its low 8 bits rounded off to nearest (ties to even), keeping FP32's
// sign and exponent and 15 stored significand bits (16-bit precision, unit
// roundoff 2^-16).  A row of ld entries (ld a
// multiple of 64) takes 3 ld bytes and is split into blocks of up to 512
// entries; a block of nb entries holds the entries' high 16 bits (2 nb
// bytes) followed by their low 8 bits (nb bytes), so that eight consecutive
// entries are one 16-byte and one 8-byte aligned load.
struct F24 {
  unsigned char b[3];
};
constexpr int kF24Block = 512;

typedef unsigned int f24_u4 __attribute__((ext_vector_type(4)));
typedef unsigned int f24_u2 __attribute__((ext_vector_type(2)));

__device__ inline unsigned f24_round(double v) {
  const unsigned u = __float_as_uint((float)v);
  return (u + 0x7fu + ((u >> 8) & 1u)) >> 8;
}

// Entry j of a 24-bit row: byte offsets of its high and low parts.
__device__ inline void f24_offsets(int j, int ld, size_t* hi, size_t* lo) {
  const int j0 = j & ~(kF24Block - 1);
  const int nb = min(kF24Block, ld - j0);
  *hi = 3 * (size_t)j0 + 2 * (size_t)(j - j0);
  *lo = 3 * (size_t)j0 + 2 * (size_t)nb + (size_t)(j - j0);
}

// Entries j .. j+7 (j a multiple of 8) of a 24-bit row of length ld, read
// once (non-temporal) and widened to T.
template <typename T>
__device__ inline void f24_load8(const F24* row, int j, int ld, T* out) {
  size_t hi, lo;
  f24_offsets(j, ld, &hi, &lo);
  const unsigned char* base = reinterpret_cast<const unsigned char*>(row);
  const f24_u4 h = __builtin_nontemporal_load(reinterpret_cast<const f24_u4*>(base + hi));
  const f24_u2 l = __builtin_nontemporal_load(reinterpret_cast<const f24_u2*>(base + lo));
  const unsigned hw[4] = {h.x, h.y, h.z, h.w};
  const unsigned lw[2] = {l.x, l.y};
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const unsigned hi16 = (hw[e >> 1] >> (16 * (e & 1))) & 0xffffu;
    const unsigned lo8 = (lw[e >> 2] >> (8 * (e & 3))) & 0xffu;
    out[e] = (T)__uint_as_float((hi16 << 16) | (lo8 << 8));
  }
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
