// Kernels of the dense demag potential operator set-up (nsim_dense.cpp):
// scattering sparse rows into dense column-major matrices, symmetrising the
// Cholesky inverse, row gathers and updates, and the final transposing store
// into the row-major storage type read by the grouped GEMV.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nsim_dense.h"
#include "nsim_dev_common.h"

namespace nsim {

namespace {

using namespace dev;

__global__ __launch_bounds__(256) void k_fill(long long count, double v, double* __restrict__ A) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
    A[i] = v;
}

// One thread per local row lr of the magnet's CSR rows [row0, row0 + nr):
// A[r + c lda] += K(row0 + lr, col0 + lc) with r = rmap[lr], c = cmap[lc]
// (identity when a map is null; entries with a negative index are skipped).
__global__ __launch_bounds__(256) void k_scatter_csr(int nr, int row0, int col0,
                                                     const int* __restrict__ rp,
                                                     const int* __restrict__ ci,
                                                     const double* __restrict__ val,
                                                     const int* __restrict__ rmap,
                                                     const int* __restrict__ cmap,
                                                     double* __restrict__ A, long long lda) {
  const int lr = blockIdx.x * 256 + threadIdx.x;
  if (lr >= nr) return;
  const int r = rmap ? rmap[lr] : lr;
  if (r < 0) return;
  for (int k = rp[row0 + lr]; k < rp[row0 + lr + 1]; k++) {
    const int lc = ci[k] - col0;
    const int c = cmap ? cmap[lc] : lc;
    if (c >= 0) A[r + c * lda] += val[k];
  }
}

// Lower triangle -> full symmetric matrix, minus c0 everywhere.
__global__ __launch_bounds__(256) void k_sym_shift(int n, double* __restrict__ A, long long lda,
                                                   double c0) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= n || j >= n || i < j) return;
  const double v = A[i + j * lda] - c0;
  A[i + j * lda] = v;
  A[j + i * lda] = v;
}

__global__ __launch_bounds__(256) void k_gather_rows(int nr, const int* __restrict__ rows,
                                                     const double* __restrict__ A, long long lda,
                                                     double* __restrict__ out, long long ldo) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  const long long j = blockIdx.y;
  if (l >= nr) return;
  out[l + j * ldo] = A[rows[l] + j * lda];
}

__global__ __launch_bounds__(256) void k_add_rows(int nr, const int* __restrict__ rows, double s,
                                                  const double* __restrict__ src, long long lds,
                                                  double* __restrict__ A, long long lda) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  const long long j = blockIdx.y;
  if (l >= nr) return;
  A[rows[l] + j * lda] += s * src[l + j * lds];
}

// out[i ld + j] = A[i + j lda] for i < n, j < ld (zero padding for j >= n),
// through a 32 x 32 LDS tile so both sides are coalesced.
template <typename TB>
__global__ __launch_bounds__(256) void k_store_rowmajor(int n, int ld,
                                                        const double* __restrict__ A,
                                                        long long lda, TB* __restrict__ out) {
  __shared__ double tile[32][33];
  const int j0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int i = i0 + tx, j = j0 + k;
    tile[k][tx] = (i < n && j < n) ? A[i + (long long)j * lda] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int i = i0 + k, j = j0 + tx;
    if (i < n && j < ld) store_b<TB>(out + (long long)i * ld + j, tile[tx][k]);
  }
}

// The same transposing store into 24-bit rows (F24): entry (i, j) is rounded
// and its high 16 and low 8 bits go to the two parts of its row block.
__global__ __launch_bounds__(256) void k_store_rowmajor_f24(int n, int ld,
                                                            const double* __restrict__ A,
                                                            long long lda,
                                                            unsigned char* __restrict__ out) {
  __shared__ double tile[32][33];
  const int j0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int i = i0 + tx, j = j0 + k;
    tile[k][tx] = (i < n && j < n) ? A[i + (long long)j * lda] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int i = i0 + k, j = j0 + tx;
    if (i < n && j < ld) {
      const unsigned u = f24_round(tile[tx][k]);
      size_t hi, lo;
      f24_offsets(j, ld, &hi, &lo);
      unsigned char* row = out + (size_t)i * ld * 3;
      *reinterpret_cast<uint16_t*>(row + hi) = (uint16_t)(u >> 8);
      row[lo] = (unsigned char)(u & 0xffu);
    }
  }
}

inline int blocks(long long n) { return (int)((n + 255) / 256); }

}  // namespace

hipError_t dense_fill(long long count, double v, double* A, hipStream_t st) {
  const int g = (int)std::min<long long>(blocks(count), 65536);
  if (count > 0) k_fill<<<g, 256, 0, st>>>(count, v, A);
  return hipGetLastError();
}

hipError_t dense_scatter_csr(int nr, int row0, int col0, const int* rp, const int* ci,
                             const double* val, const int* rmap, const int* cmap, double* A,
                             long long lda, hipStream_t st) {
  if (nr > 0)
    k_scatter_csr<<<blocks(nr), 256, 0, st>>>(nr, row0, col0, rp, ci, val, rmap, cmap, A, lda);
  return hipGetLastError();
}

hipError_t dense_sym_shift(int n, double* A, long long lda, double c0, hipStream_t st) {
  const int b = (n + 15) / 16;
  if (n > 0) k_sym_shift<<<dim3(b, b), 256, 0, st>>>(n, A, lda, c0);
  return hipGetLastError();
}

hipError_t dense_gather_rows(int nr, int ncols, const int* rows, const double* A, long long lda,
                             double* out, long long ldo, hipStream_t st) {
  if (nr > 0 && ncols > 0)
    k_gather_rows<<<dim3(blocks(nr), ncols), 256, 0, st>>>(nr, rows, A, lda, out, ldo);
  return hipGetLastError();
}

hipError_t dense_add_rows(int nr, int ncols, const int* rows, double s, const double* src,
                          long long lds, double* A, long long lda, hipStream_t st) {
  if (nr > 0 && ncols > 0)
    k_add_rows<<<dim3(blocks(nr), ncols), 256, 0, st>>>(nr, rows, s, src, lds, A, lda);
  return hipGetLastError();
}

hipError_t dense_store_rowmajor(int type, int n, int ld, const double* A, long long lda,
                                void* out, hipStream_t st) {
  const dim3 grid((ld + 31) / 32, (n + 31) / 32);
  if (n <= 0) return hipSuccess;
  if (type == kBemF64)
    k_store_rowmajor<double><<<grid, 256, 0, st>>>(n, ld, A, lda, (double*)out);
  else if (type == kBemF32)
    k_store_rowmajor<float><<<grid, 256, 0, st>>>(n, ld, A, lda, (float*)out);
  else if (type == kBemF24)
    k_store_rowmajor_f24<<<grid, 256, 0, st>>>(n, ld, A, lda, (unsigned char*)out);
  else
    k_store_rowmajor<uint16_t><<<grid, 256, 0, st>>>(n, ld, A, lda, (uint16_t*)out);
  return hipGetLastError();
}

}  // namespace nsim
