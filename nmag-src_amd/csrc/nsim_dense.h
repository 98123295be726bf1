// Launchers of the dense potential operator set-up kernels (nsim_dense.hip).
// Dense matrices are FP64 and column-major (the layout of rocBLAS/rocSOLVER).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "nsim_device.h"

namespace nsim {

hipError_t dense_fill(long long count, double v, double* A, hipStream_t st);
// A[r + c lda] += K(row0 + lr, col0 + lc) over the CSR rows [row0, row0 + nr)
// with r = rmap[lr], c = cmap[lc] (identity if null, skipped if negative)
hipError_t dense_scatter_csr(int nr, int row0, int col0, const int* rp, const int* ci,
                             const double* val, const int* rmap, const int* cmap, double* A,
                             long long lda, hipStream_t st);
// lower triangle -> symmetric, then subtract c0 from every entry
hipError_t dense_sym_shift(int n, double* A, long long lda, double c0, hipStream_t st);
// out(l, :) = A(rows[l], :)
hipError_t dense_gather_rows(int nr, int ncols, const int* rows, const double* A, long long lda,
                             double* out, long long ldo, hipStream_t st);
// A(rows[l], :) += s src(l, :)
hipError_t dense_add_rows(int nr, int ncols, const int* rows, double s, const double* src,
                          long long lds, double* A, long long lda, hipStream_t st);
// row-major copy with stride ld (zero padded) in storage type `type`
hipError_t dense_store_rowmajor(int type, int n, int ld, const double* A, long long lda,
                                void* out, hipStream_t st);

}  // namespace nsim
