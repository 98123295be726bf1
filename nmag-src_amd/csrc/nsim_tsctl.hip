// Device side of the BDF step controller (nsim_tsctl.h): the start-of-attempt
// vector pass and the one-thread controller kernel that runs after every
// Newton iteration.
#include <hip/hip_runtime.h>

#include "nsim_tsctl.h"
#include "nsim_vec.h"

namespace nsim {

namespace {

__device__ inline const double* src_of(double* const* slots, const double* fn, int s) {
  return s == kSlotF ? fn : (s >= 0 ? slots[s] : nullptr);
}

__global__ __launch_bounds__(256) void k_ts_begin(int n3, const TsState* __restrict__ s,
                                                  double* const* __restrict__ slots,
                                                  const double* __restrict__ fn,
                                                  double* __restrict__ ycur,
                                                  double* __restrict__ yp,
                                                  double* __restrict__ avec,
                                                  double* __restrict__ ewt) {
  // every field of the record and the slot pointers are read before the
  // mode is tested, so their loads are in flight together (the record is
  // always valid to read; the slot indices are always in range or negative)
  const int mode = s->mode, halt = s->halt;
  const int copy = s->copy;
  const double* yn = slots[s->sh[0]];
  double* dst = copy >= 0 ? slots[copy] : nullptr;
  // index 0 of the predictor and of the history term is y_n itself
  const double* P[kHist];
  double L[kHist];
  const double* Av[kMaxOrder];
  double A[kMaxOrder];
#pragma unroll
  for (int j = 0; j < kHist; j++) {
    P[j] = j > 0 ? src_of(slots, fn, s->p[j]) : nullptr;
    L[j] = s->L[j];
  }
#pragma unroll
  for (int j = 0; j < kMaxOrder; j++) {
    Av[j] = j > 0 ? src_of(slots, fn, s->as[j]) : nullptr;
    A[j] = s->A[j];
  }
  const double rtol = s->rtol, atol = s->atol;
  if (mode != kTsBegin || halt) return;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n3; i += gridDim.x * 256) {
    // the accepted iterate becomes y_n (same values as the ring slot gets)
    const double y = dst ? ycur[i] : yn[i];
    if (dst) dst[i] = y;
    ewt[i] = 1.0 / (rtol * fabs(y) + atol);
    double v = L[0] * y;
#pragma unroll
    for (int j = 1; j < kHist; j++)
      if (P[j]) v += L[j] * P[j][i];
    yp[i] = v;
    ycur[i] = v;
    double a = A[0] * y;
#pragma unroll
    for (int j = 1; j < kMaxOrder; j++)
      if (Av[j]) a += A[j] * Av[j][i];
    avec[i] = a;
  }
}

// One thread works on a register copy of the state (read and written back
// whole, so the loads are in flight together rather than one field at a time).
__global__ void k_ts_ctrl(TsState* s, const double* __restrict__ status, int n3) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ts_control(s, status[0], status[1], status[2], status[3], status[4], status[5], n3);
}

}  // namespace

hipError_t launch_ts_begin(hipStream_t st, int n3, const TsState* s, double* const* slots,
                           const double* fn, double* ycur, double* yp, double* avec,
                           double* ewt) {
  k_ts_begin<<<vec_reduce_grid(n3), 256, 0, st>>>(n3, s, slots, fn, ycur, yp, avec, ewt);
  return hipGetLastError();
}

hipError_t launch_ts_ctrl(hipStream_t st, TsState* s, const double* status, int n3) {
  k_ts_ctrl<<<1, 64, 0, st>>>(s, status, n3);
  return hipGetLastError();
}

}  // namespace nsim
