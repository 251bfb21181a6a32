// Batched log-sum-exp and ESS-of-log-weights reductions in FP64 (mocat/src/metrics.py:69-78:
// log ESS = 2 logsumexp(w) - logsumexp(2 w); jax.scipy.special.logsumexp of mocat's
// log-normaliser recursion, mocat/src/transport/smc.py:160,212-215).
// One workgroup per row: a max pass, then sums of exp(w - m) and exp(2 (w - m)) in FP64.
#include "common.hpp"

using namespace mocat;

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void lse_kernel(const float* __restrict__ w, int64_t n, double* __restrict__ lse,
                                                       double* __restrict__ lse2) {
  __shared__ float smax[kThreads / kWave];
  __shared__ double ssum[kThreads / kWave];
  const float* row = w + (int64_t)blockIdx.x * n;
  float m = -INFINITY;
  for (int64_t i = threadIdx.x; i < n; i += kThreads) m = fmaxf(m, row[i]);
  m = block_max(m, smax);
  double s1 = 0.0, s2 = 0.0;
  if (m > -INFINITY && m < INFINITY) {
    for (int64_t i = threadIdx.x; i < n; i += kThreads) {
      const double e = exp((double)row[i] - (double)m);
      s1 += e;
      s2 += e * e;
    }
  }
  s1 = block_sum(s1, ssum);
  s2 = block_sum(s2, ssum);
  if (threadIdx.x == 0) {
    double a, b;
    if (m == -INFINITY) {
      a = -INFINITY;
      b = -INFINITY;
    } else if (m == INFINITY || isnan(m)) {
      a = (double)m;
      b = (double)m;
    } else {
      a = (double)m + log(s1);
      b = 2.0 * (double)m + log(s2);
    }
    if (lse) lse[blockIdx.x] = a;
    if (lse2) lse2[blockIdx.x] = b;
  }
}

}  // namespace

extern "C" int mocat_logsumexp(const float* w, int64_t batch, int64_t n, double* lse, double* lse2,
                               hipStream_t stream) {
  if (batch < 0 || n < 0 || (batch > 0 && !w) || batch > 0x7FFFFFFF) return MOCAT_ERR_ARG;
  if (batch == 0) return MOCAT_OK;
  hipLaunchKernelGGL(lse_kernel, dim3((unsigned)batch), dim3(kThreads), 0, stream, w, n, lse, lse2);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}
