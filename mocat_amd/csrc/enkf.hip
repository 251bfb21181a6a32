// Ensemble Kalman filter: EnsembleKalmanFilter (mocat/src/ssm/nonlinear_gaussian.py:284-350) driven by
// run_particle_filter_for_marginals (mocat/src/ssm/filtering.py:255-324), for the linear transition
// x -> A x of TimeHomogenousLinearGaussian (linear_gaussian.py:112-116):
//   init    (key, sub) = split(key); x0_i = m0 + K0 (y_0 - H m0) + z_i @ C0,
//           z = normal(split(sub, n)[0], (n, D))                              (:312-324)
//   step t  k = split(split(key, T-1)[t-1], n)[0]; (k1, k2) = split(k)        (filtering.py:278, 294)
//           mx_i = A x_i + LQ z1_i, z1 = normal(k1, (n, D))
//           P = sum_i (mx_i - mean)(mx_i - mean)^T / (n - 1)  (spread @ spread.T)
//           K = P H^T (H P H^T + LR LR^T)^-1                                   (utils.py:476-484)
//           x_i' = mx_i + K (y_t - H mx_i - LR z2_i), z2 = normal(k2, (n, DY))  (:326-350)
// The log-weights stay zero (ESS = n: no resampling).
//
// Launches: enkf_init (x_0, the forecast mx_1 and its per-chunk moments), then per step enkf_gain
// (one workgroup: Chan's merge of the chunk means and centred second moments in FP64, K by a Cholesky
// solve) and enkf_update (step t's analysis, then step t+1's forecast and chunk moments). Chunk
// moments are two-pass within the chunk (mean, then centred products of the register-held forecasts),
// so the covariance carries no cancellation from raw second moments. The noise is the bulk of the
// arithmetic (d + dy threefry blocks per member and step); for even n a thread draws the rows i and
// i + n/2 together, which halves it.
#include "common.hpp"

using namespace mocat;

namespace {

constexpr int EB = 256;          // threads per workgroup
constexpr int EW = EB / kWave;   // waves per workgroup
constexpr int EPPT = 4;          // particles per thread
constexpr int ECH = EB * EPPT;   // particles per chunk

// params: m0[D], H[DY*D], K0[D*DY], C0[D*D], A[D*D], LQ[D*D], LR[DY*DY] (reset_covariance's square
// roots and the startup of nonlinear_gaussian.py:297-310, factored on the host)
template <int D, int DY>
struct EnkfModel {
  static constexpr int NPARAM = D + 2 * DY * D + 3 * D * D + DY * DY;
  float m0[D], H[DY * D], K0[D * DY], C0[D * D], A[D * D], LQ[D * D], LR[DY * DY];
  MOCAT_DEV void load(const float* p) {
    auto take = [&](float* dst, int cnt) {
      for (int i = 0; i < cnt; ++i) dst[i] = p[i];
      p += cnt;
    };
    take(m0, D);
    take(H, DY * D);
    take(K0, D * DY);
    take(C0, D * D);
    take(A, D * D);
    take(LQ, D * D);
    take(LR, DY * DY);
  }
};

template <int D>
constexpr int n_moments() {
  return 1 + D + D * (D + 1) / 2;  // count, mean, upper triangle of the centred second moment
}

struct EnkfArgs {
  const float* params;
  const float* y;  // [T][DY]
  const uint32_t* key;
  int64_t n, T, n_chunks;
  float* x;          // [n][D] final ensemble
  float* mx;         // [n][D] forecast of the next step
  double* part;      // [n_chunks][n_moments]
  float* gain;       // [D][DY] gain of the current step
  float* out_value;  // [T][n][D] or NULL
  float* out_gain;   // [T-1][D][DY] or NULL
};

// Particles of a thread. threefry_2x32 over S = n * K counters returns elements m and m + S/2 from one
// call; for even n the element m + S/2 of row i is the same column of row i + n/2, so a thread takes
// the rows b and b + n/2 together (EPPT/2 such pairs) and one call yields both rows' draws. For odd n
// each row is drawn on its own. A chunk holds ECH rows either way.
MOCAT_DEV void slot_rows(int64_t n, int64_t (&idx)[EPPT], bool (&live)[EPPT]) {
  const int64_t c = blockIdx.x;
  if ((n & 1) == 0) {
    const int64_t half = n >> 1;
#pragma unroll
    for (int pk = 0; pk < EPPT; ++pk) {
      const int64_t b = c * (ECH / 2) + (pk >> 1) * EB + threadIdx.x;
      idx[pk] = b + (pk & 1) * half;
      live[pk] = b < half;
    }
  } else {
#pragma unroll
    for (int pk = 0; pk < EPPT; ++pk) {
      idx[pk] = c * ECH + pk * EB + threadIdx.x;
      live[pk] = idx[pk] < n;
    }
  }
}

MOCAT_DEV int64_t chunk_rows(int64_t n) {
  const int64_t c = blockIdx.x;
  if ((n & 1) == 0) {
    const int64_t r = (n >> 1) - c * (ECH / 2);
    return 2 * (r < ECH / 2 ? r : ECH / 2);
  }
  const int64_t r = n - c * ECH;
  return r < ECH ? r : ECH;
}

// Rows idx[] of normal(k, (n, K)) (jax's halves layout over S = n * K elements).
template <int K>
MOCAT_DEV void normal_rows(Key k, int64_t n, const int64_t (&idx)[EPPT], const bool (&live)[EPPT],
                           float (&z)[EPPT][K]) {
  const uint32_t S = (uint32_t)(n * K);
  if ((n & 1) == 0) {
#pragma unroll
    for (int pk = 0; pk < EPPT; pk += 2) {
      if (!live[pk]) continue;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        uint32_t lo, hi;
        bits_pair(k, S, (uint32_t)idx[pk] * (uint32_t)K + (uint32_t)q, lo, hi);
        z[pk][q] = normal_from_bits(lo);
        z[pk + 1][q] = normal_from_bits(hi);
      }
    }
  } else {
#pragma unroll
    for (int pk = 0; pk < EPPT; ++pk) {
      if (!live[pk]) continue;
#pragma unroll
      for (int q = 0; q < K; ++q)
        z[pk][q] = normal_from_bits(bits_at(k, S, (uint32_t)idx[pk] * (uint32_t)K + (uint32_t)q));
    }
  }
}

// Workgroup sum of NV doubles per thread, fixed order; every thread gets the totals.
template <int NV>
MOCAT_DEV void block_sum_vec(double* v, double (*sh)[NV]) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[w][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double r = sh[0][k];
    for (int i = 1; i < EW; ++i) r += sh[i][k];
    v[k] = r;
  }
}

// Step t's noise keys: k1 (forecast) and k2 (observation perturbation).
MOCAT_DEV void step_keys(const uint32_t* key, int64_t T, int64_t n, int64_t t, Key& k1, Key& k2) {
  const Key root{key[0], key[1]};
  const Key ik = split_at(split_at(root, 2, 0), (uint32_t)(T - 1), (uint32_t)(t - 1));
  const Key k = split_at(ik, (uint32_t)n, 0u);
  k1 = split_at(k, 2, 0);
  k2 = split_at(k, 2, 1);
}

// Forecast mx = A x + LQ z1 for the thread's particles, stored, and the chunk's moments.
template <int D, int DY>
MOCAT_DEV void forecast_moments(const EnkfArgs& a, const EnkfModel<D, DY>& M, const float (&x)[EPPT][D],
                                const bool (&live)[EPPT], const int64_t (&idx)[EPPT], Key k1) {
  constexpr int NM = D * (D + 1) / 2;
  __shared__ double shm[EW][NM > D ? NM : D];
  float mx[EPPT][D], z[EPPT][D];
  double s[D];
#pragma unroll
  for (int q = 0; q < D; ++q) s[q] = 0.0;
  normal_rows<D>(k1, a.n, idx, live, z);
#pragma unroll
  for (int pk = 0; pk < EPPT; ++pk) {
    if (!live[pk]) continue;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float u = 0.f, v = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        u += M.A[i * D + k] * x[pk][k];
        v += M.LQ[i * D + k] * z[pk][k];
      }
      mx[pk][i] = u + v;
      a.mx[idx[pk] * D + i] = mx[pk][i];
      s[i] += (double)mx[pk][i];
    }
  }
  const int64_t c = blockIdx.x;
  const int64_t cnt = chunk_rows(a.n);
  block_sum_vec<D>(s, reinterpret_cast<double (*)[D]>(&shm[0][0]));
  double mean[D];
#pragma unroll
  for (int q = 0; q < D; ++q) mean[q] = s[q] / (double)cnt;
  double m2[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) m2[k] = 0.0;
#pragma unroll
  for (int pk = 0; pk < EPPT; ++pk) {
    if (!live[pk]) continue;
    double dv[D];
#pragma unroll
    for (int q = 0; q < D; ++q) dv[q] = (double)mx[pk][q] - mean[q];
    int k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = i; j < D; ++j) m2[k++] += dv[i] * dv[j];
  }
  block_sum_vec<NM>(m2, reinterpret_cast<double (*)[NM]>(&shm[0][0]));
  if (threadIdx.x == 0) {
    double* p = a.part + c * n_moments<D>();
    p[0] = (double)cnt;
#pragma unroll
    for (int q = 0; q < D; ++q) p[1 + q] = mean[q];
#pragma unroll
    for (int k = 0; k < NM; ++k) p[1 + D + k] = m2[k];
  }
}

template <int D, int DY>
__global__ __launch_bounds__(EB) void enkf_init_kernel(EnkfArgs a) {
  __shared__ EnkfModel<D, DY> M;
  __shared__ uint32_t sk[4];
  const int tid = threadIdx.x;
  load_shared_params(M, a.params);
  if (tid == 0) {
    const Key root{a.key[0], a.key[1]};
    const Key k0 = split_at(split_at(root, 2, 1), (uint32_t)a.n, 0u);
    sk[0] = k0.a;
    sk[1] = k0.b;
    if (a.T > 1) {
      Key k1, k2;
      step_keys(a.key, a.T, a.n, 1, k1, k2);
      sk[2] = k1.a;
      sk[3] = k1.b;
    }
  }
  __syncthreads();
  // conditioned initial mean m0 + K0 (y_0 - H m0), the same in every thread
  float mc[D];
  {
    float r[DY];
#pragma unroll
    for (int j = 0; j < DY; ++j) {
      float hm = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) hm += M.H[j * D + k] * M.m0[k];
      r[j] = a.y[j] - hm;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float g = 0.f;
#pragma unroll
      for (int j = 0; j < DY; ++j) g += M.K0[i * DY + j] * r[j];
      mc[i] = M.m0[i] + g;
    }
  }
  float x[EPPT][D], z[EPPT][D];
  bool live[EPPT];
  int64_t idx[EPPT];
  slot_rows(a.n, idx, live);
  normal_rows<D>(Key{sk[0], sk[1]}, a.n, idx, live, z);
#pragma unroll
  for (int pk = 0; pk < EPPT; ++pk) {
    if (!live[pk]) continue;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float c = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) c += z[pk][k] * M.C0[k * D + i];
      x[pk][i] = mc[i] + c;
      if (a.out_value) a.out_value[idx[pk] * D + i] = x[pk][i];
      if (a.T == 1) a.x[idx[pk] * D + i] = x[pk][i];
    }
  }
  if (a.T > 1) forecast_moments<D, DY>(a, M, x, live, idx, Key{sk[2], sk[3]});
}

// One workgroup: the forecast ensemble's covariance from the chunk moments, then the gain.
template <int D, int DY>
__global__ __launch_bounds__(EB) void enkf_gain_kernel(EnkfArgs a, int64_t t) {
  constexpr int NM = D * (D + 1) / 2, NS = n_moments<D>();
  __shared__ double shm[EW][NM > D ? NM : D];
  __shared__ float H[DY * D], LR[DY * DY];
  const int tid = threadIdx.x;
  // H and LR staged by the workgroup (the reductions below synchronise before thread 0 reads them)
  for (int i = tid; i < DY * D; i += EB) H[i] = a.params[D + i];
  for (int i = tid; i < DY * DY; i += EB) LR[i] = a.params[D + 2 * DY * D + 3 * D * D + i];
  double s[D];
#pragma unroll
  for (int q = 0; q < D; ++q) s[q] = 0.0;
  for (int64_t c = tid; c < a.n_chunks; c += EB) {
    const double* p = a.part + c * NS;
#pragma unroll
    for (int q = 0; q < D; ++q) s[q] += p[0] * p[1 + q];
  }
  block_sum_vec<D>(s, reinterpret_cast<double (*)[D]>(&shm[0][0]));
  double mean[D];
#pragma unroll
  for (int q = 0; q < D; ++q) mean[q] = s[q] / (double)a.n;
  double m2[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) m2[k] = 0.0;
  for (int64_t c = tid; c < a.n_chunks; c += EB) {
    const double* p = a.part + c * NS;
    double dv[D];
#pragma unroll
    for (int q = 0; q < D; ++q) dv[q] = p[1 + q] - mean[q];
    int k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = i; j < D; ++j, ++k) m2[k] += p[1 + D + k] + p[0] * dv[i] * dv[j];
  }
  block_sum_vec<NM>(m2, reinterpret_cast<double (*)[NM]>(&shm[0][0]));
  if (tid != 0) return;
  double P[D][D];
  {
    int k = 0;
    for (int i = 0; i < D; ++i)
      for (int j = i; j < D; ++j, ++k) P[i][j] = P[j][i] = m2[k] / (double)(a.n - 1);
  }
  // B = H P (DY x D); Sy = B H^T + LR LR^T
  double B[DY][D], Sy[DY][DY];
  for (int r = 0; r < DY; ++r)
    for (int j = 0; j < D; ++j) {
      double v = 0.0;
      for (int k = 0; k < D; ++k) v += (double)H[r * D + k] * P[k][j];
      B[r][j] = v;
    }
  for (int r = 0; r < DY; ++r)
    for (int s2 = 0; s2 < DY; ++s2) {
      double v = 0.0, w = 0.0;
      for (int k = 0; k < D; ++k) v += B[r][k] * (double)H[s2 * D + k];
      for (int k = 0; k < DY; ++k) w += (double)LR[r * DY + k] * (double)LR[s2 * DY + k];
      Sy[r][s2] = v + w;
    }
  // Cholesky Sy = L L^T, then K^T = Sy^-1 B by two triangular solves
  double L[DY][DY];
  for (int i = 0; i < DY; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = Sy[i][j];
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
      L[i][j] = (i == j) ? sqrt(v) : v / L[j][j];
    }
  for (int j = 0; j < D; ++j) {
    double u[DY];
    for (int i = 0; i < DY; ++i) {
      double v = B[i][j];
      for (int k = 0; k < i; ++k) v -= L[i][k] * u[k];
      u[i] = v / L[i][i];
    }
    for (int i = DY - 1; i >= 0; --i) {
      double v = u[i];
      for (int k = i + 1; k < DY; ++k) v -= L[k][i] * u[k];
      u[i] = v / L[i][i];
    }
    for (int i = 0; i < DY; ++i) {
      a.gain[j * DY + i] = (float)u[i];
      if (a.out_gain) a.out_gain[(t - 1) * D * DY + j * DY + i] = (float)u[i];
    }
  }
}

template <int D, int DY>
__global__ __launch_bounds__(EB) void enkf_update_kernel(EnkfArgs a, int64_t t) {
  __shared__ EnkfModel<D, DY> M;
  __shared__ float K[D * DY];
  __shared__ uint32_t sk[4];
  const int tid = threadIdx.x;
  const bool last = t + 1 >= a.T;
  load_shared_params(M, a.params);
  if (tid == 0) {
    Key k1, k2;
    step_keys(a.key, a.T, a.n, t, k1, k2);
    sk[0] = k2.a;
    sk[1] = k2.b;
    if (!last) {
      step_keys(a.key, a.T, a.n, t + 1, k1, k2);
      sk[2] = k1.a;
      sk[3] = k1.b;
    }
  }
  if (tid < D * DY) K[tid] = a.gain[tid];
  __syncthreads();
  const float* yt = a.y + t * DY;
  float x[EPPT][D], z[EPPT][DY];
  bool live[EPPT];
  int64_t idx[EPPT];
  slot_rows(a.n, idx, live);
  float mxa[EPPT][D];
#pragma unroll
  for (int pk = 0; pk < EPPT; ++pk)
#pragma unroll
    for (int q = 0; q < D; ++q) mxa[pk][q] = live[pk] ? a.mx[idx[pk] * D + q] : 0.f;
  normal_rows<DY>(Key{sk[0], sk[1]}, a.n, idx, live, z);
#pragma unroll
  for (int pk = 0; pk < EPPT; ++pk) {
    if (!live[pk]) continue;
    const float* mx = mxa[pk];
    float r[DY];
#pragma unroll
    for (int j = 0; j < DY; ++j) {
      float hx = 0.f, e = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) hx += M.H[j * D + k] * mx[k];
#pragma unroll
      for (int k = 0; k < DY; ++k) e += M.LR[j * DY + k] * z[pk][k];
      r[j] = yt[j] - (hx + e);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float g = 0.f;
#pragma unroll
      for (int j = 0; j < DY; ++j) g += K[i * DY + j] * r[j];
      x[pk][i] = mx[i] + g;
      if (a.out_value) a.out_value[(t * a.n + idx[pk]) * D + i] = x[pk][i];
      if (last) a.x[idx[pk] * D + i] = x[pk][i];
    }
  }
  if (!last) forecast_moments<D, DY>(a, M, x, live, idx, Key{sk[2], sk[3]});
}

struct EnkfLayout {
  int64_t o_mx, o_part, o_gain, bytes;
};

EnkfLayout enkf_layout(int64_t n, int d) {
  auto al = [](int64_t v) { return (v + 255) / 256 * 256; };
  const int64_t nch = (n + ECH - 1) / ECH;
  EnkfLayout L;
  L.o_mx = 0;
  L.o_part = al(n * d * 4);
  L.o_gain = al(L.o_part + nch * (1 + d + d * (d + 1) / 2) * 8);
  L.bytes = al(L.o_gain + 64 * 4);
  return L;
}

template <int D, int DY>
int launch_enkf(const EnkfArgs& a, hipStream_t st) {
  const dim3 grid((unsigned)a.n_chunks), blk(EB);
  hipLaunchKernelGGL((enkf_init_kernel<D, DY>), grid, blk, 0, st, a);
  for (int64_t t = 1; t < a.T; ++t) {
    hipLaunchKernelGGL((enkf_gain_kernel<D, DY>), dim3(1), blk, 0, st, a, t);
    hipLaunchKernelGGL((enkf_update_kernel<D, DY>), grid, blk, 0, st, a, t);
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace

extern "C" {

int mocat_enkf_param_count(int d, int dy, int64_t* count) {
  if (!count || d < 1 || d > 8 || dy < 1 || dy > 8) return MOCAT_ERR_ARG;
  *count = d + 2 * dy * d + 3 * d * d + dy * dy;
  return MOCAT_OK;
}

int mocat_enkf_workspace_bytes(int64_t n, int d, int64_t* bytes) {
  if (!bytes || n < 2 || d < 1 || d > 8) return MOCAT_ERR_ARG;
  *bytes = enkf_layout(n, d).bytes;
  return MOCAT_OK;
}

int mocat_enkf_run(int d, int dy, const float* params, const float* y, int64_t T, int64_t n, const uint32_t* key,
                   float* x_out, void* workspace, int64_t workspace_bytes, float* out_value, float* out_gain,
                   hipStream_t stream) {
  if (!params || !y || !key || !x_out || !workspace || T < 1 || n < 2 || d < 1 || d > 8 || dy < 1 || dy > 8)
    return MOCAT_ERR_ARG;
  if (n > 0x7FFFFFFFll || n * (int64_t)(d > dy ? d : dy) > 0xFFFFFFFFll || T - 1 > 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  const EnkfLayout L = enkf_layout(n, d);
  if (workspace_bytes < L.bytes) return MOCAT_ERR_SIZE;
  char* ws = static_cast<char*>(workspace);
  EnkfArgs a{params, y, key, n, T, (n + ECH - 1) / ECH, x_out, reinterpret_cast<float*>(ws + L.o_mx),
             reinterpret_cast<double*>(ws + L.o_part), reinterpret_cast<float*>(ws + L.o_gain), out_value, out_gain};
#define MOCAT_ENKF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_enkf<DD, YY>(a, stream);
  MOCAT_ENKF_CASE(1, 1) MOCAT_ENKF_CASE(2, 2) MOCAT_ENKF_CASE(4, 4) MOCAT_ENKF_CASE(4, 2) MOCAT_ENKF_CASE(5, 5)
  MOCAT_ENKF_CASE(8, 8) MOCAT_ENKF_CASE(2, 1) MOCAT_ENKF_CASE(3, 3) MOCAT_ENKF_CASE(4, 1) MOCAT_ENKF_CASE(8, 4)
#undef MOCAT_ENKF_CASE
  return MOCAT_ERR_ARG;
}

}  // extern "C"
