// Forward-filtering backward-simulation, full resampling (mocat/src/ssm/backward.py:20-39, 210-243).
//
// One backward step, for every trajectory m of M: key_m = split(key_t, M)[m] and
//   x_t[m] = particles_t[argmax_j(gumbel(key_m, (N,))[j] + log_weight_t[j] - U(particles_t[j] -> x_{t+1}[m]))]
// (random.categorical: the first maximum wins), U the model's transition potential as mocat's
// gaussian_potential evaluates it (utils.py:49-105): 0.5 |(x' - A x) @ S|^2 + nlz, S the transition
// precision square root. One wave per trajectory: the lanes take pairs (j, j + N/2), whose Gumbel bits
// come out of one threefry block (the halves layout), and keep a running maximum; a butterfly with
// the lower index winning ties combines them. Work per step: N x M potentials and N x M / 2 threefry
// blocks.
#include <cmath>

#include "common.hpp"

using namespace mocat;

namespace {

// linear Gaussian: params A[D*D], S[D*D] (transition precision square root), nlz
template <int D>
struct LgTrans {
  float A[D * D], S[D * D], nlz;
  MOCAT_DEV void load(const float* p, int /*d: D*/) {
#pragma unroll
    for (int i = 0; i < D * D; ++i) A[i] = p[i];
#pragma unroll
    for (int i = 0; i < D * D; ++i) S[i] = p[D * D + i];
    nlz = p[2 * D * D];
  }
  MOCAT_DEV float pot(const float* xp, const float* xn) const {
    float diff[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) a += A[i * D + k] * xp[k];
      diff[i] = xn[i] - a;
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      float v = 0.f;
#pragma unroll
      for (int i = 0; i < D; ++i) v += diff[i] * S[i * D + k];
      s += v * v;
    }
    return 0.5f * s + nlz;
  }
};

// stochastic volatility, unit steps (stochastic_volatility.py:78-84): params phi, 1/sigma, nlz
struct SvTrans {
  float phi, isig, nlz;
  MOCAT_DEV void load(const float* p, int /*d: 1*/) {
    phi = p[0];
    isig = p[1];
    nlz = p[2];
  }
  MOCAT_DEV float pot(const float* xp, const float* xn) const {
    const float diff = xn[0] - xp[0] * phi;
    return 0.5f * (diff * diff) * (isig * isig) + nlz;
  }
};

// linear Gaussian at a runtime d <= kRtMaxDim (the kernels' template D = 0): A and S read from params,
// the difference vector in the calling lane's LDS row, the products in LgTrans's order
constexpr int kRtMaxDim = 16;

struct LgTransRt {
  const float *A, *S;
  float nlz;
  int d;
  MOCAT_DEV void load(const float* p, int dd) {
    d = dd;
    A = p;
    S = p + dd * dd;
    nlz = p[2 * dd * dd];
  }
  MOCAT_DEV float pot(const float* xp, const float* xn) const {
    __shared__ float DF[kRtMaxDim][256];
    const int t = threadIdx.x;
    for (int i = 0; i < d; ++i) {
      float a = 0.f;
      for (int k = 0; k < d; ++k) a += A[i * d + k] * xp[k];
      DF[i][t] = xn[i] - a;
    }
    float s = 0.f;
    for (int k = 0; k < d; ++k) {
      float v = 0.f;
      for (int i = 0; i < d; ++i) v += DF[i][t] * S[i * d + k];
      s += v * v;
    }
    return 0.5f * s + nlz;
  }
};

constexpr float kTiny = 1.17549435e-38f;  // float32 tiny: the lower end of jax's gumbel uniforms

// argmax_j(gumbel_j + logit(j)) over j < n for one wave (the first maximum): the Gumbel draws are
// elements off + j of a bits array of size S under key k (off = 0, S = n: gumbel(k, (n,)); off = m n,
// S = M n: row m of gumbel(k, (M, n))). When off = 0 and S = n, the lanes take pairs (j, j + n/2),
// whose bits come out of one threefry block.
template <class Logit>
MOCAT_DEV int64_t gumbel_argmax_wave(Key k, uint32_t S, uint64_t off, int64_t n, int lane, Logit logit) {
  float best = -INFINITY;
  int64_t bi = n;  // no candidate yet
  auto consider = [&](int64_t j, uint32_t bits) {
    const float u = uniform_from_bits(bits, kTiny, 1.f);
    const float v = -logf(-logf(u)) + logit(j);
    if (bi == n || v > best || (v == best && j < bi)) {
      best = v;
      bi = j;
    }
  };
  if (off == 0 && (uint64_t)S == (uint64_t)n) {
    const uint32_t h = (S + 1u) >> 1;
    for (int64_t j = lane; j < h; j += kWave) {
      uint32_t w0, w1;
      bits_pair(k, S, (uint32_t)j, w0, w1);
      consider(j, w0);
      if (j + h < n) consider(j + h, w1);
    }
  } else {
    for (int64_t j = lane; j < n; j += kWave) consider(j, bits_at(k, S, (uint32_t)(off + j)));
  }
  // the maximum over the wave, the lower index on ties (first maximum)
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o, kWave);
    const int64_t oi = (int64_t)__shfl_xor((long long)bi, o, kWave);
    if (oi != n && (bi == n || ob > best || (ob == best && oi < bi))) {
      best = ob;
      bi = oi;
    }
  }
  return bi == n ? 0 : bi;  // n >= 1: lane 0 always ends with a candidate
}

template <class Tr, int D>
__global__ __launch_bounds__(256) void ffbsi_step_kernel(const float* __restrict__ params,
                                                         const float* __restrict__ xs, const float* __restrict__ lw,
                                                         int64_t n, const float* __restrict__ xnext, int64_t m_count,
                                                         const uint32_t* __restrict__ key, int32_t* __restrict__ idx,
                                                         float* __restrict__ xout) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= m_count) return;
  Tr tr;
  tr.load(params, D);
  const Key km = split_at(Key{key[0], key[1]}, (uint32_t)m_count, (uint32_t)m);
  float xn[D];
#pragma unroll
  for (int q = 0; q < D; ++q) xn[q] = xnext[m * D + q];
  const int64_t bi = gumbel_argmax_wave(km, (uint32_t)n, 0, n, lane,
                                        [&](int64_t j) { return lw[j] - tr.pot(xs + j * D, xn); });
  if (lane == 0 && idx) idx[m] = (int32_t)bi;
  if (lane < D) xout[m * D + lane] = xs[bi * D + lane];
}

template <class Tr, int D>
int launch_ffbsi(const float* params, const float* xs, const float* lw, int64_t n, const float* xnext, int64_t m,
                 const uint32_t* key, int32_t* idx, float* xout, hipStream_t stream) {
  const int per = 256 / kWave;
  hipLaunchKernelGGL((ffbsi_step_kernel<Tr, D>), dim3((unsigned)((m + per - 1) / per)), dim3(256), 0, stream, params,
                     xs, lw, n, xnext, m, key, idx, xout);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// The full backward step for the linear-Gaussian model at a d without a compiled kernel (d <= kRtMaxDim):
// the same draw and potential (0.5 |(x' - A x) @ S|^2 + nlz, the products in LgTrans's order) with A and
// S in LDS, each wave's x_{t+1} in LDS, and each lane's particle state and difference vector in its own
// LDS rows (indexed at run time).

__global__ __launch_bounds__(256) void ffbsi_step_rt_kernel(int d, const float* __restrict__ params,
                                                            const float* __restrict__ xs,
                                                            const float* __restrict__ lw, int64_t n,
                                                            const float* __restrict__ xnext, int64_t m_count,
                                                            const uint32_t* __restrict__ key,
                                                            int32_t* __restrict__ idx, float* __restrict__ xout) {
  __shared__ float A[kRtMaxDim * kRtMaxDim], S[kRtMaxDim * kRtMaxDim];
  __shared__ float XN[256 / kWave][kRtMaxDim];
  __shared__ float XP[kRtMaxDim][256], DF[kRtMaxDim][256];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  for (int k = t; k < d * d; k += 256) {
    A[k] = params[k];
    S[k] = params[d * d + k];
  }
  const float nlz = params[2 * d * d];
  const int64_t m = (int64_t)blockIdx.x * (256 / kWave) + w;
  if (lane < d && m < m_count) XN[w][lane] = xnext[m * d + lane];
  __syncthreads();
  if (m >= m_count) return;
  const Key km = split_at(Key{key[0], key[1]}, (uint32_t)m_count, (uint32_t)m);
  auto pot = [&](int64_t j) {
    for (int k = 0; k < d; ++k) XP[k][t] = xs[j * d + k];
    for (int i = 0; i < d; ++i) {
      float a = 0.f;
      for (int k = 0; k < d; ++k) a += A[i * d + k] * XP[k][t];
      DF[i][t] = XN[w][i] - a;
    }
    float s = 0.f;
    for (int k = 0; k < d; ++k) {
      float v = 0.f;
      for (int i = 0; i < d; ++i) v += DF[i][t] * S[i * d + k];
      s += v * v;
    }
    return 0.5f * s + nlz;
  };
  const int64_t bi = gumbel_argmax_wave(km, (uint32_t)n, 0, n, lane, [&](int64_t j) { return lw[j] - pot(j); });
  if (lane == 0 && idx) idx[m] = (int32_t)bi;
  if (lane < d) xout[m * d + lane] = xs[bi * d + lane];
}

// ---------------------------------------------------------------- rejection backward simulation
// mocat/src/ssm/backward.py:42-194 for one backward step (rejection_resampling): initial proposals
// x0[m] = categorical(k0, lw, shape=(M,))[m] with (k0, k1, k2) = split(key_t, 3), densities
// p[m] = exp(-U(x0[m] -> x1[m])), bound = max p * inflation if max p > the initial bound parameter
// (else that parameter), not_yet[m] = uniform(k1, (M,))[m] > p[m] / bound, keys[m] = split(k2, M)[m];
// then rounds while some trajectory waits and fewer than maximum_rejections rounds ran: a waiting
// trajectory splits its key into (key, choice, uniform), proposes categorical(choice, lw), and stays
// waiting iff uniform(uniform) > p / bound; a round whose largest density exceeds the bound raises it
// to that density times the inflation and sends every trajectory back to waiting. Trajectories still
// waiting at the end are drawn by full resampling with their keys. Each round is one launch over the
// trajectories (one wave each) and one single-workgroup launch for the shared bound and counts.
struct RejState {
  float bound;
  int reset;                 // the last round raised the bound: every trajectory proposes again
  int any;                   // some trajectory waits for an acceptance
  int pad;
  unsigned long long evals;  // transition-density evaluations of this step
};

struct RejBuffers {
  uint32_t* keys;   // [M][2]
  int32_t* wait;    // [M] not yet accepted
  int32_t* entered; // [M] proposed in the current round
  float* dens;      // [M] density of the current round's proposal (0 when none)
  RejState* st;
};

// D = 0: the runtime dimension rt_d (LgTransRt); otherwise rt_d is unused
template <class Tr, int D>
__global__ __launch_bounds__(256) void rej_init_kernel(const float* __restrict__ params, const float* __restrict__ xs,
                                                       const float* __restrict__ lw, int64_t n,
                                                       const float* __restrict__ xnext, int64_t M,
                                                       const uint32_t* __restrict__ key, int32_t* __restrict__ idx,
                                                       RejBuffers b, int rt_d) {
  const int64_t DD = D > 0 ? D : rt_d;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= M) return;
  Tr tr;
  tr.load(params, (int)DD);
  const Key kt{key[0], key[1]};
  const int64_t bi = gumbel_argmax_wave(split_at(kt, 3, 0), (uint32_t)(M * n), (uint64_t)m * (uint64_t)n, n, lane,
                                        [&](int64_t j) { return lw[j]; });
  if (lane == 0) {
    idx[m] = (int32_t)bi;
    b.dens[m] = expf(-tr.pot(xs + bi * DD, xnext + m * DD));
  }
}

// mode 0: the initial bound from the initial densities, the step's evaluation count set to M;
// mode 1: end of a round (raise the bound if exceeded, count the round's proposals);
// mode 2: end of the step: add n per trajectory resampled in full and write the count out.
__global__ __launch_bounds__(1024) void rej_reduce_kernel(int mode, int64_t M, int64_t n, float init_bound,
                                                          float inflation, RejBuffers b, long long* evals_out) {
  __shared__ float smax[16];
  __shared__ long long scnt[2][16];
  RejState& st = *b.st;
  if (mode == 1 && !st.any) return;
  float mx = 0.f;
  long long c0 = 0, c1 = 0;
  for (int64_t m = threadIdx.x; m < M; m += blockDim.x) {
    mx = fmaxf(mx, b.dens[m]);
    c0 += b.entered[m];
    c1 += b.wait[m];
  }
  mx = wave_max(mx);
  c0 = wave_sum(c0);
  c1 = wave_sum(c1);
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    smax[w] = mx;
    scnt[0][w] = c0;
    scnt[1][w] = c1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  mx = smax[0];
  c0 = scnt[0][0];
  c1 = scnt[1][0];
  for (int q = 1; q < nw; ++q) {
    mx = fmaxf(mx, smax[q]);
    c0 += scnt[0][q];
    c1 += scnt[1][q];
  }
  if (mode == 0) {
    st.bound = mx > init_bound ? mx * inflation : init_bound;
    st.evals = (unsigned long long)M;
    st.reset = 0;
    st.any = 1;  // decided after the acceptance flags (rej_flags_kernel)
  } else if (mode == 1) {
    const bool reset = mx > st.bound;
    if (reset) st.bound = mx * inflation;
    st.evals += (unsigned long long)c0;
    st.reset = reset ? 1 : 0;
    st.any = (reset || c1 > 0) ? 1 : 0;
  } else {
    st.evals += (unsigned long long)n * (unsigned long long)c0;
    *evals_out = (long long)st.evals;
  }
}

// acceptance flags of the initial proposals and the trajectories' keys; the waiting count in entered
__global__ __launch_bounds__(256) void rej_flags_kernel(int64_t M, const uint32_t* __restrict__ key, RejBuffers b) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const Key kt{key[0], key[1]};
  const float u = uniform_from_bits(bits_at(split_at(kt, 3, 1), (uint32_t)M, (uint32_t)m), 0.f, 1.f);
  const int wait = u > b.dens[m] / b.st->bound ? 1 : 0;
  b.wait[m] = wait;
  b.entered[m] = wait;
  const Key km = split_at(split_at(kt, 3, 2), (uint32_t)M, (uint32_t)m);
  b.keys[2 * m] = km.a;
  b.keys[2 * m + 1] = km.b;
  b.dens[m] = 0.f;
}

// whether any trajectory waits after the initial flags (the first round's condition)
__global__ __launch_bounds__(1024) void rej_any_kernel(int64_t M, RejBuffers b) {
  __shared__ int sany;
  if (threadIdx.x == 0) sany = 0;
  __syncthreads();
  int a = 0;
  for (int64_t m = threadIdx.x; m < M; m += blockDim.x) a |= b.wait[m];
  if (a) sany = 1;
  __syncthreads();
  if (threadIdx.x == 0) b.st->any = sany;
}

template <class Tr, int D>
__global__ __launch_bounds__(256) void rej_round_kernel(const float* __restrict__ params, const float* __restrict__ xs,
                                                        const float* __restrict__ lw, int64_t n,
                                                        const float* __restrict__ xnext, int64_t M,
                                                        int32_t* __restrict__ idx, RejBuffers b, int rt_d) {
  const int64_t DD = D > 0 ? D : rt_d;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= M || !b.st->any) return;
  const bool wait = b.st->reset || b.wait[m];
  if (!wait) {
    if (lane == 0) {
      b.entered[m] = 0;
      b.dens[m] = 0.f;
    }
    return;
  }
  Tr tr;
  tr.load(params, (int)DD);
  const Key k{b.keys[2 * m], b.keys[2 * m + 1]};
  const Key kr = split_at(k, 3, 0), kc = split_at(k, 3, 1), ku = split_at(k, 3, 2);
  const int64_t bi = gumbel_argmax_wave(kc, (uint32_t)n, 0, n, lane, [&](int64_t j) { return lw[j]; });
  if (lane == 0) {
    const float dd = expf(-tr.pot(xs + bi * DD, xnext + m * DD));
    const float u = uniform_from_bits(bits_at(ku, 1, 0), 0.f, 1.f);
    idx[m] = (int32_t)bi;
    b.dens[m] = dd;
    b.wait[m] = u > dd / b.st->bound ? 1 : 0;
    b.entered[m] = 1;
    b.keys[2 * m] = kr.a;
    b.keys[2 * m + 1] = kr.b;
  }
}

// trajectories still waiting: full resampling with their own keys; then the rows of the step
template <class Tr, int D>
__global__ __launch_bounds__(256) void rej_final_kernel(const float* __restrict__ params, const float* __restrict__ xs,
                                                        const float* __restrict__ lw, int64_t n,
                                                        const float* __restrict__ xnext, int64_t M,
                                                        int32_t* __restrict__ idx, float* __restrict__ xout,
                                                        RejBuffers b, int rt_d) {
  const int64_t DD = D > 0 ? D : rt_d;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= M) return;
  const bool wait = b.st->any && (b.st->reset || b.wait[m]);
  int64_t bi;
  if (wait) {
    Tr tr;
    tr.load(params, (int)DD);
    const Key k{b.keys[2 * m], b.keys[2 * m + 1]};
    if constexpr (D > 0) {
      float xn[D];
#pragma unroll
      for (int q = 0; q < D; ++q) xn[q] = xnext[m * D + q];
      bi = gumbel_argmax_wave(k, (uint32_t)n, 0, n, lane, [&](int64_t j) { return lw[j] - tr.pot(xs + j * D, xn); });
    } else {
      const float* xn = xnext + m * DD;
      bi = gumbel_argmax_wave(k, (uint32_t)n, 0, n, lane, [&](int64_t j) { return lw[j] - tr.pot(xs + j * DD, xn); });
    }
    if (lane == 0) idx[m] = (int32_t)bi;
  } else {
    bi = idx[m];
  }
  if (lane == 0) b.entered[m] = wait ? 1 : 0;
  for (int64_t q = lane; q < DD; q += kWave) xout[m * DD + q] = xs[bi * DD + q];
}

template <class Tr, int D>
int launch_rejection(const float* params, const float* xs, const float* lw, int64_t n, const float* xnext, int64_t M,
                     const uint32_t* key, int max_rejections, float init_bound, float inflation, RejBuffers b,
                     int32_t* idx, float* xout, long long* evals_out, hipStream_t st, int rt_d = D) {
  const int per = 256 / kWave;
  const dim3 waves((unsigned)((M + per - 1) / per)), thr((unsigned)((M + 255) / 256));
  hipLaunchKernelGGL((rej_init_kernel<Tr, D>), waves, dim3(256), 0, st, params, xs, lw, n, xnext, M, key, idx, b,
                     rt_d);
  hipLaunchKernelGGL(rej_reduce_kernel, dim3(1), dim3(1024), 0, st, 0, M, n, init_bound, inflation, b, evals_out);
  hipLaunchKernelGGL(rej_flags_kernel, thr, dim3(256), 0, st, M, key, b);
  hipLaunchKernelGGL(rej_any_kernel, dim3(1), dim3(1024), 0, st, M, b);
  for (int r = 1; r < max_rejections; ++r) {
    hipLaunchKernelGGL((rej_round_kernel<Tr, D>), waves, dim3(256), 0, st, params, xs, lw, n, xnext, M, idx, b, rt_d);
    hipLaunchKernelGGL(rej_reduce_kernel, dim3(1), dim3(1024), 0, st, 1, M, n, init_bound, inflation, b, evals_out);
  }
  hipLaunchKernelGGL((rej_final_kernel<Tr, D>), waves, dim3(256), 0, st, params, xs, lw, n, xnext, M, idx, xout, b,
                     rt_d);
  hipLaunchKernelGGL(rej_reduce_kernel, dim3(1), dim3(1024), 0, st, 2, M, n, init_bound, inflation, b, evals_out);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int64_t rej_layout(int64_t M, RejBuffers* b, char* base) {
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = o;
    o = (o + bytes + 255) / 256 * 256;
    return base ? base + at : nullptr;
  };
  char* keys = take(M * 8);
  char* wait = take(M * 4);
  char* entered = take(M * 4);
  char* dens = take(M * 4);
  char* st = take(sizeof(RejState));
  if (b) *b = RejBuffers{reinterpret_cast<uint32_t*>(keys), reinterpret_cast<int32_t*>(wait),
                         reinterpret_cast<int32_t*>(entered), reinterpret_cast<float*>(dens),
                         reinterpret_cast<RejState*>(st)};
  return o;
}

}  // namespace

extern "C" {

int mocat_ffbsi_rejection_workspace_bytes(int64_t m, int64_t* bytes) {
  if (m <= 0 || !bytes) return MOCAT_ERR_ARG;
  *bytes = rej_layout(m, nullptr, nullptr);
  return MOCAT_OK;
}

int mocat_ffbsi_rejection_step(int model, int d, const float* params, const float* particles, const float* log_weight,
                               int64_t n, const float* x_next, int64_t m, const uint32_t* key, int max_rejections,
                               float init_bound, float bound_inflation, void* workspace, int64_t workspace_bytes,
                               int32_t* idx, float* x_out, long long* evals, hipStream_t stream) {
  if (!params || !particles || !log_weight || !x_next || !key || !workspace || !idx || !x_out || !evals || n <= 0 ||
      m <= 0 || n * m >= 0xFFFFFFFFll || max_rejections < 1)
    return MOCAT_ERR_ARG;
  if (workspace_bytes < rej_layout(m, nullptr, nullptr)) return MOCAT_ERR_SIZE;
  RejBuffers b;
  rej_layout(m, &b, static_cast<char*>(workspace));
#define MOCAT_REJ(TR, DD)                                                                                          \
  return launch_rejection<TR, DD>(params, particles, log_weight, n, x_next, m, key, max_rejections, init_bound,    \
                                  bound_inflation, b, idx, x_out, evals, stream)
  if (model == 1) {
    if (d != 1) return MOCAT_ERR_ARG;
    MOCAT_REJ(SvTrans, 1);
  }
  if (model != 0) return MOCAT_ERR_ARG;
  switch (d) {
    case 1: MOCAT_REJ(LgTrans<1>, 1);
    case 2: MOCAT_REJ(LgTrans<2>, 2);
    case 3: MOCAT_REJ(LgTrans<3>, 3);
    case 4: MOCAT_REJ(LgTrans<4>, 4);
    case 5: MOCAT_REJ(LgTrans<5>, 5);
    case 8: MOCAT_REJ(LgTrans<8>, 8);
    default:
      if (d < 1 || d > kRtMaxDim) return MOCAT_ERR_ARG;
      return launch_rejection<LgTransRt, 0>(params, particles, log_weight, n, x_next, m, key, max_rejections,
                                            init_bound, bound_inflation, b, idx, x_out, evals, stream, d);
  }
#undef MOCAT_REJ
}


int mocat_ffbsi_step(int model, int d, const float* params, const float* particles, const float* log_weight,
                     int64_t n, const float* x_next, int64_t m, const uint32_t* key, int32_t* idx, float* x_out,
                     hipStream_t stream) {
  if (!params || !particles || !log_weight || !x_next || !key || !x_out || n <= 0 || n >= 0xFFFFFFFFll || m <= 0 ||
      m >= 0xFFFFFFFFll)
    return MOCAT_ERR_ARG;
  if (model == 1) {
    if (d != 1) return MOCAT_ERR_ARG;
    return launch_ffbsi<SvTrans, 1>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
  }
  if (model != 0) return MOCAT_ERR_ARG;
  switch (d) {
    case 1: return launch_ffbsi<LgTrans<1>, 1>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 2: return launch_ffbsi<LgTrans<2>, 2>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 3: return launch_ffbsi<LgTrans<3>, 3>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 4: return launch_ffbsi<LgTrans<4>, 4>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 5: return launch_ffbsi<LgTrans<5>, 5>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 8: return launch_ffbsi<LgTrans<8>, 8>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    default: {
      if (d < 1 || d > kRtMaxDim) return MOCAT_ERR_ARG;
      const int per = 256 / kWave;
      hipLaunchKernelGGL(ffbsi_step_rt_kernel, dim3((unsigned)((m + per - 1) / per)), dim3(256), 0, stream, d, params,
                         particles, log_weight, n, x_next, m, key, idx, x_out);
      MOCAT_LAUNCH_CHECK();
      return MOCAT_OK;
    }
  }
}

}  // extern "C"
