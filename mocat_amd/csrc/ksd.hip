// Kernel Stein discrepancy (mocat/src/metrics.py:88-130) for mocat's Gaussian and IMQ kernels
// (mocat/src/kernels.py:83-111, 164-215), over the full ensemble or over random minibatches.
//
// k_0(x_i, x_j) depends on the pair only through four sums over the coordinates,
//   r2 = |x_i - x_j|^2, a = (x_i - x_j).g_j, b = (x_i - x_j).g_i, c = g_i.g_j,
// so a workgroup takes a 64 x 64 tile of pairs, stages 32-coordinate chunks of both sides in LDS and
// gives each thread a 4 x 4 block of pairs (FP32 sums over the coordinates; the weighted k_0 row
// sums in FP64). Row partial sums over each workgroup's j range go to part[js][i]; the host adds
// them, weights by w_i and divides by (sum w)^2.
#include "common.hpp"

using namespace mocat;

namespace {

constexpr int kTile = 64, kChunk = 32, kStride = kTile + 4, kKsdThreads = 256;

struct KsdParams {
  int kernel;  // 0 Gaussian, 1 IMQ
  float h, c, beta;
  int d;
};

// mocat's k_0 (metrics.py:116-124) in terms of (r2, a, b, c)
MOCAT_DEV double k0_pair(const KsdParams& p, float r2, float a, float b, float c) {
  const float h2 = p.h * p.h;
  if (p.kernel == 0) {
    // Gaussian: diag_grad_xy summed = (d h^2 - r2) k / h^4, grad_x k = (y - x) k / h^2, grad_y k = (x - y) k / h^2
    const float k = __expf(-0.5f * r2 / h2);
    return (double)k * ((double)((float)p.d * h2 - r2) / ((double)h2 * h2) + (double)(b - a) / h2 + (double)c);
  }
  // IMQ as mocat writes it: base = c + |x - y|^2 / (2 h^2); k = base^beta;
  // grad_x = diff beta base^(beta-1); grad_y = -diff / h beta base^(beta-1);
  // diag_grad_xy = (-beta base^(beta-1) + beta (beta-1) diff^2 base^(beta-2)) / h^2, diff = (x - y) / h
  const double dd = (double)r2 / h2;
  const double base = (double)p.c + 0.5 * dd;
  const double kb = pow(base, (double)p.beta), kb1 = kb / base, kb2 = kb1 / base;
  const double beta = p.beta;
  return (-(double)p.d * beta * kb1 + beta * (beta - 1.0) * dd * kb2) / h2 + beta * kb1 * a / p.h -
         beta * kb1 * b / h2 + kb * c;
}

__global__ __launch_bounds__(kKsdThreads) void ksd_rows_kernel(const float* __restrict__ x,
                                                               const float* __restrict__ g,
                                                               const float* __restrict__ w, int n, KsdParams p,
                                                               int js_count, double* __restrict__ part) {
  __shared__ alignas(16) float xi[kChunk][kStride], gi[kChunk][kStride], xj[kChunk][kStride], gj[kChunk][kStride];
  __shared__ double red[kKsdThreads / 16][kTile];
  const int t = threadIdx.x, ti = t & 15, tj = t >> 4;
  const int i0 = blockIdx.x * kTile, js = blockIdx.y;
  const int d = p.d;
  const int ntiles = (n + kTile - 1) / kTile;
  double rows[4] = {0.0, 0.0, 0.0, 0.0};
  for (int jt = js; jt < ntiles; jt += js_count) {
    const int j0 = jt * kTile;
    float r2[4][4], sa[4][4], sb[4][4], sc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) r2[u][v] = sa[u][v] = sb[u][v] = sc[u][v] = 0.f;
    for (int k0 = 0; k0 < d; k0 += kChunk) {
      __syncthreads();
      for (int e = t; e < kTile * kChunk; e += kKsdThreads) {
        const int row = e / kChunk, k = e - row * kChunk;
        const bool kin = k0 + k < d;
        const int ii = i0 + row, jj = j0 + row;
        xi[k][row] = (kin && ii < n) ? x[(int64_t)ii * d + k0 + k] : 0.f;
        gi[k][row] = (kin && ii < n) ? g[(int64_t)ii * d + k0 + k] : 0.f;
        xj[k][row] = (kin && jj < n) ? x[(int64_t)jj * d + k0 + k] : 0.f;
        gj[k][row] = (kin && jj < n) ? g[(int64_t)jj * d + k0 + k] : 0.f;
      }
      __syncthreads();
      const int kn = min(kChunk, d - k0);
      for (int k = 0; k < kn; ++k) {
        const float4 xa = *reinterpret_cast<const float4*>(&xi[k][ti * 4]);
        const float4 ga = *reinterpret_cast<const float4*>(&gi[k][ti * 4]);
        const float4 xb = *reinterpret_cast<const float4*>(&xj[k][tj * 4]);
        const float4 gb = *reinterpret_cast<const float4*>(&gj[k][tj * 4]);
        const float xv[4] = {xa.x, xa.y, xa.z, xa.w}, gv[4] = {ga.x, ga.y, ga.z, ga.w};
        const float yv[4] = {xb.x, xb.y, xb.z, xb.w}, hv[4] = {gb.x, gb.y, gb.z, gb.w};
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float df = xv[u] - yv[v];
            r2[u][v] = fmaf(df, df, r2[u][v]);
            sa[u][v] = fmaf(df, hv[v], sa[u][v]);
            sb[u][v] = fmaf(df, gv[u], sb[u][v]);
            sc[u][v] = fmaf(gv[u], hv[v], sc[u][v]);
          }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = j0 + tj * 4 + v;
      if (j >= n) continue;
      const double wj = w ? (double)w[j] : 1.0;
#pragma unroll
      for (int u = 0; u < 4; ++u) rows[u] += wj * k0_pair(p, r2[u][v], sa[u][v], sb[u][v], sc[u][v]);
    }
  }
  // fixed-order sum over the 16 thread columns of each row
#pragma unroll
  for (int u = 0; u < 4; ++u) red[tj][ti * 4 + u] = rows[u];
  __syncthreads();
  if (t < kTile && i0 + t < n) {
    double s = 0.0;
    for (int q = 0; q < kKsdThreads / 16; ++q) s += red[q][t];
    part[(int64_t)js * n + i0 + t] = s;
  }
}

// Minibatch rows: one wave per particle i; lane b takes batch entries b, b + 64, ... and sums the
// four coordinate sums of its pair in coordinate order (FP32, as the tile kernel), then k_0 in FP64;
// the lanes' FP64 partial sums are combined by a fixed butterfly.
__global__ __launch_bounds__(256) void ksd_batch_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                                        const float* __restrict__ w, int n, int64_t batch,
                                                        const int32_t* __restrict__ inds, KsdParams p,
                                                        double* __restrict__ row_k, double* __restrict__ row_w) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= n) return;
  const int d = p.d;
  const float* xi = x + i * d;
  const float* gi = g + i * d;
  double sk = 0.0, sw = 0.0;
  for (int64_t b = lane; b < batch; b += kWave) {
    const int64_t j = inds[i * batch + b];
    const float* xj = x + j * d;
    const float* gj = g + j * d;
    float r2 = 0.f, a = 0.f, bb = 0.f, c = 0.f;
    for (int k = 0; k < d; ++k) {
      const float df = xi[k] - xj[k];
      r2 = fmaf(df, df, r2);
      a = fmaf(df, gj[k], a);
      bb = fmaf(df, gi[k], bb);
      c = fmaf(gi[k], gj[k], c);
    }
    const double wj = w ? (double)w[j] : 1.0;
    sk += wj * k0_pair(p, r2, a, bb, c);
    sw += wj;
  }
  sk = wave_sum(sk);
  sw = wave_sum(sw);
  if (lane == 0) {
    row_k[i] = sk;
    row_w[i] = sw;
  }
}

}  // namespace

extern "C" {

int mocat_ksd_rows(const float* x, const float* grad_potential, const float* weights, int64_t n, int64_t d,
                   int kernel, float bandwidth, float c, float beta, int js_count, double* part,
                   hipStream_t stream) {
  if (!x || !grad_potential || !part || n <= 0 || n > (1 << 24) || d <= 0 || d > (1 << 16) || js_count <= 0 ||
      js_count > 65535 || (kernel != 0 && kernel != 1) || !(bandwidth > 0.f) || (kernel == 1 && !(c > 0.f)))
    return MOCAT_ERR_ARG;
  const KsdParams p{kernel, bandwidth, c, beta, (int)d};
  hipLaunchKernelGGL(ksd_rows_kernel, dim3((unsigned)((n + kTile - 1) / kTile), (unsigned)js_count),
                     dim3(kKsdThreads), 0, stream, x, grad_potential, weights, (int)n, p, js_count, part);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_ksd_batch_rows(const float* x, const float* grad_potential, const float* weights, int64_t n, int64_t d,
                         int64_t batch, const int32_t* inds, int kernel, float bandwidth, float c, float beta,
                         double* row_k, double* row_w, hipStream_t stream) {
  if (!x || !grad_potential || !inds || !row_k || !row_w || n <= 0 || n > (1 << 24) || d <= 0 || d > (1 << 16) ||
      batch <= 0 || batch > (1ll << 31) / n || (kernel != 0 && kernel != 1) || !(bandwidth > 0.f) ||
      (kernel == 1 && !(c > 0.f)))
    return MOCAT_ERR_ARG;
  const KsdParams p{kernel, bandwidth, c, beta, (int)d};
  const int per = 256 / kWave;
  hipLaunchKernelGGL(ksd_batch_kernel, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, stream, x, grad_potential,
                     weights, (int)n, batch, inds, p, row_k, row_w);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
