// Stein variational gradient descent (mocat/src/transport/svgd.py) with the Gaussian RBF kernel
// (mocat/src/kernels.py:82-116) and a diagonal Gaussian-mixture target.
//
// phi_i = (1/N) sum_j [ -k_ij gradU_j + (x_i - x_j) k_ij / h^2 ],  k_ij = exp(-|x_i - x_j|^2 / (2 h^2))
//       = (1/N) [ -(K G)_i + (s_i x_i - (K X)_i) / h^2 ],  s_i = sum_j k_ij        (svgd.py:18-32)
// computed flash-attention style per tile of 32 particles i: S = X_J X_I^T on MFMA (bf16 hi/lo
// split, FP32 accumulation) for 32 particles j at a time, k_ij formed in registers from
// |x_i|^2 + |x_j|^2 - 2 S (particles centred first: distances are translation invariant and the
// centring keeps the cancellation small), then K [G | X]_J accumulated on MFMA with the k values as
// the A operand (chains of the accumulator rows = the j rows of the packed B layout).
// The j range is split over workgroups; partial sums are combined by the update kernel, which also
// applies jax.example_libraries.optimizers.adagrad(step, momentum=0.9) with gradient -phi
// (svgd.py:104-106, 138-139):  G += g^2; m = 0.1 g / sqrt(G) + 0.9 m; x -= step m.
//
// Bandwidth: median (or mean) of all N^2 pairwise distances, diagonal zeros included, divided by
// sqrt(2 log N) (kernels.py:220-229). Squared distances are written once (N x N f32) and the two
// middle order statistics found exactly by an 8-bit radix select over their bit patterns.
#include <cmath>
#include <cstdlib>

#include "common.hpp"

using namespace mocat;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int CT = 32;

MOCAT_DEV int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

MOCAT_DEV void split_bf16(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

template <int DP>
struct SvgdLayout {
  static constexpr int KB = DP / 16;       // K blocks of the distance GEMM
  static constexpr int NB = (2 * DP) / 32;  // N blocks of [G | X] (width 2 DP)
  static constexpr int A_ELEMS = 2 * KB * 512;
  static constexpr int B_ELEMS = 2 * 2 * NB * 512;
  static constexpr int BLOCK_BF16 = A_ELEMS + B_ELEMS;
  static constexpr int BLOCK_BYTES = BLOCK_BF16 * 2 + 32 * 4;
};

// Column means, FP64, in a fixed summation order, in two passes: workgroup (column group cg, row slice
// rs) sums its slice's rows (lane = column, coalesced row segments; wave w takes rows w, w+4, ...;
// the four wave sums in wave order) into partial[rs][column]; then every packing workgroup adds the
// slices of each column in order (one thread per column), so all of them hold the same means.
constexpr int kMeanSlices = 64;
__global__ __launch_bounds__(256) void colmean_part_kernel(const float* __restrict__ x, int n, int d, int slices,
                                                           double* __restrict__ partial) {
  __shared__ double red[4][kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int k = blockIdx.x * kWave + lane, rs = blockIdx.y;
  const int per = (n + slices - 1) / slices, r0 = rs * per, r1 = min(n, r0 + per);
  double acc0 = 0.0, acc1 = 0.0;
  if (k < d) {
    int i = r0 + w;
    for (; i + 4 < r1; i += 8) {
      acc0 += (double)x[(int64_t)i * d + k];
      acc1 += (double)x[(int64_t)(i + 4) * d + k];
    }
    for (; i < r1; i += 4) acc0 += (double)x[(int64_t)i * d + k];
  }
  red[w][lane] = acc0 + acc1;
  __syncthreads();
  if (w == 0 && k < d) partial[(int64_t)rs * d + k] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
}

// Pack the centred particles (A layout), [grad | centred x] (B layout) and |x_j - mean|^2 per
// 32-row block, one workgroup per block: a thread writes 8 consecutive bf16 (16 bytes) of the hi and of
// the lo part; the norms add the A threads' 8-column partial sums of a row in column order. The means
// come from the column-mean partials (slices <= kMeanSlices); workgroup 0 also stores them in mean.
template <int DP>
__global__ __launch_bounds__(256) void svgd_pack_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                                        const double* __restrict__ cpart, int slices,
                                                        float* __restrict__ mean, int n, int d,
                                                        char* __restrict__ out) {
  using L = SvgdLayout<DP>;
  constexpr int KB = L::KB, NB = L::NB;
  static_assert(DP <= 256, "one thread per column for the means");
  __shared__ float nsq[32][2 * KB];
  __shared__ float smean[DP];
  const int rb = blockIdx.x, t = threadIdx.x;
  char* blk = out + (int64_t)rb * L::BLOCK_BYTES;
  bf16x8* A = reinterpret_cast<bf16x8*>(blk);
  bf16x8* B = A + L::A_ELEMS / 8;
  // A layout [2][KB][64][8]: lane l holds row l%32, columns kb*16 + (l/32)*8 + s
  // a thread takes tasks t, t + 256, ... of each layout (whole rounds); all its loads of both layouts
  // are issued before the first conversion
  constexpr int PA = KB * 64 / 256, PB = 2 * NB * 64 / 256;
  static_assert(KB * 64 % 256 == 0 && 2 * NB * 64 % 256 == 0, "whole task rounds per thread");
  float va[PA][8], vb[PB][8];
#pragma unroll
  for (int p = 0; p < PA; ++p) {
    const int task = p * 256 + t, kb = task / 64, lane = task % 64;
    const int row = rb * 32 + (lane & 31), c0 = kb * 16 + (lane >> 5) * 8;
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      const int c = c0 + s2;
      va[p][s2] = (row < n && c < d) ? x[(int64_t)row * d + c] : 0.f;
    }
  }
#pragma unroll
  for (int p = 0; p < PB; ++p) {
    const int task = p * 256 + t, kb2 = task / (NB * 64), rest = task % (NB * 64), lane = rest % 64;
    const int c = (rest / 64) * 32 + (lane & 31);
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      const int row = rb * 32 + acc_row(kb2 * 8 + s2, lane);
      float v = 0.f;
      if (row < n) {
        if (c < DP) {
          if (c < d) v = g[(int64_t)row * d + c];
        } else if (c - DP < d) {
          v = x[(int64_t)row * d + (c - DP)];
        }
      }
      vb[p][s2] = v;
    }
  }
  // the column means: the slices added in order, as one sum (loads in two batches of 32)
  if (t < d) {
    double s = 0.0;
#pragma unroll 1
    for (int u0 = 0; u0 < kMeanSlices; u0 += 32) {
      double v[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) v[u] = u0 + u < slices ? cpart[(int64_t)(u0 + u) * d + t] : 0.0;
#pragma unroll
      for (int u = 0; u < 32; ++u) s += v[u];
    }
    const float m = (float)(s / (double)n);
    smean[t] = m;
    if (rb == 0) mean[t] = m;
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PA; ++p) {
    const int task = p * 256 + t, kb = task / 64, lane = task % 64;
    const int rl = lane & 31, row = rb * 32 + rl, c0 = kb * 16 + (lane >> 5) * 8;
    bf16x8 h8, l8;
    float ss = 0.f;
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      const int c = c0 + s2;
      const float v = (row < n && c < d) ? va[p][s2] - smean[c] : 0.f;
      ss += v * v;
      __bf16 h, l;
      split_bf16(v, h, l);
      h8[s2] = h;
      l8[s2] = l;
    }
    A[(0 * KB + kb) * 64 + lane] = h8;
    A[(1 * KB + kb) * 64 + lane] = l8;
    nsq[rl][kb * 2 + (lane >> 5)] = ss;
  }
  // B layout [2][2][NB][64][8]: lane l holds column nb*32 + l%32 (grad for < DP, centred x above), rows
  // acc_row(kb2*8 + s, l)
#pragma unroll
  for (int p = 0; p < PB; ++p) {
    const int task = p * 256 + t, kb2 = task / (NB * 64), rest = task % (NB * 64), nb = rest / 64, lane = rest % 64;
    const int c = nb * 32 + (lane & 31);
    const bool centred = c >= DP && c - DP < d;
    bf16x8 h8, l8;
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      const int row = rb * 32 + acc_row(kb2 * 8 + s2, lane);
      const float v = (centred && row < n) ? vb[p][s2] - smean[c - DP] : vb[p][s2];
      __bf16 h, l;
      split_bf16(v, h, l);
      h8[s2] = h;
      l8[s2] = l;
    }
    B[((0 * 2 + kb2) * NB + nb) * 64 + lane] = h8;
    B[((1 * 2 + kb2) * NB + nb) * 64 + lane] = l8;
  }
  __syncthreads();
  // norms in accumulator-row order: entry qq = g*16 + r is row 8*(r>>2) + 4*g + (r&3); padding rows -1
  if (t < 32) {
    const int gg = t >> 4, r = t & 15, rl = 8 * (r >> 2) + 4 * gg + (r & 3), row = rb * 32 + rl;
    float nrm = -1.f;
    if (row < n) {
      nrm = 0.f;
      for (int q = 0; q < 2 * KB; ++q) nrm += nsq[rl][q];
    }
    reinterpret_cast<float*>(blk + L::BLOCK_BF16 * 2)[t] = nrm;
  }
}

// The distance GEMM's B operand for the 32 particles of packed block rb, read from its A layout (the
// two operand layouts of the 32x32x16 product coincide: lane l holds particle l % 32, columns
// kb * 16 + (l / 32) * 8 + s), and the norm of particle rb * 32 + lane % 32 (norms are stored in
// accumulator-row order). Blocks past the end give zero operands.
template <int DP>
MOCAT_DEV float load_i_operands(const char* __restrict__ packed, int rb, int nrb, int lane, bf16x8* bih,
                                bf16x8* bil) {
  using L = SvgdLayout<DP>;
  constexpr int KB = L::KB;
  if (rb >= nrb) {
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) {
        bih[kb][s2] = (__bf16)0.f;
        bil[kb][s2] = (__bf16)0.f;
      }
    return -1.f;
  }
  const char* blk = packed + (int64_t)rb * L::BLOCK_BYTES;
  const bf16x8* A = reinterpret_cast<const bf16x8*>(blk);
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    bih[kb] = A[kb * 64 + lane];
    bil[kb] = A[(KB + kb) * 64 + lane];
  }
  const int col = lane & 31;
  return reinterpret_cast<const float*>(blk + L::BLOCK_BF16 * 2)[((col >> 2) & 1) * 16 + 4 * (col >> 3) + (col & 3)];
}

// Tile of 32 particles i: centred coordinates in LDS, their squared norms, and the bf16 hi/lo
// B operand of the distance GEMM (lane: particle l%32, columns kb*16 + (l/32)*8 + s).
template <int DP>
struct ITile {
  alignas(16) float xi[CT * DP];
  alignas(16) bf16x8 op[2][DP / 16][64];
  float ni[CT];
};

template <int DP, int NT>
MOCAT_DEV void load_itile(ITile<DP>& it, const float* __restrict__ x, const float* __restrict__ mean, int n, int d,
                          int i0) {
  const int t = threadIdx.x;
  for (int e = t; e < CT * DP; e += NT) {
    const int c = e / DP, k = e - c * DP;
    it.xi[e] = (i0 + c < n && k < d) ? x[(int64_t)(i0 + c) * d + k] - mean[k] : 0.f;
  }
  __syncthreads();
  for (int e = t; e < (DP / 16) * 64; e += NT) {
    const int kb = e / 64, ln = e % 64;
    const float* src = it.xi + (ln & 31) * DP + kb * 16 + (ln >> 5) * 8;
    bf16x8 h8, l8;
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2) {
      __bf16 h, l;
      split_bf16(src[s2], h, l);
      h8[s2] = h;
      l8[s2] = l;
    }
    it.op[0][kb][ln] = h8;
    it.op[1][kb][ln] = l8;
  }
  if (t < CT) {
    float a = 0.f;
    for (int k = 0; k < d; ++k) a += it.xi[t * DP + k] * it.xi[t * DP + k];
    it.ni[t] = a;
  }
  __syncthreads();
}

// S = X_J X_I^T for one 32-row block of j (accumulator: lane column = particle i, rows = j).
template <int DP>
MOCAT_DEV floatx16 gram_block(const bf16x8* __restrict__ A, const ITile<DP>& it, int lane) {
  constexpr int KB = DP / 16;
  floatx16 z;
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = 0.f;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    const bf16x8 ah = A[(0 * KB + kb) * 64 + lane];
    const bf16x8 al = A[(1 * KB + kb) * 64 + lane];
    const bf16x8 bh = it.op[0][kb][lane], bl = it.op[1][kb][lane];
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, z, 0, 0, 0);
  }
  return z;
}

// ---------------------------------------------------------------- strict upper triangle of the distance matrix
// The squared-distance matrix is symmetric with a zero diagonal, so its order statistics and its mean
// follow from the strict upper triangle (row < column): the full n^2 entries are the n diagonal zeros
// and every upper entry twice. Only tiles on or above the diagonal are computed and read, each pair
// once (the matrix the selection sees is exactly symmetric).

constexpr int kRowUnroll = 4;  // row elements loaded per thread before they are used
// Workgroups of the selection passes: each ends with global atomics on a few shared counters or
// histogram bins, which serialise across workgroups, so one workgroup per compute unit.
constexpr int kSelectBlocks = 256;
constexpr int kSelThreads = 1024;  // 16 waves per compute unit to cover the load latency
// The bracketed selection: 512 threads per workgroup (three per compute unit, so kSelectBlocks of them
// fit on a whole device with room to spare; the grid is also capped by what the device holds); eight
// row loads in flight per thread on a pass over all pairs.
constexpr int kBracketThreads = 512, kBracketUnroll = 8;

// Lanes of a wave that share an 8-bit digit (ballots of the digit's bits); the lowest lane of each
// group adds the group's size to h[digit]: one LDS atomic per distinct digit of the wave, with no
// serialisation on a popular bin. Wave-uniform control flow is required.
MOCAT_DEV void wave_multisplit_count(bool act, uint32_t dig, uint32_t* h, int lane) {
  const uint64_t actm = __ballot(act);
  if (actm == 0) return;
  uint64_t same = actm;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (dig >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    same &= bit ? bb : ~bb;
  }
  if (act && __ffsll((unsigned long long)same) - 1 == lane) atomicAdd(&h[dig], (uint32_t)__popcll(same));
}

// One radix pass over the strict upper triangle of the row-major n x n matrix v; rows are taken
// round-robin by the workgroups, columns row+1..n-1 by the threads (block-uniform trip counts).
__global__ __launch_bounds__(kSelThreads) void radix_hist_upper_kernel(const uint32_t* __restrict__ v, int n, int shift,
                                                               const uint32_t* __restrict__ state,
                                                               uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[2][256];
  const uint32_t mask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
  const uint32_t p0 = state[0] & mask, p1 = state[1] & mask;
  const bool shared_prefix = p0 == p1;
  const int lane = threadIdx.x & (kWave - 1);
  for (int b = threadIdx.x; b < 512; b += kSelThreads) (&lh[0][0])[b] = 0u;
  __syncthreads();
  for (int row = blockIdx.x; row < n - 1; row += gridDim.x) {
    const uint32_t* vr = v + (int64_t)row * n;
    for (int c0 = row + 1; c0 < n; c0 += kRowUnroll * kSelThreads) {
      uint32_t uu[kRowUnroll];
#pragma unroll
      for (int k = 0; k < kRowUnroll; ++k) {
        const int c = c0 + k * kSelThreads + (int)threadIdx.x;
        uu[k] = c < n ? vr[c] : 0u;
      }
#pragma unroll
      for (int k = 0; k < kRowUnroll; ++k) {
        const bool valid = c0 + k * kSelThreads + (int)threadIdx.x < n;
        const uint32_t u = uu[k];
        const uint32_t dig = (u >> shift) & 255u;
        wave_multisplit_count(valid && (u & mask) == p0, dig, lh[0], lane);
        if (!shared_prefix) wave_multisplit_count(valid && (u & mask) == p1, dig, lh[1], lane);
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 512; b += kSelThreads) {
    const uint32_t c = shared_prefix ? lh[0][b & 255] : (&lh[0][0])[b];
    if (c) atomicAdd(&hist[b], c);
  }
}

// sum of sqrt over the strict upper triangle, FP64, one partial per workgroup
__global__ __launch_bounds__(256) void sqrt_sum_upper_kernel(const float* __restrict__ d2, int n,
                                                             double* __restrict__ partial) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int row = blockIdx.x; row < n - 1; row += gridDim.x) {
    const float* r = d2 + (int64_t)row * n;
    for (int c0 = row + 1; c0 < n; c0 += kRowUnroll * 256) {
      float vv[kRowUnroll];
#pragma unroll
      for (int k = 0; k < kRowUnroll; ++k) {
        const int c = c0 + k * 256 + (int)threadIdx.x;
        vv[k] = c < n ? r[c] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < kRowUnroll; ++k) s += (double)sqrtf(vv[k]);
    }
  }
  s = block_sum(s, sc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------- phi (partial over a j range)
// One workgroup = 4 waves = 128 particles i, 32 per wave, whose distance-GEMM operands (bf16 hi/lo
// of the centred coordinates) stay in registers. The packed 32-particle j blocks of the workgroup's
// j range are staged through LDS, double buffered: each block is read from L2 once per workgroup and
// used by all four waves. Per j block and wave: S = X_J X_I^T (MFMA), K = exp(-|x_i - x_j|^2 / 2h^2)
// in registers (the accumulator reused as the next A operand), then [K G_J | K X_J] += on MFMA.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kPhiWaves = 4, kPhiThreads = kPhiWaves * kWave, kPhiI = kPhiWaves * 32;

template <int DP>
__global__ __launch_bounds__(kPhiThreads) void svgd_phi_kernel(const char* __restrict__ packed, int n,
                                                               float h_host, const float* __restrict__ h_dev,
                                                               int js_count, float* __restrict__ part_o,
                                                               float* __restrict__ part_s) {
  using L = SvgdLayout<DP>;
  constexpr int KB = L::KB, NB = L::NB;
  static_assert(L::BLOCK_BYTES % 16 == 0, "packed blocks are copied in 16-byte chunks");
  constexpr int CHUNKS = L::BLOCK_BYTES / 16, PER = (CHUNKS + kPhiThreads - 1) / kPhiThreads;
  __shared__ u32x4 stage[2][CHUNKS];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave, g = lane >> 5, col = lane & 31;
  const int i0 = blockIdx.x * kPhiI + w * 32;
  const int nrb = (n + 31) / 32;
  const int js = blockIdx.y;
  const int rb0 = (int)((int64_t)nrb * js / js_count), rb1 = (int)((int64_t)nrb * (js + 1) / js_count);
  const float hb = h_dev ? *h_dev : h_host;
  const float inv2h2 = 0.5f / (hb * hb);

  // B operand of the distance GEMM: lane holds particle i0 + col, columns kb*16 + g*8 + s
  bf16x8 bih[KB], bil[KB];
  const float ni = load_i_operands<DP>(packed, i0 / 32, nrb, lane, bih, bil);

  floatx16 gacc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int r = 0; r < 16; ++r) gacc[nb][r] = 0.f;
  float spart = 0.f;

  // j blocks staged global -> LDS directly (16 bytes per lane, no registers): lane l of wave w copies
  // chunk p * kPhiThreads + w * 64 + l to the same LDS chunk; the copy into the other buffer overlaps
  // the current block's products and is waited for before the step's closing barrier
  auto stage_block = [&](int buf, int rb) {
    const u32x4* src = reinterpret_cast<const u32x4*>(packed + (int64_t)rb * L::BLOCK_BYTES);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int c = p * kPhiThreads + t;
      if (c < CHUNKS)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + c),
                                         (__attribute__((address_space(3))) void*)(&stage[buf][c - lane]), 16, 0,
                                         0);
    }
  };
  const int nsteps = rb1 - rb0;
  if (nsteps > 0) stage_block(0, rb0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    if (st + 1 < nsteps) stage_block((st + 1) & 1, rb0 + st + 1);
    const char* blk = reinterpret_cast<const char*>(stage[st & 1]);
    const bf16x8* A = reinterpret_cast<const bf16x8*>(blk);
    const bf16x8* B = A + L::A_ELEMS / 8;
    floatx16 z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.f;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const bf16x8 ah = A[(0 * KB + kb) * 64 + lane];
      const bf16x8 al = A[(1 * KB + kb) * 64 + lane];
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bih[kb], z, 0, 0, 0);
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bil[kb], z, 0, 0, 0);
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bih[kb], z, 0, 0, 0);
    }
    const floatx4* nj4 = reinterpret_cast<const floatx4*>(blk + L::BLOCK_BF16 * 2 + g * 64);
    float nj[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const floatx4 v = nj4[q];
      nj[4 * q] = v.x;
      nj[4 * q + 1] = v.y;
      nj[4 * q + 2] = v.z;
      nj[4 * q + 3] = v.w;
    }
    bf16x8 kh[2], kl[2];
    const int jbase = (rb0 + st) * 32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      // the self-distance is exactly 0 (the norms-minus-Gram form leaves cancellation residue there)
      const float D = (jbase + acc_row(r, lane) == i0 + col) ? 0.f : fmaxf(ni + nj[r] - 2.f * z[r], 0.f);
      const float kv = nj[r] >= 0.f ? __expf(-D * inv2h2) : 0.f;
      spart += kv;
      __bf16 h, l;
      split_bf16(kv, h, l);
      kh[r >> 3][r & 7] = h;
      kl[r >> 3][r & 7] = l;
    }
    // consecutive products go to different accumulators (a chain of dependent accumulations stalls on
    // each result); each accumulator still receives its six products in the same order
#pragma unroll
    for (int kb2 = 0; kb2 < 2; ++kb2) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const bf16x8 xh = B[((0 * 2 + kb2) * NB + nb) * 64 + lane];
        gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kh[kb2], xh, gacc[nb], 0, 0, 0);
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const bf16x8 xl = B[((1 * 2 + kb2) * NB + nb) * 64 + lane];
        gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kh[kb2], xl, gacc[nb], 0, 0, 0);
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const bf16x8 xh = B[((0 * 2 + kb2) * NB + nb) * 64 + lane];
        gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kl[kb2], xh, gacc[nb], 0, 0, 0);
      }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  // accumulator rows = particles i (acc_row), columns = features; row sums of K from both lane halves
  spart += __shfl_xor(spart, 32, kWave);
  if (g == 0 && i0 + col < n) part_s[(int64_t)js * n + i0 + col] = spart;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = i0 + acc_row(r, lane);
      if (i < n) part_o[((int64_t)js * n + i) * 2 * DP + nb * 32 + col] = gacc[nb][r];
    }
}

// phi of one coordinate from its sums, then the adagrad step: returns the new coordinate, updates G
// and m. Contraction off and the fused multiply-adds written out, so that every kernel using it rounds
// alike whatever the surrounding code.
struct AdagradOut {
  float x, G, m, phi;
};
MOCAT_DEV AdagradOut svgd_adagrad(float xv, float xc, float kg, float kx, float s, float G, float mprev, float inv_n,
                                  float inv_h2, float step, float momentum) {
#pragma clang fp contract(off)
  const float phi = inv_n * fmaf(fmaf(s, xc, -kx), inv_h2, -kg);
  const float gr = -phi;
  const float G2 = fmaf(gr, gr, G);
  const float ginv = G2 > 0.f ? 1.f / sqrtf(G2) : 0.f;
  const float m = fmaf(momentum, mprev, (1.f - momentum) * (gr * ginv));
  return {fmaf(-step, m, xv), G2, m, phi};
}

// phi from the partial sums, then the adagrad step (x, G, m updated in place); phi_out optional.
__global__ void svgd_update_kernel(const float* __restrict__ part_o, const float* __restrict__ part_s, int js_count,
                                   int n, int d, int DP, const float* __restrict__ mean, float inv_n, float h_host,
                                   const float* __restrict__ h_dev, float step, float momentum,
                                   float* __restrict__ x, float* __restrict__ gsq, float* __restrict__ mom,
                                   float* __restrict__ phi_out) {
  const int64_t total = (int64_t)n * d;
  const float hb = h_dev ? *h_dev : h_host;
  const float inv_h2 = 1.f / (hb * hb);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / d;
    const int k = (int)(e - i * d);
    float kg = 0.f, kx = 0.f, s = 0.f;
    for (int js = 0; js < js_count; ++js) {
      const float* o = part_o + ((int64_t)js * n + i) * 2 * DP;
      kg += o[k];
      kx += o[DP + k];
      s += part_s[(int64_t)js * n + i];
    }
    const float xv = x[e];
    const AdagradOut u = svgd_adagrad(xv, xv - mean[k], kg, kx, s, gsq[e], mom[e], inv_n, inv_h2, step, momentum);
    if (phi_out) phi_out[e] = u.phi;
    gsq[e] = u.G;
    mom[e] = u.m;
    x[e] = u.x;
  }
}

// ---------------------------------------------------------------- Gaussian mixture (diagonal)
// U(x) = -logsumexp_k [ logc_k - 0.5 sum_j lam_kj (x_j - mu_kj)^2 ],  grad = sum_k r_k lam_k (x - mu_k)
// (mocat/src/scenarios/toy_examples.py:54-116, evaluated in log-sum-exp form). One wave per particle.
__global__ __launch_bounds__(256) void gmm_kernel(const float* __restrict__ x, int n, int d, int K,
                                                  const float* __restrict__ mu, const float* __restrict__ lam,
                                                  const float* __restrict__ logc, float* __restrict__ pot,
                                                  float* __restrict__ grad) {
  // contraction off and the fused multiply-adds written out (as in svgd_update_gmm_kernel, which must
  // round alike)
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= n) return;
  const float* xi = x + i * d;
  float q[8];
  float qmax = -INFINITY;
  for (int k = 0; k < K; ++k) {
    float a = 0.f;
    for (int j = lane; j < d; j += kWave) {
      const float df = xi[j] - mu[k * d + j];
      a = fmaf(lam[k * d + j] * df, df, a);
    }
    a = wave_sum(a);
    q[k] = fmaf(-0.5f, a, logc[k]);
    qmax = fmaxf(qmax, q[k]);
  }
  float ssum = 0.f;
  for (int k = 0; k < K; ++k) ssum += __expf(q[k] - qmax);
  if (lane == 0) pot[i] = -(qmax + __logf(ssum));
  for (int j = lane; j < d; j += kWave) {
    float gsum = 0.f;
    for (int k = 0; k < K; ++k) gsum = fmaf(__expf(q[k] - qmax) / ssum * lam[k * d + j], xi[j] - mu[k * d + j], gsum);
    grad[i * d + j] = gsum;
  }
}

// Full covariances: U(x) = -logsumexp_k [ logc_k - 0.5 r_k^T S_k r_k ], grad = sum_k softmax_k S_k r_k with
// r_k = x - mu_k and S_k symmetric ([K][d][d]: the precision, or mocat's L^T L form,
// mocat/src/scenarios/toy_examples.py:95-100). A wave evaluates 4 particles; lane l holds columns l and
// l + 64 (d <= 128); for each component the rows of S_k are read once per wave (coalesced, L2-resident)
// and applied to the 4 particles, the row's coefficient r_i broadcast from its lane by readlane.
constexpr int kGmmFullPer = 4;
template <int K>
__global__ __launch_bounds__(256) void gmm_full_kernel(const float* __restrict__ x, int n, int d,
                                                       const float* __restrict__ mu, const float* __restrict__ S,
                                                       const float* __restrict__ logc, float* __restrict__ pot,
                                                       float* __restrict__ grad) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t p0 = ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kGmmFullPer;
  if (p0 >= n) return;
  float xv[kGmmFullPer][2];
#pragma unroll
  for (int p = 0; p < kGmmFullPer; ++p)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = lane + c * kWave;
      xv[p][c] = (p0 + p < n && col < d) ? x[(p0 + p) * d + col] : 0.f;
    }
  float q[K][kGmmFullPer], sv[K][kGmmFullPer][2];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float r[kGmmFullPer][2], s[kGmmFullPer][2];
#pragma unroll
    for (int p = 0; p < kGmmFullPer; ++p)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int col = lane + c * kWave;
        r[p][c] = col < d ? xv[p][c] - mu[k * d + col] : 0.f;
        s[p][c] = 0.f;
      }
    const float* Sk = S + (int64_t)k * d * d;
    for (int i = 0; i < d; ++i) {
      float srow[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int col = lane + c * kWave;
        srow[c] = col < d ? Sk[(int64_t)i * d + col] : 0.f;
      }
      const int src = i & (kWave - 1);
#pragma unroll
      for (int p = 0; p < kGmmFullPer; ++p) {
        const float rv = i < kWave ? r[p][0] : r[p][1];
        const float ri = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rv), src));
#pragma unroll
        for (int c = 0; c < 2; ++c) s[p][c] = fmaf(srow[c], ri, s[p][c]);
      }
    }
#pragma unroll
    for (int p = 0; p < kGmmFullPer; ++p) {
      float a = fmaf(r[p][0], s[p][0], r[p][1] * s[p][1]);
      a = wave_sum(a);
      q[k][p] = fmaf(-0.5f, a, logc[k]);
#pragma unroll
      for (int c = 0; c < 2; ++c) sv[k][p][c] = s[p][c];
    }
  }
#pragma unroll
  for (int p = 0; p < kGmmFullPer; ++p) {
    if (p0 + p >= n) break;
    float qmax = -INFINITY;
#pragma unroll
    for (int k = 0; k < K; ++k) qmax = fmaxf(qmax, q[k][p]);
    float ssum = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k) ssum += __expf(q[k][p] - qmax);
    if (lane == 0) pot[p0 + p] = -(qmax + __logf(ssum));
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = lane + c * kWave;
      if (col >= d) continue;
      float gsum = 0.f;
#pragma unroll
      for (int k = 0; k < K; ++k) gsum = fmaf(__expf(q[k][p] - qmax) / ssum, sv[k][p][c], gsum);
      grad[(p0 + p) * d + col] = gsum;
    }
  }
}

// ---------------------------------------------------------------- bracketed median (fused SVGD run)
// Between SVGD iterations the median pair distance moves little, so each selection first counts the
// pairs below a bracket [lo, hi) around the previous median and copies those inside it to a candidate
// buffer (in the fused run the distance kernel does this as it writes the distances). When the wanted
// ranks fall inside the bracket (and the candidates fit), the radix passes
// run over the candidates only, with the ranks shifted by the count below; otherwise over all pairs.
// Either way the result is the exact order statistic. The bracket half-width (relative) shrinks
// after hits and grows after misses.
// The count below the bracket is striped over counters: every distance workgroup adds its count at
// its end, and one shared counter would serialise those device-scope atomics.
constexpr int kBelowSlots = 32;
struct Bracket {
  uint32_t lo, hi;     // bit patterns of non-negative floats (ordered as unsigned ints)
  uint32_t below[kBelowSlots];  // pairs < lo (workgroup k of the distance pass adds to slot k % kBelowSlots)
  uint32_t inside;     // pairs in [lo, hi) (the candidate count, possibly above the capacity)
  float delta;         // relative half-width of the next bracket
  uint32_t overflow;   // some workgroup had more candidates than its LDS buffer
  float prev;          // the previous selected value (0: none yet)
  uint32_t bar[2];     // grid barrier of the digit passes: arrivals, abort flag
};
static_assert(sizeof(Bracket) <= 256, "the bracket state has 256 bytes of the run workspace");

constexpr uint32_t kBracketCap = 1u << 20;  // candidate capacity (4 MB)
constexpr uint32_t kSmallCap = 32768;       // candidates selected by one workgroup in LDS (128 KB)
constexpr uint32_t kTargetCand = 4096;      // the bracket width aims at about this many candidates

// ---------------------------------------------------------------- pairwise squared distances
// As the Stein kernel: a workgroup owns 128 particles i (32 per wave, operands in registers) and a
// share of the j blocks, each staged once through LDS (A layout and norms only: 16.5 KB at d <= 128,
// double buffered) for its four waves. d2[j][i] = |x_i|^2 + |x_j|^2 - 2 x_i.x_j (>= 0, exactly 0 on
// the diagonal). upper != 0: only j blocks up to the last diagonal block of the workgroup.
// One workgroup per task: a group of 128 particles i and a run of kDistJB j blocks. With upper != 0
// group gi has 4 (gi + 1) j blocks (on and above its diagonal blocks), gi + 1 tasks: the triangle is
// cut into equal tasks, numbered row by row.
constexpr int kDistJB = kPhiI / 32;

// |x_i - x_j|^2 from the norms and the Gram entry, clamped at 0; one explicit multiply-add (contraction
// off), so that every use rounds alike
MOCAT_DEV float pair_d2(float ni, float nj, float z) {
#pragma clang fp contract(off)
  return fmaxf(fmaf(-2.f, z, ni + nj), 0.f);
}

// COUNT: also the bracket count of the pairs i > j (the strict upper triangle of the selection):
// pairs below br->lo counted, pairs in [lo, hi) gathered in LDS and appended to cand with one global
// atomic per workgroup (a workgroup with more than its LDS holds marks the bracket unusable).
template <int DP, bool COUNT>
__global__ __launch_bounds__(kPhiThreads) void sqdist_kernel(const char* __restrict__ packed, int n, int upper,
                                                             float* __restrict__ d2, Bracket* br,
                                                             uint32_t* __restrict__ cand) {
  using L = SvgdLayout<DP>;
  constexpr int KB = L::KB;
  constexpr int ABYTES = L::A_ELEMS * 2;  // A layout of a block; the norms follow the B layout
  constexpr int CHUNKS = ABYTES / 16 + 8;  // + the 32 norms (128 B)
  constexpr int PER = (CHUNKS + kPhiThreads - 1) / kPhiThreads;
  __shared__ u32x4 stage[2][CHUNKS];
  constexpr int kLocal = COUNT ? 2048 : 1;
  __shared__ uint32_t lbuf[kLocal];
  __shared__ uint32_t lcount, lbelow, lbase;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave, g = lane >> 5, col = lane & 31;
  uint32_t lo = 0u, hi = 0u, below = 0u;
  if constexpr (COUNT) {
    lo = br->lo;
    hi = br->hi;
    if (t == 0) {
      lcount = 0u;
      lbelow = 0u;
    }
  }
  const int nrb = (n + 31) / 32;
  const int tpg = (nrb + kDistJB - 1) / kDistJB;  // tasks per group, full matrix
  int gi, task;
  const int k = blockIdx.x;
  if (upper) {
    // k = gi (gi + 1) / 2 + task, task <= gi
    gi = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((int64_t)(gi + 1) * (gi + 2) / 2 <= k) ++gi;
    while ((int64_t)gi * (gi + 1) / 2 > k) --gi;
    task = k - gi * (gi + 1) / 2;
  } else {
    gi = k / tpg;
    task = k - gi * tpg;
  }
  const int i0 = gi * kPhiI + w * 32;
  const int rb_all = upper ? min(nrb, (gi + 1) * (kPhiI / 32)) : nrb;
  const int rb0 = min(rb_all, task * kDistJB), rb1 = min(rb_all, rb0 + kDistJB);

  bf16x8 bih[KB], bil[KB];
  const float ni = load_i_operands<DP>(packed, i0 / 32, nrb, lane, bih, bil);

  // chunk c < ABYTES / 16: A layout; the last 8 chunks: the norms
  auto src_chunk = [&](int rb, int c) -> const u32x4* {
    const char* blk = packed + (int64_t)rb * L::BLOCK_BYTES;
    return c < ABYTES / 16 ? reinterpret_cast<const u32x4*>(blk) + c
                           : reinterpret_cast<const u32x4*>(blk + L::BLOCK_BF16 * 2) + (c - ABYTES / 16);
  };
  // blocks copied global -> LDS directly (16 bytes per lane: lane l of wave w copies chunk
  // p * kPhiThreads + w * 64 + l to the same LDS chunk); the next block's copy overlaps the current
  // block's products and is waited for before the step's closing barrier
  auto stage_block = [&](int buf, int rb) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int c = p * kPhiThreads + t;
      if (c < CHUNKS)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src_chunk(rb, c),
                                         (__attribute__((address_space(3))) void*)(&stage[buf][c - lane]), 16, 0,
                                         0);
    }
  };
  const int nsteps = rb1 - rb0;
  if (nsteps > 0) stage_block(0, rb0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    if (st + 1 < nsteps) stage_block((st + 1) & 1, rb0 + st + 1);
    const char* blk = reinterpret_cast<const char*>(stage[st & 1]);
    const bf16x8* A = reinterpret_cast<const bf16x8*>(blk);
    floatx16 z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.f;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const bf16x8 ah = A[(0 * KB + kb) * 64 + lane];
      const bf16x8 al = A[(1 * KB + kb) * 64 + lane];
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bih[kb], z, 0, 0, 0);
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bil[kb], z, 0, 0, 0);
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bih[kb], z, 0, 0, 0);
    }
    const float* nj = reinterpret_cast<const float*>(blk + ABYTES) + g * 16;
    const int jbase = (rb0 + st) * 32, i = i0 + col;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int j = jbase + acc_row(r, lane);
      const float D = (i == j) ? 0.f : pair_d2(ni, nj[r], z[r]);
      if (i < n && j < n) d2[(int64_t)j * n + i] = D;
      if constexpr (COUNT) {
        const uint32_t u = __float_as_uint(D);
        const bool valid = i > j && i < n;
        below += (valid && u < lo) ? 1u : 0u;
        const bool in = valid && u >= lo && u < hi;
        const uint64_t m = __ballot(in);
        if (m) {  // one LDS atomic per wave for the pairs of this row of the tile
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(&lcount, (uint32_t)__popcll(m));
          base = (uint32_t)__shfl((int)base, 0, kWave);
          if (in) {
            const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (idx < (uint32_t)kLocal) lbuf[idx] = u;
          }
        }
      }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  if constexpr (COUNT) {
    below = wave_sum(below);
    if (lane == 0 && below) atomicAdd(&lbelow, below);
    __syncthreads();
    const uint32_t cnt = lcount;
    if (t == 0) {
      if (lbelow) atomicAdd(&br->below[blockIdx.x % kBelowSlots], lbelow);
      lbase = cnt ? atomicAdd(&br->inside, cnt) : 0u;
      if (cnt > (uint32_t)kLocal) atomicOr(&br->overflow, 1u);
    }
    __syncthreads();
    const uint32_t gbase = lbase, keep = cnt < (uint32_t)kLocal ? cnt : (uint32_t)kLocal;
    for (uint32_t k = t; k < keep; k += kPhiThreads)
      if (gbase + k < kBracketCap) cand[gbase + k] = lbuf[k];
  }
}

// ---------------------------------------------------------------- exact radix select (2 ranks)
// state[0..1] prefix bits, state[2..3] remaining rank, hist[2][256]
// One radix pass of the two-rank selection: histogram of the byte at `shift` over the values whose
// higher bytes equal each rank's prefix so far. When both ranks share the prefix (always on the top
// byte, and usually after it for adjacent ranks) one histogram is counted and written to both.
__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t* __restrict__ v, int64_t m, int shift,
                                                         const uint32_t* __restrict__ state,
                                                         uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[2][256];
  const uint32_t mask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
  const uint32_t p0 = state[0] & mask, p1 = state[1] & mask;
  const bool shared_prefix = p0 == p1;
  const int lane = threadIdx.x & (kWave - 1);
  for (int b = threadIdx.x; b < 512; b += 256) (&lh[0][0])[b] = 0u;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // block-uniform trip count, so every lane of a wave reaches the ballots
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < m; base += stride) {
    const int64_t e = base + threadIdx.x;
    const bool valid = e < m;
    const uint32_t u = valid ? v[e] : 0u;
    const uint32_t dig = (u >> shift) & 255u;
    const bool a0 = valid && (u & mask) == p0;
    wave_multisplit_count(a0, dig, lh[0], lane);
    if (!shared_prefix) wave_multisplit_count(valid && (u & mask) == p1, dig, lh[1], lane);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 512; b += 256) {
    const uint32_t c = shared_prefix ? lh[0][b & 255] : (&lh[0][0])[b];
    if (c) atomicAdd(&hist[b], c);
  }
}

// Picks each rank's digit from the pass histogram (block prefix sums over the 256 bins), records it in
// the rank's prefix, reduces the remaining rank, and clears the histogram for the next pass.
__global__ __launch_bounds__(256) void radix_choose_kernel(int shift, uint32_t* __restrict__ state,
                                                           uint32_t* __restrict__ hist) {
  __shared__ uint32_t wtot[2][4];
  const int b = threadIdx.x, lane = b & (kWave - 1), w = b / kWave;
  uint32_t k[2], c[2], inc[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    k[t] = state[2 + t];
    c[t] = hist[t * 256 + b];
    inc[t] = wave_inclusive_scan(c[t]);
    if (lane == kWave - 1) wtot[t][w] = inc[t];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    uint32_t off = 0;
    for (int q = 0; q < w; ++q) off += wtot[t][q];
    const uint32_t lo = off + inc[t] - c[t], hi = off + inc[t];
    // exactly one bin holds rank k (the counts sum to the number of candidates, which exceeds k)
    if (k[t] >= lo && k[t] < hi) {
      state[t] = (state[t] & ~(255u << shift)) | ((uint32_t)b << shift);
      state[2 + t] = k[t] - lo;
    }
    hist[t * 256 + b] = 0u;
  }
}

__global__ void select_init_kernel(uint32_t* __restrict__ ws, uint32_t rank0, uint32_t rank1) {
  for (int b = threadIdx.x; b < 4 + 512; b += blockDim.x) ws[b] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    ws[2] = rank0;
    ws[3] = rank1;
  }
}

// mean of sqrt(d2) over all entries, FP64 accumulation (one value per block, summed on the host side
// by a second single-block pass)
__global__ __launch_bounds__(256) void sqrt_sum_kernel(const float* __restrict__ d2, int64_t m,
                                                       double* __restrict__ partial) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    s += (double)sqrtf(d2[e]);
  s = block_sum(s, sc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Selection state of the fused run: st[k] (k = 0..4) = {prefix0, prefix1, rank0, rank1} before radix
// pass k (st[4]: the selected bit patterns), then one 2 x 256-bin histogram per pass.
constexpr int kSelStateWords = 5 * 4;
constexpr int kSelWords = kSelStateWords + 4 * 512;

// The digit choice of a finished pass, by threads 0..255 of the calling workgroup (bin = thread):
// each rank's bin from prefix sums over the pass histogram, recorded in the rank's prefix and the rank
// reduced (exactly one bin holds each rank). out (LDS) receives the next state; every thread of the
// workgroup calls this (it holds barriers).
MOCAT_DEV void radix_choose_block(int shift, const uint32_t* __restrict__ st_in, const uint32_t* __restrict__ hist,
                                  uint32_t* out, uint32_t (*wtot)[4]) {
  const int b = threadIdx.x, lane = b & (kWave - 1), w = b / kWave;
  uint32_t c[2] = {0u, 0u}, inc[2] = {0u, 0u};
  if (b < 4) out[b] = st_in[b];
  if (b < 256) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      c[t] = hist[t * 256 + b];
      inc[t] = wave_inclusive_scan(c[t]);
      if (lane == kWave - 1) wtot[t][w] = inc[t];
    }
  }
  __syncthreads();
  if (b < 256) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      uint32_t off = 0;
      for (int q = 0; q < w; ++q) off += wtot[t][q];
      const uint32_t lo = off + inc[t] - c[t], hi = off + inc[t], k = st_in[2 + t];
      if (k >= lo && k < hi) {
        out[t] = (st_in[t] & ~(255u << shift)) | ((uint32_t)b << shift);
        out[2 + t] = k - lo;
      }
    }
  }
  __syncthreads();
}

// Both order statistics of the m <= kSmallCap bracket candidates by one workgroup: the four 8-bit digit
// passes over the candidates (read from global memory in every pass, eight loads in flight per thread;
// they stay in L2) with LDS histograms and the same digit choice as the multi-workgroup passes. One LDS
// atomic per candidate: on one compute unit a wave multisplit's ballots cost more vector issue than
// the same-bin conflicts they avoid. Digits inside the common prefix of lo and hi - 1 are shared by
// every candidate (they lie in [lo, hi)) and are taken without a scan. st (LDS, 2 x 4 words) starts as
// {0, 0, rank0, rank1} in st[0..3]; the result (the two bit patterns, then the remaining ranks) ends in
// st[0..3]. Every thread of the workgroup calls this.
MOCAT_DEV void select_small_block(const uint32_t* __restrict__ cand, uint32_t m, uint32_t lo, uint32_t hi,
                                  uint32_t* lh, uint32_t* st, uint32_t (*wtot)[4]) {
  const uint32_t diff = lo ^ (hi - 1u);
  const int common = diff ? __clz(diff) : 32;  // leading bits shared by every candidate
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    const uint32_t mask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
    uint32_t* in = st + (pass & 1) * 4;
    uint32_t* out = st + ((pass + 1) & 1) * 4;
    if (common >= 32 - shift) {
      // the digit is lo's for every candidate: prefixes extended, ranks unchanged
      if (threadIdx.x < 2) out[threadIdx.x] = in[threadIdx.x] | (lo & (255u << shift));
      if (threadIdx.x >= 2 && threadIdx.x < 4) out[threadIdx.x] = in[threadIdx.x];
      __syncthreads();
      continue;
    }
    for (int b = threadIdx.x; b < 512; b += blockDim.x) lh[b] = 0u;
    __syncthreads();
    const uint32_t p0 = in[0] & mask, p1 = in[1] & mask;
    const bool shared_prefix = p0 == p1;
    for (uint32_t k0 = 0; k0 < m; k0 += 8 * blockDim.x) {
      uint32_t u[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t k = k0 + q * blockDim.x + threadIdx.x;
        u[q] = k < m ? cand[k] : 0u;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t k = k0 + q * blockDim.x + threadIdx.x;
        if (k < m) {
          const uint32_t dig = (u[q] >> shift) & 255u;
          if ((u[q] & mask) == p0) atomicAdd(&lh[dig], 1u);
          if (!shared_prefix && (u[q] & mask) == p1) atomicAdd(&lh[256 + dig], 1u);
        }
      }
    }
    __syncthreads();
    if (shared_prefix) {
      for (int b = threadIdx.x; b < 256; b += blockDim.x) lh[256 + b] = lh[b];
      __syncthreads();
    }
    radix_choose_block(shift, in, lh, out, wtot);
  }
  // four passes: the result is back in st[0..3]
}

// Bounded grid-wide barrier of the bracketed passes: every workgroup adds one to the arrival counter
// bar[0] after a release fence; barrier k is passed once the counter reaches k * gridDim.x, followed by
// an acquire fence. A workgroup that polls max_polls times without seeing every arrival sets the abort
// flag bar[1], and every workgroup leaves at its next barrier once the flag is set. The grid (at most
// kSelectBlocks workgroups of kBracketThreads threads, 2 KB of LDS each, and no more than the device's
// compute units times the kernel's occupancy) is resident at once, so a barrier only ever waits on
// running workgroups.
constexpr uint32_t kBracketBarrierPolls = 1u << 22;

MOCAT_DEV bool bracket_grid_sync(uint32_t* bar, uint32_t k, uint32_t max_polls) {
  __shared__ int pass_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = k * gridDim.x;
    int ok = 1;
    for (uint32_t polls = 0;; ++polls) {
      if (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = 0;
        break;
      }
      if (polls >= max_polls) {
        __hip_atomic_store(&bar[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      if (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    pass_ok = ok;
  }
  __syncthreads();
  return pass_ok != 0;
}

// The bracketed selection of the two middle order statistics in one launch, then the bandwidth and
// the next bracket. The decision (every workgroup alike, from the finished count pass): on the
// candidates, with the ranks reduced by the count below, when both ranks lie inside the bracket and the
// candidates fit, else on all pairs (upper triangle). With at most kSmallCap candidates (the usual
// case) workgroup 0 selects them alone and every other workgroup returns at once; otherwise the four
// digit passes (shift 24 - 8 pass) run over the whole grid (at most kSelectBlocks workgroups, all
// resident), with a grid barrier before each later pass and one after the last, each workgroup applying
// a finished pass's digit choice itself from the pass's global histogram. Workgroup 0 then stores the
// median bandwidth (FP64 square roots of the two order statistics, their mean, divided by
// sqrt(2 log n) computed on the host, rounded to f32; NaN when a barrier timed out) and the next bracket.
// The bracket counters, the barrier and the histograms are cleared by the next iteration's update
// kernel (svgd_update_gmm_kernel), which runs after every workgroup of this launch has ended: a
// workgroup of this launch may still be reading them when workgroup 0 finishes.
__global__ __launch_bounds__(kBracketThreads) void bracket_select_kernel(const uint32_t* __restrict__ v, int n,
                                                                     const uint32_t* __restrict__ cand, Bracket* br,
                                                                     uint32_t* __restrict__ sel, uint32_t rank0,
                                                                     uint32_t rank1, double norm, float* h,
                                                                     float* trace, uint32_t max_polls) {
  __shared__ uint32_t lh[2][256];
  __shared__ uint32_t wtot[2][4];
  __shared__ uint32_t st[2][4];
  __shared__ Bracket bs;
  if (threadIdx.x == 0) bs = *br;
  __syncthreads();
  uint32_t below = 0u;
  for (int k = 0; k < kBelowSlots; ++k) below += bs.below[k];
  const uint32_t inside = bs.inside;
  const bool use_cand = !bs.overflow && below <= rank0 && (uint64_t)rank1 < (uint64_t)below + inside &&
                        inside <= kBracketCap;
  const bool small = use_cand && inside <= kSmallCap;
  if (threadIdx.x == 0) {
    st[0][0] = 0u;
    st[0][1] = 0u;
    st[0][2] = use_cand ? rank0 - below : rank0;
    st[0][3] = use_cand ? rank1 - below : rank1;
  }
  __syncthreads();
  bool ok = true;
  if (small) {
    if (blockIdx.x != 0) return;
    select_small_block(cand, inside, bs.lo, bs.hi, &lh[0][0], &st[0][0], wtot);  // result in st[0]
  } else {
    const int lane = threadIdx.x & (kWave - 1);
    for (int pass = 0; pass < 4 && ok; ++pass) {
      const int shift = 24 - 8 * pass;
      uint32_t* hist = sel + kSelStateWords + pass * 512;
      uint32_t* cur = st[pass & 1];
      if (pass > 0) {
        if (!bracket_grid_sync(br->bar, (uint32_t)pass, max_polls)) {
          ok = false;
          break;
        }
        radix_choose_block(shift + 8, st[(pass - 1) & 1], hist - 512, cur, wtot);
      }
      const uint32_t mask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
      const uint32_t p0 = cur[0] & mask, p1 = cur[1] & mask;
      const bool shared_prefix = p0 == p1;
      for (int b = threadIdx.x; b < 512; b += kBracketThreads) (&lh[0][0])[b] = 0u;
      __syncthreads();
      if (use_cand) {
        const int64_t m = inside;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < m; base += stride) {
          const int64_t e = base + threadIdx.x;
          const bool valid = e < m;
          const uint32_t u = valid ? cand[e] : 0u;
          const uint32_t dig = (u >> shift) & 255u;
          wave_multisplit_count(valid && (u & mask) == p0, dig, lh[0], lane);
          if (!shared_prefix) wave_multisplit_count(valid && (u & mask) == p1, dig, lh[1], lane);
        }
      } else {
        for (int row = blockIdx.x; row < n - 1; row += gridDim.x) {
          const uint32_t* vr = v + (int64_t)row * n;
          for (int c0 = row + 1; c0 < n; c0 += kBracketUnroll * kBracketThreads) {
            uint32_t uu[kBracketUnroll];
#pragma unroll
            for (int k = 0; k < kBracketUnroll; ++k) {
              const int c = c0 + k * kBracketThreads + (int)threadIdx.x;
              uu[k] = c < n ? vr[c] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kBracketUnroll; ++k) {
              const bool valid = c0 + k * kBracketThreads + (int)threadIdx.x < n;
              const uint32_t u = uu[k];
              const uint32_t dig = (u >> shift) & 255u;
              wave_multisplit_count(valid && (u & mask) == p0, dig, lh[0], lane);
              if (!shared_prefix) wave_multisplit_count(valid && (u & mask) == p1, dig, lh[1], lane);
            }
          }
        }
      }
      __syncthreads();
      for (int b = threadIdx.x; b < 512; b += kBracketThreads) {
        const uint32_t c = shared_prefix ? lh[0][b & 255] : (&lh[0][0])[b];
        if (c) atomicAdd(&hist[b], c);
      }
    }
    if (ok && !bracket_grid_sync(br->bar, 4u, max_polls)) ok = false;
    if (blockIdx.x != 0) return;
    // the last digit choice: the two order statistics in st[0]
    if (ok) radix_choose_block(0, st[1], sel + kSelStateWords + 3 * 512, st[0], wtot);
  }
  if (threadIdx.x != 0) return;
  for (int t = 0; t < 4; ++t) sel[16 + t] = st[0][t];
  const double v0 = (double)__uint_as_float(st[0][0]), v1 = (double)__uint_as_float(st[0][1]);
  const float hv = ok ? (float)(0.5 * (sqrt(v0) + sqrt(v1)) / norm) : __builtin_nanf("");
  *h = hv;
  if (trace) *trace = hv;
  // half-width: narrower when the candidates overflowed or are many, wider after a miss or when very
  // few (aiming at about kTargetCand); the bracket is centred on the linear extrapolation of the last
  // two values (the median drifts smoothly from one SVGD iteration to the next: measured relative
  // extrapolation errors of a few 1e-6, against drifts of about 1e-4 per iteration)
  float delta = bs.delta;
  if (bs.overflow || inside > kBracketCap)
    delta = fmaxf(delta * 0.5f, 1e-7f);
  else if (!use_cand)
    delta = fminf(delta * 4.f, 1.f);
  else if (inside > 2 * kTargetCand)
    delta = fmaxf(delta * 0.7f, 1e-7f);
  else if (inside < kTargetCand / 4)
    delta = fminf(delta * 1.25f, 1.f);
  const float center = ok ? __uint_as_float(st[0][1]) : bs.prev, prev = bs.prev;
  float pred = prev > 0.f ? center + (center - prev) : center;
  if (!(pred > 0.f)) pred = center;
  br->prev = center;
  const float lo = pred * (1.f - delta), hi = pred * (1.f + delta);
  br->lo = __float_as_uint(fmaxf(lo, 0.f));
  br->hi = hi > pred ? __float_as_uint(hi) : 0x7F800001u;  // above every finite value
  br->delta = delta;
}

// Clears the bracket counters, the barrier and the selection histograms for the next selection (by
// workgroup 0 of the update kernel, which runs between one selection and the next count pass).
MOCAT_DEV void bracket_clear_block(Bracket* br, uint32_t* sel) {
  for (int b = threadIdx.x; b < kSelWords - kSelStateWords; b += blockDim.x) sel[kSelStateWords + b] = 0u;
  if (threadIdx.x == 0) {
    for (int k = 0; k < kBelowSlots; ++k) br->below[k] = 0u;
    br->inside = 0u;
    br->overflow = 0u;
    br->bar[0] = 0u;
    br->bar[1] = 0u;
  }
}

// First bracket: empty (no candidates: the first selection runs over all pairs); histograms cleared.
__global__ __launch_bounds__(256) void bracket_init_kernel(Bracket* br, uint32_t* __restrict__ sel) {
  for (int b = threadIdx.x; b < kSelWords; b += blockDim.x) sel[b] = 0u;
  if (threadIdx.x == 0) {
    br->lo = 0u;
    br->hi = 0u;
    for (int k = 0; k < kBelowSlots; ++k) br->below[k] = 0u;
    br->inside = 0u;
    br->overflow = 0u;
    br->prev = 0.f;
    br->delta = 2e-3f;
    br->bar[0] = 0u;
    br->bar[1] = 0u;
  }
}

// ---------------------------------------------------------------- fused SVGD run helpers
// The adagrad step of svgd_update_kernel and the mixture potential and gradient of gmm_kernel (times
// the temperature T) in one pass, one wave per particle: the same operations in the same order, on
// the updated row held in registers. Workgroup 0 also advances the key: key <- split(key, num)[idx]
// (SVGD.update's key advance), and clears the median's bracket state when br is given.
__global__ __launch_bounds__(256) void svgd_update_gmm_kernel(
    const float* __restrict__ part_o, const float* __restrict__ part_s, int js_count, int n, int d, int DP,
    const float* __restrict__ mean, float inv_n, const float* __restrict__ h_dev, float step, float momentum,
    float* __restrict__ x, float* __restrict__ gsq, float* __restrict__ mom, int K, const float* __restrict__ mu,
    const float* __restrict__ lam, const float* __restrict__ logc, float T, float* __restrict__ pot,
    float* __restrict__ grad, uint32_t* key, uint32_t knum, uint32_t kidx, Bracket* br, uint32_t* sel) {
#pragma clang fp contract(off)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const Key k = split_at(Key{key[0], key[1]}, knum, kidx);
    key[0] = k.a;
    key[1] = k.b;
  }
  if (blockIdx.x == 0 && br) bracket_clear_block(br, sel);
  // the mixture parameters (K * d <= 1024 each), staged once per workgroup
  __shared__ float smu[8 * 128], slam[8 * 128];
  {
    const int kd = K * d;
    float a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = u * 256 + (int)threadIdx.x;
      a[u] = e < kd ? mu[e] : 0.f;
      b[u] = e < kd ? lam[e] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = u * 256 + (int)threadIdx.x;
      if (e < kd) {
        smu[e] = a[u];
        slam[e] = b[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= n) return;
  const float hb = *h_dev;
  const float inv_h2 = 1.f / (hb * hb);
  // the row's coordinates, adagrad state and centring mean (columns lane and lane + 64, d <= 128),
  // loaded ahead of the partial sums
  float xv[2], gv[2], mv[2], cv[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k = lane + q * kWave;
    const bool in = k < d;
    const int64_t e = i * d + k;
    xv[q] = in ? x[e] : 0.f;
    gv[q] = in ? gsq[e] : 0.f;
    mv[q] = in ? mom[e] : 0.f;
    cv[q] = in ? mean[k] : 0.f;
  }
  // the j-split partial sums, added in split order; loads issued 8 splits ahead of the additions
  constexpr int U = 8;
  float kg[2] = {0.f, 0.f}, kx[2] = {0.f, 0.f}, s = 0.f;
  for (int j0 = 0; j0 < js_count; j0 += U) {
    float vg[U][2], vx[U][2], vs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = j0 + u < js_count;
      const float* o = part_o + ((int64_t)(j0 + u) * n + i) * 2 * DP;
      vs[u] = in ? part_s[(int64_t)(j0 + u) * n + i] : 0.f;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int k = lane + q * kWave;
        vg[u][q] = in && k < d ? o[k] : 0.f;
        vx[u][q] = in && k < d ? o[DP + k] : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j0 + u >= js_count) break;
      s += vs[u];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        kg[q] += vg[u][q];
        kx[q] += vx[u][q];
      }
    }
  }
  float xn[2] = {0.f, 0.f};  // d <= 128: columns lane and lane + 64
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k = lane + q * kWave;
    if (k >= d) continue;
    const int64_t e = i * d + k;
    const AdagradOut u =
        svgd_adagrad(xv[q], xv[q] - cv[q], kg[q], kx[q], s, gv[q], mv[q], inv_n, inv_h2, step, momentum);
    gsq[e] = u.G;
    mom[e] = u.m;
    xn[q] = u.x;
    x[e] = u.x;
  }
  float qk[8];
  float qmax = -INFINITY;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c >= K) break;
    float a = 0.f;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int j = lane + q * kWave;
      if (j < d) {
        const float df = xn[q] - smu[c * d + j];
        a = fmaf(slam[c * d + j] * df, df, a);
      }
    }
    a = wave_sum(a);
    qk[c] = fmaf(-0.5f, a, logc[c]);
    qmax = fmaxf(qmax, qk[c]);
  }
  float ssum = 0.f;
  for (int c = 0; c < K; ++c) ssum += __expf(qk[c] - qmax);
  if (lane == 0) {
    const float u = -(qmax + __logf(ssum));
    pot[i] = T != 1.f ? u * T : u;
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = lane + q * kWave;
    if (j >= d) continue;
    float gsum = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (c >= K) break;
      gsum = fmaf(__expf(qk[c] - qmax) / ssum * slam[c * d + j], xn[q] - smu[c * d + j], gsum);
    }
    grad[i * d + j] = T != 1.f ? gsum * T : gsum;
  }
}

// mean bandwidth from the FP64 partial sums of the pair distances: 2 sum / n^2 / sqrt(2 log n)
__global__ void mean_bandwidth_kernel(const double* partial, int n_partial, double inv_nn, double norm, float* h,
                                      float* trace) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < n_partial; b += blockDim.x) s += partial[b];
  s = block_sum(s, sc);
  if (threadIdx.x == 0) {
    const float hv = (float)(2.0 * s * inv_nn / norm);
    *h = hv;
    if (trace) *trace = hv;
  }
}

__global__ void copy_bandwidth_kernel(const float* h, float* trace) {
  if (threadIdx.x == 0) *trace = *h;
}

// centring mean and the packed operands of x (A layout, norms) and [grad | x] (B layout)
template <int DP>
void svgd_prepare_dp(const float* x, const float* grad, int n, int d, void* packed, float* mean, hipStream_t st) {
  const int nrb = (n + 31) / 32;
  // column-mean partials live in the partial-sum region of the step workspace, which the Stein kernel
  // writes only after the packing: slices * d doubles <= n * d * 8 bytes < its n * (2 * DP + 1) * 4
  const int slices = std::max(1, std::min(kMeanSlices, (n + 63) / 64));
  double* cpart = reinterpret_cast<double*>(reinterpret_cast<char*>(mean) + 512);
  hipLaunchKernelGGL(colmean_part_kernel, dim3((d + kWave - 1) / kWave, slices), dim3(256), 0, st, x, n, d, slices,
                     cpart);
  hipLaunchKernelGGL(svgd_pack_kernel<DP>, dim3((unsigned)nrb), dim3(256), 0, st, x, grad, (const double*)cpart,
                     slices, mean, n, d, (char*)packed);
}

// Stein direction from the prepared operands of x_in, then the adagrad step of x
template <int DP>
void svgd_phi_update_dp(const float* x_in, int n, int d, float bandwidth, const float* bandwidth_dev, int js_count,
                        const void* packed, const float* mean, float* part_o, float* part_s, float step,
                        float momentum, float* x, float* gsq, float* mom, float* phi_out, hipStream_t st) {
  hipLaunchKernelGGL((svgd_phi_kernel<DP>), dim3((unsigned)((n + kPhiI - 1) / kPhiI), (unsigned)js_count),
                     dim3(kPhiThreads), 0, st, (const char*)packed, n, bandwidth, bandwidth_dev, js_count, part_o,
                     part_s);
  const int64_t tot2 = (int64_t)n * d;
  const unsigned ub = (unsigned)std::min<int64_t>((tot2 + 255) / 256, 65535);
  hipLaunchKernelGGL(svgd_update_kernel, dim3(ub), dim3(256), 0, st, part_o, part_s, js_count, n, d, DP, mean,
                     1.f / (float)n, bandwidth, bandwidth_dev, step, momentum, x, gsq, mom, phi_out);
}

template <int DP>
int svgd_step_dp(const float* x_in, const float* grad, int n, int d, float bandwidth, const float* bandwidth_dev,
                 int js_count, void* packed, float* mean, float* part_o, float* part_s, float step, float momentum,
                 float* x, float* gsq, float* mom, float* phi_out, hipStream_t st) {
  svgd_prepare_dp<DP>(x_in, grad, n, d, packed, mean, st);
  svgd_phi_update_dp<DP>(x_in, n, d, bandwidth, bandwidth_dev, js_count, packed, mean, part_o, part_s, step,
                         momentum, x, gsq, mom, phi_out, st);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// squared distances from prepared operands (upper != 0: the tiles on and above the diagonal)
template <int DP>
void sqdist_packed_dp(const void* packed, int n, int upper, float* d2, Bracket* br, uint32_t* cand, hipStream_t st) {
  const int groups = (n + kPhiI - 1) / kPhiI, nrb = (n + 31) / 32;
  const int64_t tasks = upper ? (int64_t)groups * (groups + 1) / 2
                              : (int64_t)groups * ((nrb + kDistJB - 1) / kDistJB);
  if (br)
    hipLaunchKernelGGL((sqdist_kernel<DP, true>), dim3((unsigned)tasks), dim3(kPhiThreads), 0, st,
                       (const char*)packed, n, upper, d2, br, cand);
  else
    hipLaunchKernelGGL((sqdist_kernel<DP, false>), dim3((unsigned)tasks), dim3(kPhiThreads), 0, st,
                       (const char*)packed, n, upper, d2, (Bracket*)nullptr, (uint32_t*)nullptr);
}

template <int DP>
int sqdist_dp(const float* x, int n, int d, void* packed, float* mean, int upper, float* d2, hipStream_t st) {
  // the packed B layout needs a gradient: pass x again (only the A layout and norms are read here)
  svgd_prepare_dp<DP>(x, x, n, d, packed, mean, st);
  sqdist_packed_dp<DP>(packed, n, upper, d2, nullptr, nullptr, st);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int64_t packed_bytes(int n, int d) {
  const int64_t nrb = (n + 31) / 32;
  return nrb * (d <= 64 ? SvgdLayout<64>::BLOCK_BYTES : SvgdLayout<128>::BLOCK_BYTES);
}

}  // namespace

extern "C" {

int mocat_svgd_workspace_bytes(int64_t n, int64_t d, int js_count, int64_t* bytes) {
  if (n <= 0 || d <= 0 || d > 128 || js_count <= 0 || !bytes) return MOCAT_ERR_ARG;
  const int DP = d <= 64 ? 64 : 128;
  const int64_t pk = (packed_bytes((int)n, (int)d) + 255) / 256 * 256;
  const int64_t part = (int64_t)js_count * n * (2 * DP + 1) * 4;
  *bytes = pk + 512 + part;  // packed | mean (<= 128 floats) | partial sums
  return MOCAT_OK;
}

int mocat_svgd_step(const float* x_in, const float* grad, int64_t n, int64_t d, float bandwidth,
                    const float* bandwidth_dev, int js_count, void* workspace, float step, float momentum, float* x,
                    float* gsq, float* mom, float* phi_out, hipStream_t stream) {
  if (!x_in || !grad || !workspace || !x || !gsq || !mom || n <= 0 || n > (1 << 26) || d <= 0 || d > 128 ||
      js_count <= 0 || js_count > 64 || (!bandwidth_dev && !(bandwidth > 0.f)))
    return MOCAT_ERR_ARG;
  const int DP = d <= 64 ? 64 : 128;
  char* ws = (char*)workspace;
  const int64_t pk = (packed_bytes((int)n, (int)d) + 255) / 256 * 256;
  float* mean = (float*)(ws + pk);
  float* part_o = (float*)(ws + pk + 512);
  float* part_s = part_o + (int64_t)js_count * n * 2 * DP;
  if (DP == 64)
    return svgd_step_dp<64>(x_in, grad, (int)n, (int)d, bandwidth, bandwidth_dev, js_count, ws, mean, part_o, part_s,
                            step, momentum, x, gsq, mom, phi_out, stream);
  return svgd_step_dp<128>(x_in, grad, (int)n, (int)d, bandwidth, bandwidth_dev, js_count, ws, mean, part_o, part_s,
                           step, momentum, x, gsq, mom, phi_out, stream);
}

int mocat_gmm_potential_grad(const float* x, int64_t n, int64_t d, int K, const float* means, const float* lam,
                             const float* logc, float* potential, float* grad, hipStream_t stream) {
  if (!x || !means || !lam || !logc || !potential || !grad || n <= 0 || d <= 0 || K <= 0 || K > 8) return MOCAT_ERR_ARG;
  const int per = 256 / kWave;
  hipLaunchKernelGGL(gmm_kernel, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, stream, x, (int)n, (int)d, K,
                     means, lam, logc, potential, grad);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_gmm_full_potential_grad(const float* x, int64_t n, int64_t d, int K, const float* means,
                                  const float* S, const float* logc, float* potential, float* grad,
                                  hipStream_t stream) {
  if (!x || !means || !S || !logc || !potential || !grad || n <= 0 || d <= 0 || d > 128 || K <= 0 || K > 8)
    return MOCAT_ERR_ARG;
  const int64_t per_block = (256 / kWave) * kGmmFullPer;
  const dim3 grid((unsigned)((n + per_block - 1) / per_block));
  switch (K) {
#define MOCAT_GMM_FULL(KK)                                                                                   \
  case KK:                                                                                                 \
    hipLaunchKernelGGL(gmm_full_kernel<KK>, grid, dim3(256), 0, stream, x, (int)n, (int)d, means, S, logc,   \
                       potential, grad);                                                                    \
    break;
    MOCAT_GMM_FULL(1) MOCAT_GMM_FULL(2) MOCAT_GMM_FULL(3) MOCAT_GMM_FULL(4)
    MOCAT_GMM_FULL(5) MOCAT_GMM_FULL(6) MOCAT_GMM_FULL(7) MOCAT_GMM_FULL(8)
#undef MOCAT_GMM_FULL
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_pairwise_sqdist(const float* x, int64_t n, int64_t d, void* workspace, float* d2, hipStream_t stream) {
  if (!x || !workspace || !d2 || n <= 0 || n > (1 << 16) || d <= 0 || d > 128) return MOCAT_ERR_ARG;
  char* ws = (char*)workspace;
  const int64_t pk = (packed_bytes((int)n, (int)d) + 255) / 256 * 256;
  float* mean = (float*)(ws + pk);
  return d <= 64 ? sqdist_dp<64>(x, (int)n, (int)d, ws, mean, 0, d2, stream)
                 : sqdist_dp<128>(x, (int)n, (int)d, ws, mean, 0, d2, stream);
}

int mocat_pairwise_sqdist_upper(const float* x, int64_t n, int64_t d, void* workspace, float* d2, hipStream_t stream) {
  if (!x || !workspace || !d2 || n <= 0 || n > (1 << 16) || d <= 0 || d > 128) return MOCAT_ERR_ARG;
  char* ws = (char*)workspace;
  const int64_t pk = (packed_bytes((int)n, (int)d) + 255) / 256 * 256;
  float* mean = (float*)(ws + pk);
  return d <= 64 ? sqdist_dp<64>(x, (int)n, (int)d, ws, mean, 1, d2, stream)
                 : sqdist_dp<128>(x, (int)n, (int)d, ws, mean, 1, d2, stream);
}

// Exact order statistics of m non-negative floats (bit patterns ordered as unsigned ints):
// out_bits[t] = bit pattern of the rank[t]-th smallest value (0-based), t = 0, 1.
// workspace: 4 + 512 uint32 (state: prefixes, remaining ranks; then two 256-bin histograms).
int mocat_select2(const float* v, int64_t m, int64_t rank0, int64_t rank1, uint32_t* workspace, hipStream_t stream) {
  if (!v || !workspace || m <= 0 || rank0 < 0 || rank1 < 0 || rank0 >= m || rank1 >= m || m > 0xFFFFFFFFll)
    return MOCAT_ERR_ARG;
  hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(256), 0, stream, workspace, (uint32_t)rank0, (uint32_t)rank1);
  const int64_t blocks = std::min<int64_t>((m + 255) / 256, kSelectBlocks);
  for (int shift = 24; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint32_t*)v, m, shift,
                       workspace, workspace + 4);
    hipLaunchKernelGGL(radix_choose_kernel, dim3(1), dim3(256), 0, stream, shift, workspace, workspace + 4);
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_select2_upper(const float* d2, int64_t n, int64_t rank0, int64_t rank1, uint32_t* workspace,
                        hipStream_t stream) {
  const int64_t m = n * (n - 1) / 2;
  if (!d2 || !workspace || n < 2 || n > (1 << 16) || rank0 < 0 || rank1 < 0 || rank0 >= m || rank1 >= m)
    return MOCAT_ERR_ARG;
  hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(256), 0, stream, workspace, (uint32_t)rank0, (uint32_t)rank1);
  const unsigned blocks = (unsigned)std::min<int64_t>(n - 1, kSelectBlocks);
  for (int shift = 24; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(radix_hist_upper_kernel, dim3(blocks), dim3(kSelThreads), 0, stream, (const uint32_t*)d2, (int)n, shift,
                       workspace, workspace + 4);
    hipLaunchKernelGGL(radix_choose_kernel, dim3(1), dim3(256), 0, stream, shift, workspace, workspace + 4);
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_sqrt_sum_upper(const float* d2, int64_t n, double* partial, int n_partial, hipStream_t stream) {
  if (!d2 || !partial || n < 2 || n > (1 << 16) || n_partial <= 0 || n_partial > 4096) return MOCAT_ERR_ARG;
  hipLaunchKernelGGL(sqrt_sum_upper_kernel, dim3((unsigned)n_partial), dim3(256), 0, stream, d2, (int)n, partial);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_sqrt_mean(const float* d2, int64_t m, double* partial, int n_partial, hipStream_t stream) {
  if (!d2 || !partial || m <= 0 || n_partial <= 0 || n_partial > 4096) return MOCAT_ERR_ARG;
  hipLaunchKernelGGL(sqrt_sum_kernel, dim3((unsigned)n_partial), dim3(256), 0, stream, d2, m, partial);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// Workspace of mocat_svgd_gmm_run: the step workspace, the n x n distance matrix, the selection
// state, FP64 partial sums, 256-byte aligned.
static int64_t svgd_run_layout(int64_t n, int64_t d, int js_count, int64_t* o_d2, int64_t* o_sel, int64_t* o_part) {
  int64_t step_bytes = 0;
  mocat_svgd_workspace_bytes(n, d, js_count, &step_bytes);
  int64_t o = (step_bytes + 255) / 256 * 256;
  *o_d2 = o;
  o += (n * n * 4 + 255) / 256 * 256;
  *o_sel = o;
  o += ((int64_t)kSelWords * 4 + 255) / 256 * 256;
  *o_part = o;
  o += 1024 * 8;
  o = (o + 255) / 256 * 256;  // bracket state (256 B) then the candidate buffer
  o += 256 + (int64_t)kBracketCap * 4;
  return o;
}

int mocat_svgd_gmm_run_workspace_bytes(int64_t n, int64_t d, int js_count, int64_t* bytes) {
  if (n <= 1 || n > (1 << 16) || d <= 0 || d > 128 || js_count <= 0 || js_count > 64 || !bytes) return MOCAT_ERR_ARG;
  int64_t a, b, c;
  *bytes = svgd_run_layout(n, d, js_count, &a, &b, &c);
  return MOCAT_OK;
}

int mocat_svgd_gmm_run(float* x, int64_t n, int64_t d, int K, const float* means, const float* lam, const float* logc,
                       float temperature, float step, float momentum, int bandwidth_mode, float* bandwidth,
                       int n_iter, int js_count, void* workspace, int64_t workspace_bytes, float* gsq, float* mom,
                       float* potential, float* grad, uint32_t* key, float* bandwidth_trace, hipStream_t stream) {
  if (!x || !means || !lam || !logc || !bandwidth || !workspace || !gsq || !mom || !potential || !grad || !key ||
      n <= 1 || n > (1 << 16) || d <= 0 || d > 128 || K <= 0 || K > 8 || n_iter < 0 || js_count <= 0 ||
      js_count > 64 || bandwidth_mode < 0 || bandwidth_mode > 2)
    return MOCAT_ERR_ARG;
  if (bandwidth_mode == 2 && n < 3) return MOCAT_ERR_ARG;
  int64_t o_d2, o_sel, o_part;
  if (workspace_bytes < svgd_run_layout(n, d, js_count, &o_d2, &o_sel, &o_part)) return MOCAT_ERR_SIZE;
  char* ws = static_cast<char*>(workspace);
  float* d2 = reinterpret_cast<float*>(ws + o_d2);
  uint32_t* sel = reinterpret_cast<uint32_t*>(ws + o_sel);
  double* part = reinterpret_cast<double*>(ws + o_part);
  const int64_t o_br = (o_part + 1024 * 8 + 255) / 256 * 256;
  Bracket* br = reinterpret_cast<Bracket*>(ws + o_br);
  uint32_t* cand = reinterpret_cast<uint32_t*>(ws + o_br + 256);
  if (bandwidth_mode == 2) hipLaunchKernelGGL(bracket_init_kernel, dim3(1), dim3(256), 0, stream, br, sel);
  // the barrier wait limit of the digit passes (the MOCAT_SVGD_BARRIER_POLLS environment variable sets
  // another; 0 makes every barrier time out, which tests use to check the NaN bandwidth)
  uint32_t polls = kBracketBarrierPolls;
  if (const char* e = std::getenv("MOCAT_SVGD_BARRIER_POLLS")) polls = (uint32_t)std::strtoul(e, nullptr, 10);
  // the selection's grid: at most one workgroup per row, kSelectBlocks, and what the device holds at once
  // (its grid barriers wait on every workgroup; a partitioned device has fewer compute units)
  unsigned sel_rows = 1;
  if (bandwidth_mode == 2) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(bracket_select_kernel),
                                                     kBracketThreads, 0) != hipSuccess)
      return MOCAT_ERR_HIP;
    if (per_cu < 1 || cus < 1) return MOCAT_ERR_SIZE;
    sel_rows = (unsigned)std::min<int64_t>(std::min<int64_t>(n - 1, kSelectBlocks), (int64_t)cus * per_cu);
  }
  const double norm = std::sqrt(2.0 * std::log((double)n));
  const int64_t m = n * n;
  const int64_t r0 = (m - 1) / 2, r1 = m / 2;  // both >= n for n >= 3
  // the step workspace: packed operands | centring mean | partial sums (as mocat_svgd_step)
  const int DP = d <= 64 ? 64 : 128;
  const int64_t pk = (packed_bytes((int)n, (int)d) + 255) / 256 * 256;
  float* mean = reinterpret_cast<float*>(ws + pk);
  float* part_o = reinterpret_cast<float*>(ws + pk + 512);
  float* part_s = part_o + (int64_t)js_count * n * 2 * DP;
  // operands of the current particles and gradients; refreshed after every update, where they serve the
  // distance matrix of the bandwidth update and the next Stein direction
  auto prepare = [&]() {
    if (DP == 64)
      svgd_prepare_dp<64>(x, grad, (int)n, (int)d, ws, mean, stream);
    else
      svgd_prepare_dp<128>(x, grad, (int)n, (int)d, ws, mean, stream);
  };
  prepare();
  const unsigned phi_groups = (unsigned)((n + kPhiI - 1) / kPhiI), upd_blocks = (unsigned)((n + 3) / 4);
  for (int it = 0; it < n_iter; ++it) {
    // SVGD.update: Stein direction, then the adagrad step in place with the key advance and the
    // potential of the updated particles
    if (DP == 64)
      hipLaunchKernelGGL((svgd_phi_kernel<64>), dim3(phi_groups, (unsigned)js_count), dim3(kPhiThreads), 0, stream,
                         (const char*)ws, (int)n, 1.f, bandwidth, js_count, part_o, part_s);
    else
      hipLaunchKernelGGL((svgd_phi_kernel<128>), dim3(phi_groups, (unsigned)js_count), dim3(kPhiThreads), 0, stream,
                         (const char*)ws, (int)n, 1.f, bandwidth, js_count, part_o, part_s);
    hipLaunchKernelGGL(svgd_update_gmm_kernel, dim3(upd_blocks), dim3(256), 0, stream, part_o, part_s, js_count,
                       (int)n, (int)d, DP, mean, 1.f / (float)n, bandwidth, step, momentum, x, gsq, mom, K, means, lam,
                       logc, temperature, potential, grad, key, (uint32_t)(n + 2), (uint32_t)n,
                       bandwidth_mode == 2 ? br : nullptr, sel);
    prepare();
    float* tr = bandwidth_trace ? bandwidth_trace + it : nullptr;
    if (bandwidth_mode == 0) {
      if (tr) hipLaunchKernelGGL(copy_bandwidth_kernel, dim3(1), dim3(64), 0, stream, bandwidth, tr);
      continue;
    }
    // distances (and, for the median, the bracket count of the pairs as they are written)
    Bracket* count_br = bandwidth_mode == 2 ? br : nullptr;
    if (DP == 64)
      sqdist_packed_dp<64>(ws, (int)n, 1, d2, count_br, cand, stream);
    else
      sqdist_packed_dp<128>(ws, (int)n, 1, d2, count_br, cand, stream);
    if (bandwidth_mode == 2) {
      hipLaunchKernelGGL(bracket_select_kernel, dim3(sel_rows), dim3(kBracketThreads), 0, stream, (const uint32_t*)d2,
                         (int)n, (const uint32_t*)cand, br, sel, (uint32_t)((r0 - n) / 2), (uint32_t)((r1 - n) / 2),
                         norm, bandwidth, tr, polls);
    } else {
      const int rc = mocat_sqrt_sum_upper(d2, n, part, 1024, stream);
      if (rc != MOCAT_OK) return rc;
      hipLaunchKernelGGL(mean_bandwidth_kernel, dim3(1), dim3(256), 0, stream, part, 1024,
                         1.0 / ((double)n * (double)n), norm, bandwidth, tr);
    }
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
