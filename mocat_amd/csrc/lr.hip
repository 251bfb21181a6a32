// Bayesian logistic regression on the GPU: fused potential + hand-written adjoint, and the fused
// HMC / MALA / underdamped-Langevin iteration built on it.
//
// Model (a user Scenario in the reference's README pattern, README.md:35-49):
//   prior_potential(theta)      = 0.5 * prec * |theta|^2 + 0.5 * d * (log 2pi - log prec)
//   likelihood_potential(theta) = sum_i softplus(x_i . theta) - y_i * (x_i . theta)
//   grad_likelihood             = X^T (sigmoid(X theta) - y)
// Chains/particles are processed in tiles of 32: the contraction over the data rows is the hot loop.
//
// Three engines evaluate (U, grad U) for a tile:
//   Fp32Engine - data rows staged through LDS in chunks, FP32 FMAs on the vector ALU.
//   Fp64Engine - the same in FP64 (the reference for the error of the other two).
//   Bf16Engine - MFMA v_mfma_f32_32x32x16_bf16 with split precision: every operand is
//                hi + lo bf16 and each product is hi*hi + hi*lo + lo*hi in FP32 accumulation.
//                Z = X Theta^T is accumulated with rows of X on the M side and the 32 chains on
//                N, so each lane's accumulator column is one chain. sigmoid(Z) - y is formed in
//                registers and reused directly as the A operand (chains x rows) of
//                G = R X: the K slots of that operand are the rows the lane already holds, and
//                X's second packed layout lists its rows in the same order.
//
// The iteration kernel follows mocat's Underdamped + Metropolis/RMMetropolis update
// (mocat/src/mcmc/standard_mcmc.py:107-153, mocat/src/utils.py:108-146,
//  mocat/src/mcmc/metropolis.py:48-161), one chain per (tile, slot), key tree per chain as mocat's
// single-chain run (mocat/src/mcmc/sampler.py:92-111).
#include <type_traits>

#include "common.hpp"

using namespace mocat;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int CT = 32;  // chains per tile
constexpr float kLog2Pi = 1.83787706640934548356f;

MOCAT_DEV int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// --- per-row likelihood pieces -------------------------------------------------------------
// r = sigmoid(z) - y ; c = softplus(z) - y z. Rows with y < 0 are padding: r = c = 0.
// Hardware exp2/log2/rcp (v_exp_f32, v_log_f32, v_rcp_f32): |error| of c per row <= ~1e-7 absolute.
MOCAT_DEV void lr_row(float z, float y, float& r, float& c) {
  const float e = __expf(-fabsf(z));
  const float ope = 1.f + e;
  const float inv = __builtin_amdgcn_rcpf(ope);
  const float sig = z >= 0.f ? inv : e * inv;
  const bool valid = y >= 0.f;
  r = valid ? sig - y : 0.f;
  c = valid ? fmaxf(z, 0.f) + __logf(ope) - y * z : 0.f;
}

// MFMA-engine rows on the hardware transcendentals only (v_exp_f32 = 2^x, v_log_f32 = log2,
// v_rcp_f32): t = 2^(-z log2 e) = exp(-z), r = 1 / (1 + t) - y; softplus(z) = max(z + ln2 log2(1 + t), 0)
// with the exponent clamped below overflow (1 + t >= 1, so no denormal inputs reach v_log_f32).
// Padding rows hold x = 0 and y = 0.5, so r = 0 there.
constexpr float kLog2E = 1.44269504088896340736f;
constexpr float kLn2 = 0.69314718055994530942f;
// Both take zs = -z log2(e), the engine's accumulator (X is packed premultiplied by -log2(e)).
MOCAT_DEV float lr_row_grad_hw(float zs, float y) {
  const float t = __builtin_amdgcn_exp2f(zs);
  return __builtin_amdgcn_rcpf(1.f + t) - y;
}
MOCAT_DEV float lr_row_full_hw(float zs, float y, float& c) {
  const float t = __builtin_amdgcn_exp2f(fminf(zs, 126.f));
  const float ope = 1.f + t;
  const float z = zs * -kLn2;
  c = fmaxf(z + kLn2 * __builtin_amdgcn_logf(ope), 0.f) - y * z;
  return __builtin_amdgcn_rcpf(ope) - y;
}

// Split two floats into packed bf16 pairs, v = hi + lo: hi keeps the upper 16 bits of v (truncation,
// one v_perm_b32 packs both), lo = v - hi exactly in FP32, rounded to bf16 by v_cvt_pk_bf16_f32.
// |v - hi - lo| <= 2^-16 |v|, the same order as a rounded split.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
MOCAT_DEV void split_pair(float v0, float v1, uint32_t& hi, uint32_t& lo) {
  const uint32_t u0 = __float_as_uint(v0) & 0xffff0000u, u1 = __float_as_uint(v1) & 0xffff0000u;
  hi = __builtin_amdgcn_perm(u1, u0, 0x07060302u);
  const floatx2 l = {v0 - __uint_as_float(u0), v1 - __uint_as_float(u1)};
  const bf16x2 lb = __builtin_convertvector(l, bf16x2);
  lo = __builtin_bit_cast(uint32_t, lb);
}

// ------------------------------------------------------------------------------------------
// FP32 engine.  LDS: X chunk [RC][DP+4], R chunk [CT][RC+4].
template <int DP, int NT>
struct Fp32Engine {
  static constexpr int RC = 64;
  static constexpr int XS = DP + 4;
  static constexpr int RS = RC + 4;
  struct alignas(16) Smem {
    alignas(16) float x[RC * XS];
    alignas(16) float r[CT * RS];
    float y[RC];
    float red[NT / kWave][CT];
  };

  // theta: LDS [CT][DP] (zero beyond d). Outputs: lik[CT], glik[CT][DP] in LDS.
  static MOCAT_DEV void eval(Smem& s, const float* __restrict__ X, const float* __restrict__ y, const void*,
                             int n, int d, const float* theta, float* lik, float* glik, bool /*want_pot*/) {
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), w = t / kWave;
    constexpr int NW = NT / kWave;
    constexpr int P1 = CT / NW;         // chains per thread in phase 1 (row = lane)
    constexpr int KPT = DP / kWave;     // k columns per thread in phase 2 (1 or 2)
    float gacc[P1][KPT];
    float cacc[P1];
#pragma unroll
    for (int m = 0; m < P1; ++m) {
      cacc[m] = 0.f;
#pragma unroll
      for (int q = 0; q < KPT; ++q) gacc[m][q] = 0.f;
    }
    for (int i0 = 0; i0 < n; i0 += RC) {
      __syncthreads();
      // stage X rows [i0, i0+RC) and y
      for (int e = t; e < RC * DP; e += NT) {
        const int i = e / DP, k = e - i * DP;
        const int gi = i0 + i;
        s.x[i * XS + k] = (gi < n && k < d) ? X[(int64_t)gi * d + k] : 0.f;
      }
      for (int i = t; i < RC; i += NT) s.y[i] = (i0 + i < n) ? y[i0 + i] : -1.f;
      __syncthreads();
      // phase 1: z[c][i] for row i = lane, chains c = w + NW*m
      {
        float z[P1];
#pragma unroll
        for (int m = 0; m < P1; ++m) z[m] = 0.f;
        const float* xr = &s.x[lane * XS];
#pragma unroll 4
        for (int k = 0; k < DP; k += 4) {
          const floatx4 xv = *reinterpret_cast<const floatx4*>(xr + k);
#pragma unroll
          for (int m = 0; m < P1; ++m) {
            const floatx4 tv = *reinterpret_cast<const floatx4*>(theta + (w + NW * m) * DP + k);
            z[m] += xv.x * tv.x + xv.y * tv.y + xv.z * tv.z + xv.w * tv.w;
          }
        }
        const float yv = s.y[lane];
#pragma unroll
        for (int m = 0; m < P1; ++m) {
          float r, c;
          lr_row(z[m], yv, r, c);
          cacc[m] += c;
          s.r[(w + NW * m) * RS + lane] = r;
        }
      }
      __syncthreads();
      // phase 2: g[c][k] += sum_i r[c][i] x[i][k], k = lane (+64), chains c = w + NW*m
#pragma unroll 2
      for (int i = 0; i < RC; i += 4) {
        float xv[4][KPT];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
          for (int q = 0; q < KPT; ++q) xv[ii][q] = s.x[(i + ii) * XS + lane + kWave * q];
#pragma unroll
        for (int m = 0; m < P1; ++m) {
          const floatx4 rv = *reinterpret_cast<const floatx4*>(&s.r[(w + NW * m) * RS + i]);
#pragma unroll
          for (int q = 0; q < KPT; ++q)
            gacc[m][q] += rv.x * xv[0][q] + rv.y * xv[1][q] + rv.z * xv[2][q] + rv.w * xv[3][q];
        }
      }
    }
    // outputs: each (chain, k) owned by exactly one thread; potentials reduced over lanes
#pragma unroll
    for (int m = 0; m < P1; ++m) {
      const int c = w + NW * m;
#pragma unroll
      for (int q = 0; q < KPT; ++q) glik[c * DP + lane + kWave * q] = gacc[m][q];
      const float cs = wave_sum(cacc[m]);
      if (lane == 0) lik[c] = cs;
    }
    __syncthreads();
  }
};

// ------------------------------------------------------------------------------------------
// FP64 engine: the same contraction with every product, sum and row function in double (the
// reference run that error statements of the split-precision engine are made against). The tile
// state stays f32: theta is widened exactly, the outputs are rounded once. LDS: X chunk
// [RC][DP+1] and R chunk [CT][RC+1] in double.
template <int DP, int NT>
struct Fp64Engine {
  static constexpr int RC = 32;
  static constexpr int XS = DP + 1;
  static constexpr int RS = RC + 1;
  struct alignas(16) Smem {
    double x[RC * XS];
    double r[CT * RS];
    double y[RC];
  };

  static MOCAT_DEV void eval(Smem& s, const float* __restrict__ X, const float* __restrict__ y, const void*,
                             int n, int d, const float* theta, float* lik, float* glik, bool /*want_pot*/) {
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), w = t / kWave;
    constexpr int NW = NT / kWave;
    constexpr int G1 = NW * 2;          // phase 1: chain groups (row = lane % 32)
    constexpr int P1 = CT / G1;         // chains per thread in phase 1
    constexpr int P2 = CT / NW;         // chains per thread in phase 2 (column = lane (+64))
    constexpr int KPT = DP / kWave;
    static_assert(CT % G1 == 0 && CT % NW == 0, "tile shape");
    const int row = lane & 31, grp = w * 2 + (lane >> 5);
    double gacc[P2][KPT];
    double cacc[P1];
#pragma unroll
    for (int m = 0; m < P1; ++m) cacc[m] = 0.0;
#pragma unroll
    for (int m = 0; m < P2; ++m)
#pragma unroll
      for (int q = 0; q < KPT; ++q) gacc[m][q] = 0.0;
    for (int i0 = 0; i0 < n; i0 += RC) {
      __syncthreads();
      for (int e = t; e < RC * DP; e += NT) {
        const int i = e / DP, k = e - i * DP;
        const int gi = i0 + i;
        s.x[i * XS + k] = (gi < n && k < d) ? (double)X[(int64_t)gi * d + k] : 0.0;
      }
      for (int i = t; i < RC; i += NT) s.y[i] = (i0 + i < n) ? (double)y[i0 + i] : -1.0;
      __syncthreads();
      {
        double z[P1];
#pragma unroll
        for (int m = 0; m < P1; ++m) z[m] = 0.0;
        const double* xr = &s.x[row * XS];
        for (int k = 0; k < d; ++k) {
          const double xv = xr[k];
#pragma unroll
          for (int m = 0; m < P1; ++m) z[m] = fma(xv, (double)theta[(grp + G1 * m) * DP + k], z[m]);
        }
        const double yv = s.y[row];
#pragma unroll
        for (int m = 0; m < P1; ++m) {
          double r = 0.0, c = 0.0;
          if (yv >= 0.0) {
            const double e = exp(-fabs(z[m]));
            const double sig = z[m] >= 0.0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
            r = sig - yv;
            c = fmax(z[m], 0.0) + log1p(e) - yv * z[m];
          }
          cacc[m] += c;
          s.r[(grp + G1 * m) * RS + row] = r;
        }
      }
      __syncthreads();
      for (int i = 0; i < RC; ++i) {
        double xv[KPT];
#pragma unroll
        for (int q = 0; q < KPT; ++q) xv[q] = s.x[i * XS + lane + kWave * q];
#pragma unroll
        for (int m = 0; m < P2; ++m) {
          const double rv = s.r[(w + NW * m) * RS + i];
#pragma unroll
          for (int q = 0; q < KPT; ++q) gacc[m][q] = fma(rv, xv[q], gacc[m][q]);
        }
      }
    }
    __syncthreads();
    // potentials: the two half-waves hold different chains; sum each chain's rows over its 32 lanes
#pragma unroll
    for (int m = 0; m < P1; ++m) {
      double v = cacc[m];
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
      if (row == 0) lik[grp + G1 * m] = (float)v;
    }
#pragma unroll
    for (int m = 0; m < P2; ++m)
#pragma unroll
      for (int q = 0; q < KPT; ++q) glik[(w + NW * m) * DP + lane + kWave * q] = (float)gacc[m][q];
    __syncthreads();
  }
};

// ------------------------------------------------------------------------------------------
// BF16 split-precision MFMA engine.
// Packed data per block of 32 rows (rb):
//   A layout [2 parts][KB][64 lanes][8]   lane l: row l%32, columns kb*16 + (l/32)*8 + s
//   B layout [2 parts][2][NB][64][8]      lane l: column nb*32 + l%32, rows acc_row(kb2*8+s, l)
//   y layout [2][16] f32                  g = l/32: y[acc_row(r, l)], 0.5 on padding rows (x = 0 there)
template <int DP>
struct PackLayout {
  // The engine reads X premultiplied by -log2(e) in the GEMM-1 (A) layout, so that its accumulators
  // hold z' = -z log2(e) and exp(-z) = 2^z' needs no multiply; the B layout is plain X.
  static constexpr float kAScale = -kLog2E;
  static constexpr int KB = DP / 16;
  static constexpr int NB = DP / 32;
  static constexpr int A_ELEMS = 2 * KB * 512;
  static constexpr int B_ELEMS = 2 * 2 * NB * 512;
  static constexpr int BLOCK_BF16 = A_ELEMS + B_ELEMS;
  static constexpr int BLOCK_BYTES = BLOCK_BF16 * 2 + 32 * 4;
};

MOCAT_DEV void split_bf16(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

// ------------------------------------------------------------------------------------------
// BF16 split-precision MFMA engine (every product hi*hi + hi*lo + lo*hi, FP32 accumulation).
// Z = X Theta^T is accumulated with rows of X on the M side and the 32 chains on N, so each lane's
// accumulator column is one chain; R = sigmoid(Z) - y is formed in registers and reused directly as
// the A operand (chains x rows) of G = R X: the K slots of that operand are the rows the lane already
// holds, and X's second packed layout lists its rows in the same order. With NW waves, wave w takes
// the 32-row blocks rb = w + NW j; two waves per SIMD (NW = 8) let one wave's row functions on the
// vector ALU overlap the other's MFMAs.
template <int DP, int NT>
struct Bf16Engine {
  using L = PackLayout<DP>;
  static constexpr int KB = L::KB, NB = L::NB, NW = NT / kWave;
  static_assert(NW % 4 == 0, "whole SIMDs");
  static_assert(L::kAScale == -kLog2E, "the rows expect the scaled A layout");
  static constexpr bool kThetaRegs = DP <= 64;
  static constexpr int GS = DP + 4;  // row stride of the partial gradients (spreads LDS banks)
  struct alignas(16) Smem {
    alignas(16) float gred[NW][CT * GS];
    alignas(16) bf16x8 thl[2][KB][64];  // theta^T operand, hi and lo parts, per lane
    float ured[NW][2][CT];
  };
  struct ABuf {
    bf16x8 h[KB], l[KB];
  };
  struct BBuf {
    floatx4 y[4];
    bf16x8 h[2][NB], l[2][NB];
  };

  static MOCAT_DEV void eval(Smem& s, const float* X, const float* y, const void* packed, int n, int d,
                             const float* theta, float* lik, float* glik, bool want_pot) {
    // the potential terms are needed after the last leapfrog step only: two compiled passes, so that
    // each block's row functions and MFMAs form one straight-line region the scheduler can interleave
    if (want_pot)
      pass<true>(s, packed, n, theta, lik, glik);
    else
      pass<false>(s, packed, n, theta, lik, glik);
  }

  template <bool WANT_POT>
  static MOCAT_DEV void pass(Smem& s, const void* packed, int n, const float* theta, float* lik, float* glik) {
    const int t = threadIdx.x;
    for (int e = t; e < KB * 64; e += NT) {
      const int kb = e / 64, ln = e % 64;
      const float* src = theta + (ln & 31) * DP + kb * 16 + (ln >> 5) * 8;
      bf16x8 h8, l8;
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) {
        __bf16 h, l;
        split_bf16(src[s2], h, l);
        h8[s2] = h;
        l8[s2] = l;
      }
      s.thl[0][kb][ln] = h8;
      s.thl[1][kb][ln] = l8;
    }
    __syncthreads();
    partials<WANT_POT>(s, packed, n);
    __syncthreads();
    for (int e = t; e < CT * DP; e += NT) {
      const int c = e / DP, k = e - c * DP;
      float acc = s.gred[0][c * GS + k];
#pragma unroll
      for (int q = 1; q < NW; ++q) acc += s.gred[q][c * GS + k];
      glik[e] = acc;
    }
    if (t < CT) {
      float acc = 0.f;
      for (int q = 0; q < NW; ++q) acc += s.ured[q][0][t] + s.ured[q][1][t];
      lik[t] = acc;
    }
    __syncthreads();
  }

  // The block loop of one gradient evaluation. Theta's operands are read from s.thl (written by the
  // caller before a barrier); each wave leaves its partial gradient in s.gred[w] ([chain][GS]) and its
  // potential partials in s.ured[w]. The caller synchronises before reading them.
  template <bool WANT_POT>
  static MOCAT_DEV void partials(Smem& s, const void* packed, int n) {
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), w = t / kWave, g = lane >> 5, col = lane & 31;
    const int nrb = (n + 31) / 32;
    const int nmine = nrb > w ? (nrb - w + NW - 1) / NW : 0;  // blocks rb = w + j * NW, j < nmine
    const char* base = reinterpret_cast<const char*>(packed);
    bf16x8 th[kThetaRegs ? KB : 1], tl[kThetaRegs ? KB : 1];
    if constexpr (kThetaRegs) {
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        th[kb] = s.thl[0][kb][lane];
        tl[kb] = s.thl[1][kb][lane];
      }
    }
    floatx16 gacc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) gacc[nb][r] = 0.f;
    float upart = 0.f;

    auto blk = [&](int j) { return base + (int64_t)(w + j * NW) * L::BLOCK_BYTES; };
    auto load_a = [&](ABuf& a, int j) {
      const bf16x8* A = reinterpret_cast<const bf16x8*>(blk(j));
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        a.h[kb] = A[(0 * KB + kb) * 64 + lane];
        a.l[kb] = A[(1 * KB + kb) * 64 + lane];
      }
    };
    auto load_b = [&](BBuf& b, int j) {
      const char* bp = blk(j);
      const floatx4* yv4 = reinterpret_cast<const floatx4*>(bp + L::BLOCK_BF16 * 2 + g * 64);
#pragma unroll
      for (int q = 0; q < 4; ++q) b.y[q] = yv4[q];
      const bf16x8* B = reinterpret_cast<const bf16x8*>(bp) + L::A_ELEMS / 8;
#pragma unroll
      for (int kb2 = 0; kb2 < 2; ++kb2)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          b.h[kb2][nb] = B[((0 * 2 + kb2) * NB + nb) * 64 + lane];
          b.l[kb2][nb] = B[((1 * 2 + kb2) * NB + nb) * 64 + lane];
        }
    };
    auto gemm1 = [&](const ABuf& a, floatx16& z) {
#pragma unroll
      for (int r = 0; r < 16; ++r) z[r] = 0.f;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        bf16x8 h8, l8;
        if constexpr (kThetaRegs) {
          h8 = th[kb];
          l8 = tl[kb];
        } else {
          h8 = s.thl[0][kb][lane];
          l8 = s.thl[1][kb][lane];
        }
        z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.h[kb], h8, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.h[kb], l8, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.l[kb], h8, z, 0, 0, 0);
      }
    };
    // one block: rows of Z_j (zc) with labels and B operands bc, then G += R_j X_j
    auto step = [&](int j, const floatx16& zc, const BBuf& bc) {
      float yv[16];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        yv[4 * q] = bc.y[q].x;
        yv[4 * q + 1] = bc.y[q].y;
        yv[4 * q + 2] = bc.y[q].z;
        yv[4 * q + 3] = bc.y[q].w;
      }
      float rv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if constexpr (WANT_POT) {
          float cc;
          rv[r] = lr_row_full_hw(zc[r], yv[r], cc);
          upart += cc;  // padding rows are taken out after the loop
        } else {
          rv[r] = lr_row_grad_hw(zc[r], yv[r]);
        }
      }
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      u32x4 ph[2], pl[2];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        uint32_t hw, lw;
        split_pair(rv[2 * q], rv[2 * q + 1], hw, lw);
        ph[q >> 2][q & 3] = hw;
        pl[q >> 2][q & 3] = lw;
      }
#pragma unroll
      for (int kb2 = 0; kb2 < 2; ++kb2) {
        const bf16x8 rh = __builtin_bit_cast(bf16x8, ph[kb2]), rl = __builtin_bit_cast(bf16x8, pl[kb2]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, bc.h[kb2][nb], gacc[nb], 0, 0, 0);
          gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, bc.l[kb2][nb], gacc[nb], 0, 0, 0);
          gacc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, bc.h[kb2][nb], gacc[nb], 0, 0, 0);
        }
      }
    };
    for (int j = 0; j < nmine; ++j) {
      ABuf a;
      BBuf b;
      floatx16 z;
      load_a(a, j);
      // At d <= 64 (two waves per SIMD) a wave issuing its first GEMM and the next loads runs at raised
      // priority, so that the other wave's row functions fill the vector ALU meanwhile (measured 3%
      // faster per HMC iteration at the bench shape; at d > 64, one wave per SIMD, it does not apply)
      if constexpr (DP == 64) __builtin_amdgcn_s_setprio(1);
      gemm1(a, z);
      load_b(b, j);
      if constexpr (DP == 64) __builtin_amdgcn_s_setprio(0);
      step(j, z, b);
    }
    if constexpr (WANT_POT) {
      // padding rows of the last block (x = 0, y = 0.5: z = 0) each added the same term
      const int rbl = nrb - 1;
      if (rbl % NW == w) {
        float cpad;
        lr_row_full_hw(0.f, 0.5f, cpad);
        int npad = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) npad += (rbl * 32 + acc_row(r, lane) >= n);
        upart -= (float)npad * cpad;
      }
    }
    // every wave leaves its partials; readers sum them over the waves in order (deterministic)
    s.ured[w][g][col] = upart;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) s.gred[w][acc_row(r, lane) * GS + nb * 32 + col] = gacc[nb][r];
  }
};

// Pack X (n x d, row-major f32) and y into the Bf16Engine layout, zero padded to DP columns and a
// multiple of 32 rows. One thread per packed element.
template <int DP>
__global__ void lr_pack_kernel(const float* __restrict__ X, const float* __restrict__ y, int n, int d,
                               char* __restrict__ out) {
  using L = PackLayout<DP>;
  const int nrb = (n + 31) / 32;
  const int64_t per = L::BLOCK_BF16 + 32;
  const int64_t total = (int64_t)nrb * per;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int rb = (int)(e / per);
    const int q = (int)(e - rb * per);
    char* blk = out + (int64_t)rb * L::BLOCK_BYTES;
    if (q < L::A_ELEMS) {
      const int s = q & 7, lane = (q >> 3) & 63, pk = q >> 9;  // pk = part*KB + kb
      const int part = pk / L::KB, kb = pk - part * L::KB;
      const int row = rb * 32 + (lane & 31), colm = kb * 16 + (lane >> 5) * 8 + s;
      const float v = (row < n && colm < d) ? X[(int64_t)row * d + colm] * L::kAScale : 0.f;
      __bf16 h, l;
      split_bf16(v, h, l);
      reinterpret_cast<__bf16*>(blk)[q] = part ? l : h;
    } else if (q < L::BLOCK_BF16) {
      const int qq = q - L::A_ELEMS;
      const int s = qq & 7, lane = (qq >> 3) & 63, rest = qq >> 9;  // rest = (part*2 + kb2)*NB + nb
      const int nb = rest % L::NB, pk2 = rest / L::NB, kb2 = pk2 & 1, part = pk2 >> 1;
      const int row = rb * 32 + acc_row(kb2 * 8 + s, lane), colm = nb * 32 + (lane & 31);
      const float v = (row < n && colm < d) ? X[(int64_t)row * d + colm] : 0.f;
      __bf16 h, l;
      split_bf16(v, h, l);
      reinterpret_cast<__bf16*>(blk)[q] = part ? l : h;
    } else {
      const int qq = q - L::BLOCK_BF16;  // 0..31 = g*16 + r
      const int gg = qq >> 4, r = qq & 15;
      const int row = rb * 32 + 8 * (r >> 2) + 4 * gg + (r & 3);
      reinterpret_cast<float*>(blk + L::BLOCK_BF16 * 2)[qq] = row < n ? y[row] : 0.5f;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Diagonal Gaussian likelihood engine (mocat/src/scenarios/toy_examples.py:19-51 with a diagonal
// covariance): lik = 0.5 * sum_k lam_k (theta_k - mu_k)^2, grad = lam * (theta - mu).
// Data: X -> mu[d], y -> lam[d] (precisions); n unused.
template <int DP, int NT>
struct GaussEngine {
  struct alignas(16) Smem {
    float part[CT * DP];
  };
  static MOCAT_DEV void eval(Smem& s, const float* __restrict__ mu, const float* __restrict__ lam, const void*, int,
                             int d, const float* theta, float* lik, float* glik, bool /*want_pot*/) {
    const int t = threadIdx.x;
    for (int e = t; e < CT * DP; e += NT) {
      const int k = e % DP;
      float g = 0.f, c = 0.f;
      if (k < d) {
        const float df = theta[e] - mu[k];
        g = lam[k] * df;
        c = 0.5f * lam[k] * df * df;
      }
      glik[e] = g;
      s.part[e] = c;
    }
    __syncthreads();
    if (t < CT) {
      float acc = 0.f;
      for (int k = 0; k < d; ++k) acc += s.part[t * DP + k];
      lik[t] = acc;
    }
    __syncthreads();
  }
};

// ------------------------------------------------------------------------------------------
// Tile state shared by the evaluation and iteration kernels.
template <int DP>
struct alignas(16) TileState {
  alignas(16) float theta[CT * DP];
  float mom[CT * DP];
  float grad[CT * DP];  // grad of the tempered potential
  float glik[CT * DP];
  float lik[CT];
  float pot[CT];
  float prior[CT];
  float likv[CT];
  float pot0[CT];
  float prior0[CT];
  float lik0[CT];
  float step[CT];
  float alpha[CT];
  float kin0[CT];
  float sq[CT];
  uint32_t key[CT][2];
};

// out[c] = sum_k v[c][k]^2 for the CT chains of a tile (one wave per chain group), then a barrier.
template <int DP, int NT>
MOCAT_DEV void chain_sqnorm(const float* v, int d, float* out) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  __syncthreads();
  for (int c = w; c < CT; c += NW) {
    float a = 0.f;
    for (int k = lane; k < d; k += kWave) a += v[c * DP + k] * v[c * DP + k];
    a = wave_sum(a);
    if (lane == 0) out[c] = a;
  }
  __syncthreads();
}

struct LrArgs {
  const float* X;
  const float* y;
  const void* packed;
  int n, d;
  float prior_prec, temperature;
};

// Evaluates prior/likelihood potentials and gradients of the tile's theta (in LDS).
// want_pot = false (intermediate leapfrog steps): the gradient only; the potentials are left as they were.
template <class Engine, int DP, int NT>
MOCAT_DEV void tile_eval(typename Engine::Smem& es, TileState<DP>& ts, const LrArgs& a, bool want_pot = true) {
  Engine::eval(es, a.X, a.y, a.packed, a.n, a.d, ts.theta, ts.lik, ts.glik, want_pot);
  const int t = threadIdx.x;
  for (int e = t; e < CT * DP; e += NT) ts.grad[e] = a.prior_prec * ts.theta[e] + a.temperature * ts.glik[e];
  if (want_pot) {
    chain_sqnorm<DP, NT>(ts.theta, a.d, ts.sq);
    if (t < CT) {
      // prec == 0: improper flat prior, potential 0 (mocat's default uniform prior, core.py:183-189)
      const float pr = a.prior_prec > 0.f
                           ? 0.5f * a.prior_prec * ts.sq[t] + 0.5f * (float)a.d * (kLog2Pi - logf(a.prior_prec))
                           : 0.f;
      ts.prior[t] = pr;
      ts.likv[t] = ts.lik[t];
      ts.pot[t] = pr + a.temperature * ts.lik[t];
    }
  }
  __syncthreads();
}

template <class Engine, int DP, int NT>
__global__ __launch_bounds__(NT) void lr_eval_kernel(LrArgs a, const float* __restrict__ theta, int64_t C,
                                                     float* prior_pot, float* lik_pot, float* grad_prior,
                                                     float* grad_lik) {
  __shared__ TileState<DP> ts;
  __shared__ typename Engine::Smem es;
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * CT;
  for (int e = t; e < CT * DP; e += NT) {
    const int c = e / DP, k = e - c * DP;
    ts.theta[e] = (c0 + c < C && k < a.d) ? theta[(c0 + c) * a.d + k] : 0.f;
  }
  __syncthreads();
  tile_eval<Engine, DP, NT>(es, ts, a);
  for (int e = t; e < CT * DP; e += NT) {
    const int c = e / DP, k = e - c * DP;
    if (c0 + c < C && k < a.d) {
      if (grad_lik) grad_lik[(c0 + c) * a.d + k] = ts.glik[e];
      if (grad_prior) grad_prior[(c0 + c) * a.d + k] = a.prior_prec * ts.theta[e];
    }
  }
  if (t < CT && c0 + t < C) {
    if (prior_pot) prior_pot[c0 + t] = ts.prior[t];
    if (lik_pot) lik_pot[c0 + t] = ts.likv[t];
  }
}

// ------------------------------------------------------------------------------------------
struct HmcArgs {
  LrArgs lr;
  int64_t C;
  float* theta;
  float* mom;
  float* grad;
  float* pot;
  float* prior;
  float* lik;
  float* step;
  float* alpha;
  uint32_t* keys;
  int64_t iter0;
  int n_iter;
  int leapfrog_steps;
  float friction;
  int correction;  // 0 none, 1 Metropolis, 2 Robbins-Monro-adapted Metropolis (log update)
  float target;
  float rm_scale, rm_neg_exponent;
  double adapt_cut_off;
  float* tr_value;
  float* tr_pot;
  float* tr_alpha;
  float* tr_step;
  float* tr_mom;
};

template <class Engine, int DP, int NT>
__global__ __launch_bounds__(NT) void hmc_lr_kernel(HmcArgs h) {
  __shared__ TileState<DP> ts;
  __shared__ typename Engine::Smem es;
  const int t = threadIdx.x;
  const int d = h.lr.d;
  const int64_t c0 = (int64_t)blockIdx.x * CT;
  const int nvalid = (int)((h.C - c0) < CT ? (h.C - c0) : CT);

  // load the tile (padding chains are zero and never written back)
  for (int e = t; e < CT * DP; e += NT) {
    const int c = e / DP, k = e - c * DP;
    const bool v = c < nvalid && k < d;
    ts.theta[e] = v ? h.theta[(c0 + c) * d + k] : 0.f;
    ts.mom[e] = v ? h.mom[(c0 + c) * d + k] : 0.f;
    ts.grad[e] = v ? h.grad[(c0 + c) * d + k] : 0.f;
  }
  if (t < CT) {
    const bool v = t < nvalid;
    ts.pot[t] = v ? h.pot[c0 + t] : 0.f;
    ts.prior[t] = v ? h.prior[c0 + t] : 0.f;
    ts.likv[t] = v ? h.lik[c0 + t] : 0.f;
    ts.step[t] = v ? h.step[c0 + t] : 1.f;  // padding chains: finite dummy stepsize
    ts.alpha[t] = v ? h.alpha[c0 + t] : 0.f;
    ts.key[t][0] = v ? h.keys[2 * (c0 + t)] : 0u;
    ts.key[t][1] = v ? h.keys[2 * (c0 + t) + 1] : 0u;
  }
  __syncthreads();

  const uint32_t hd = (uint32_t)(d + 1) >> 1;  // threefry blocks per normal(sub, (d,))
  for (int it = 0; it < h.n_iter; ++it) {
    const int64_t iter = h.iter0 + it + 1;
    // ---- always: p <- -p ; (key, sub) = split(key) ; OU refresh (standard_mcmc.py:107-123)
    __shared__ uint32_t sub[CT][2];
    if (t < CT) {
      Key k{ts.key[t][0], ts.key[t][1]};
      Key nk = split_at(k, 2, 0), sk = split_at(k, 2, 1);
      ts.key[t][0] = nk.a;
      ts.key[t][1] = nk.b;
      sub[t][0] = sk.a;
      sub[t][1] = sk.b;
      // keys = split(key, L+1); key = keys[0] (proposal, standard_mcmc.py:129-131)
      Key pk = split_at(nk, (uint32_t)h.leapfrog_steps + 1u, 0);
      ts.key[t][0] = pk.a;
      ts.key[t][1] = pk.b;
    }
    __syncthreads();
    for (int e = t; e < CT * (int)hd; e += NT) {
      const int c = e / hd;
      const uint32_t i = (uint32_t)(e - c * hd);
      uint32_t w0, w1;
      bits_pair(Key{sub[c][0], sub[c][1]}, (uint32_t)d, i, w0, w1);
      const float hstep = ts.step[c];
      const float decay = __expf(-h.friction * hstep);
      const float noise = sqrtf(1.f - __expf(-2.f * h.friction * hstep));
      float* p = &ts.mom[c * DP];
      p[i] = -p[i] * decay + noise * normal_from_bits(w0);
      if (i + hd < (uint32_t)d) p[i + hd] = -p[i + hd] * decay + noise * normal_from_bits(w1);
    }
    __syncthreads();
    // keep the refreshed momenta for rejection in the global momenta array (position and gradient of
    // the pre-proposal state stay in global memory until a proposal is accepted)
    for (int e = t; e < CT * DP; e += NT) {
      const int c = e / DP, k = e - c * DP;
      if (c < nvalid && k < d) h.mom[(c0 + c) * d + k] = ts.mom[e];
    }
    chain_sqnorm<DP, NT>(ts.mom, d, ts.sq);
    if (t < CT) {
      ts.pot0[t] = ts.pot[t];
      ts.prior0[t] = ts.prior[t];
      ts.lik0[t] = ts.likv[t];
      ts.kin0[t] = 0.5f * ts.sq[t] + 0.5f * (float)d * kLog2Pi;
    }
    // the pre-proposal position/gradient stay in global memory (unchanged until accepted)
    __syncthreads();
    // ---- leapfrog (utils.py:117-146)
    for (int s = 0; s < h.leapfrog_steps; ++s) {
      for (int e = t; e < CT * DP; e += NT) {
        const int c = e / DP;
        const float hh = ts.step[c];
        const float ph = ts.mom[e] - (hh / 2.f) * ts.grad[e];
        ts.mom[e] = ph;
        ts.theta[e] = ((e - c * DP) < d) ? ts.theta[e] + hh * ph : 0.f;
      }
      __syncthreads();
      tile_eval<Engine, DP, NT>(es, ts, h.lr, s == h.leapfrog_steps - 1);
      for (int e = t; e < CT * DP; e += NT) {
        const int c = e / DP;
        ts.mom[e] = ts.mom[e] - (ts.step[c] / 2.f) * ts.grad[e];
      }
      __syncthreads();
    }
    // ---- correction (metropolis.py:48-70, 118-161)
    __shared__ int accept[CT];
    chain_sqnorm<DP, NT>(ts.mom, d, ts.sq);
    if (t < CT) {
      const float kin1 = 0.5f * ts.sq[t] + 0.5f * (float)d * kLog2Pi;  // |-p|^2 == |p|^2
      float a = 1.f;
      int acc = 1;
      if (h.correction > 0) {
        a = fminf(1.f, __expf(-ts.pot[t] + ts.pot0[t] - kin1 + ts.kin0[t]));
        a = isnan(a) ? 0.f : a;
        Key k{ts.key[t][0], ts.key[t][1]};
        Key nk = split_at(k, 2, 0), sk = split_at(k, 2, 1);
        ts.key[t][0] = nk.a;
        ts.key[t][1] = nk.b;
        const float u = uniform_from_bits(bits_at(sk, 1, 0), 0.f, 1.f);
        acc = u < a;
      }
      accept[t] = acc;
      ts.alpha[t] = a;
      if (!acc) {
        ts.pot[t] = ts.pot0[t];
        ts.prior[t] = ts.prior0[t];
        ts.likv[t] = ts.lik0[t];
      }
      if (h.correction == 2 && (double)iter <= h.adapt_cut_off) {
        const float rm = h.rm_scale * powf((float)iter, -h.rm_neg_exponent);
        ts.step[t] = expf(logf(ts.step[t]) + rm * (a - h.target));
      }
    }
    __syncthreads();
    // accepted: momenta <- -p'; rejected: restore position/grad from global, momenta = refreshed
    for (int e = t; e < CT * DP; e += NT) {
      const int c = e / DP, k = e - c * DP;
      if (accept[c]) {
        ts.mom[e] = -ts.mom[e];
        if (c < nvalid && k < d) {
          h.theta[(c0 + c) * d + k] = ts.theta[e];
          h.grad[(c0 + c) * d + k] = ts.grad[e];
        }
      } else {
        ts.mom[e] = (c < nvalid && k < d) ? h.mom[(c0 + c) * d + k] : 0.f;
        if (c < nvalid && k < d) {
          ts.theta[e] = h.theta[(c0 + c) * d + k];
          ts.grad[e] = h.grad[(c0 + c) * d + k];
        } else {
          ts.theta[e] = 0.f;
          ts.grad[e] = 0.f;
        }
      }
      if (c < nvalid && k < d) {
        if (h.tr_value) h.tr_value[((int64_t)it * h.C + c0 + c) * d + k] = ts.theta[e];
        if (h.tr_mom) h.tr_mom[((int64_t)it * h.C + c0 + c) * d + k] = ts.mom[e];
      }
    }
    if (t < nvalid) {
      const int64_t o = (int64_t)it * h.C + c0 + t;
      if (h.tr_pot) h.tr_pot[o] = ts.pot[t];
      if (h.tr_alpha) h.tr_alpha[o] = ts.alpha[t];
      if (h.tr_step) h.tr_step[o] = ts.step[t];
    }
    __syncthreads();
  }
  // write back the final per-chain state
  for (int e = t; e < CT * DP; e += NT) {
    const int c = e / DP, k = e - c * DP;
    if (c < nvalid && k < d) h.mom[(c0 + c) * d + k] = ts.mom[e];
  }
  if (t < nvalid) {
    h.pot[c0 + t] = ts.pot[t];
    h.prior[c0 + t] = ts.prior[t];
    h.lik[c0 + t] = ts.likv[t];
    h.step[c0 + t] = ts.step[t];
    h.alpha[c0 + t] = ts.alpha[t];
    h.keys[2 * (c0 + t)] = ts.key[t][0];
    h.keys[2 * (c0 + t) + 1] = ts.key[t][1];
  }
}

constexpr int kThreads = 512;

// The same iteration as hmc_lr_kernel for the bf16x3 engine, with the tile state in
// registers: thread t owns chain c = t % 32 and the EPT = DP / (NT / 32) columns k0 = EPT * (t / 32) ..
// of position, momentum and gradient. A leapfrog step is then: update the owned elements and write
// their bf16 hi/lo split into the engine's theta operand (barrier), the engine's block loop (barrier),
// and the owned gradient elements summed from every wave's partials. The refreshed momenta, position
// and gradient of the pre-proposal state stay in global memory until a proposal is accepted (as in
// hmc_lr_kernel); per-chain scalars and keys live in LDS.
template <int DP, int NT>
__global__ __launch_bounds__(NT) void hmc_lr_bf16_kernel(HmcArgs h) {
  using E = Bf16Engine<DP, NT>;
  constexpr int NG = NT / CT;  // column groups
  constexpr int EPT = DP / NG;
  static_assert(NG * EPT == DP && EPT % 4 == 0, "thread ownership");
  __shared__ typename E::Smem es;
  __shared__ float sq0[NG][CT], sq1[NG][CT], sqt[NG][CT];
  __shared__ float step[CT], pot[CT], prior[CT], likv[CT], alpha[CT];
  __shared__ uint32_t sub[CT][2];
  __shared__ int accept[CT];
  const int t = threadIdx.x, c = t % CT, kg = t / CT, k0 = kg * EPT;
  const int d = h.lr.d;
  const int64_t c0 = (int64_t)blockIdx.x * CT;
  const int nvalid = (int)((h.C - c0) < CT ? (h.C - c0) : CT);
  const bool cv = c < nvalid;
  const int64_t row = (c0 + c) * (int64_t)d;
  const float prec = h.lr.prior_prec, temp = h.lr.temperature;

  float th[EPT], p[EPT], g[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int k = k0 + q;
    const bool v = cv && k < d;
    th[q] = v ? h.theta[row + k] : 0.f;
    p[q] = v ? h.mom[row + k] : 0.f;
    g[q] = v ? h.grad[row + k] : 0.f;
  }
  Key key{0u, 0u};
  if (t < CT) {
    const bool v = t < nvalid;
    pot[t] = v ? h.pot[c0 + t] : 0.f;
    prior[t] = v ? h.prior[c0 + t] : 0.f;
    likv[t] = v ? h.lik[c0 + t] : 0.f;
    step[t] = v ? h.step[c0 + t] : 1.f;  // padding chains: finite dummy stepsize
    alpha[t] = v ? h.alpha[c0 + t] : 0.f;
    key = v ? Key{h.keys[2 * (c0 + t)], h.keys[2 * (c0 + t) + 1]} : Key{0u, 0u};
  }
  const uint32_t L1 = (uint32_t)h.leapfrog_steps + 1u;
  for (int it = 0; it < h.n_iter; ++it) {
    const int64_t iter = h.iter0 + it + 1;
    // ---- always: p <- -p ; (key, sub) = split(key) ; OU refresh (standard_mcmc.py:107-123); then the
    // proposal's keys = split(key, L+1), key = keys[0] (standard_mcmc.py:129-131)
    float pot0 = 0.f, prior0 = 0.f, lik0 = 0.f;
    if (t < CT) {
      const Key nk = split_at(key, 2, 0), sk = split_at(key, 2, 1);
      sub[t][0] = sk.a;
      sub[t][1] = sk.b;
      key = split_at(nk, L1, 0);
      pot0 = pot[t];
      prior0 = prior[t];
      lik0 = likv[t];
    }
    __syncthreads();
    const float hh = step[c];
    {
      const float decay = __expf(-h.friction * hh);
      const float noise = sqrtf(1.f - __expf(-2.f * h.friction * hh));
      const Key sk{sub[c][0], sub[c][1]};
      float s2 = 0.f;
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int k = k0 + q;
        if (k < d) {
          p[q] = -p[q] * decay + noise * normal_from_bits(bits_at(sk, (uint32_t)d, (uint32_t)k));
          if (cv) h.mom[row + k] = p[q];  // the refreshed momenta, kept for a rejection
        }
        s2 += p[q] * p[q];
      }
      sq0[kg][c] = s2;
    }
    // ---- leapfrog (utils.py:117-146)
    for (int s = 0; s < h.leapfrog_steps; ++s) {
      const bool last = s == h.leapfrog_steps - 1;
      float s2 = 0.f;
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const float ph = p[q] - (hh / 2.f) * g[q];
        p[q] = ph;
        th[q] = (k0 + q < d) ? th[q] + hh * ph : 0.f;
        s2 += th[q] * th[q];
      }
      if (last) sqt[kg][c] = s2;
      typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int j = 0; j < EPT / 4; ++j) {
        bf16x4 h4, l4;
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          __bf16 hb, lb;
          split_bf16(th[4 * j + s4], hb, lb);
          h4[s4] = hb;
          l4[s4] = lb;
        }
        // columns colb .. colb + 3 of chain c: operand (kb, lane c + 32 * half), elements colb % 8 ..
        const int colb = k0 + 4 * j;
        const int ln = c + 32 * ((colb / 8) & 1);
        reinterpret_cast<bf16x4*>(&es.thl[0][colb / 16][ln])[(colb / 4) & 1] = h4;
        reinterpret_cast<bf16x4*>(&es.thl[1][colb / 16][ln])[(colb / 4) & 1] = l4;
      }
      __syncthreads();
      if (last)
        E::template partials<true>(es, h.lr.packed, h.lr.n);
      else
        E::template partials<false>(es, h.lr.packed, h.lr.n);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < EPT; q += 4) {
        floatx4 acc = *reinterpret_cast<const floatx4*>(&es.gred[0][c * E::GS + k0 + q]);
#pragma unroll
        for (int w = 1; w < E::NW; ++w) acc += *reinterpret_cast<const floatx4*>(&es.gred[w][c * E::GS + k0 + q]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          g[q + e] = prec * th[q + e] + temp * acc[e];
          p[q + e] = p[q + e] - (hh / 2.f) * g[q + e];
        }
      }
      if (last && t < CT) {
        float lk = 0.f, st = 0.f;
        for (int w = 0; w < E::NW; ++w) lk += es.ured[w][0][t] + es.ured[w][1][t];
        for (int q = 0; q < NG; ++q) st += sqt[q][t];
        // prec == 0: improper flat prior, potential 0 (mocat's default uniform prior, core.py:183-189)
        const float pr = prec > 0.f ? 0.5f * prec * st + 0.5f * (float)d * (kLog2Pi - logf(prec)) : 0.f;
        prior[t] = pr;
        likv[t] = lk;
        pot[t] = pr + temp * lk;
      }
    }
    {
      float s2 = 0.f;
#pragma unroll
      for (int q = 0; q < EPT; ++q) s2 += p[q] * p[q];
      sq1[kg][c] = s2;
    }
    __syncthreads();
    // ---- correction (metropolis.py:48-70, 118-161)
    if (t < CT) {
      float k0s = 0.f, k1s = 0.f;
      for (int q = 0; q < NG; ++q) {
        k0s += sq0[q][t];
        k1s += sq1[q][t];
      }
      const float kin0 = 0.5f * k0s + 0.5f * (float)d * kLog2Pi;
      const float kin1 = 0.5f * k1s + 0.5f * (float)d * kLog2Pi;  // |-p|^2 == |p|^2
      float a = 1.f;
      int acc = 1;
      if (h.correction > 0) {
        a = fminf(1.f, __expf(-pot[t] + pot0 - kin1 + kin0));
        a = isnan(a) ? 0.f : a;
        const Key nk = split_at(key, 2, 0), sk = split_at(key, 2, 1);
        key = nk;
        const float u = uniform_from_bits(bits_at(sk, 1, 0), 0.f, 1.f);
        acc = u < a;
      }
      accept[t] = acc;
      alpha[t] = a;
      if (!acc) {
        pot[t] = pot0;
        prior[t] = prior0;
        likv[t] = lik0;
      }
      if (h.correction == 2 && (double)iter <= h.adapt_cut_off) {
        const float rm = h.rm_scale * powf((float)iter, -h.rm_neg_exponent);
        step[t] = expf(logf(step[t]) + rm * (a - h.target));
      }
    }
    __syncthreads();
    // accepted: momenta <- -p' and the proposal is stored; rejected: the pre-proposal state is read back
    const bool acc = accept[c] != 0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int k = k0 + q;
      const bool v = cv && k < d;
      if (acc) {
        p[q] = -p[q];
        if (v) {
          h.theta[row + k] = th[q];
          h.grad[row + k] = g[q];
        }
      } else {
        p[q] = v ? h.mom[row + k] : 0.f;
        th[q] = v ? h.theta[row + k] : 0.f;
        g[q] = v ? h.grad[row + k] : 0.f;
      }
      if (v) {
        if (h.tr_value) h.tr_value[((int64_t)it * h.C + c0 + c) * d + k] = th[q];
        if (h.tr_mom) h.tr_mom[((int64_t)it * h.C + c0 + c) * d + k] = p[q];
      }
    }
    if (t < nvalid) {
      const int64_t o = (int64_t)it * h.C + c0 + t;
      if (h.tr_pot) h.tr_pot[o] = pot[t];
      if (h.tr_alpha) h.tr_alpha[o] = alpha[t];
      if (h.tr_step) h.tr_step[o] = step[t];
    }
  }
  // write back the final per-chain state (position and gradient are already in place)
#pragma unroll
  for (int q = 0; q < EPT; ++q)
    if (cv && k0 + q < d) h.mom[row + k0 + q] = p[q];
  if (t < nvalid) {
    h.pot[c0 + t] = pot[t];
    h.prior[c0 + t] = prior[t];
    h.lik[c0 + t] = likv[t];
    h.step[c0 + t] = step[t];
    h.alpha[c0 + t] = alpha[t];
    h.keys[2 * (c0 + t)] = key.a;
    h.keys[2 * (c0 + t) + 1] = key.b;
  }
}

template <template <int, int> class E>
int dispatch_eval(const LrArgs& a, const float* theta, int64_t C, float* pp, float* lp, float* gp, float* gl,
                  hipStream_t st) {
  const unsigned blocks = (unsigned)((C + CT - 1) / CT);
  if (a.d <= 64)
    hipLaunchKernelGGL((lr_eval_kernel<E<64, kThreads>, 64, kThreads>), dim3(blocks), dim3(kThreads), 0, st, a,
                       theta, C, pp, lp, gp, gl);
  else
    hipLaunchKernelGGL((lr_eval_kernel<E<128, kThreads>, 128, kThreads>), dim3(blocks), dim3(kThreads), 0, st, a,
                       theta, C, pp, lp, gp, gl);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

template <template <int, int> class E>
int dispatch_hmc(const HmcArgs& h, hipStream_t st) {
  const unsigned blocks = (unsigned)((h.C + CT - 1) / CT);
  if (h.lr.d <= 64)
    hipLaunchKernelGGL((hmc_lr_kernel<E<64, kThreads>, 64, kThreads>), dim3(blocks), dim3(kThreads), 0, st, h);
  else
    hipLaunchKernelGGL((hmc_lr_kernel<E<128, kThreads>, 128, kThreads>), dim3(blocks), dim3(kThreads), 0, st, h);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// bf16x3: two waves per SIMD at d <= 64; at d <= 128 one wave per SIMD (the operand registers of a
// 128-column block, and the LDS of eight waves' partial gradients, do not fit two waves)
constexpr int kBf16Threads64 = 512, kBf16Threads128 = 256;

int dispatch_eval_bf16(const LrArgs& a, const float* theta, int64_t C, float* pp, float* lp, float* gp, float* gl,
                       hipStream_t st) {
  const unsigned blocks = (unsigned)((C + CT - 1) / CT);
  if (a.d <= 64)
    hipLaunchKernelGGL((lr_eval_kernel<Bf16Engine<64, kBf16Threads64>, 64, kBf16Threads64>), dim3(blocks),
                       dim3(kBf16Threads64), 0, st, a, theta, C, pp, lp, gp, gl);
  else
    hipLaunchKernelGGL((lr_eval_kernel<Bf16Engine<128, kBf16Threads128>, 128, kBf16Threads128>), dim3(blocks),
                       dim3(kBf16Threads128), 0, st, a, theta, C, pp, lp, gp, gl);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int dispatch_hmc_bf16(const HmcArgs& h, hipStream_t st) {
  const unsigned blocks = (unsigned)((h.C + CT - 1) / CT);
  if (h.lr.d <= 64)
    hipLaunchKernelGGL((hmc_lr_bf16_kernel<64, kBf16Threads64>), dim3(blocks), dim3(kBf16Threads64), 0, st, h);
  else
    hipLaunchKernelGGL((hmc_lr_bf16_kernel<128, kBf16Threads128>), dim3(blocks), dim3(kBf16Threads128), 0, st, h);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int check_lr(const float* X, const float* y, const void* packed, int64_t n, int64_t d, int precision) {
  if (n <= 0 || d <= 0 || d > 128 || n > (1ll << 30)) return MOCAT_ERR_ARG;
  if (precision == 0 || precision == 2 || precision == 3) return (X && y) ? MOCAT_OK : MOCAT_ERR_ARG;
  if (precision == 1) return packed ? MOCAT_OK : MOCAT_ERR_ARG;
  return MOCAT_ERR_ARG;
}

}  // namespace

extern "C" {

int mocat_lr_packed_bytes(int64_t n, int64_t d, int64_t* bytes) {
  if (n <= 0 || d <= 0 || d > 128 || !bytes) return MOCAT_ERR_ARG;
  const int64_t nrb = (n + 31) / 32;
  *bytes = nrb * (d <= 64 ? PackLayout<64>::BLOCK_BYTES : PackLayout<128>::BLOCK_BYTES);
  return MOCAT_OK;
}

int mocat_lr_pack(const float* X, const float* y, int64_t n, int64_t d, void* packed, hipStream_t stream) {
  if (!X || !y || !packed || n <= 0 || d <= 0 || d > 128 || n > (1ll << 30)) return MOCAT_ERR_ARG;
  const int64_t nrb = (n + 31) / 32;
  const int threads = 256;
  if (d <= 64) {
    const int64_t total = nrb * (PackLayout<64>::BLOCK_BF16 + 32);
    const unsigned blocks = (unsigned)std::min<int64_t>((total + threads - 1) / threads, 65535);
    hipLaunchKernelGGL(lr_pack_kernel<64>, dim3(blocks), dim3(threads), 0, stream, X, y, (int)n, (int)d,
                       (char*)packed);
  } else {
    const int64_t total = nrb * (PackLayout<128>::BLOCK_BF16 + 32);
    const unsigned blocks = (unsigned)std::min<int64_t>((total + threads - 1) / threads, 65535);
    hipLaunchKernelGGL(lr_pack_kernel<128>, dim3(blocks), dim3(threads), 0, stream, X, y, (int)n, (int)d,
                       (char*)packed);
  }
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_lr_potential_grad(const float* X, const float* y, const void* packed, int64_t n, int64_t d, int precision,
                            float prior_prec, float temperature, const float* theta, int64_t C, float* prior_pot,
                            float* lik_pot, float* grad_prior, float* grad_lik, hipStream_t stream) {
  int rc = check_lr(X, y, packed, n, d, precision);
  if (rc) return rc;
  if (C < 0 || (C > 0 && !theta)) return MOCAT_ERR_ARG;
  if (C == 0) return MOCAT_OK;
  LrArgs a{X, y, packed, (int)n, (int)d, prior_prec, temperature};
  if (precision == 2)
    return dispatch_eval<GaussEngine>(a, theta, C, prior_pot, lik_pot, grad_prior, grad_lik, stream);
  if (precision == 3)
    return dispatch_eval<Fp64Engine>(a, theta, C, prior_pot, lik_pot, grad_prior, grad_lik, stream);
  return precision == 0 ? dispatch_eval<Fp32Engine>(a, theta, C, prior_pot, lik_pot, grad_prior, grad_lik, stream)
                        : dispatch_eval_bf16(a, theta, C, prior_pot, lik_pot, grad_prior, grad_lik, stream);
}

int mocat_hmc_lr_iterate(const float* X, const float* y, const void* packed, int64_t n, int64_t d, int precision,
                         float prior_prec, float temperature, int64_t C, float* theta, float* momenta, float* grad,
                         float* potential, float* prior_pot, float* lik_pot, float* stepsize, float* alpha,
                         uint32_t* keys, int64_t iter0, int n_iter, int leapfrog_steps, float friction,
                         int correction, float target, float rm_scale, float rm_neg_exponent, double adapt_cut_off,
                         float* trace_value, float* trace_potential, float* trace_alpha, float* trace_stepsize,
                         float* trace_momenta, hipStream_t stream) {
  int rc = check_lr(X, y, packed, n, d, precision);
  if (rc) return rc;
  if (C < 0 || n_iter < 0 || leapfrog_steps < 1 || correction < 0 || correction > 2) return MOCAT_ERR_ARG;
  if (C == 0 || n_iter == 0) return MOCAT_OK;
  if (!theta || !momenta || !grad || !potential || !prior_pot || !lik_pot || !stepsize || !alpha || !keys)
    return MOCAT_ERR_ARG;
  HmcArgs h{{X, y, packed, (int)n, (int)d, prior_prec, temperature},
            C, theta, momenta, grad, potential, prior_pot, lik_pot, stepsize, alpha, keys, iter0, n_iter,
            leapfrog_steps, friction, correction, target, rm_scale, rm_neg_exponent, adapt_cut_off,
            trace_value, trace_potential, trace_alpha, trace_stepsize, trace_momenta};
  if (precision == 2) return dispatch_hmc<GaussEngine>(h, stream);
  if (precision == 3) return dispatch_hmc<Fp64Engine>(h, stream);
  return precision == 0 ? dispatch_hmc<Fp32Engine>(h, stream) : dispatch_hmc_bf16(h, stream);
}

}  // extern "C"
