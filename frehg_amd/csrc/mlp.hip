// Learned initial guess: feature extraction, pair logging, BF16 MFMA forward / backward, Adam;
// FP16 MFMA inference (k_mlp_guess).
//
// Every dense layer runs on v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] B[32x16], FP32
// accumulate). Lane l of a wave supplies A[l%16][8(l/16) .. 8(l/16)+7] and
// B[8(l/16) .. 8(l/16)+7][l%16], and receives D[4(l/16)+r][l%16], r = 0..3. Operands are staged
// in LDS as FP32 and rounded to BF16 (round to nearest even) when a lane packs them, so a
// transposed operand is just a different LDS stride.
//
// A block of 256 threads (4 waves) handles 64 rows at a time; wave w owns rows 16w..16w+15 in
// the forward pass and the 16-row block w of the weight gradients.
//   forward:  Z1 = X W1^T + b1, A1 = tanh Z1;  Z2 = A1 W2^T + b2, A2 = tanh Z2;  y = A2 w3 + b3
//   backward: dy = 2 (y - t) / B;  dZ2 = dy w3 (1 - A2^2);  dZ1 = (dZ2 W2) (1 - A1^2)
//             dW2 = dZ2^T A1, dW1 = dZ1^T X, dw3 = dy^T A2, db = column sums
// Gradients are written as per-block partials and summed in a fixed order by one kernel, which
// also applies the Adam update, so training is deterministic for a given seed.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/frehg_gpu.h"
#include "mlp.hpp"
#include "re.hpp"

namespace frehg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MT = 256;   // threads per block
constexpr int RB = 64;    // rows per block iteration
constexpr int XP = 33;    // LDS row pitch of X (32 padded inputs + 1)
constexpr int HP = 65;    // LDS row pitch of 64-wide activations
constexpr int TRAIN_BLOCKS = 256;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// 8 consecutive FP32 values (stride `st` floats) packed as a BF16 MFMA operand.
__device__ __forceinline__ bf16x8 pack8(const float* p, int st) {
  bf16x8 v;
#pragma unroll
  for (int t = 0; t < 8; ++t) v[t] = (__bf16)p[t * st];
  return v;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// tanh x = 1 - 2 / (1 + e^{2x}) with the hardware exp2 and reciprocal (absolute error ~1e-7, far
// below the BF16 rounding of the activations); saturates correctly for large |x|.
__device__ __forceinline__ float act_tanh(float x) {
  const float e = __expf(2.0f * x);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

// Orders this wave's LDS accesses in the compiler's view (a wave executes them in order).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inputs of the feature function: the Richards state before the solve.
struct FeatSrc {
  Geo g;
  const double* hn;
  const double* wcn;
  SoilView soil;
  const double* depth;  // surface padded, or null
  // Base increment per cell, inc_a - inc_b (the previous step's solve increment; inc_b null:
  // inc_a alone; inc_a null: zero): the network predicts the increment minus it.
  const double* inc_a;
  const double* inc_b;
  // Inference gate: with a previous increment available and gate > 0, a cell whose |previous
  // increment| <= gate takes the extrapolated start (no network correction).
  double gate;
  // Set to 1 by the inference kernel when a network output is not finite (that cell then takes
  // the extrapolated start); may be null.
  double* bad;
  // Cells of the chunks the network ran on (counted per chunk); may be null.
  unsigned long long* inferred;
  // Gated inference in two passes: masks[ch] = the cells of chunk ch whose correction the network
  // supplies (k_guess_base writes them, and the base start into the output); null: one pass.
  const unsigned long long* masks;
  // The next call's inferred-cell counter, cleared by this call's inference kernel (the counters
  // alternate, so that no memset sits between launches); may be null.
  unsigned long long* inferred_next;
};

__device__ __forceinline__ double base_increment(const FeatSrc& F, long long c) {
  if (!F.inc_a) return 0.0;
  return F.inc_b ? F.inc_a[c] - F.inc_b[c] : F.inc_a[c];
}

// ln K and C at head h for the features, from the hardware base-2 exponential and logarithm
// (the features enter the network in FP16). Van Genuchten classes: with p = (alpha |h|)^n,
// Se = (1 + p)^-m, K = Ks Se^(1/2) t^2 with t = 1 - (p / (1 + p))^m = 1 - exp(-m ln(1 + 1/p)),
// C = capc p Se / (alpha |h| (1 + p)); K = Ks and C = 0 for h > 0. ln(1 + x) and 1 - exp(-x) are
// taken from their series below 1e-3. Other classes through the FP64 law.
__device__ __forceinline__ float log2_1p(float x) {  // log2(1 + x), x >= 0
  constexpr float kLog2e = 1.44269504f;
  return x < 1e-3f ? x * kLog2e * (1.0f - 0.5f * x) : __builtin_amdgcn_logf(1.0f + x);
}
template <bool VG_ONLY>
__device__ inline void feature_kc(const SoilClass& s, double ksz, double h, float& lnk, float& cap) {
  constexpr float kLnFloor = -69.0775528f;  // ln(1e-30)
  constexpr float kLn2 = 0.693147181f, kLog2e = 1.44269504f;
  if (!VG_ONLY && (!s.use_vg || s.use_mvg)) {
    const VGState v = vg_state(s, h);
    lnk = logf((float)fmax(ksz * v.kr, 1e-30));
    cap = (float)v.cap;
    return;
  }
  const float lks = fmaxf(kLn2 * __builtin_amdgcn_logf((float)ksz), kLnFloor);
  const float a = fabsf((float)s.alpha * (float)h);
  if (h > 0.0 || a == 0.0f) {
    lnk = lks;
    cap = 0.0f;
    return;
  }
  const float m = (float)s.m;
  const float l2p = (float)s.n * __builtin_amdgcn_logf(a);  // log2 p
  const float p = __builtin_amdgcn_exp2f(l2p);
  const float l2_1p = log2_1p(p);                                       // log2(1 + p)
  const float md = m * kLn2 * log2_1p(__builtin_amdgcn_exp2f(-l2p));  // m ln(1 + 1/p)
  const float t = md < 1e-3f ? md * (1.0f - 0.5f * md) : 1.0f - __builtin_amdgcn_exp2f(-kLog2e * md);
  lnk = fmaxf(lks + fminf(kLn2 * (2.0f * __builtin_amdgcn_logf(t) - 0.5f * m * l2_1p), 0.0f), kLnFloor);
  cap = (float)s.capc * p * __builtin_amdgcn_exp2f(-m * l2_1p) / (a * (1.0f + p));
}

// VG_ONLY: every soil class is a plain van Genuchten class (no FP64 fallback code in the kernel).
template <bool VG_ONLY>
__device__ inline void cell_features(const FeatSrc& F, int i, int j, int k, float* f) {
  const Geo& g = F.g;
  const long long c = g.idx(i, j, k);
  const double hc = F.hn[c];
  f[0] = (float)hc;
  f[1] = (float)(F.hn[c - 1] - hc);
  f[2] = (float)(F.hn[c + 1] - hc);
  f[3] = (float)(F.hn[c - g.sy] - hc);
  f[4] = (float)(F.hn[c + g.sy] - hc);
  f[5] = (float)(F.hn[c - g.sz] - hc);
  f[6] = (float)(F.hn[c + g.sz] - hc);
  feature_kc<VG_ONLY>(F.soil.at(c), F.soil.Ksz(c), hc, f[7], f[8]);
  f[9] = (float)F.wcn[c];
  f[10] = F.depth ? (float)F.depth[g.sidx(i, j)] : 0.0f;
}

// Shared-memory image of the network and one block of 64 rows.
struct MlpLds {
  float w1[kMlpH][XP];   // W1 padded to 32 inputs
  float w2[kMlpH][HP];
  float w3[kMlpH];
  float b1[kMlpH], b2[kMlpH];
  float x[RB][XP];       // standardised inputs (zero beyond 11)
  float a1[RB][HP];
  float a2[RB][HP];
  float t[RB];           // standardised targets
  float dy[RB];
};

struct MlpBwdLds {
  float dz1[RB][HP];
  float dz2[RB][HP];
};

__device__ void load_weights(MlpLds& L, const float* P) {
  for (int q = threadIdx.x; q < kMlpH * 32; q += MT) {
    const int n = q / 32, f = q % 32;
    L.w1[n][f] = f < kMlpIn ? P[kOffW1 + n * kMlpIn + f] : 0.0f;
  }
  for (int q = threadIdx.x; q < kMlpH * kMlpH; q += MT) L.w2[q / kMlpH][q % kMlpH] = P[kOffW2 + q];
  if (threadIdx.x < kMlpH) {
    L.w3[threadIdx.x] = P[kOffW3 + threadIdx.x];
    L.b1[threadIdx.x] = P[kOffB1 + threadIdx.x];
    L.b2[threadIdx.x] = P[kOffB2 + threadIdx.x];
  }
}

// Forward pass of wave w's 16 rows; returns y for rows 16w + 4(l/16) + r in lanes l%16 == 0.
// Called by every thread of the block (it contains block barriers).
__device__ f32x4 forward16(MlpLds& L, float b3) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = 16 * w, ra = lane & 15, kg = 8 * (lane >> 4), rd = 4 * (lane >> 4);
  // layer 1: K = 32 (11 inputs, zero padded)
  const bf16x8 xa = pack8(&L.x[row0 + ra][kg], 1);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(xa, pack8(&L.w1[16 * nb + ra][kg], 1), acc);
    const int col = 16 * nb + ra;
#pragma unroll
    for (int r = 0; r < 4; ++r) L.a1[row0 + rd + r][col] = act_tanh(acc[r] + L.b1[col]);
  }
  __syncthreads();
  // layer 2: K = 64
  bf16x8 aa[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) aa[ks] = pack8(&L.a1[row0 + ra][32 * ks + kg], 1);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) acc = mfma(aa[ks], pack8(&L.w2[16 * nb + ra][32 * ks + kg], 1), acc);
    const int col = 16 * nb + ra;
#pragma unroll
    for (int r = 0; r < 4; ++r) L.a2[row0 + rd + r][col] = act_tanh(acc[r] + L.b2[col]);
  }
  __syncthreads();
  // layer 3: B = w3 in column 0, zero elsewhere
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    bf16x8 b;
    if (ra == 0) {
      b = pack8(&L.w3[32 * ks + kg], 1);
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t) b[t] = (__bf16)0.0f;
    }
    acc = mfma(pack8(&L.a2[row0 + ra][32 * ks + kg], 1), b, acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] += b3;
  return acc;
}

// Stage 64 rows (11 raw features + raw target each) into LDS, standardised. Rows past `valid`
// are zero and carry no weight.
__device__ void stage_rows(MlpLds& L, const float* rows, const long long* idx, int valid, const float* norm) {
  for (int q = threadIdx.x; q < RB * 32; q += MT) {
    const int s = q / 32, f = q % 32;
    float v = 0.0f;
    if (s < valid && f < kMlpIn) v = (rows[idx[s] * kMlpRow + f] - norm[f]) / norm[kMlpRow + f];
    L.x[s][f] = v;
  }
  if (threadIdx.x < RB) {
    const int s = threadIdx.x;
    L.t[s] = s < valid ? (rows[idx[s] * kMlpRow + kMlpIn] - norm[kMlpIn]) / norm[kMlpRow + kMlpIn] : 0.0f;
  }
}

// Loss and gradient partials over `nrows` rows; row q of the batch is log row
// (sequential ? q : hash(seed, q) mod nlog). part[block][kMlpParams + 1] (last = sum of squares).
__global__ __launch_bounds__(MT) void k_mlp_grad(const float* P, const float* norm, const float* rows,
                                                  long long nlog, long long nrows, int sequential, uint64_t seed,
                                                  float inv_batch, float* part) {
  __shared__ MlpLds L;
  __shared__ MlpBwdLds B;
  __shared__ long long idx[RB];
  __shared__ float red[MT / 64];
  load_weights(L, P);
  const float b3 = P[kOffB3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ra = lane & 15, kg = 8 * (lane >> 4), rd = 4 * (lane >> 4);
  f32x4 gW2[4], gW1[2], gW3;
  for (int q = 0; q < 4; ++q) gW2[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  gW1[0] = gW1[1] = gW3 = f32x4{0.f, 0.f, 0.f, 0.f};
  float gb = 0.0f;  // thread n < 64: db1[n]; 64 <= thread < 128: db2[n-64]; thread 128: db3
  float sq = 0.0f;
  const long long nchunks = (nrows + RB - 1) / RB;
  for (long long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const long long q0 = ch * RB;
    const int valid = (int)(nrows - q0 < RB ? nrows - q0 : RB);
    if (threadIdx.x < RB) {
      const long long q = q0 + threadIdx.x;
      idx[threadIdx.x] = threadIdx.x < valid ? (sequential ? q : (long long)(splitmix64(seed ^ (uint64_t)q) % (uint64_t)nlog)) : 0;
    }
    __syncthreads();
    stage_rows(L, rows, idx, valid, norm);
    __syncthreads();
    const f32x4 y = forward16(L, b3);
    if (ra == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int s = 16 * w + rd + r;
        const float e = s < valid ? y[r] - L.t[s] : 0.0f;
        sq += e * e;
        L.dy[s] = 2.0f * e * inv_batch;
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < RB * kMlpH; q += MT) {
      const int s = q / kMlpH, n = q % kMlpH;
      const float a = L.a2[s][n];
      B.dz2[s][n] = L.dy[s] * L.w3[n] * (1.0f - a * a);
    }
    __syncthreads();
    // dA1 = dZ2 W2 for wave w's rows, then dZ1 = dA1 (1 - A1^2)
    {
      const int row0 = 16 * w;
      bf16x8 za[2];
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) za[ks] = pack8(&B.dz2[row0 + ra][32 * ks + kg], 1);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) acc = mfma(za[ks], pack8(&L.w2[32 * ks + kg][16 * nb + ra], HP), acc);
        const int col = 16 * nb + ra;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float a = L.a1[row0 + rd + r][col];
          B.dz1[row0 + rd + r][col] = acc[r] * (1.0f - a * a);
        }
      }
    }
    __syncthreads();
    // weight gradients: K = the block's 64 rows
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int s0 = 32 * ks + kg;
      const bf16x8 dz2a = pack8(&B.dz2[s0][16 * w + ra], HP);  // A[i][s] = dZ2[s][i]
      const bf16x8 dz1a = pack8(&B.dz1[s0][16 * w + ra], HP);
#pragma unroll
      for (int nj = 0; nj < 4; ++nj) gW2[nj] = mfma(dz2a, pack8(&L.a1[s0][16 * nj + ra], HP), gW2[nj]);
#pragma unroll
      for (int nj = 0; nj < 2; ++nj) gW1[nj] = mfma(dz1a, pack8(&L.x[s0][16 * nj + ra], XP), gW1[nj]);
      bf16x8 dya;
      if (ra == 0) {
        dya = pack8(&L.dy[s0], 1);
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) dya[t] = (__bf16)0.0f;
      }
      gW3 = mfma(dya, pack8(&L.a2[s0][16 * w + ra], HP), gW3);
    }
    if (threadIdx.x < kMlpH) {
      for (int s = 0; s < RB; ++s) gb += B.dz1[s][threadIdx.x];
    } else if (threadIdx.x < 2 * kMlpH) {
      for (int s = 0; s < RB; ++s) gb += B.dz2[s][threadIdx.x - kMlpH];
    } else if (threadIdx.x == 2 * kMlpH) {
      for (int s = 0; s < RB; ++s) gb += L.dy[s];
    }
    __syncthreads();
  }
  // partials of this block
  float* out = part + (size_t)blockIdx.x * (kMlpParams + 1);
#pragma unroll
  for (int nj = 0; nj < 4; ++nj)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[kOffW2 + (16 * w + rd + r) * kMlpH + 16 * nj + ra] = gW2[nj][r];
#pragma unroll
  for (int nj = 0; nj < 2; ++nj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = 16 * nj + ra;
      if (f < kMlpIn) out[kOffW1 + (16 * w + rd + r) * kMlpIn + f] = gW1[nj][r];
    }
  if (lane < 16) out[kOffW3 + 16 * w + ra] = gW3[0];  // row 0 of the tile: lanes 0..15, r = 0
  if (threadIdx.x < kMlpH) out[kOffB1 + threadIdx.x] = gb;
  else if (threadIdx.x < 2 * kMlpH) out[kOffB2 + threadIdx.x - kMlpH] = gb;
  else if (threadIdx.x == 2 * kMlpH) out[kOffB3] = gb;
  // sum of squared errors
  float v = sq;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[kMlpParams] = red[0] + red[1] + red[2] + red[3];
}

// Sum the partials in block order; optionally apply one Adam step.
__global__ __launch_bounds__(MT) void k_mlp_reduce_adam(const float* part, int nblocks, float* grad, float* P,
                                                         float* m, float* v, int adam, float lr, float b1t,
                                                         float b2t) {
  const int p = blockIdx.x * MT + threadIdx.x;
  if (p > kMlpParams) return;
  float g = 0.0f;
  for (int b = 0; b < nblocks; ++b) g += part[(size_t)b * (kMlpParams + 1) + p];
  grad[p] = g;
  if (!adam || p == kMlpParams) return;
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.0f - b1t), vh = vv / (1.0f - b2t);
  P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Forward predictions (raw increments) for rows of a host batch.
__global__ __launch_bounds__(MT) void k_mlp_predict_rows(const float* P, const float* norm, const float* rows,
                                                          long long nrows, float* out) {
  __shared__ MlpLds L;
  __shared__ long long idx[RB];
  load_weights(L, P);
  const float b3 = P[kOffB3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, ra = lane & 15, rd = 4 * (lane >> 4);
  const long long nchunks = (nrows + RB - 1) / RB;
  for (long long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const long long q0 = ch * RB;
    const int valid = (int)(nrows - q0 < RB ? nrows - q0 : RB);
    if (threadIdx.x < RB) idx[threadIdx.x] = threadIdx.x < valid ? q0 + threadIdx.x : 0;
    __syncthreads();
    stage_rows(L, rows, idx, valid, norm);
    __syncthreads();
    const f32x4 y = forward16(L, b3);
    if (ra == 0)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int s = 16 * w + rd + r;
        if (s < valid) out[q0 + s] = y[r] * norm[kMlpRow + kMlpIn] + norm[kMlpIn];
      }
    __syncthreads();
  }
}

// Learned guess on every interior cell: h_out = h^n + increment.
//
// Inference form of the network (k_mlp_prep_infer). With s = 2 log2(e) and r = 1 / (1 + 2^(s z)),
// tanh z = 1 - 2 r: a hidden layer outputs r (one exp2, one add, one reciprocal per neuron) and
// the next layer absorbs the affine map, W (1 - 2 r) + b = (-2 W) r + (b + W 1); the scale s of
// the pre-activations is carried by the weights and biases. The layers run on
// v_mfma_f32_16x16x32_f16 with the biases as the accumulator's initial value; weights are rounded
// to FP16 (nearest even), the standardised features and the r in [0, 1] to FP16 (toward zero,
// two values per conversion).
//
// The layers run transposed (Z^T = W X^T: neurons are MFMA rows, cells are columns) so that each
// layer's output registers already hold the next layer's B operand, with no LDS round trip
// between layers. After layer 1, lane l (g = l/16, c = l%16) holds neurons 16 nb + 4 g + r
// (nb, r = 0..3) of cell c. The next layer's B operand asks lane l for reduction indices
// kappa = 32 ks + 8 g + t (ks = 0..1, t = 0..7); the bijection
//   pi(32 ks + 8 g + t) = 16 (2 ks + t/4) + 4 g + t%4
// maps every kappa a lane needs to a neuron the lane holds, so the weight operands are stored
// with their columns permuted by pi (a relabelling of the hidden units, exact).
// Each wave works alone on chunks of 64 consecutive cells of one grid row; every lane computes
// one cell's features into the wave's LDS rows, which are the only LDS traffic.
constexpr int GW = 4;   // waves per block of the guess kernel
constexpr int FP = 16;  // LDS pitch of a feature row (11 used, zero padded)

// Inference parameter vector (FP32): W1s[64][16] (inputs zero padded), B1s[64], W2s[64][64],
// B2s[64], W3s[64], B3s.
constexpr int kIW1 = 0, kIB1 = kIW1 + kMlpH * FP, kIW2 = kIB1 + kMlpH, kIB2 = kIW2 + kMlpH * kMlpH,
              kIW3 = kIB2 + kMlpH, kIB3 = kIW3 + kMlpH, kInfer = kIB3 + 1;

__global__ __launch_bounds__(MT) void k_mlp_prep_infer(const float* P, float* Q) {
  const float s = 2.0f * 1.44269504088896341f;  // 2 log2(e)
  for (int q = threadIdx.x; q < kMlpH * FP; q += MT) {
    const int n = q / FP, f = q % FP;
    Q[kIW1 + q] = f < kMlpIn ? s * P[kOffW1 + n * kMlpIn + f] : 0.0f;
  }
  for (int q = threadIdx.x; q < kMlpH * kMlpH; q += MT) Q[kIW2 + q] = (-2.0f * s) * P[kOffW2 + q];
  if (threadIdx.x < kMlpH) {
    const int n = threadIdx.x;
    float sum = 0.0f;
    for (int k = 0; k < kMlpH; ++k) sum += P[kOffW2 + n * kMlpH + k];
    Q[kIB1 + n] = s * P[kOffB1 + n];
    Q[kIB2 + n] = s * (P[kOffB2 + n] + sum);
    Q[kIW3 + n] = -2.0f * P[kOffW3 + n];
  }
  if (threadIdx.x == 0) {
    float sum = 0.0f;
    for (int k = 0; k < kMlpH; ++k) sum += P[kOffW3 + k];
    Q[kIB3] = P[kOffB3] + sum;
  }
}

__device__ __forceinline__ int hidden_perm(int kappa) {
  const int ks = kappa >> 5, g = (kappa >> 3) & 3, t = kappa & 7;
  return 16 * (2 * ks + (t >> 2)) + 4 * g + (t & 3);
}

// Two FP32 values as one FP16 pair (round toward zero).
__device__ __forceinline__ unsigned pk16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(a, b));
}
__device__ __forceinline__ f16x8 pk16x8(float v0, float v1, float v2, float v3, float v4, float v5, float v6,
                                        float v7) {
  return __builtin_bit_cast(f16x8, u32x4{pk16(v0, v1), pk16(v2, v3), pk16(v4, v5), pk16(v6, v7)});
}
__device__ __forceinline__ f32x4 mfma16(f16x8 a, f16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
// r = 1 / (1 + 2^z) (z = s times the pre-activation); 0 for large z, 1 for very negative z.
__device__ __forceinline__ float act_r(float z) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z)); }

// Three waves per SIMD for the van Genuchten variant (at most 168 VGPRs, the weights in registers).
template <bool VG_ONLY>
__global__ __launch_bounds__(64 * GW) __attribute__((amdgpu_waves_per_eu(VG_ONLY ? 3 : 1))) void k_mlp_guess(const float* Q, const float* norm, FeatSrc F, double* h_out) {
  __shared__ __attribute__((aligned(16))) float xs[GW][RB][FP];
  __shared__ float ys[GW][RB];  // network outputs of the chunk's cells
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ra = lane & 15, g = lane >> 4, kg = 8 * g;
  // A operands (rows = neurons): W1s[16 nb + ra][kg + t] (kg < 16); W2s[16 nb + ra][pi(32 ks + kg + t)];
  // W3s[pi(32 ks + kg + t)] in row 0 only. Accumulator initial values: the biases of the neurons
  // this lane's outputs hold.
  f16x8 a1w[4], a2w[4][2], a3w[2];
  f32x4 b1r[4], b2r[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const int n = 16 * nb + ra;
#pragma unroll
    for (int t = 0; t < 8; ++t) a1w[nb][t] = (_Float16)(kg < FP ? Q[kIW1 + n * FP + kg + t] : 0.0f);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int t = 0; t < 8; ++t) a2w[nb][ks][t] = (_Float16)Q[kIW2 + n * kMlpH + hidden_perm(32 * ks + kg + t)];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      b1r[nb][r] = Q[kIB1 + 16 * nb + 4 * g + r];
      b2r[nb][r] = Q[kIB2 + 16 * nb + 4 * g + r];
    }
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int t = 0; t < 8; ++t) a3w[ks][t] = (_Float16)(ra == 0 ? Q[kIW3 + hidden_perm(32 * ks + kg + t)] : 0.0f);
  const f32x4 b3r = {Q[kIB3], 0.f, 0.f, 0.f};
  float nmu[kMlpIn], nisd[kMlpIn];
#pragma unroll
  for (int t = 0; t < kMlpIn; ++t) {
    nmu[t] = norm[t];
    nisd[t] = 1.0f / norm[kMlpRow + t];
  }
  const float ymu = norm[kMlpIn], ysd = norm[kMlpRow + kMlpIn];
  const Geo& g3 = F.g;
  const int iblocks = (g3.nx + RB - 1) / RB;
  const int nchunks = iblocks * g3.ny * g3.nz;  // fits in 32 bits (checked on the host)
  if (F.inferred_next && blockIdx.x == 0 && threadIdx.x == 0) *F.inferred_next = 0ull;
  float(*xw)[FP] = xs[w];
  unsigned long long inferred = 0;  // cells of the chunks this wave ran the network on
  // Wave wid takes chunks wid, wid + nw, wid + 2 nw, ... With masks, lane l first reads the mask
  // of the l-th of the next 64 of them, and the wave visits only those with a cell to correct.
  const int nw = (int)gridDim.x * GW, wid = (int)blockIdx.x * GW + w;
  const int span = F.masks ? 64 : 1;
  for (int it0 = wid; it0 < nchunks; it0 += nw * span) {
    unsigned long long lm = 0ull, todo = 1ull;
    if (F.masks) {
      const long long cq = (long long)it0 + (long long)lane * nw;
      lm = cq < nchunks ? F.masks[cq] : 0ull;
      todo = __ballot(lm != 0ull);
    }
    while (todo) {
      const int l = __ffsll((long long)todo) - 1;
      todo &= todo - 1ull;
      const int ch = it0 + l * nw;
      const unsigned long long mask = F.masks ? __shfl(lm, l) : 0ull;
      const int row = ch / iblocks, ib = ch - row * iblocks;
      const int j = row % g3.ny, k = row / g3.ny;
      const int i = ib * RB + lane;
      float f[kMlpIn];
#pragma unroll
      for (int t = 0; t < kMlpIn; ++t) f[t] = 0.0f;
      long long c = 0;
      double base = 0.0;  // h^n + base increment of this lane's cell
      bool need = false;  // this lane's cell takes the network's correction
      if (i < g3.nx) {
        c = g3.idx(i, j, k);
        if (F.masks) {  // k_guess_base wrote h^n + base increment here (in place of the increment's inputs)
          base = h_out[c];
          need = ((mask >> lane) & 1ull) != 0ull;
        } else {
          const double inc = base_increment(F, c);
          base = F.hn[c] + inc;
          need = !(F.gate > 0.0 && F.inc_a) || fabs(inc) > F.gate;
        }
      }
      if (!__any(need)) {  // the whole chunk takes the extrapolated start (wave-uniform)
        if (i < g3.nx) h_out[c] = base;
        continue;
      }
      inferred += (unsigned long long)min(RB, g3.nx - ib * RB);  // per wave; one atomic per block below
      if (i < g3.nx) {
        cell_features<VG_ONLY>(F, i, j, k, f);
        // standardised features, clamped far outside any training distribution so that the FP16
        // operands below stay finite (heads of -1e5 m and beyond occur where the corrector drains a
        // cell to theta_r)
#pragma unroll
        for (int t = 0; t < kMlpIn; ++t) f[t] = fminf(fmaxf((f[t] - nmu[t]) * nisd[t], -1024.0f), 1024.0f);
      }
      wave_lds_fence();
#pragma unroll
      for (int t = 0; t < kMlpIn; ++t) xw[lane][t] = f[t];
#pragma unroll
      for (int t = kMlpIn; t < FP; ++t) xw[lane][t] = 0.0f;
      wave_lds_fence();
#pragma unroll 1
      for (int tile = 0; tile < 4; ++tile) {
        // layer 1: B = X^T, lane supplies features kg..kg+7 of cell 16 tile + ra (zero for kg >= 16)
        f16x8 xb = pk16x8(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
        if (g < 2) {
          const float* xr = &xw[16 * tile + ra][kg];
          xb = pk16x8(xr[0], xr[1], xr[2], xr[3], xr[4], xr[5], xr[6], xr[7]);
        }
        float a[4][4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
          const f32x4 z = mfma16(a1w[nb], xb, b1r[nb]);
#pragma unroll
          for (int r = 0; r < 4; ++r) a[nb][r] = act_r(z[r]);
        }
        // layer 2: B[kappa][cell] = r1[pi(kappa)] from this lane's registers
        f16x8 hb[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          hb[ks] = pk16x8(a[2 * ks][0], a[2 * ks][1], a[2 * ks][2], a[2 * ks][3], a[2 * ks + 1][0],
                          a[2 * ks + 1][1], a[2 * ks + 1][2], a[2 * ks + 1][3]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
          f32x4 z = b2r[nb];
#pragma unroll
          for (int ks = 0; ks < 2; ++ks) z = mfma16(a2w[nb][ks], hb[ks], z);
#pragma unroll
          for (int r = 0; r < 4; ++r) a[nb][r] = act_r(z[r]);
        }
        // layer 3: row 0 of W3s[pi] times R2^T
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          hb[ks] = pk16x8(a[2 * ks][0], a[2 * ks][1], a[2 * ks][2], a[2 * ks][3], a[2 * ks + 1][0],
                          a[2 * ks + 1][1], a[2 * ks + 1][2], a[2 * ks + 1][3]);
        f32x4 y = b3r;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) y = mfma16(a3w[ks], hb[ks], y);
        if (lane < 16) ys[w][16 * tile + ra] = y[0];  // D row 0 (g = 0, r = 0) holds y of cell ra
      }
      wave_lds_fence();
      if (i < g3.nx) {
        const double y = (double)(ys[w][lane] * ysd + ymu);
        const bool ok = isfinite(y);
        if (!ok && F.bad) *F.bad = 1.0;
        h_out[c] = base + ((need && ok) ? y : 0.0);
      }
    }
  }
  if (F.inferred) {
    __shared__ unsigned long long wsum[GW];
    if (lane == 0) wsum[w] = inferred;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int q = 0; q < GW; ++q) t += wsum[q];
      if (t) atomicAdd(F.inferred, t);
    }
  }
}

// The gated start's first pass: every interior cell gets the extrapolated start h^n + base
// increment, and masks[ch] the cells of chunk ch (RB cells along x, one wave) whose |previous
// increment| exceeds the gate. A light streaming pass at full occupancy, so that the inference's
// register-heavy waves visit only the chunks with a cell to correct. The increment's inputs may be
// the output array itself, so the inference kernel takes the base from the output and the cells
// from the mask.
__global__ __launch_bounds__(MT) void k_guess_base(FeatSrc F, double* h_out, unsigned long long* masks) {
  const int lane = threadIdx.x & 63;
  const Geo& g3 = F.g;
  const int iblocks = (g3.nx + RB - 1) / RB;
  const long long ch = (long long)blockIdx.x * (MT / 64) + (threadIdx.x >> 6);
  if (ch >= (long long)iblocks * g3.ny * g3.nz) return;  // wave-uniform
  const int row = (int)(ch / iblocks), ib = (int)(ch - (long long)row * iblocks);
  const int j = row % g3.ny, k = row / g3.ny;
  const int i = ib * RB + lane;
  bool need = false;
  if (i < g3.nx) {
    const long long c = g3.idx(i, j, k);
    const double inc = base_increment(F, c);
    h_out[c] = F.hn[c] + inc;
    need = fabs(inc) > F.gate;
  }
  const unsigned long long m = __ballot(need);
  if (lane == 0) masks[ch] = m;
}

// Log (features, h - h^n - base increment) of interior cells q with q % stride == phase into the
// ring buffer.
template <bool VG_ONLY>
__global__ __launch_bounds__(MT) void k_mlp_log(FeatSrc F, const double* h, int stride, int phase, long long base,
                                                 long long cap, float* log) {
  const Geo& g = F.g;
  const long long N = (long long)g.nx * g.ny * g.nz;
  const long long qq = (long long)blockIdx.x * MT + threadIdx.x;  // index among sampled cells
  const long long q = qq * stride + phase;
  if (q >= N) return;
  const int i = (int)(q % g.nx), j = (int)((q / g.nx) % g.ny), k = (int)(q / ((long long)g.nx * g.ny));
  float f[kMlpIn];
  cell_features<VG_ONLY>(F, i, j, k, f);
  float* row = log + ((base + qq) % cap) * kMlpRow;
#pragma unroll
  for (int t = 0; t < kMlpIn; ++t) row[t] = f[t];
  const long long c = g.idx(i, j, k);
  row[kMlpIn] = (float)(h[c] - F.hn[c] - base_increment(F, c));
}

// Per-block sums and sums of squares of the 12 logged columns (FP64).
__global__ __launch_bounds__(MT) void k_mlp_stats(const float* log, long long n, double* part) {
  double s[kMlpRow], s2[kMlpRow];
  for (int t = 0; t < kMlpRow; ++t) s[t] = s2[t] = 0.0;
  for (long long r = (long long)blockIdx.x * MT + threadIdx.x; r < n; r += (long long)gridDim.x * MT)
    for (int t = 0; t < kMlpRow; ++t) {
      const double v = log[r * kMlpRow + t];
      s[t] += v;
      s2[t] += v * v;
    }
  __shared__ double lds[MT / 64];
  for (int t = 0; t < kMlpRow; ++t) {
    const double a = block_sum<MT>(s[t], lds);
    const double b = block_sum<MT>(s2[t], lds);
    if (threadIdx.x == 0) {
      part[(size_t)blockIdx.x * 2 * kMlpRow + t] = a;
      part[(size_t)blockIdx.x * 2 * kMlpRow + kMlpRow + t] = b;
    }
  }
}

FeatSrc feat_src(ReState& re, const double* depth) {
  FeatSrc F;
  F.g = re.geo();
  F.hn = re.d_hn();
  F.wcn = re.d_wcn();
  F.soil = re.soil_view();
  F.depth = depth;
  F.inc_a = nullptr;
  F.inc_b = nullptr;
  F.gate = 0.0;
  F.bad = nullptr;
  F.inferred = nullptr;
  F.masks = nullptr;
  F.inferred_next = nullptr;
  return F;
}

constexpr int kStatBlocks = 256;

}  // namespace

Mlp::Mlp(hipStream_t s, Profiler* prof) : s_(s), prof_(prof) {
  FREHG_HIP(hipMalloc(&params_, kMlpParams * sizeof(float)));
  FREHG_HIP(hipMalloc(&infer_, kInfer * sizeof(float)));
  FREHG_HIP(hipMalloc(&m_, kMlpParams * sizeof(float)));
  FREHG_HIP(hipMalloc(&v_, kMlpParams * sizeof(float)));
  FREHG_HIP(hipMalloc(&norm_, kMlpNorm * sizeof(float)));
  FREHG_HIP(hipMalloc(&grad_, (kMlpParams + 1) * sizeof(float)));
  FREHG_HIP(hipMalloc(&part_, (size_t)TRAIN_BLOCKS * (kMlpParams + 1) * sizeof(float)));
  FREHG_HIP(hipMalloc(&stat_, (size_t)kStatBlocks * 2 * kMlpRow * sizeof(double)));
  FREHG_HIP(hipHostMalloc(&hbuf_, (kMlpParams + 1) * sizeof(float), hipHostMallocDefault));
  FREHG_HIP(hipMalloc(&inferred_, 2 * sizeof(unsigned long long)));
  FREHG_HIP(hipMemsetAsync(inferred_, 0, 2 * sizeof(unsigned long long), s_));
  // The guess kernel's blocks loop over chunks of cells: one round of resident blocks, so that
  // no partial last round idles most of the chip and each block loads the weights once.
  int dev = 0, cus = 0, per_cu = 0;
  FREHG_HIP(hipGetDevice(&dev));
  FREHG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  FREHG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_guess<true>, 64 * GW, 0));
  guess_blocks_ = (unsigned)std::max(1, cus * std::max(1, per_cu));
  init(1);
}

Mlp::~Mlp() {
  for (void* p : {(void*)params_, (void*)infer_, (void*)m_, (void*)v_, (void*)norm_, (void*)grad_, (void*)part_, (void*)log_,
                  (void*)scratch_, (void*)stat_, (void*)inferred_, (void*)masks_})
    if (p) (void)hipFree(p);
  if (hbuf_) (void)hipHostFree(hbuf_);
}

long long Mlp::inferred_cells() {
  unsigned long long n = 0;
  FREHG_HIP(hipMemcpyAsync(&n, inferred_ + last_par_, sizeof(n), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  return (long long)n;
}

void Mlp::init(uint64_t seed) {
  std::vector<float> P(kMlpParams, 0.0f), N(kMlpNorm, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 12345;
  auto uni = [&]() {  // splitmix64 -> [0, 1)
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (int q = 0; q < fan_out * fan_in; ++q) P[off + q] = (float)((2.0 * uni() - 1.0) * a);
  };
  fill(kOffW1, kMlpH, kMlpIn);
  fill(kOffW2, kMlpH, kMlpH);
  fill(kOffW3, 1, kMlpH);
  for (int t = 0; t < kMlpRow; ++t) N[kMlpRow + t] = 1.0f;
  FREHG_HIP(hipMemcpy(params_, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  FREHG_HIP(hipMemset(m_, 0, kMlpParams * sizeof(float)));
  FREHG_HIP(hipMemset(v_, 0, kMlpParams * sizeof(float)));
  FREHG_HIP(hipMemcpy(norm_, N.data(), kMlpNorm * sizeof(float), hipMemcpyHostToDevice));
  adam_t_ = 0;
  have_norm_ = false;
  infer_stale_ = true;
  ready_ = false;
}

void Mlp::set_logging(int stride, long long capacity) {
  if (stride < 0) throw ArgError("frehg_mlp_log: stride must be >= 0");
  if (stride > 0 && capacity < 1) throw ArgError("frehg_mlp_log: capacity must be >= 1");
  if (stride > 0 && capacity != cap_) {
    if (log_) FREHG_HIP(hipFree(log_));
    log_ = nullptr;
    FREHG_HIP(hipMalloc(&log_, (size_t)capacity * kMlpRow * sizeof(float)));
    cap_ = capacity;
    count_ = 0;
  }
  stride_ = stride;
}

void Mlp::log_pair(ReState& re, const double* hn, const double* h, const double* depth, const double* inc0) {
  if (!logging()) return;
  FeatSrc F = feat_src(re, depth);
  F.hn = hn;
  F.inc_a = inc0;
  const Geo& g = F.g;
  const long long N = (long long)g.nx * g.ny * g.nz;
  const int phase = (int)(log_steps_ % stride_);
  const long long nsamp = phase < N ? (N - phase + stride_ - 1) / stride_ : 0;
  if (nsamp > 0) {
    const dim3 grid((unsigned)ceil_div(nsamp, (long long)MT));
    if (re.soil_plain_vg())
      hipLaunchKernelGGL(k_mlp_log<true>, grid, dim3(MT), 0, s_, F, h, stride_, phase, count_ % cap_, cap_, log_);
    else
      hipLaunchKernelGGL(k_mlp_log<false>, grid, dim3(MT), 0, s_, F, h, stride_, phase, count_ % cap_, cap_, log_);
    FREHG_LAUNCH_CHECK();
  }
  count_ += nsamp;
  ++log_steps_;
}

double Mlp::run_grad(const float* rows, long long nrows, int batch, uint64_t seed, int sequential) {
  const long long n = sequential ? nrows : batch;
  const int nb = (int)std::min<long long>(TRAIN_BLOCKS, (n + RB - 1) / RB);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(MT), 0, s_, (const float*)params_, (const float*)norm_, rows, nrows,
                     n, sequential, seed, 1.0f / (float)n, part_);
  FREHG_LAUNCH_CHECK();
  return (double)nb;
}

double Mlp::train(int steps, int batch, double lr, uint64_t seed, double* first) {
  const long long nlog = logged();
  if (nlog < 1) throw ArgError("frehg_mlp_train: no logged pairs (enable logging and run plain steps first)");
  if (steps < 1 || batch < 1) throw ArgError("frehg_mlp_train: steps and batch must be >= 1");
  if (!have_norm_) {
    hipLaunchKernelGGL(k_mlp_stats, dim3(kStatBlocks), dim3(MT), 0, s_, (const float*)log_, nlog, stat_);
    FREHG_LAUNCH_CHECK();
    std::vector<double> part((size_t)kStatBlocks * 2 * kMlpRow);
    FREHG_HIP(hipMemcpyAsync(part.data(), stat_, part.size() * sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    std::vector<float> N(kMlpNorm);
    for (int t = 0; t < kMlpRow; ++t) {
      double s = 0, s2 = 0;
      for (int b = 0; b < kStatBlocks; ++b) {
        s += part[(size_t)b * 2 * kMlpRow + t];
        s2 += part[(size_t)b * 2 * kMlpRow + kMlpRow + t];
      }
      const double mu = s / (double)nlog;
      const double var = std::max(0.0, s2 / (double)nlog - mu * mu);
      const double sd = std::sqrt(var);
      N[t] = (float)mu;
      N[kMlpRow + t] = sd > 1e-12 * std::max(1.0, std::fabs(mu)) ? (float)sd : 1.0f;
    }
    FREHG_HIP(hipMemcpyAsync(norm_, N.data(), kMlpNorm * sizeof(float), hipMemcpyHostToDevice, s_));
    have_norm_ = true;
  }
  double loss_first = 0.0, loss = 0.0;
  for (int it = 0; it < steps; ++it) {
    const int nb = (int)run_grad(log_, nlog, batch, seed + 0x100000001ull * (uint64_t)(adam_t_ + 1), 0);
    ++adam_t_;
    const float b1t = (float)std::pow(0.9, adam_t_), b2t = (float)std::pow(0.999, adam_t_);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(ceil_div(kMlpParams + 1, MT)), dim3(MT), 0, s_, (const float*)part_, nb,
                       grad_, params_, m_, v_, 1, (float)lr, b1t, b2t);
    FREHG_LAUNCH_CHECK();
    if (it == 0 || it == steps - 1) {
      FREHG_HIP(hipMemcpyAsync(hbuf_, grad_ + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, s_));
      FREHG_HIP(hipStreamSynchronize(s_));
      const double l = hbuf_[0] / (double)batch;
      if (it == 0) loss_first = l;
      loss = l;
    }
  }
  if (first) *first = loss_first;
  ready_ = true;
  infer_stale_ = true;
  return loss;
}

void Mlp::predict_guess(ReState& re, double* h_out, const double* depth, const double* inc_a, const double* inc_b,
                        const double* hn, const double* wcn, double* bad) {
  if (!ready_) throw ArgError("learned guess requested but the model is not trained or loaded");
  FeatSrc F = feat_src(re, depth);
  F.inc_a = inc_a;
  F.inc_b = inc_b;
  F.gate = gate_;
  F.bad = bad;
  // The counters alternate between two slots from call to call: this call uses slot `par` (cleared
  // by the previous call's inference kernel, or at set-up) and its inference kernel clears the other
  // one for the next call, so no memset sits between the launches.
  const int par = 1 - last_par_;
  last_par_ = par;
  F.inferred = inferred_ + par;
  F.inferred_next = inferred_ + (1 - par);
  if (hn) F.hn = hn;
  if (wcn) F.wcn = wcn;
  const long long nchunks = (long long)((F.g.nx + RB - 1) / RB) * F.g.ny * F.g.nz;
  if (nchunks >= (1LL << 31) - (long long)guess_blocks_ * GW) throw ArgError("learned guess: grid too large");
  const unsigned nb = (unsigned)std::min<long long>(guess_blocks_, (nchunks + GW - 1) / GW);
  if (infer_stale_) {
    hipLaunchKernelGGL(k_mlp_prep_infer, dim3(1), dim3(MT), 0, s_, (const float*)params_, infer_);
    FREHG_LAUNCH_CHECK();
    infer_stale_ = false;
  }
  hipEvent_t ev0 = prof_->start(FREHG_KT_GW_GUESS);
  hipEvent_t ev1 = prof_->stop();
  const bool gated = F.gate > 0.0 && F.inc_a != nullptr;
  if (gated) {  // the base pass marks the cells that need the network; the inference visits only their chunks
    chunk_masks(nchunks);
    hipExtLaunchKernelGGL(k_guess_base, dim3(ceil_div(nchunks, MT / 64)), dim3(MT), 0, s_, ev0, nullptr, 0, F, h_out,
                          masks_);
    FREHG_LAUNCH_CHECK();
    ev0 = nullptr;
    F.masks = masks_;
  }
  if (re.soil_plain_vg())
    hipExtLaunchKernelGGL(k_mlp_guess<true>, dim3(nb), dim3(64 * GW), 0, s_, ev0, ev1, 0, (const float*)infer_,
                          (const float*)norm_, F, h_out);
  else
    hipExtLaunchKernelGGL(k_mlp_guess<false>, dim3(nb), dim3(64 * GW), 0, s_, ev0, ev1, 0, (const float*)infer_,
                          (const float*)norm_, F, h_out);
  FREHG_LAUNCH_CHECK();
}

unsigned long long* Mlp::chunk_masks(long long nchunks) {
  if (masks_cap_ < nchunks) {
    if (masks_) FREHG_HIP(hipFree(masks_));
    masks_ = nullptr;
    masks_cap_ = 0;
    FREHG_HIP(hipMalloc(&masks_, (size_t)nchunks * sizeof(unsigned long long)));
    masks_cap_ = nchunks;
  }
  return masks_;
}

void Mlp::infer_masked(ReState& re, double* h_out, const double* depth, double* bad) {
  if (!ready_) throw ArgError("learned guess requested but the model is not trained or loaded");
  FeatSrc F = feat_src(re, depth);
  F.gate = gate_;
  F.bad = bad;
  const int par = 1 - last_par_;  // the counter slots alternate as in predict_guess
  last_par_ = par;
  F.inferred = inferred_ + par;
  F.inferred_next = inferred_ + (1 - par);
  F.masks = masks_;
  const long long nchunks = (long long)((F.g.nx + RB - 1) / RB) * F.g.ny * F.g.nz;
  if (!masks_ || masks_cap_ < nchunks) throw ArgError("learned guess: chunk masks not set up");
  if (nchunks >= (1LL << 31) - (long long)guess_blocks_ * GW) throw ArgError("learned guess: grid too large");
  const unsigned nb = (unsigned)std::min<long long>(guess_blocks_, (nchunks + GW - 1) / GW);
  if (infer_stale_) {
    hipLaunchKernelGGL(k_mlp_prep_infer, dim3(1), dim3(MT), 0, s_, (const float*)params_, infer_);
    FREHG_LAUNCH_CHECK();
    infer_stale_ = false;
  }
  hipEvent_t ev0 = prof_->start(FREHG_KT_GW_GUESS);
  hipEvent_t ev1 = prof_->stop();
  if (re.soil_plain_vg())
    hipExtLaunchKernelGGL(k_mlp_guess<true>, dim3(nb), dim3(64 * GW), 0, s_, ev0, ev1, 0, (const float*)infer_,
                          (const float*)norm_, F, h_out);
  else
    hipExtLaunchKernelGGL(k_mlp_guess<false>, dim3(nb), dim3(64 * GW), 0, s_, ev0, ev1, 0, (const float*)infer_,
                          (const float*)norm_, F, h_out);
  FREHG_LAUNCH_CHECK();
}

void Mlp::get_params(float* params, float* norm) {
  if (params) FREHG_HIP(hipMemcpyAsync(params, params_, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s_));
  if (norm) FREHG_HIP(hipMemcpyAsync(norm, norm_, kMlpNorm * sizeof(float), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

void Mlp::set_params(const float* params, const float* norm) {
  if (!params || !norm) throw ArgError("frehg_mlp_set_params: null argument");
  for (int t = 0; t < kMlpRow; ++t)
    if (!(norm[kMlpRow + t] > 0.0f)) throw ArgError("frehg_mlp_set_params: standard deviations must be > 0");
  FREHG_HIP(hipMemcpyAsync(params_, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(norm_, norm, kMlpNorm * sizeof(float), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipMemsetAsync(m_, 0, kMlpParams * sizeof(float), s_));
  FREHG_HIP(hipMemsetAsync(v_, 0, kMlpParams * sizeof(float), s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  adam_t_ = 0;
  have_norm_ = true;
  ready_ = true;
  infer_stale_ = true;
}

void Mlp::ensure_scratch(int nrows) {
  if (nrows <= scratch_rows_) return;
  if (scratch_) FREHG_HIP(hipFree(scratch_));
  scratch_ = nullptr;
  FREHG_HIP(hipMalloc(&scratch_, (size_t)nrows * kMlpRow * sizeof(float)));
  scratch_rows_ = nrows;
}

double Mlp::loss_grad(const float* X, const float* Y, int n, float* grad) {
  if (n < 1 || !X || !Y) throw ArgError("frehg_mlp_loss_grad: need n >= 1 rows");
  ensure_scratch(n);
  std::vector<float> rows((size_t)n * kMlpRow);
  for (int r = 0; r < n; ++r) {
    for (int t = 0; t < kMlpIn; ++t) rows[(size_t)r * kMlpRow + t] = X[(size_t)r * kMlpIn + t];
    rows[(size_t)r * kMlpRow + kMlpIn] = Y[r];
  }
  FREHG_HIP(hipMemcpyAsync(scratch_, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, s_));
  const int nb = (int)run_grad(scratch_, n, n, 0, 1);
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(ceil_div(kMlpParams + 1, MT)), dim3(MT), 0, s_, (const float*)part_, nb,
                     grad_, params_, m_, v_, 0, 0.0f, 0.0f, 0.0f);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(hbuf_, grad_, (kMlpParams + 1) * sizeof(float), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  if (grad) std::memcpy(grad, hbuf_, kMlpParams * sizeof(float));
  return hbuf_[kMlpParams] / (double)n;
}

void Mlp::predict_batch(const float* X, int n, float* out) {
  if (n < 1 || !X || !out) throw ArgError("frehg_mlp_predict: need n >= 1 rows");
  ensure_scratch(n);
  std::vector<float> rows((size_t)n * kMlpRow, 0.0f);
  for (int r = 0; r < n; ++r)
    for (int t = 0; t < kMlpIn; ++t) rows[(size_t)r * kMlpRow + t] = X[(size_t)r * kMlpIn + t];
  FREHG_HIP(hipMemcpyAsync(scratch_, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, s_));
  DeviceBuffer<float> tmp((size_t)n);
  const unsigned nb = (unsigned)std::min<long long>(1024, ((long long)n + RB - 1) / RB);
  hipLaunchKernelGGL(k_mlp_predict_rows, dim3(nb), dim3(MT), 0, s_, (const float*)params_, (const float*)norm_,
                     (const float*)scratch_, (long long)n, tmp.get());
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(out, tmp.get(), (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

}  // namespace frehg
