// CPU reference of the learned-guess MLP -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// The MLP has no counterpart in the reference; this file is the checker for the GPU kernels
// of frehg_amd/csrc/mlp.hip. It evaluates the same network (11 -> 64 -> 64 -> 1, tanh hidden
// layers, standardised inputs and target, loss = mean squared error of the standardised
// target) and its gradient by explicit back-propagation, in one of two modes:
//   bf16 = 0: everything in FP64 (pinned by central finite differences in the tests);
//   bf16 = 1: the GPU's arithmetic -- standardisation, biases, tanh and elementwise products in
//             FP32, every matrix-product operand rounded to BF16 (round to nearest even), the
//             products accumulated exactly in FP64 (the GPU accumulates in FP32).
// Predictions have a third mode, 2, the arithmetic of the GPU's inference kernel: with
// s = 2 log2(e) and r = 1 / (1 + 2^(s z)) = (1 - tanh z) / 2, the hidden layers output r and the
// next layer takes weights -2 W and biases b + sum_k W[.][k]; the scaled weights and biases are
// formed in FP32, weights rounded to FP16 (nearest even), standardised inputs and every r to FP16
// (toward zero), products accumulated in FP64 (the GPU accumulates in FP32).
// Parameter layout: W1[64][11], b1[64], W2[64][64], b2[64], w3[64], b3.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int NI = 11, NH = 64, ROW = 12;
constexpr int W1 = 0, B1 = W1 + NH * NI, W2 = B1 + NH, B2 = W2 + NH * NH, W3 = B2 + NH, B3 = W3 + NH;
constexpr int NP = B3 + 1;

// float -> bf16 -> float, round to nearest even (NaN/Inf not expected here).
double bf(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

// double -> FP16 value (11-bit significand, normal exponents down to -14, subnormals below),
// round to nearest even or toward zero; no overflow expected here.
double f16(double v, bool rtz) {
  if (v == 0.0) return 0.0;
  int e;
  (void)std::frexp(std::fabs(v), &e);  // |v| = m 2^e, m in [0.5, 1)
  const double ulp = std::ldexp(1.0, std::max(e - 1, -14) - 10);
  const double q = std::fabs(v) / ulp;
  const double r = rtz ? std::floor(q) : std::nearbyint(q);
  return std::copysign(r * ulp, v);
}

struct Net {
  const float* P;
  const float* N;
  bool b16;
  double op(double v) const { return b16 ? bf(v) : v; }          // matrix-product operand
  double fl(double v) const { return b16 ? (double)(float)v : v; } // FP32 elementwise value
};

struct Fwd {
  double x[32], a1[NH], a2[NH], y, t;
};

void forward(const Net& n, const float* row, Fwd& f, bool with_target) {
  for (int q = 0; q < 32; ++q) f.x[q] = 0.0;
  for (int q = 0; q < NI; ++q)
    f.x[q] = n.b16 ? (double)(((float)row[q] - n.N[q]) / n.N[ROW + q]) : ((double)row[q] - n.N[q]) / n.N[ROW + q];
  f.t = 0.0;
  if (with_target)
    f.t = n.b16 ? (double)(((float)row[NI] - n.N[NI]) / n.N[ROW + NI]) : ((double)row[NI] - n.N[NI]) / n.N[ROW + NI];
  for (int h = 0; h < NH; ++h) {
    double z = 0.0;
    for (int q = 0; q < NI; ++q) z += n.op(f.x[q]) * n.op(n.P[W1 + h * NI + q]);
    z = n.fl(n.fl(z) + n.P[B1 + h]);
    f.a1[h] = n.fl(std::tanh(z));
  }
  for (int h = 0; h < NH; ++h) {
    double z = 0.0;
    for (int k = 0; k < NH; ++k) z += n.op(f.a1[k]) * n.op(n.P[W2 + h * NH + k]);
    z = n.fl(n.fl(z) + n.P[B2 + h]);
    f.a2[h] = n.fl(std::tanh(z));
  }
  double y = 0.0;
  for (int k = 0; k < NH; ++k) y += n.op(f.a2[k]) * n.op(n.P[W3 + k]);
  f.y = n.fl(n.fl(y) + n.P[B3]);
}

// Prediction of the inference kernel's arithmetic (mode 2 above), standardised output.
double forward_infer(const float* P, const float* N, const float* row) {
  const float s = 2.0f * 1.44269504088896341f;
  double x[NI], r1[NH], r2[NH];
  for (int q = 0; q < NI; ++q) x[q] = f16((double)(((float)row[q] - N[q]) * (1.0f / N[ROW + q])), true);
  for (int h = 0; h < NH; ++h) {
    double z = (double)(s * P[B1 + h]);
    for (int q = 0; q < NI; ++q) z += f16((double)(s * P[W1 + h * NI + q]), false) * x[q];
    z = (double)(float)z;
    r1[h] = f16((double)(float)(1.0 / (1.0 + std::exp2(z))), true);
  }
  for (int h = 0; h < NH; ++h) {
    float sum = 0.0f;
    for (int k = 0; k < NH; ++k) sum += P[W2 + h * NH + k];
    double z = (double)(s * (P[B2 + h] + sum));
    for (int k = 0; k < NH; ++k) z += f16((double)((-2.0f * s) * P[W2 + h * NH + k]), false) * r1[k];
    z = (double)(float)z;
    r2[h] = f16((double)(float)(1.0 / (1.0 + std::exp2(z))), true);
  }
  float sum = 0.0f;
  for (int k = 0; k < NH; ++k) sum += P[W3 + k];
  double y = (double)(P[B3] + sum);
  for (int k = 0; k < NH; ++k) y += f16((double)(-2.0f * P[W3 + k]), false) * r2[k];
  return (double)(float)y;
}

}  // namespace

extern "C" {

// Raw predictions (increments) for x[n][11]; bf16 is the mode (0, 1 or 2, see above).
int orc_mlp_predict(const float* params, const float* norm, const float* x, int n, int bf16, double* out) {
  const Net net{params, norm, bf16 != 0};
  std::vector<float> row(ROW, 0.0f);
  for (int r = 0; r < n; ++r) {
    for (int q = 0; q < NI; ++q) row[q] = x[(size_t)r * NI + q];
    if (bf16 == 2) {
      out[r] = (double)((float)forward_infer(params, norm, row.data()) * norm[ROW + NI] + norm[NI]);
      continue;
    }
    Fwd f;
    forward(net, row.data(), f, false);
    out[r] = net.b16 ? (double)((float)f.y * norm[ROW + NI] + norm[NI]) : f.y * norm[ROW + NI] + norm[NI];
  }
  return 0;
}

// Loss (mean squared error of the standardised target) and its gradient for x[n][11], y[n].
int orc_mlp_loss_grad(const float* params, const float* norm, const float* x, const float* y, int n, int bf16,
                      double* grad, double* loss) {
  const Net net{params, norm, bf16 != 0};
  std::vector<double> g(NP, 0.0);
  std::vector<float> row(ROW);
  double sq = 0.0;
  for (int r = 0; r < n; ++r) {
    for (int q = 0; q < NI; ++q) row[q] = x[(size_t)r * NI + q];
    row[NI] = y[r];
    Fwd f;
    forward(net, row.data(), f, true);
    const double e = net.fl(f.y - f.t);
    sq += e * e;
    const double dy = net.fl(net.fl(2.0 * e) * net.fl(1.0 / (double)n));
    double dz2[NH], dz1[NH];
    for (int h = 0; h < NH; ++h) dz2[h] = net.fl(net.fl(dy * params[W3 + h]) * net.fl(1.0 - net.fl(f.a2[h] * f.a2[h])));
    for (int k = 0; k < NH; ++k) {
      double s = 0.0;
      for (int h = 0; h < NH; ++h) s += net.op(dz2[h]) * net.op(params[W2 + h * NH + k]);
      dz1[k] = net.fl(net.fl(s) * net.fl(1.0 - net.fl(f.a1[k] * f.a1[k])));
    }
    for (int k = 0; k < NH; ++k) g[W3 + k] += net.op(dy) * net.op(f.a2[k]);
    g[B3] += dy;
    for (int h = 0; h < NH; ++h) {
      for (int k = 0; k < NH; ++k) g[W2 + h * NH + k] += net.op(dz2[h]) * net.op(f.a1[k]);
      g[B2 + h] += dz2[h];
      for (int q = 0; q < NI; ++q) g[W1 + h * NI + q] += net.op(dz1[h]) * net.op(f.x[q]);
      g[B1 + h] += dz1[h];
    }
  }
  for (int p = 0; p < NP; ++p) grad[p] = g[p];
  if (loss) *loss = sq / (double)n;
  return 0;
}

}  // extern "C"
