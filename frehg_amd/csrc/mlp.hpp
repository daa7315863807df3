// Learned Krylov initial guess for the Richards step: an 11 -> 64 -> 64 -> 1 MLP with tanh hidden
// layers that predicts each cell's head increment h^{n+1} - h^n from local stencil features, so
// that the solve starts from x0 = h^n + prediction instead of x0 = 0. Trained on the GPU by this
// file's own kernels (BF16 MFMA forward and backward with FP32 accumulation, Adam on FP32 master
// weights) from (features, converged increment) pairs logged while the plain solver runs. The
// solve itself is unchanged: it converges to the same tolerance and falls back to x0 = 0 when the
// guess's residual is not below the zero guess's.
//
// Features of cell c (all from the state before the solve): h^n(c); h^n of the six face
// neighbours minus h^n(c); ln K(h^n) (Ks_z * kr); C(h^n); theta^n; surface water depth of the
// column (0 without a surface model). Features and target are standardised with the mean and
// standard deviation of the logged data.
#pragma once

#include <cstdint>

#include "common.hpp"
#include "profile.hpp"

namespace frehg {

class ReState;

constexpr int kMlpIn = 11;
constexpr int kMlpH = 64;
// Parameter vector (FP32): W1[64][11], b1[64], W2[64][64], b2[64], w3[64], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH;
constexpr int kMlpParams = kOffB3 + 1;  // 4993
constexpr int kMlpRow = kMlpIn + 1;     // logged row: 11 features + target
constexpr int kMlpNorm = 2 * kMlpRow;   // mean[12], sd[12]

class Mlp {
 public:
  Mlp(hipStream_t s, Profiler* prof);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;

  // Glorot-uniform weights, zero biases, identity normalisation (deterministic in `seed`).
  void init(uint64_t seed);
  bool ready() const { return ready_; }

  // Pair logging: every `stride`-th interior cell (the phase rotates each logged step) into a
  // ring buffer of `capacity` rows. stride 0 switches logging off; the buffer is kept.
  void set_logging(int stride, long long capacity);
  bool logging() const { return stride_ > 0; }
  long long logged() const { return count_ < cap_ ? count_ : cap_; }
  void clear_log() { count_ = 0; }
  // The logged target is h - h^n - inc0 (inc0: the base increment the guess adds, or null).
  void log_pair(ReState& re, const double* hn, const double* h, const double* depth, const double* inc0);

  // `steps` Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of `batch` rows
  // drawn with replacement from the log. The first call computes the normalisation from the log.
  // Returns the mean standardised squared error of the last minibatch (first one in *first).
  double train(int steps, int batch, double lr, uint64_t seed, double* first);

  // h_out (interior) = h^n + (inc_a - inc_b) + predicted increment (inc_a null: no base
  // increment; inc_b may be h_out itself, each cell's value is read before it is written);
  // hn / wcn default to the solver's h^n / theta^n. A cell whose network output is not finite,
  // or (with a base increment and gate > 0) whose |base increment| <= gate, takes no predicted
  // increment; the first case sets *bad = 1 (bad may be null).
  void predict_guess(ReState& re, double* h_out, const double* depth, const double* inc_a, const double* inc_b,
                     const double* hn = nullptr, const double* wcn = nullptr, double* bad = nullptr);
  // The gated start split for a caller that takes the base pass itself: chunk_masks(nchunks) is
  // the buffer of per-chunk masks (64-cell row chunks along x, then rows, then layers; bit l:
  // the network corrects cell l of the chunk) that the caller fills while writing h^n + base
  // increment to h_out; infer_masked then runs the network on the chunks with a set bit, as the
  // second pass of predict_guess does.
  unsigned long long* chunk_masks(long long nchunks);
  void infer_masked(ReState& re, double* h_out, const double* depth, double* bad);
  // Inference gate in metres (0: the network corrects every cell).
  void set_gate(double gate) { gate_ = gate; }
  double gate() const { return gate_; }
  // Cells of the chunks the network ran on in the last predict_guess (synchronises the stream).
  long long inferred_cells();

  void get_params(float* params, float* norm);
  void set_params(const float* params, const float* norm);

  // Test hooks on a host batch of raw rows (n x 11 features, n targets): loss and its gradient
  // with respect to the parameter vector, and forward predictions (raw increments).
  double loss_grad(const float* X, const float* Y, int n, float* grad);
  void predict_batch(const float* X, int n, float* out);

 private:
  void ensure_scratch(int nrows);
  double run_grad(const float* rows, long long nrows, int batch, uint64_t seed, int sequential);

  hipStream_t s_;
  Profiler* prof_;
  bool ready_ = false;
  // Inference form of the parameters (k_mlp_prep_infer), rebuilt before the next guess whenever
  // the parameters or the normalisation change.
  float* infer_ = nullptr;
  bool infer_stale_ = true;
  unsigned guess_blocks_ = 4096;  // guess kernel grid: resident blocks of one round
  double gate_ = 0.0;
  unsigned long long* inferred_ = nullptr;  // device counters of inferred cells (two alternating slots)
  unsigned long long* masks_ = nullptr;      // per chunk: the cells the gated inference corrects
  long long masks_cap_ = 0;
  int last_par_ = 1;                         // slot of the last guess's counters
  int stride_ = 0;
  long long cap_ = 0, count_ = 0, log_steps_ = 0;
  int adam_t_ = 0;
  bool have_norm_ = false;
  float* params_ = nullptr;  // FP32 master weights
  float* m_ = nullptr;       // Adam moments
  float* v_ = nullptr;
  float* norm_ = nullptr;    // mean[12], sd[12]
  float* grad_ = nullptr;    // reduced gradient + loss
  float* part_ = nullptr;    // per-block partial gradients
  float* log_ = nullptr;     // ring buffer, cap_ x 12
  float* scratch_ = nullptr; // host-batch rows for the test hooks
  long long scratch_rows_ = 0;
  double* stat_ = nullptr;   // normalisation partials
  float* hbuf_ = nullptr;    // pinned host mirror (loss)
};

}  // namespace frehg
