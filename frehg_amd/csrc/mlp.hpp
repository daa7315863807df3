// Learned Picard/Krylov initial guess: a 11 -> 64 -> 64 -> 1 tanh MLP over local stencil
// features, trained on (state, converged head increment) pairs logged from the plain solver.
#pragma once

#include "common.hpp"

namespace frehg {

class ReState;

class Mlp {
 public:
  virtual ~Mlp() = default;
  virtual bool ready() const = 0;
  virtual bool logging() const = 0;
  // h_out (interior) = h^n + predicted increment.
  virtual void predict_guess(ReState& re, double* h_out) = 0;
  // Record (features(h^n), h^{n+1} - h^n) pairs after a converged solve.
  virtual void log_pair(ReState& re, const double* hn, const double* h) = 0;
};

}  // namespace frehg
