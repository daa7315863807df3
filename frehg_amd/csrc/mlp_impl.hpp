// Learned initial guess implementation (placeholder until the trained network is wired in).
#pragma once

#include "mlp.hpp"

namespace frehg {

class MlpGuess : public Mlp {
 public:
  bool ready() const override { return false; }
  bool logging() const override { return false; }
  void predict_guess(ReState&, double*) override {}
  void log_pair(ReState&, const double*, const double*) override {}
};

}  // namespace frehg
