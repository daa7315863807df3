// SW <-> GW exchange and water-volume accounting on the GPU.
#pragma once

#include "re.hpp"
#include "swe.hpp"

namespace frehg {

class Coupler {
 public:
  explicit Coupler(hipStream_t s);
  ~Coupler();
  // Coupling::exchange (src/coupling/Coupling.cpp:145-164); returns sum(q*dt).
  double exchange(SweState& swe, ReState& re, const frehg_coupling_params& cp, double dt, double* q_out);
  // The same exchange without waiting for it: sum(q*dt) arrives in host memory with the stream's
  // next synchronisation and is read by exchanged_volume().
  void exchange_async(SweState& swe, ReState& re, const frehg_coupling_params& cp, double dt);
  double exchanged_volume();
  // Orchestrator::computeWaterVolume (src/driver/Orchestrator.cpp:898-936).
  void volume(SweState* swe, ReState* re, double* surface, double* subsurface);
  // Keep the per-column rates of every exchange on the device (nx * ny, x fastest; m^3/s,
  // + = seepage), for the solute exchange that follows the water exchange.
  void set_capture_rates(bool on) { capture_q_ = on; }
  const double* rates() const { return qdev_; }

 private:
  void ensure(unsigned int nblocks, long long ncol);
  hipStream_t s_;
  double* part_ = nullptr;
  unsigned int nblk_ = 0;
  double* qdev_ = nullptr;
  long long nq_ = 0;
  double* dres_ = nullptr;
  double* hres_ = nullptr;
  unsigned int* counter_ = nullptr;
  bool want_q_ = false;  // exchange_async also stores the per-column rates
  bool capture_q_ = false;
};

}  // namespace frehg
