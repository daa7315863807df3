// Device-resident passive solute (replaces src/solute/* and the solute parts of
// src/driver/Orchestrator.cpp): one concentration per surface cell and per soil cell, moved by
// the flow solvers' own velocities and Darcy fluxes after each flow step (operator splitting).
//
// Per flow step (Orchestrator::applySolute, :938-1056): rain mixing on cells wet at the start of
// the step (not on the last y row, which the surface step gives no rain), then substeps of
// first-order decay, explicit conservative advection (upwind or minmod MUSCL, refused above the
// Courant ceiling: the substep count doubles until it passes), and implicit diffusion
// (I - dt D L) C' = C through the Jacobi-CG solver. In a coupled step the exchange between the
// surface water and the top soil cell carries solute with the exchanged water
// (applyCouplingSoluteExchange, :1270-1299; applyInterfaceExchange, src/solute/SourceSink.cpp).
#pragma once

#include <memory>
#include <string>

#include "../../include/frehg_gpu.h"
#include "common.hpp"
#include "krylov.hpp"

namespace frehg {

class SweState;
class ReState;

class Solute {
 public:
  Solute(const frehg_solute_params& p, SweState* swe, ReState* re, hipStream_t s, Profiler* prof);
  ~Solute();
  Solute(const Solute&) = delete;
  Solute& operator=(const Solute&) = delete;

  // "surface" ((nx+2)(ny+2)) / "subsurface" ((nx+2)(ny+2)(nz+2)), the ABI's packed layout.
  void set_field(const std::string& name, const double* src);
  void get_field(const std::string& name, double* dst);
  // Interior values (x fastest); either may be null (that domain keeps its concentration).
  void set_initial(const double* surface, const double* subsurface);

  // After the water exchange of a coupled step: q = the per-column exchange rates on the device.
  void interface_exchange(const double* q, double dt, double area);
  // applySolute for the flow step [t, t + dt] with surface rain rate `rain`.
  void apply(double t, double dt, double rain);
  // computeSoluteMass: sum C max(eta - bottom, 0) dx dy + sum C theta dx dy dz.
  double mass();

  const frehg_solute_stats& totals() const { return tot_; }
  void reset_totals() { tot_ = frehg_solute_stats{}; }
  const frehg_solute_params& params() const { return p_; }

 private:
  double max_courant(double dt);
  void decay(double fac);
  void advect();
  void diffuse(double dt);
  void substep(double dt);

  frehg_solute_params p_;
  SweState* swe_;
  ReState* re_;
  hipStream_t s_;
  Profiler* prof_;
  Geo gs_{}, gg_{};
  double dx_ = 1, dy_ = 1, dz_ = 1;
  long long ns_ = 0, ng_ = 0;
  double *cs_ = nullptr, *cg_ = nullptr;      // concentrations (padded)
  double *cs_old_ = nullptr, *cg_old_ = nullptr;  // advection source copies, then the diffusion rhs
  double *Ds_ = nullptr, *axs_ = nullptr, *ays_ = nullptr;  // surface diffusion operator
  double *Dg_ = nullptr, *axg_ = nullptr, *ayg_ = nullptr, *azg_ = nullptr;  // soil diffusion operator
  double* part_ = nullptr;  // block partials (max Courant, sums)
  double* dres_ = nullptr;
  double* hres_ = nullptr;
  unsigned int nb_ = 0;
  std::unique_ptr<Krylov> kry_s_, kry_g_;
  double cur_dt_ = 0.0;
  frehg_solute_stats tot_{};
};

}  // namespace frehg
