// Device-resident Richards-equation PCA step (replaces src/re/ReSolver.{hpp,cpp}).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/frehg_gpu.h"
#include "common.hpp"
#include "krylov.hpp"

namespace frehg {

class Mlp;  // learned initial guess (mlp.hpp)

class ReState {
 public:
  ReState(int nx, int ny, int nz, const frehg_re_params& p, hipStream_t s, Profiler* prof);
  ~ReState();

  void set_params(const frehg_re_params& p);
  void set_soil(const frehg_soil_params* classes, int ncls, const int32_t* class_idx, int idx3d,
                const double* ks_cell);
  void set_top_flux(const double* q_interior);
  void set_initial(const double* h_interior, const double* wc_interior);
  void set_field(const std::string& name, const double* src);
  void get_field(const std::string& name, double* dst);
  void step(frehg_re_stats* st);
  double dt() const { return p_.dt; }
  void prime_dt(double dt) {
    p_.dt = dt;
    dtn_ = dt;
  }

  const Geo& geo() const { return g_; }
  const frehg_re_params& params() const { return p_; }
  const SoilClass& class0() const { return host_classes_[0]; }
  double dt_used() const { return dt_used_; }

  // Device views used by the coupling / learned-guess kernels.
  double* d_h() { return h_; }
  double* d_wc() { return wc_; }
  double* d_hn() { return hn_; }
  double* d_wcn() { return wcn_; }
  double* d_kr() { return kr_; }
  // Darcy fluxes of the last step (call materialize_fluxes() first).
  double* d_qx() { return qx_; }
  double* d_qy() { return qy_; }
  double* d_qz() { return qz_; }
  double* d_ch() { return ch_; }
  // Per-cell "head changed since kr and C were last evaluated" flags for a kernel that changes h
  // between steps (the SW exchange), or null when the next step re-evaluates every cell anyway.
  unsigned char* d_hmod() { return kc_fresh_ ? hmod_ : nullptr; }
  SoilView soil_view() const { return SoilView{classes_, cls_, ks_, host_classes_[0]}; }
  // Every soil class is plain van Genuchten (neither the exponential law nor modified VG).
  bool soil_plain_vg() const {
    for (const SoilClass& c : host_classes_)
      if (!c.use_vg || c.use_mvg) return false;
    return true;
  }
  Krylov& krylov() { return *kry_; }
  // Stores Kx..qz of the last step if they are pending (see flux_pending_).
  void materialize_fluxes();
  // Stores the system of the last Picard iterate if it is pending (see op_pending_).
  void materialize_operator();
  // Learned increment h^{n+1} - h^n the model predicts for a step taken from the current state
  // (features from h, K/C at h, theta), interior cells in i + j nx + k nx ny order.
  void learned_increment(double* dst);
  // The learned start the next step would take (h^n + d^{n-1} + the gated correction), interior.
  void learned_start(double* dst);
  // Learned-guess model and the surface depth it reads (either may be null).
  void set_mlp(Mlp* m, const double* surface_depth) {
    mlp_ = m;
    depth_ = surface_depth;
  }

 private:
  void build(const frehg_re_params& p);
  void release();
  double* field_ptr(const std::string& name);

  Geo g_;
  frehg_re_params p_;
  hipStream_t s_;
  Profiler* prof_;
  long long n_ = 0;
  // GwFields (src/re/GwFields.hpp) that cross the ABI, halo padded, FP64.
  double *h_ = nullptr, *hn_ = nullptr, *wc_ = nullptr, *wcn_ = nullptr, *wch_ = nullptr,
         *hwc_ = nullptr, *ch_ = nullptr, *Kx_ = nullptr, *Ky_ = nullptr, *Kz_ = nullptr,
         *qx_ = nullptr, *qy_ = nullptr, *qz_ = nullptr;
  // Lazy flux fields (default; FREHG_LAZY_FLUX=0 stores them in the step). The step's flux +
  // corrector kernel evaluates the face conductivities and fluxes but stores only one byte of
  // nonzero-face bits per cell (kf_), which is all the reallocation reads; finalize writes the
  // final head to a second buffer, so the post-solve head survives in hp_. Until the next step,
  // Kx..qz follow bit for bit from hp_, kr_ (post-solve) and wcn_, and are stored from them when
  // read (get_field) or before anything changes those inputs (flux_pending_).
  double* hp_ = nullptr;
  // After a step, hp_ holds its post-solve head and hn_ its starting head (h_ after the next
  // step's swap): inc_valid_ says so, and their difference is the increment the extrapolated and
  // learned starts add to h^n. Any change of the state clears it. dhp_ holds a copy of that
  // increment for pair logging (allocated on first use).
  bool inc_valid_ = false;
  double* dhp_ = nullptr;
  void ensure_dhp();
  unsigned char* kf_ = nullptr;
  bool lazy_flux_ = true;
  bool flux_pending_ = false;
  // A Picard loop that ended on the residual test of its last iterate h^m taken before assembly
  // (FREHG_PICARD_CHECK=0 assembles and solves instead) leaves D..rhs at the previous iterate.
  // Until the next step the system at h^m follows bit for bit from kr_ and ch_ (at h^m), wcn_,
  // hn_, the soil, the top flux and the parameters with the step's dt, and is stored from them
  // when read or before anything changes those inputs.
  bool picard_check_ = true;
  // The assembly takes the first sweep of the mixed-precision solve after it
  // (FREHG_ASSEMBLE_SWEEP=0: the solve takes it).
  bool assemble_sweep_ = true;
  // The Picard test before assembly shares its pass with the fluxes and corrector at the same
  // iterate (FREHG_CHECK_FLUX=0: the test alone, the fluxes after the loop).
  bool check_flux_ = true;
  // The extrapolated and gated learned starts take the post-solve pass at h^0 in their own
  // kernels (FREHG_FUSED_START=0: a pass of its own after the start).
  bool fused_start_ = true;
  bool op_pending_ = false;
  // Both lazy stores, before a change of the state.
  void flush_lazy() {
    materialize_fluxes();
    materialize_operator();
  }
  // Relative conductivity K/Ks per cell (incl. face-adjacent ghosts).
  double* kr_ = nullptr;
  // Assembled system: D = Gct, ax/ay/az = +x/+y/+z couplings, rhs = Grhs; FP32 copies.
  double *D_ = nullptr, *ax_ = nullptr, *ay_ = nullptr, *az_ = nullptr, *rhs_ = nullptr;
  float *D32_ = nullptr, *ax32_ = nullptr, *ay32_ = nullptr, *az32_ = nullptr;
  // Saturation-reallocation flags (uint8, padded).
  unsigned char *sat0_ = nullptr, *cand_ = nullptr, *satA_ = nullptr, *satB_ = nullptr;
  // kr and C at h after a step are the post-solve values (ReSolver.cpp:963) except at cells whose
  // head the reallocation (or the SW exchange) replaced afterwards; those are flagged in hmod_.
  // While kc_fresh_ holds, the next step re-evaluates only the flagged cells; any other change of
  // the state (fields, soil, initial condition, parameters) clears it.
  unsigned char* hmod_ = nullptr;
  bool kc_fresh_ = false;
  static constexpr int kSweepFlags = 64;
  int* dflag_ = nullptr;   // device: per-sweep "saturation state changed" flags
  int* hflag_ = nullptr;   // pinned host mirror
  double* red_ = nullptr;  // reduction partials (adaptive dt)
  double* red2_ = nullptr; // first-stage sums of the ledger reduction
  double* led_ = nullptr;  // water-ledger partials: 5 sums + max|dh| (+ pad) per block
  // device results: [0..1] Picard max |dh| or adaptive-dt max/min, [2..6] ledger sums, [7..8] max
  // |h^{n+1} - h^n| (and its unused min), [10] learned guess had a non-finite output
  double* dres_ = nullptr;
  double* hres_ = nullptr;
  unsigned int* dcount_ = nullptr;
  // Soil.
  SoilClass* classes_ = nullptr;
  std::vector<SoilClass> host_classes_;
  int32_t* cls_ = nullptr;
  double* ks_ = nullptr;
  double* qtop_ = nullptr;  // per-column top flux (surface padded) or null
  std::unique_ptr<Krylov> kry_;
  Mlp* mlp_ = nullptr;
  const double* depth_ = nullptr;
  double dtn_ = 0.0, dt_used_ = 0.0;
};

}  // namespace frehg
