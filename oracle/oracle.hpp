// FP64 CPU oracle for the Frehg2 coupled step -- TEST INFRASTRUCTURE ONLY.
//
// This library is the checker that the HIP backend is compared against, and the CPU baseline
// that bench.py times beside it. Nothing in the product path links or calls it. It restates
// the reference algorithm (zLi90/frehg, "Frehg2") function by function; each function names
// the reference file:line it follows. Arithmetic follows the reference expressions term for
// term (same operand order) so that the oracle reproduces the reference's rounding wherever
// the reference is deterministic.
//
// Deliberate departures from the reference (each one documented at its definition):
//   * The Krylov solve is this file's own Jacobi-preconditioned CG on the assembled stencil
//     (PETSc's KSP is third-party and absent). Convergence test = PETSc's default for CG:
//     preconditioned residual 2-norm <= max(rtol * ||B b||, atol), x0 = 0.
//   * The relative-viscosity/density arrays r_* of GwFields are identically 1.0 in the
//     non-baroclinic path; multiplying by 1.0 is exact in IEEE-754, so they are omitted.
//   * actv is 1 on every interior cell (ReSolver::buildColumnGeometry, ReSolver.cpp:121-130);
//     the inactive-cell branches are therefore never taken and are omitted.
//   * A per-cell saturated conductivity array (ks_cell) may replace Ks_{x,y,z} of the cell's
//     class. This is the reference's "one soil class per cell" SoilMap
//     (src/soil/SoilMap.hpp:33-46) without materialising N3 class records.
//   * "free drainage" (bc_type[4] == 3) is a new bottom BC: unit gradient flux K(h) out of
//     the bottom face. The reference code already produces exactly this for any bottom BC
//     value other than 0/1/2 (ReSolver.cpp:281,516,47-57), but its capability gate rejects such
//     values (Orchestrator.cpp:164-167); parity for it is therefore pinned by the oracle only.
#ifndef FREHG_ORACLE_HPP
#define FREHG_ORACLE_HPP

#include <cmath>
#include <cstdint>
#include <string>
#include <array>
#include <vector>

namespace frehg_oracle {

using real = double;

// include/frehg2/re/SoilParams.hpp:13-25 (defaults kept).
struct Soil {
  real alpha = 1.43;
  real n = 1.56;
  real theta_s = 0.33;
  real theta_r = 0.0;
  real Ks_x = 0.0;
  real Ks_y = 0.0;
  real Ks_z = 2.89e-6;
  real Ss = 1.0e-5;
  int use_vg = 1;
  int use_mvg = 0;
  real air_entry = -0.02;
};

// Van Genuchten / modified VG relations, src/re/VanGenuchten.hpp:18-137.
struct VG {
  static real m(real n) { return 1.0 - 1.0 / n; }  // :19

  static real wcm(const Soil& s) {  // :21-28
    const real mm = m(s.n);
    if (s.use_mvg)
      return s.theta_r +
             (s.theta_s - s.theta_r) * std::pow(1.0 + std::pow(std::fabs(s.air_entry) * s.alpha, s.n), mm);
    return s.theta_s;
  }

  // theta(h), VanGenuchten.hpp:31-47.
  static real theta(const Soil& s, real h) {
    real wc;
    if (s.use_vg) {
      const real mm = m(s.n);
      const real wcm_val = wcm(s);
      const real se = std::pow(1.0 + std::pow(std::fabs(s.alpha * h), s.n), -mm);
      if (h > s.air_entry)
        wc = s.theta_s;
      else
        wc = s.theta_r + (wcm_val - s.theta_r) * se;
    } else {
      wc = s.theta_r + (s.theta_s - s.theta_r) * std::exp(0.1634 * h);
    }
    if (wc > s.theta_s) wc = s.theta_s;
    if (wc < s.theta_r) wc = s.theta_r;
    return wc;
  }

  // h(theta), VanGenuchten.hpp:50-65.
  static real head(const Soil& s, real wc) {
    const real eps = 1.0e-7;
    if (s.use_vg) {
      const real mm = m(s.n);
      const real wcm_val = wcm(s);
      real th = wc;
      if (th - s.theta_r < eps) th = s.theta_r + eps;
      if (th < s.theta_s)
        return -(1.0 / s.alpha) *
               std::pow(std::pow((wcm_val - s.theta_r) / (th - s.theta_r), 1.0 / mm) - 1.0, 1.0 / s.n);
      return 0.0;
    }
    return std::log((wc - s.theta_r) / (s.theta_s - s.theta_r)) / 0.1634;
  }

  // C(h) = d theta / d h, VanGenuchten.hpp:68-81.
  static real capacity(const Soil& s, real h) {
    const real mm = m(s.n);
    const real wcm_val = wcm(s);
    const real nume =
        s.alpha * s.n * mm * (wcm_val - s.theta_r) * std::pow(std::fabs(s.alpha * h), s.n - 1.0);
    const real deno = std::pow(1.0 + std::pow(std::fabs(s.alpha * h), s.n), mm + 1.0);
    real c = nume / deno;
    if (s.use_mvg) {
      if (h > s.air_entry) c = 0.0;
    } else if (h > 0.0) {
      c = 0.0;
    }
    return c;
  }

  // K(h), VanGenuchten.hpp:84-111.
  static real conductivity(const Soil& s, real h, real Ks) {
    if (!s.use_vg) return Ks * std::exp(0.1634 * h);
    const real mm = m(s.n);
    const real se = std::pow(1.0 + std::pow(std::fabs(s.alpha * h), s.n), -mm);
    const real wcm_val = wcm(s);
    real Keff;
    if (s.use_mvg) {
      const real nume =
          1.0 - std::pow(1.0 - std::pow(se * (s.theta_s - s.theta_r) / (wcm_val - s.theta_r), 1.0 / mm), mm);
      const real deno =
          1.0 - std::pow(1.0 - std::pow((s.theta_s - s.theta_r) / (wcm_val - s.theta_r), 1.0 / mm), mm);
      Keff = (deno == 0.0) ? Ks : Ks * std::pow(se, 0.5) * std::pow(nume / deno, 2.0);
    } else {
      Keff = Ks * std::pow(se, 0.5) * std::pow(1.0 - std::pow(1.0 - std::pow(se, 1.0 / mm), mm), 2.0);
    }
    if (Keff > Ks) Keff = Ks;
    if (s.use_mvg) {
      if (h > s.air_entry) Keff = Ks;
    } else if (h > 0.0) {
      Keff = Ks;
    }
    return Keff;
  }

  // dK/dtheta, VanGenuchten.hpp:114-136.
  static real dKdwc(const Soil& s, real wc, real Ks) {
    real th = wc;
    if (th > 0.9999 * s.theta_s && th < s.theta_s) th = 0.9999 * s.theta_s;
    const real mm = m(s.n);
    const real se = (th - s.theta_r) / (s.theta_s - s.theta_r);
    if (s.use_mvg) {
      const real wcm_val = wcm(s);
      const real c2 = (s.theta_s - s.theta_r) / (wcm_val - s.theta_r);
      const real c1 = 1.0 / std::pow(1.0 - std::pow(1.0 - std::pow(c2, 1.0 / mm), mm), 2.0);
      const real term0 = std::pow(1.0 - std::pow(c2 * se, 1.0 / mm), mm);
      const real term1 = 0.5 * c1 * Ks * std::pow(se, -0.5) * (1.0 - term0) * (1.0 - term0);
      const real term2 = 2.0 * c1 * Ks * c2 * std::pow(se, 0.5) * std::pow(c2 * se, 1.0 / mm - 1.0) *
                         (1.0 - term0) * std::pow(1.0 - std::pow(c2 * se, 1.0 / mm), mm - 1.0);
      return (term1 + term2) / (s.theta_s - s.theta_r);
    }
    const real term0 = std::pow(1.0 - std::pow(se, 1.0 / mm), mm);
    const real term1 = 0.5 * Ks * std::pow(se, -0.5) * (1.0 - term0) * (1.0 - term0);
    const real term2 = 2.0 * Ks * std::pow(se, (2.0 - mm) / (2.0 * mm)) * (1.0 - term0) *
                       std::pow(1.0 - std::pow(se, 1.0 / mm), mm - 1.0);
    return (term1 + term2) / (s.theta_s - s.theta_r);
  }
};

// Halo-padded index convention, src/core/Grid.hpp:42-60: one ghost cell on every side,
// i fastest, then j, then k; k = 0 is the TOP layer; k = -1 and k = nz are ghosts.
// Loops over fewer cells than this run on one thread: thread start-up and barriers would cost more
// than the work (and stall badly on a busy host). Results do not depend on it: the parallel loops
// write one cell each, and the sums use fixed blocks whatever the thread count.
inline bool omp_wide(long long cells) { return cells >= 32768; }

struct Grid {
  int nx = 1, ny = 1, nz = 1;
  long long plane() const { return (long long)(nx + 2) * (ny + 2); }
  long long idx(int i, int j, int k) const {
    return (i + 1) + (long long)(j + 1) * (nx + 2) + (long long)(k + 1) * plane();
  }
  long long sidx(int i, int j) const { return (i + 1) + (long long)(j + 1) * (nx + 2); }
  long long ncell() const { return plane() * (nz + 2); }
  long long nsurf() const { return plane(); }
};

// Solver configuration, include/frehg2/linear/SolverConfig.hpp:14-20.
struct SolverCfg {
  real rtol = 1e-8;
  real atol = 1e-50;
  int max_it = 10000;
  int ksp = 0;  // 0 "cg", 1 "bcgs"
};

struct SolveStats {
  int iterations = 0;
  real residual_norm = 0.0;   // final preconditioned residual 2-norm
  real rhs_norm = 0.0;        // ||D^-1 b||
  int converged = 0;
};

// Deterministic sums: fixed 4096-element blocks summed in order (independent of thread count).
real det_sum(const std::vector<real>& partials);

// ---------------------------------------------------------------------------------------------
// Stencil operator on the halo-padded grid (the oracle's stand-in for PETSc MATMPIAIJ).
// Row c (interior cell) holds its diagonal and one coefficient per neighbour direction, in the
// reference's per-row form (ReSolver.cpp:576-610 maps G* to columns; SweSolver.cpp:542-569 maps
// -S* to columns). Neighbours outside the owned domain have no column (rowW/E/S/N/B/T == -1).
struct StencilOp {
  Grid g;
  bool three_d = true;
  std::vector<real> diag, xm, xp, ym, yp, zm, zp;  // size g.ncell() (3D) or g.nsurf() (2D)
  std::vector<unsigned char> fixed;                // 2D only: decoupled (identity-like) rows
  void resize();
  void apply(const std::vector<real>& x, std::vector<real>& y) const;  // interior rows only
};

// Jacobi-PCG, x0 = 0, PETSc default convergence test (see header comment).
// For 2D operators with `fixed` rows (decoupled dry rows), those unknowns are solved exactly
// and eliminated into the right-hand side first, so CG runs on the symmetric remainder.
// cfg.ksp == 1 selects Jacobi-preconditioned BiCGSTAB ("bcgs") instead of CG.
SolveStats pcg_solve(const StencilOp& A, const std::vector<real>& b, std::vector<real>& x,
                     const SolverCfg& cfg);

// ---------------------------------------------------------------------------------------------
// Richards-equation PCA step: src/re/ReSolver.{hpp,cpp}.
struct ReParams {  // ReSolver.hpp:24-46 (defaults kept)
  Soil soil;
  real dx = 1.0, dy = 1.0, dz = 0.01, botz = -1.0;
  real dt = 1.0e-4, dt_min = 1.0e-4, dt_max = 2.0, co_max = 2.0;
  int adaptive_dt = 1, use_corrector = 1, use_full3d = 0, sim_shallowwater = 0;
  int bc_type[6] = {0, 0, 0, 0, 0, 1};
  real htop = 0.0, hbot = 0.0, qtop = 0.0, qbot = 0.0;
  SolverCfg solver{1e-10, 1e-50, 10000};  // Orchestrator.cpp:487-490 (GW rtol 1e-10)
  // Picard outer loop (no counterpart in the reference: picard_max = 1 is its predictor).
  int picard_max = 1;
  real picard_tol = 1e-8;
};

// Water ledger of one step (m^3), as reported by frehg_re_stats: theta_c is the corrector's theta
// before reallocation and clamp.
struct ReStepStats {
  SolveStats solve;         // the last Picard iteration's solve (iterations: all of them)
  int picard_iterations = 1;
  bool picard_converged = true;  // the max |dh| test passed (picard_max = 1: no test, true)
  real dt_used = 0.0;
  real dt_next = 0.0;
  real max_dh = 0.0;
  real mass_before = 0.0;       // sum theta^n V
  real mass_after = 0.0;        // sum theta^{n+1} V
  real mass_boundary_in = 0.0;  // dt * sum(q_bottom_face - q_top_face)
  real mass_storage_ss = 0.0;   // -sum theta_c V Ss (h - h^n) / theta_s
  real mass_adjust = 0.0;       // sum (theta^{n+1} - theta_c) V
};

class ReSolver {
 public:
  ReSolver(int nx, int ny, int nz, const ReParams& p);
  void setParams(const ReParams& p) { p_ = p; }
  const ReParams& params() const { return p_; }
  // Soil map: class table + optional class index (2D per-column if nz_idx == 0, else 3D) +
  // optional per-cell Ks (i + j*nx + k*nx*ny).
  void setSoilMap(const std::vector<Soil>& classes, const std::vector<int>& class_idx, bool idx3d,
                  const std::vector<real>& ks_cell);
  void setTopFluxField(const std::vector<real>& q_interior);  // ReSolver.cpp:68-80
  void initializeGeometry();                                  // ReSolver.cpp:104-133
  void setInitialWaterContent(const std::vector<real>& wc);   // :135-153
  void setInitialHead(const std::vector<real>& h);            // :155-173
  void finalizeInitialState();                                // :175-178
  void advanceStep(ReStepStats* st = nullptr);                // :931-969
  // The step's tail on the fields as they stand (reallocateWaterContent, clampWaterContent,
  // enforceMoistureBc: ReSolver.cpp:788-878), for hand-built reallocation cases.
  void reallocateAndClamp() {
    reallocateWaterContent();
    clampWaterContent();
    enforceMoistureBc();
  }

  Grid g;
  // GwFields (src/re/GwFields.hpp:14-50), halo padded.
  std::vector<real> h, hn, wc, wcn, wch, hwc, ch, Kx, Ky, Kz, qx, qy, qz;
  std::vector<real> Gxp, Gxm, Gyp, Gym, Gzp, Gzm, Gct, Grhs;
  std::vector<real> qtop_field;
  bool has_qtop_field = false;
  real dtn = 0.0, dt_used = 0.0;
  // Which sweep order reallocateWaterContent uses: 0 = the reference's sequential i->j->k sweep.
  int realloc_mode = 0;

  real soilKs(int i, int j, int k, int dir) const;
  Soil soilAt(int i, int j, int k) const;  // ReSolver.cpp:87-96 (ghost indices clamped)

 private:
  ReParams p_;
  std::vector<Soil> classes_;
  std::vector<int> class_idx_;
  bool idx3d_ = false;
  std::vector<real> ks_cell_;

  real dzCell() const { return p_.dz; }
  real V() const { return p_.dx * p_.dy * p_.dz; }
  void updateDiagnostics();
  void computeKFace();
  void matCoeff();
  void rhs();
  void assembleAndSolve(SolveStats* st);
  void enforceHeadBc();
  void groundwaterFlux();
  void updateWaterContent();
  void reallocateWaterContent();
  void clampWaterContent();
  void enforceMoistureBc();
  void adaptiveTimeStep();
};

// ---------------------------------------------------------------------------------------------
// Semi-implicit shallow-water step: src/swe/SweSolver.{hpp,cpp} (serial path).
struct SweParams {  // SweSolver.hpp:28-39
  real gravity = 9.81, manning = 0.0, min_depth = 1.0e-8, visc_x = 0.0, visc_y = 0.0;
  real hD = 0.1, hE = 0.0, wtfh = 1.0e-8, dt = 1.0, offset = 0.0;
  SolverCfg solver{1e-12, 1e-50, 10000};  // Orchestrator.cpp:434-437 (SW rtol 1e-12)
};

class SweSolver {
 public:
  SweSolver(int nx, int ny, real dx, real dy, const SweParams& p);
  void setBathymetry(const std::vector<real>& physical);       // SweSolver.cpp:144-157
  void initializeState(real init_eta);                          // :926-934
  void setInitialEtaPhysical(const std::vector<real>& eta);     // :855-870
  void setInitialVelocityConstant(real u, real v);              // :872-884
  void finalizeInitialState();                                  // :904-924
  void advanceStep(real rain, real evap, SolveStats* st = nullptr);  // :803-853
  real maxCfl() const;                                          // :954-975
  void updateDepth();     // :207-304 (public for the polygon boundary refresh)
  void updateGeometry();  // :306-369
  const SweParams& params() const { return p_; }

  Grid g;  // nz unused
  real dx, dy;
  std::vector<real> eta, etan, bottom, bottomXP, bottomYP, dept, deptx, depty, uu, vv, un, vn, uy,
      vx, Vs, Vsn, Asx, Asy, Asz, Aszx, Aszy, Vsx, Vsy, CDx, CDy, Dx, Dy, Ex, Ey, cflx, cfly, Fu,
      Fv, rain, evap, Sxp, Sxm, Syp, Sym, Sct, Srhs, cfl_active, wtfx, wtfy;

 private:
  SweParams p_;
  long long s(int i, int j) const { return g.sidx(i, j); }
  void ghostsZeroGradient(std::vector<real>& f);
  void updateBottomFaces();
  void enforceSurfBc();
  void momentumSource();
  void shallowwaterRhs();
  void shallowwaterMatCoeff();
  void assembleAndSolve(SolveStats* st);
  void cflLimiter();
  void evaprain(real rain_rate, real evap_rate);
  void updateDragCoef();
  void waterfallLocation();
  void updateVelocity();
  void enforceVeloBc();
  void interpVelocity();
};

// ---------------------------------------------------------------------------------------------
// Polygon boundary regions of the surface: src/bc/Polygon.hpp (even-odd membership),
// src/bc/PolygonIndex.cpp:9-33 (cell-centre index, first match wins), src/bc/PolygonBC.cpp:16-78
// (applySurface). kind: 0 discharge (rate = Q m^3/s, > 0 out), 1 depth (rate = h m),
// 2 critical (rate unused).
struct BcPolygonRegion {
  int kind = 0;
  real rate = 0.0;
  std::vector<std::array<double, 2>> vertices;
};
struct PolygonBc {
  std::vector<BcPolygonRegion> regions;
  std::vector<int> index;  // per surface storage cell: region or -1
  std::vector<int> count;  // interior cells per region
  void build(const SweSolver& swe, double x0, double y0);
  // Net outflow volume [m^3] of one application over dt; refreshes depth and geometry.
  real apply(SweSolver& swe, real dt) const;
};

// ---------------------------------------------------------------------------------------------
// SW <-> GW exchange: src/coupling/Coupling.{hpp,cpp}.
struct CouplingParams {  // Coupling.hpp:26-31
  real dx = 1.0, dy = 1.0, visc = 1.0, min_depth = 1.0e-8;
};
// Returns the exchanged volume sum(q*dt) (Coupling.cpp:145-164); q_out (nx*ny) may be null.
real couple_exchange(const CouplingParams& cp, SweSolver& swe, ReSolver& re, real dt,
                     std::vector<real>* q_out);

// Orchestrator::step for the coupled case (Orchestrator.cpp:1322-1336): SW advance, exchange,
// GW catch-up while gw_time < t_new (Orchestrator.cpp:1243-1258).
struct CoupledState {
  real t = 0.0, gw_time = 0.0, exchange_volume = 0.0;
  int gw_substeps = 0;
  const PolygonBc* bc = nullptr;  // surface boundary regions, applied after the SW step
  real boundary_outflow = 0.0;    // accumulated net outflow volume
};
void coupled_step(const CouplingParams& cp, SweSolver& swe, ReSolver& re, CoupledState& cs,
                  real dt, real rain, real evap, real t_end_hint);

// ---------------------------------------------------------------------------------------------
// Passive solute: src/solute/{SoluteParams,Advection,Diffusion,SourceSink,SoluteStepper}.cpp and
// Orchestrator::applySolute / applyCouplingSoluteExchange / computeSoluteMass
// (src/driver/Orchestrator.cpp:852-896, 938-1056, 1270-1299).
struct SoluteParams {  // SoluteParams.hpp:16-30, plus substeps / start_time (Orchestrator.cpp:759-760)
  real c_rain = 0.0, k_decay = 0.0, D = 1.0e-9, cfl_max = 1.0, min_depth = 1.0e-8, start_time = 0.0;
  int muscl = 0, implicit_diffusion = 1, substeps = 1;
  SolverCfg solver{1e-12, 1e-50, 10000};  // Orchestrator.cpp:767-772
};
struct SoluteTotals {
  real rain_added = 0.0, decay_loss = 0.0, max_courant = 0.0;
  long long substeps = 0;
  int refused_steps = 0, diffusion_iterations = 0;
};
class SoluteOracle {
 public:
  SoluteOracle(const SoluteParams& p, SweSolver* swe, ReSolver* re);
  std::vector<real> cs, cg;  // surface / subsurface concentration, halo padded
  SoluteTotals tot;
  // applyCouplingSoluteExchange with the per-column exchange rates q (nx*ny, m^3/s, + = seepage).
  void interfaceExchange(const std::vector<real>& q, real dt, real area);
  void apply(real t, real dt, real rain);  // applySolute for the flow step [t, t + dt]
  real mass() const;                       // computeSoluteMass
  // One SoluteStepper::step with rain 0 (decay, advection, diffusion); false if refused.
  bool step(real dt, real* cfl);
  const SoluteParams& params() const { return p_; }

 private:
  SoluteParams p_;
  SweSolver* swe_;
  ReSolver* re_;
  real dx_ = 1.0, dy_ = 1.0, dz_ = 1.0;
};

// coupled_step with solute: the exchange carries solute, then applySolute after the GW catch-up.
void coupled_step_solute(const CouplingParams& cp, SweSolver& swe, ReSolver& re, CoupledState& cs,
                         SoluteOracle* sol, real dt, real rain, real evap, real t_end_hint);

}  // namespace frehg_oracle

#endif  // FREHG_ORACLE_HPP
