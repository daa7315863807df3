// Matrix-free Jacobi-preconditioned CG and BiCGSTAB on 7-point (3D) / 5-point (2D) stencil
// operators in the halo-padded layout, FP64 or FP32-inside-FP64-refinement. Replaces PETSc KSP
// "cg" / "bcgs" as driven by src/linear/backends/PetscLinearSolver.cpp:186-220.
#pragma once

#include "common.hpp"
#include "frehg_gpu.h"
#include "profile.hpp"

namespace frehg {

// Device-resident control block of one Krylov solve, so that whole iterations run without a host
// round trip; kernels return immediately once `done` is set. CG: each of the two kernels of an
// iteration reduces the other's per-block partials itself (every block, in the same order), and
// block 0 records the scalars here. BiCGSTAB: one-block kernels reduce the partials in between.
struct KrylovCtl {
  double rz, alpha, beta, zz, ttol2, bb, pq, rr_guess;
  double omega;  // BiCGSTAB: rz holds rho = rh.r, pq holds rh.v
  double rzb[2];  // CG: r.z of iterate `it` in rzb[it & 1]
  int iters, done, converged, breakdown, max_it, guess_rejected;
  int half;  // BiCGSTAB: converged on the intermediate residual s
};

// Stencil operator: diagonal plus the coupling to the +x/+y/+z neighbour of each row. When
// axm/aym/azm are null the operator is symmetric and the -x/-y/-z couplings are read from the
// neighbour (ax(c-1) etc.); otherwise they are the row's own -x/-y/-z couplings (BiCGSTAB
// only). Off-domain entries and every ghost entry are zero. az/azm are null for 2D operators.
template <typename T>
struct StencilView {
  const T* D;
  const T* ax;
  const T* ay;
  const T* az;
  const T* axm = nullptr;
  const T* aym = nullptr;
  const T* azm = nullptr;
  bool symmetric() const { return axm == nullptr; }
};

struct KrylovConfig {
  int ksp = 0;              // FREHG_KSP_CG or FREHG_KSP_BICGSTAB
  double rtol = 1e-10, atol = 1e-50;
  int max_it = 10000;
  int precision = 0;        // 0 FP64, 1 FP32 inner + FP64 refinement
  double inner_rtol = 1e-5;
  int max_refine = 8;
};

// Work vectors and control blocks of an FP64 CG solve carved from caller-owned device memory
// (Krylov::carve_workspace); n = padded vector length.
struct KrylovWorkspace {
  double *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
  double *part = nullptr, *part_u = nullptr;
  KrylovCtl* ctl = nullptr;
  long long n = 0;
};

struct SolveResult {
  int iterations = 0;
  int refinements = 0;
  double residual_norm = 0.0, rhs_norm = 0.0;
  bool converged = false;
  bool fallback_fp64 = false;
  bool guess_rejected = false;
};

// Validated KrylovConfig from the ABI solver configuration (include/frehg2/linear/SolverConfig.hpp:14-20).
inline KrylovConfig krylov_config(const frehg_solver_config& c, const char* who) {
  if (c.ksp_type != FREHG_KSP_CG && c.ksp_type != FREHG_KSP_BICGSTAB)
    throw ArgError(std::string(who) + ": ksp_type must be FREHG_KSP_CG or FREHG_KSP_BICGSTAB");
  if (c.pc_type != FREHG_PC_JACOBI) throw ArgError(std::string(who) + ": pc_type must be FREHG_PC_JACOBI");
  if (c.precision != FREHG_PREC_FP64 && c.precision != FREHG_PREC_MIXED)
    throw ArgError(std::string(who) + ": precision must be FREHG_PREC_FP64 or FREHG_PREC_MIXED");
  if (!(c.rtol >= 0.0) || !(c.atol >= 0.0) || c.max_it < 1)
    throw ArgError(std::string(who) + ": need rtol >= 0, atol >= 0, max_it >= 1");
  KrylovConfig k;
  k.ksp = c.ksp_type;
  k.rtol = c.rtol;
  k.atol = c.atol;
  k.max_it = c.max_it;
  k.precision = c.precision == FREHG_PREC_MIXED ? 1 : 0;
  k.inner_rtol = c.inner_rtol > 0 ? c.inner_rtol : 1e-5;
  k.max_refine = c.max_refine > 0 ? c.max_refine : 8;
  return k;
}

class Krylov {
 public:
  Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof);
  // A solver that allocates nothing: FP64 CG only, on the caller's workspace (which it does not
  // free), with a fixed SpMV shape instead of the autotuner. Its work vectors must hold finite
  // values (the SpMV meets their ghost entries with zero coefficients).
  Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof, const KrylovWorkspace& ws);
  ~Krylov();
  // Device bytes of a borrowed workspace for this geometry, and its layout at `base` (256-byte
  // aligned pieces).
  static size_t workspace_bytes(const Geo& g, bool three_d);
  static KrylovWorkspace carve_workspace(void* base, const Geo& g, bool three_d);
  Krylov(const Krylov&) = delete;
  Krylov& operator=(const Krylov&) = delete;

  // Solve A x = b on interior rows. x (padded; ghosts finite) holds the initial guess unless
  // x0_zero. bn (may be null = b) is the right-hand side whose ||D^-1 bn|| normalises the
  // convergence test ||D^-1 r|| <= max(rtol ||D^-1 bn||, atol) (PETSc's default test for CG
  // with x0 = 0). fixed (2D, may be null) marks rows whose x is final and whose couplings are
  // already eliminated from b. guard_guess: fall back to x0 = 0 when the guess's residual is
  // not below ||D^-1 b||. A32 (may be null) is the FP32 copy of A for precision == 1.
  // cfg.ksp selects CG (symmetric A only) or BiCGSTAB.
  // swept_blocks > 0: the caller's kernel has already taken the first sweep of the mixed-precision
  // solve (3D, symmetric, no fixed rows, bn = b, CG as the inner solver, x0 = x) into the buffers
  // of first_sweep_target(), with swept_blocks blocks of partials.
  SolveResult solve(const StencilView<double>& A, const StencilView<float>* A32, const double* b,
                    const double* bn, double* x, const unsigned char* fixed, bool x0_zero,
                    bool guard_guess, const KrylovConfig& cfg, unsigned int swept_blocks = 0);
  // Where a caller's kernel writes the first sweep of the next mixed-precision solve: r32 =
  // (float)(b - A x), z32 = r32 / D32, xw = x, and five partials per block at stride nblocks in
  // part (||D^-1 r||^2, ||D^-1 b||^2, r32.z32, z32.z32, ||D32^-1 r32||^2), nblocks <= cap.
  struct SweepTarget {
    float *r32, *z32;
    double *xw, *part;
    unsigned int cap;
  };
  SweepTarget first_sweep_target();

  // FP64 CG split for graph capture: enqueue_cg queues the start and `budget` iterations without
  // any host synchronisation; finish_cg (same budget) waits, continues past the budget if needed
  // and returns the result as solve() would (x0 = 0 or x, no guess guard).
  void enqueue_cg(const StencilView<double>& A, const double* b, const double* bn, double* x,
                  const unsigned char* fixed, bool x0_zero, const KrylovConfig& cfg, int budget);
  SolveResult finish_cg(const StencilView<double>& A, double* x, const KrylovConfig& cfg, int budget);
  int last_iterations() const { return last_iters_; }

  // CG SpMV launch shape per precision (0 FP64, 1 FP32): cells per lane (1: one column per
  // thread on 64 x 4 tiles; 2: two adjacent columns per lane on 128 x 4 tiles, paired aligned
  // loads, x-neighbours by wave shifts) and layers per block of the z-march (kc).
  struct SpmvShape {
    int cpl = 1, kc = 1;
    dim3 grid;
    unsigned int nb = 0;
  };
  const SpmvShape& spmv_shape(int prec) const { return shape_[prec]; }
  void set_spmv_shape(int prec, int cpl, int kc);
  // Times the CG SpMV for every candidate shape on the operator at hand and keeps the fastest
  // (run automatically at the first 3D solve of each precision; cached per instance).
  template <typename T>
  void autotune(const StencilView<T>& A);
  bool tuned(int prec) const { return tuned_[prec]; }

 private:
  void setup_shapes();
  void build();
  void release();
  template <typename T>
  int run_pcg(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
              bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl,
              T* r, T* z, T* p0, T* p1, T* q, bool* rejected, double* res_norm, double* rhs_norm,
              bool* converged, int kt_spmv);
  template <typename T>
  void cg_start(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed, bool zero,
                bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl, T* r, T* z, T* p0);
  // x_fresh: x0 = 0 without x having been cleared (the first update writes x = alpha p; a solve
  // that ends before its first update clears x in cg_finish). Only without fixed rows.
  template <typename T>
  void cg_iterations(const StencilView<T>& A, T* x, T* r, T* z, T* p0, T* p1, T* q, KrylovCtl* ctl, int from,
                     int count, int kt_spmv, bool x_fresh = false);
  template <typename T>
  int cg_finish(const StencilView<T>& A, T* x, T* r, T* z, T* p0, T* p1, T* q, KrylovCtl* ctl, int max_it,
                int launched, int batch, double* res_norm, double* rhs_norm, bool* converged, int kt_spmv,
                bool x_fresh = false);
  template <typename T>
  void zero_interior(T* x, const unsigned char* fixed);
  template <typename T>
  void launch_spmv(const StencilView<T>& A, const T* z, const T* pold, T* pnew, T* q, KrylovCtl* ctl, int it,
                   hipEvent_t ev0, hipEvent_t ev1);
  template <typename T>
  int run_bicgstab(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
                   bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl, bool* rejected,
                   double* res_norm, double* rhs_norm, bool* converged, int kt_spmv);
  template <typename T>
  int run(int ksp, const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
          bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl, bool* rejected,
          double* res_norm, double* rhs_norm, bool* converged, int kt_spmv);
  void read_ctl(KrylovCtl* dev, KrylovCtl* host);
  void ensure_mixed();
  void ensure_bicgstab(bool fp32);

  Geo g_;
  bool three_d_;
  hipStream_t s_;
  Profiler* prof_;
  SpmvShape shape_[2];
  int kc_f_ = 1;
  dim3 grid_f_;
  unsigned int nb_f_ = 0, part_cap_ = 0;
  bool enq_fresh_ = false;  // enqueue_cg started from an uncleared x0 = 0
  // CG update: 128 x 4 tiles of cell pairs, kc_u_ layers per block; its partials (r.z, z.z)
  int kc_u_ = 1;
  dim3 grid_u_;
  unsigned int nb_u_ = 0;
  double* part_u_ = nullptr;
  bool tuned_[2] = {false, false};
  long long n_ = 0;  // padded length
  double *r_ = nullptr, *z_ = nullptr, *p0_ = nullptr, *p1_ = nullptr, *q_ = nullptr, *xw_ = nullptr;
  float *b32_ = nullptr, *r32_ = nullptr, *z32_ = nullptr, *p032_ = nullptr, *p132_ = nullptr,
        *q32_ = nullptr, *d32_ = nullptr;
  // BiCGSTAB work vectors (shadow residual, ping-pong v, s, t), allocated on first use
  double *rh_ = nullptr, *v0_ = nullptr, *v1_ = nullptr, *sv_ = nullptr, *tv_ = nullptr;
  float *rh32_ = nullptr, *v032_ = nullptr, *v132_ = nullptr, *sv32_ = nullptr, *tv32_ = nullptr;
  double* part_ = nullptr;
  double* stage_ = nullptr;  // first-stage sums of a caller's sweep partials (5 x kRedStage)
  KrylovCtl* ctl_ = nullptr;  // [0] main / outer, [1] inner
  KrylovCtl* host_ctl_ = nullptr;
  bool borrowed_ = false;  // workspace owned by the caller (no allocation, no autotuning)
  KrylovCtl host_ctl_own_{};  // host mirror of the control block in borrowed mode
  int last_iters_ = 8;
  int last_inner_iters_ = 8;
  int last_bcg_iters_ = 8;
  int last_bcg_inner_iters_ = 8;
};

}  // namespace frehg
