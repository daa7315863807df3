// Matrix-free Jacobi-preconditioned CG on 7-point (3D) / 5-point (2D) stencil operators in the
// halo-padded layout, FP64 or FP32-inside-FP64-refinement. Replaces PETSc KSP "cg" as driven by
// src/linear/backends/PetscLinearSolver.cpp:186-220.
#pragma once

#include "common.hpp"
#include "profile.hpp"

namespace frehg {

// Device-resident control block of one Krylov solve. Each kernel's last block finalises the
// reductions it produced and writes the scalars the next kernel reads, so a whole iteration
// runs without a host round trip; kernels return immediately once `done` is set.
struct KrylovCtl {
  double rz, alpha, beta, zz, ttol2, bb, pq, rr_guess;
  int iters, done, converged, breakdown, max_it, guess_rejected;
  unsigned int counter[4];
};

// Symmetric stencil operator: diagonal plus the coupling to the +x/+y/+z neighbour of each
// row; the -x/-y/-z couplings are read from the neighbour (ax(c-1) etc.). Off-domain entries
// and every ghost entry are zero. az is null for 2D operators.
template <typename T>
struct StencilView {
  const T* D;
  const T* ax;
  const T* ay;
  const T* az;
};

struct KrylovConfig {
  double rtol = 1e-10, atol = 1e-50;
  int max_it = 10000;
  int precision = 0;        // 0 FP64, 1 FP32 inner + FP64 refinement
  double inner_rtol = 1e-5;
  int max_refine = 8;
};

struct SolveResult {
  int iterations = 0;
  int refinements = 0;
  double residual_norm = 0.0, rhs_norm = 0.0;
  bool converged = false;
  bool fallback_fp64 = false;
  bool guess_rejected = false;
};

class Krylov {
 public:
  Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof);
  ~Krylov();
  Krylov(const Krylov&) = delete;
  Krylov& operator=(const Krylov&) = delete;

  // Solve A x = b on interior rows. x (padded; ghosts finite) holds the initial guess unless
  // x0_zero. bn (may be null = b) is the right-hand side whose ||D^-1 bn|| normalises the
  // convergence test ||D^-1 r|| <= max(rtol ||D^-1 bn||, atol) (PETSc's default test for CG
  // with x0 = 0). fixed (2D, may be null) marks rows whose x is final and whose couplings are
  // already eliminated from b. guard_guess: fall back to x0 = 0 when the guess's residual is
  // not below ||D^-1 b||. A32 (may be null) is the FP32 copy of A for precision == 1.
  SolveResult solve(const StencilView<double>& A, const StencilView<float>* A32, const double* b,
                    const double* bn, double* x, const unsigned char* fixed, bool x0_zero,
                    bool guard_guess, const KrylovConfig& cfg);

  int kchunk() const { return kchunk_; }
  void set_kchunk(int kc);

 private:
  template <typename T>
  int run_pcg(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
              bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl,
              T* r, T* z, T* p0, T* p1, T* q, bool* rejected, double* res_norm, double* rhs_norm,
              bool* converged, int kt_spmv);
  void read_ctl(KrylovCtl* dev, KrylovCtl* host);
  void ensure_mixed();

  Geo g_;
  bool three_d_;
  hipStream_t s_;
  Profiler* prof_;
  int kchunk_ = 0;
  dim3 grid_;
  unsigned int nblocks_ = 0;
  long long n_ = 0;  // padded length
  double *r_ = nullptr, *z_ = nullptr, *p0_ = nullptr, *p1_ = nullptr, *q_ = nullptr, *xw_ = nullptr;
  float *b32_ = nullptr, *r32_ = nullptr, *z32_ = nullptr, *p032_ = nullptr, *p132_ = nullptr,
        *q32_ = nullptr, *d32_ = nullptr;
  double* part_ = nullptr;
  KrylovCtl* ctl_ = nullptr;  // [0] main / outer, [1] inner
  KrylovCtl* host_ctl_ = nullptr;
  int last_iters_ = 8;
  int last_inner_iters_ = 8;
};

}  // namespace frehg
