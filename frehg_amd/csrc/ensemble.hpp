// Batched ensemble of independent Richards boxes of uneven size, all advanced together: every
// kernel of the PCA step and of the Krylov solve covers all boxes in one launch (a block finds
// its box from a prefix table of tile counts), each box keeps its own Krylov scalars and
// convergence flag, and a step runs from captured HIP graphs. The reference advances such an
// ensemble as independent ReSolver instances (src/re/ReSolver.cpp:931-969), one after another.
#pragma once

#include <string>
#include <vector>

#include "../../include/frehg_gpu.h"
#include "common.hpp"
#include "krylov.hpp"
#include "profile.hpp"
#include "re_device.hpp"

namespace frehg {

struct EnsBox {
  redev::ReK P;    // geometry, soil views and physics of the box
  long long off;   // start of the box's halo-padded storage in the field pools
  int ntx, nty;    // 64 x 4 column tiles per layer
  int kc;          // SpMV layers per block
  int pad;
};

// Per-box Krylov state: outer (FP64 residual / FP64 CG) and inner (FP32 CG) control blocks.
struct EnsCtl {
  KrylovCtl o, i;
  int refinements, inner_total, parity;
  int inner_max[2];  // most inner iterations of the first refinement / of any later one
  int pad;
};

class Ensemble {
 public:
  Ensemble(int nbox, const int32_t* dims, const frehg_re_params* params, hipStream_t s, Profiler* prof);
  ~Ensemble();
  Ensemble(const Ensemble&) = delete;
  Ensemble& operator=(const Ensemble&) = delete;

  void set_soil(int box, const frehg_soil_params* classes, int ncls, const int32_t* class_idx, int idx3d,
                const double* ks_cell);
  void set_initial(int box, const double* h_interior, const double* wc_interior);
  void get_field(int box, const std::string& name, double* dst);
  void step(int use_graph, frehg_ens_stats* st, frehg_re_stats* per_box);
  long long cells() const { return cells_; }
  int boxes() const { return nbox_; }

 private:
  struct Pool {
    double *h, *hn, *wc, *wcn, *wch, *hwc, *ch, *kr, *Kx, *Ky, *Kz, *qx, *qy, *qz, *D, *ax, *ay, *az, *rhs;
    double *r, *z, *p0, *p1, *q, *xa;
    float *D32, *ax32, *ay32, *az32, *b32, *r32, *z32, *p032, *p132, *q32, *d32;
    unsigned char *sat0, *cand, *satA, *satB;
    double* hp;  // extrapolated start only (see Pools::hp)
  };
  void build(int nbox, const int32_t* dims, const frehg_re_params* params);
  void release();
  double* field(const std::string& name);
  void upload_boxes();
  void autotune();
  // Kernel sequences; each can be captured into a graph.
  void enq_prepare();                  // save h^n, theta^n; K/C; assembly; initial guess
  // Krylov budget (all boxes, early exit per box): FP64 iterations, or mixed refinements with
  // inner_first iterations in the first and inner_iters in each later inner solve.
  void enq_solve(int inner_iters, int refinements, int inner_first = -1);
  void enq_update(int sweeps);         // head BCs, fluxes, corrector, reallocation sweeps
  void enq_finish(const unsigned char* fin);  // reallocation branches, clamp, ledger
  void enq_status();
  void read_status();
  void run_graph(int which, int budget_a, int budget_b, int budget_c = 0);

  int nbox_;
  hipStream_t s_;
  Profiler* prof_;
  bool mixed_ = false;
  int guess_ = FREHG_GUESS_ZERO;  // FREHG_GUESS_ZERO, _PREVIOUS or _EXTRAPOLATED (all boxes)
  bool have_inc_ = false;         // hp holds the last step's post-solve head (extrapolated start)
  // the start of the next step: 0 zero, 1 h^n, 2 h^n + the last step's solve increment
  int prep_mode() const {
    if (guess_ == FREHG_GUESS_ZERO) return 0;
    return guess_ == FREHG_GUESS_EXTRAPOLATED && have_inc_ ? 2 : 1;
  }
  KrylovConfig kc_;
  std::vector<int32_t> dims_;
  std::vector<frehg_re_params> params_;
  std::vector<EnsBox> hbox_;
  std::vector<unsigned int> tile0_, spmv0_;  // prefix tables (nbox + 1)
  long long cells_ = 0, storage_ = 0;
  unsigned int ntile_ = 0, nspmv_ = 0;
  Pool F_{};
  EnsBox* dbox_ = nullptr;
  unsigned int *dtile0_ = nullptr, *dspmv0_ = nullptr;
  EnsCtl* ctl_ = nullptr;
  double* part_ = nullptr;   // per-block partials (up to 5 sums per tile)
  double* boxsum_ = nullptr; // per-box ledger sums (8 per box)
  int* flags_ = nullptr;     // device status: [0] all outer converged, [1] realloc changed, [2..] sweep flags
  int* hflags_ = nullptr;
  double* hboxsum_ = nullptr;
  EnsCtl* hctl_ = nullptr;
  std::vector<SoilClass*> classes_;
  std::vector<int32_t*> cls_;
  std::vector<double*> ks_;
  bool tuned_ = false;
  int budget_inner_ = 0, budget_inner0_ = 0, budget_refine_ = 0, budget_fp64_ = 0, budget_sweeps_ = 4;
  hipGraphExec_t graph_[3] = {nullptr, nullptr, nullptr};
  int graph_budget_[3] = {-1, -1, -1};
  int graph_budget2_[3] = {-1, -1, -1};
  int graph_budget3_[3] = {-1, -1, -1};
  int graph_mode_[3] = {-1, -1, -1};
  static constexpr int kMaxSweepFlags = 64;
  int variant_ = 0;  // soil variant code of the per-cell kernels (with_soil_variant_code)
};

}  // namespace frehg
