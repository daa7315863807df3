/*
 * frehg_gpu.h -- C ABI of the MI355X (gfx950) backend for the Frehg2 coupled step.
 *
 * One context = one HIP device + one HIP stream + the device-resident state of one Richards
 * (groundwater, "GW") solver, one semi-implicit shallow-water (surface, "SW") solver and their
 * exchange. Fields use the reference's halo-padded layout (src/core/Grid.hpp:42-60):
 *   3D index (i+1) + (j+1)(nx+2) + (k+1)(nx+2)(ny+2), k = 0 is the TOP layer, k = -1 / nz ghosts;
 *   2D index (i+1) + (j+1)(nx+2).
 * All arrays crossing this ABI are host pointers; the library copies them. Every call is
 * synchronous with respect to the host pointers it receives. Calls on one context are not
 * thread-safe; distinct contexts may be driven from distinct host threads.
 *
 * Return codes: 0 = ok, > 0 = informational (FREHG_INFO_*), < 0 = error (FREHG_ERR_*); the text
 * of the last error is frehg_last_error(ctx). The reference raises std::runtime_error for the
 * same conditions (src/linear/backends/PetscLinearSolver.cpp:206-211, src/re/ReSolver.cpp:71-74).
 */
#ifndef FREHG_GPU_H
#define FREHG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FREHG_ABI_VERSION 5

enum {
  FREHG_OK = 0,
  FREHG_INFO_GUESS_FALLBACK = 1, /* learned initial guess rejected, zero guess used */
  FREHG_INFO_PICARD_NOT_CONVERGED = 2, /* Picard loop stopped at picard_max before its max |dh| test
                                          passed; the step is taken from the last iterate */
  FREHG_ERR_INVALID_ARG = -1,
  FREHG_ERR_OOM = -2,
  FREHG_ERR_HIP = -3,
  FREHG_ERR_NOT_CONVERGED = -4, /* ~ KSPConvergedReason < 0 */
  FREHG_ERR_NONFINITE = -5,
  FREHG_ERR_STATE = -6, /* call order violated (e.g. step before init) */
  FREHG_ERR_NO_DEVICE = -7
};

/* Krylov method / preconditioner / precision selectors. */
/* KSP types: "cg" and "bcgs" of SolverConfig::ksp_type (include/frehg2/linear/SolverConfig.hpp:15),
 * set per module by surface_water.solver.ksp_type / groundwater.solver.ksp_type
 * (src/driver/Orchestrator.cpp:435,488). */
enum { FREHG_KSP_CG = 0, FREHG_KSP_BICGSTAB = 1 };
enum { FREHG_PC_JACOBI = 0 };
enum {
  FREHG_PREC_FP64 = 0,      /* FP64 Jacobi-PCG (reference arithmetic) */
  FREHG_PREC_MIXED = 1      /* FP32 Jacobi-PCG inside FP64 iterative refinement */
};
enum {
  FREHG_GUESS_ZERO = 0,     /* x0 = 0 (PETSc default, the reference) */
  FREHG_GUESS_PREVIOUS = 1, /* x0 = h^n */
  FREHG_GUESS_LEARNED = 2,  /* x0 = h^n + the previous step's solve increment + MLP(local stencil
                               features), zero-guess fallback; the network is trained on the
                               increment minus the previous one */
  FREHG_GUESS_EXTRAPOLATED = 3 /* x0 = h^n + the previous step's solve increment (h^n on the first step) */
};

/* Surface boundary region rules: BcKind (src/bc/PolygonRegion.hpp:20-24). */
enum { FREHG_BC_DISCHARGE = 0, FREHG_BC_DEPTH = 1, FREHG_BC_CRITICAL = 2 };

/* Soil class: include/frehg2/re/SoilParams.hpp:13-25. */
typedef struct {
  double alpha, n, theta_s, theta_r;
  double Ks_x, Ks_y, Ks_z, Ss;
  int32_t use_vg, use_mvg;
  double air_entry;
} frehg_soil_params;

/* Krylov solve configuration: include/frehg2/linear/SolverConfig.hpp:14-20, plus precision. */
typedef struct {
  int32_t ksp_type;  /* FREHG_KSP_* */
  int32_t pc_type;   /* FREHG_PC_* */
  double rtol;       /* preconditioned residual 2-norm <= rtol * ||D^-1 b|| (PETSc CG default) */
  double atol;
  int32_t max_it;
  int32_t precision;  /* FREHG_PREC_* */
  double inner_rtol;  /* FP32 inner solve relative tolerance (mixed only) */
  int32_t max_refine; /* FP64 refinement sweeps before falling back to FP64 PCG (mixed only) */
} frehg_solver_config;

/* Richards (PCA) step parameters: src/re/ReSolver.hpp:24-46 (+ solver, Picard and guess). */
typedef struct {
  frehg_soil_params soil;
  double dx, dy, dz, botz;
  double dt, dt_min, dt_max, co_max;
  int32_t adaptive_dt, use_corrector, use_full3d, sim_shallowwater;
  int32_t bc_type[6]; /* x-, x+, y-, y+, z+ (bottom), z- (top); bottom 3 = free drainage */
  double htop, hbot, qtop, qbot;
  frehg_solver_config solver;
  int32_t picard_max; /* 1 = reference PCA predictor; > 1 = Picard outer iteration */
  double picard_tol;  /* max |h^{m+1} - h^m| stopping test [m] */
  int32_t initial_guess; /* FREHG_GUESS_* */
} frehg_re_params;

/* Shallow-water parameters: src/swe/SweSolver.hpp:28-39 (+ solver). */
typedef struct {
  double gravity, manning, min_depth, visc_x, visc_y, hD, hE, wtfh, dt, offset;
  frehg_solver_config solver;
} frehg_swe_params;

/* SW<->GW exchange parameters: src/coupling/Coupling.hpp:26-31. */
typedef struct {
  double dx, dy, visc, min_depth;
} frehg_coupling_params;

typedef struct {
  int32_t iterations;   /* Krylov iterations (inner FP32 + FP64 for mixed) */
  int32_t refinements;  /* FP64 refinement sweeps (mixed) */
  double residual_norm; /* final ||D^-1 r|| */
  double rhs_norm;      /* ||D^-1 b|| */
  int32_t converged;
  int32_t fallback;     /* mixed solve fell back to FP64 PCG */
} frehg_solve_stats;

typedef struct {
  frehg_solve_stats solve; /* accumulated over Picard iterations */
  double dt_used, dt_next;
  int32_t picard_iterations;
  int32_t guess_fallback;
  double max_dh;           /* max |h^{n+1} - h^n| over interior cells */
  /* FP64 water ledger of the step (m^3; sums over interior cells in a fixed order). With the
   * corrector on, mass_after - mass_before = boundary_in + storage_ss + adjust up to round-off,
   * and mass_residual is that difference. theta_c is the corrector's theta (ReSolver.cpp:762-786),
   * before reallocation and clamp (:788-853). Lateral faces are closed, so only the top and bottom
   * faces exchange water. Without the corrector theta_c = theta(h) and the residual is the PCA's
   * own imbalance. */
  double mass_before;      /* sum theta^n V (after any SW-GW exchange) */
  double mass_after;       /* sum theta^{n+1} V */
  double mass_boundary_in; /* dt * sum over columns of (q_bottom_face - q_top_face), + = into the soil */
  double mass_storage_ss;  /* -sum theta_c V Ss (h - h^n) / theta_s: specific storage, outside theta V */
  double mass_adjust;      /* sum (theta^{n+1} - theta_c) V: reallocation and clamp */
  double mass_residual;    /* mass_after - mass_before - boundary_in - storage_ss - adjust */
  int32_t picard_converged; /* 1 if the Picard test passed (always 1 with picard_max = 1) */
} frehg_re_stats;

typedef struct {
  frehg_solve_stats sw;
  frehg_re_stats gw;       /* last GW substep; gw.solve.iterations summed over substeps */
  int32_t gw_substeps;
  double exchange_volume;  /* sum(q*dt) of this step (+ = net GW->SW seepage) */
  double t, gw_time;
  double boundary_outflow; /* net volume the surface boundary regions removed this step [m^3] */
} frehg_coupled_stats;

typedef struct frehg_ctx_s* frehg_ctx_t;

/* ---- context -------------------------------------------------------------------------------- */
int frehg_abi_version(void);
/* Create a context on HIP device `device`. Fails with FREHG_ERR_NO_DEVICE without a GPU. */
int frehg_create(frehg_ctx_t* out, int device);
void frehg_destroy(frehg_ctx_t ctx);
const char* frehg_last_error(frehg_ctx_t ctx);
int frehg_synchronize(frehg_ctx_t ctx);

/* ---- Richards solver: replaces ReSolver (src/re/ReSolver.hpp:48-105) ------------------------ */
/* ReSolver(Grid) + setParams + initializeGeometry (uniform dz, all cells active). */
int frehg_re_init(frehg_ctx_t ctx, int nx, int ny, int nz, const frehg_re_params* p);
/* ReSolver::setParams (ReSolver.hpp:53). Grid size is fixed at init. */
int frehg_re_set_params(frehg_ctx_t ctx, const frehg_re_params* p);
/* ReSolver::setSoilMap (ReSolver.hpp:58) with SoilMap::setClasses/setClassIndex{,3D}
 * (src/soil/SoilMap.hpp:33-46). class_idx: nx*ny (idx3d = 0) or nx*ny*nz (idx3d = 1), may be
 * NULL (all class 0); ks_cell: nx*ny*nz isotropic per-cell Ks replacing the class Ks (one class
 * per cell), may be NULL. Order i + j*nx (+ k*nx*ny). */
int frehg_re_set_soil(frehg_ctx_t ctx, const frehg_soil_params* classes, int ncls,
                      const int32_t* class_idx, int idx3d, const double* ks_cell);
/* ReSolver::setTopFluxField (ReSolver.hpp:105), nx*ny values, i + j*nx. */
int frehg_re_set_top_flux(frehg_ctx_t ctx, const double* q_interior);
/* ReSolver::setInitialHead / setInitialWaterContent + finalizeInitialState (:155-178):
 * interior values i + j*nx + k*nx*ny; exactly one of h / wc non-NULL. */
int frehg_re_set_initial(frehg_ctx_t ctx, const double* h_interior, const double* wc_interior);
/* Direct access to any GwFields array by name (src/re/GwFields.hpp:62-74), halo padded.
 * Names: h hn wc wcn wch hwc ch Kx Ky Kz qx qy qz Gct Grhs. Gct/Grhs are the system of the
 * step's last Picard iterate (the last assembly; a loop that ended on the residual test of an
 * iterate without storing its system forms it when it is read). */
int frehg_re_set_field(frehg_ctx_t ctx, const char* name, const double* src);
int frehg_re_get_field(frehg_ctx_t ctx, const char* name, double* dst);
/* ReSolver::advanceStep (ReSolver.cpp:931-969). */
int frehg_re_step(frehg_ctx_t ctx, frehg_re_stats* stats);
/* ReSolver::dt / primeAdaptiveDt (ReSolver.hpp:86-98). */
double frehg_re_dt(frehg_ctx_t ctx);
int frehg_re_prime_dt(frehg_ctx_t ctx, double dt);

/* ---- Shallow-water solver: replaces SweSolver (src/swe/SweSolver.hpp:41-94) ----------------- */
int frehg_swe_init(frehg_ctx_t ctx, int nx, int ny, double dx, double dy, const frehg_swe_params* p);
int frehg_swe_set_params(frehg_ctx_t ctx, const frehg_swe_params* p);
/* SweSolver::setBathymetry (physical, i + j*nx) -> updateBottomFaces (:144-205). */
int frehg_swe_set_bathymetry(frehg_ctx_t ctx, const double* bed_interior);
/* SweSolver::initializeState (:926-934): eta = init + offset clamped to bed, u = v = 0. */
int frehg_swe_initialize_state(frehg_ctx_t ctx, double init_eta);
/* SweSolver::setInitialEtaPhysical + setInitialVelocityConstant(0, 0) + finalizeInitialState
 * (src/swe/SweSolver.cpp:855-924): a non-uniform initial surface, nx*ny physical values
 * (i + j*nx, the offset is added), clamped to the bed, zero velocity. */
int frehg_swe_set_initial_eta(frehg_ctx_t ctx, const double* eta_interior);
/* Any SweFields array by name (src/swe/SweFields.hpp:131-146), halo padded. */
int frehg_swe_set_field(frehg_ctx_t ctx, const char* name, const double* src);
int frehg_swe_get_field(frehg_ctx_t ctx, const char* name, double* dst);
/* SweSolver::advanceStep(rain, evap) (:803-853). */
int frehg_swe_step(frehg_ctx_t ctx, double rain, double evap, frehg_solve_stats* stats);
/* PolygonBC(regions, PolygonIndex) (src/bc/PolygonBC.hpp:28-46, src/bc/PolygonIndex.cpp:9-56): nregions
 * polygons, region r with nvert[r] vertices (x, y pairs in xy, regions concatenated; rings closed),
 * kind[r] = FREHG_BC_* and rate[r] (discharge Q m^3/s, > 0 out; depth h m; critical: unused). A cell
 * whose centre (x0 + (i + 1/2) dx, y0 + (j + 1/2) dy) lies inside a polygon (even-odd rule, first
 * match wins) takes that region's rule. nregions = 0 removes the regions. frehg_coupled_step
 * applies them after the surface step, before the exchange (Orchestrator.cpp:1240, :1301-1308). */
int frehg_swe_set_boundary_polygons(frehg_ctx_t ctx, int nregions, const int32_t* kind, const double* rate,
                                    const int32_t* nvert, const double* xy, double x0, double y0);
/* PolygonBC::applySurface (src/bc/PolygonBC.cpp:16-78) over dt: *outflow = net volume removed [m^3]. */
int frehg_swe_apply_boundaries(frehg_ctx_t ctx, double dt, double* outflow);
/* SweSolver::maxCfl (:954-975). */
int frehg_swe_max_cfl(frehg_ctx_t ctx, double* out);

/* ---- Coupling: replaces Coupling::exchange (src/coupling/Coupling.cpp:145-164) -------------- */
/* q_out (nx*ny, i + j*nx, m^3/s, + = seepage) may be NULL. */
int frehg_couple_exchange(frehg_ctx_t ctx, const frehg_coupling_params* cp, double dt,
                          double* volume_out, double* q_out);
/* Orchestrator::step, coupled branch (src/driver/Orchestrator.cpp:1322-1336, 1243-1252):
 * SW advance at `rain`, exchange over dt, GW substeps until gw_time >= t + dt. */
int frehg_coupled_step(frehg_ctx_t ctx, const frehg_coupling_params* cp, double dt, double rain,
                       double evap, double t_end, frehg_coupled_stats* stats);
/* Water volume sum(max(0, eta-b)) dx dy + sum(theta) dx dy dz (Orchestrator.cpp:898-936),
 * FP64, fixed-order reduction. */
int frehg_water_volume(frehg_ctx_t ctx, double* surface, double* subsurface);

/* ---- LinearSolver seam: replaces LinearSolver::solve (include/frehg2/linear/LinearSolver.hpp:18-33)
 * for operators assembled by Decomp3D / Decomp2D (STAR stencil). The operator is given per
 * owned row as a diagonal plus one coefficient per neighbour direction (x-, x+, y-, y+, z-, z+;
 * 2D: the first four), exactly the values addRow would insert for those columns
 * (src/re/ReSolver.cpp:576-610, src/swe/SweSolver.cpp:542-569). b_owned / x_owned use the
 * reference's owned ordering: 3D z-fastest ((j*nx + i)*nz + k), 2D x-fastest (j*nx + i)
 * (src/linear/DomainDecomposition.cpp:23-33,109-119). */
int frehg_ls_solve(frehg_ctx_t ctx, int ndim, int nx, int ny, int nz, const double* diag,
                   const double* const* offdiag, const double* b_owned, double* x_owned,
                   const frehg_solver_config* cfg, frehg_solve_stats* stats);

/* Allocation-free variant of frehg_ls_solve. No counterpart in the reference: it is for callers
 * that cannot allocate device memory at solve time (tight serving loops, harnesses that forbid
 * hipMalloc). diag, offdiag[0..5] (2D: [0..3]), b_owned and x_owned are DEVICE arrays in the same
 * owned ordering and meaning as frehg_ls_solve (offdiag itself is a host array of device
 * pointers); x_owned receives the solution. FP64 CG (cfg->ksp_type FREHG_KSP_CG, precision
 * FREHG_PREC_FP64) on a symmetric operator, x0 = 0. The call allocates nothing: the padded
 * operator and the CG vectors live in a 64 MiB device array per GPU, held by the library until
 * the process ends, so an operator of more than about 760k padded cells is rejected with
 * FREHG_ERR_INVALID_ARG. Calls that overlap on one GPU (other threads or processes using the same
 * GPU at the same time) race on that array; this process's threads are serialised. */
int frehg_ls_solve_noalloc(frehg_ctx_t ctx, int ndim, int nx, int ny, int nz, const double* diag,
                           const double* const* offdiag, const double* b_owned, double* x_owned,
                           const frehg_solver_config* cfg, frehg_solve_stats* stats);

/* ---- Batched ensemble of independent Richards boxes (configs of uneven box sizes or parameter
 * ensembles). The reference advances such an ensemble as independent ReSolver instances
 * (src/re/ReSolver.cpp:931-969); here every kernel of a step covers all boxes in one launch, each
 * box keeps its own Krylov scalars and convergence test, and steps run from captured HIP graphs.
 * Restrictions: fixed dt (adaptive_dt = 0), the same solver configuration in every box,
 * initial guess zero, previous-step or extrapolated (h^n + the last step's solve increment; the
 * first step after frehg_ens_set_initial starts from h^n), one Picard iterate, closed lateral
 * sides. --------------------------------------------------------------------------------------- */
typedef struct {
  int32_t max_iterations;        /* largest Krylov iteration count over the boxes (inner FP32 for mixed) */
  int32_t max_refinements;       /* largest number of FP64 refinement sweeps (mixed) */
  int32_t converged;             /* 1 if every box converged */
  int32_t graph;                 /* 1 if the step ran from captured graphs */
  int32_t extra_launches;        /* eager launches needed beyond the captured budget */
  int32_t realloc_sweeps;        /* saturation fixed-point sweeps used */
  double max_rel_mass_residual;  /* max over boxes of |mass_residual| / mass_before */
  double mass_before, mass_after; /* sums over boxes */
} frehg_ens_stats;
/* dims: nboxes x (nx, ny, nz); params: one frehg_re_params per box. */
int frehg_ens_init(frehg_ctx_t ctx, int nboxes, const int32_t* dims, const frehg_re_params* params);
int frehg_ens_set_soil(frehg_ctx_t ctx, int box, const frehg_soil_params* classes, int ncls,
                       const int32_t* class_idx, int idx3d, const double* ks_cell);
int frehg_ens_set_initial(frehg_ctx_t ctx, int box, const double* h_interior, const double* wc_interior);
/* Any GwFields array of one box (names as frehg_re_get_field), halo padded. */
int frehg_ens_get_field(frehg_ctx_t ctx, int box, const char* name, double* dst);
/* One step of every box. use_graph: run from captured graphs (captured at first use). per_box may
 * be NULL, else nboxes entries with each box's solve and ledger. */
int frehg_ens_step(frehg_ctx_t ctx, int use_graph, frehg_ens_stats* stats, frehg_re_stats* per_box);

/* ---- Learned initial guess (no counterpart in the reference: it changes only the Krylov
 * x0 of the Richards solve, src/re/ReSolver.cpp:560-636, whose PETSc KSP always starts from 0).
 * An 11 -> 64 -> 64 -> 1 tanh MLP over local features of the state before the solve (h^n(c), the
 * six h^n(neighbour) - h^n(c), ln K(h^n), C(h^n), theta^n, surface depth of the column) predicts
 * the cell's solve increment h^{n+1}(c) - h^n(c) minus the previous step's, d^{n-1}(c). BF16 MFMA
 * forward and backward with FP32 accumulation, Adam on FP32 master weights, all on the GPU. With
 * frehg_re_params.initial_guess = FREHG_GUESS_LEARNED the solve (and the Picard loop, with
 * picard_max > 1) starts from h^n + d^{n-1} + prediction (d^{n-1} = 0 on the first step after
 * set-up) and falls back to x0 = 0 (stats.guess_fallback = 1) when the guess's residual is not
 * below the zero guess's. ---------------------------------------------------------------------- */
#define FREHG_MLP_INPUTS 11
#define FREHG_MLP_PARAMS 4993 /* W1[64][11], b1[64], W2[64][64], b2[64], w3[64], b3 */
#define FREHG_MLP_NORM 24     /* mean[12], sd[12] of the 11 features and the target */
/* Glorot-uniform weights, zero biases, identity normalisation. */
int frehg_mlp_init(frehg_ctx_t ctx, uint64_t seed);
/* Log (features, converged increment minus the previous step's) of every `stride`-th interior
 * cell after each Richards solve into a ring buffer of `capacity` rows; stride 0 stops logging.
 * The first step after set-up (no previous increment) logs nothing. */
int frehg_mlp_log(frehg_ctx_t ctx, int stride, int64_t capacity);
int frehg_mlp_logged(frehg_ctx_t ctx, int64_t* n);
/* `steps` Adam steps (lr; beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of `batch` logged
 * rows drawn with replacement (seeded). The first call fixes the normalisation from the log.
 * Loss = mean squared error of the standardised target, first and last minibatch. */
int frehg_mlp_train(frehg_ctx_t ctx, int steps, int batch, double lr, uint64_t seed, double* loss_first,
                    double* loss_last);
int frehg_mlp_get_params(frehg_ctx_t ctx, float* params, float* norm);
int frehg_mlp_set_params(frehg_ctx_t ctx, const float* params, const float* norm);
/* Loss and d loss / d params on a host batch: x[n][11] raw features, y[n] raw increments. */
int frehg_mlp_loss_grad(frehg_ctx_t ctx, const float* x, const float* y, int n, float* grad, double* loss);
/* Predicted raw increments for x[n][11]. */
int frehg_mlp_predict(frehg_ctx_t ctx, const float* x, int n, float* out);
/* Inference gate (metres, default 0): with a previous solve increment d^{n-1} available, a cell
 * whose |d^{n-1}| <= gate starts from h^n + d^{n-1} without the network's correction, and a chunk
 * of cells that all do so skips the network. 0 applies the network to every cell. */
int frehg_mlp_gate(frehg_ctx_t ctx, double gate);
/* Cells in the chunks the network ran on in the last learned guess (the rest took h^n + d^{n-1}). */
int frehg_mlp_inferred_cells(frehg_ctx_t ctx, int64_t* cells);
/* The network's prediction for a step taken from the current Richards state (features from h,
 * K(h), C(h), theta): the part of the learned guess's increment on top of the previous step's
 * solve increment. nx*ny*nz values, i + j*nx + k*nx*ny. */
int frehg_mlp_increment(frehg_ctx_t ctx, double* dst);
/* The learned start the next frehg_re_step (or coupled step) would take from the current state:
 * h^n + d^{n-1} + the network's correction where |d^{n-1}| exceeds the gate, h^n + d^{n-1}
 * elsewhere (h^n + the correction when no previous increment is known). nx*ny*nz values,
 * i + j*nx + k*nx*ny. Inspection of the learned start; no counterpart in the reference. */
int frehg_mlp_start(frehg_ctx_t ctx, double* dst);

/* ---- measurement --------------------------------------------------------------------------- */
/* Kernel-duration accounting with HIP events on the context's stream. kernel ids below. */
enum {
  FREHG_KT_GW_SPMV = 0,     /* groundwater FP64 Krylov SpMV+dot launches (7-point) */
  FREHG_KT_GW_SPMV32 = 1,   /* groundwater FP32 inner Krylov SpMV+dot launches (mixed) */
  FREHG_KT_GW_UPDATE = 2,   /* groundwater Krylov vector update + reductions (FP64 and FP32) */
  FREHG_KT_GW_ASSEMBLE = 3, /* K, C at h^n + 7-point assembly (one span per step) */
  FREHG_KT_GW_CORRECT = 4,  /* ghosts, K/C at h^{n+1}, fluxes, corrector (one span per step) */
  FREHG_KT_SW_SPMV = 5,     /* surface Krylov SpMV (reserved: the launch-bound 2D solve is not timed) */
  FREHG_KT_SW_UPDATE = 6,   /* surface Krylov update (reserved, not timed) */
  FREHG_KT_GW_REFINE = 7,   /* groundwater FP64 residual sweep of iterative refinement */
  FREHG_KT_GW_GUESS = 8,    /* learned-guess inference (features + MLP forward) */
  FREHG_KT_COUNT = 9
};
int frehg_profile_enable(frehg_ctx_t ctx, int enable);
/* Kernel types timed while profiling is enabled: bit kt of kernel_mask (default: all). A timed
 * launch costs the host a few microseconds more to issue, so a measurement of the whole step
 * times only the kernel types it reports. */
int frehg_profile_select(frehg_ctx_t ctx, uint32_t kernel_mask);
/* Total milliseconds and launch count recorded for kernel `kt` since the last reset. */
int frehg_profile_read(frehg_ctx_t ctx, int kt, double* total_ms, int64_t* launches);
int frehg_profile_reset(frehg_ctx_t ctx);
/* Groundwater CG SpMV launch shape in use for `precision` (FREHG_PREC_FP64: the FP64 solve,
 * FREHG_PREC_MIXED: the FP32 inner solve): cells per lane (1: one column per thread; 2: two
 * adjacent columns per lane with paired aligned loads) and layers per block. Set by the autotuner
 * at the first solve, or pinned by the environment variable FREHG_SPMV_SHAPE="cpl:kchunk" read at
 * frehg_re_init. No counterpart in the reference (PETSc's MatMult has no launch shape). */
int frehg_spmv_shape(frehg_ctx_t ctx, int precision, int* cells_per_lane, int* kchunk);
/* Streaming probe on `device`: the CG SpMV's HBM access mix (six FP64 reads and two FP64 writes
 * per cell, 16-byte accesses) with no neighbours and no reduction, over eight arrays of `cells`
 * doubles that it allocates and frees; for each of a few grid sizes, `reps` timed launches after
 * one warm-up. Returns the rate in GB/s of those 64 bytes per cell and the mean launch time of the
 * fastest grid size. A measurement aid with no
 * counterpart in the reference: the ceiling the SpMV's access mix streams at on this device. */
int frehg_stream_probe(int device, int64_t cells, int reps, double* gbytes_per_s, double* ms_per_launch);

/* ---- Passive solute: SoluteParams / SoluteStepper / Orchestrator::applySolute
 * (src/solute/SoluteParams.hpp:16-30, src/solute/SoluteStepper.cpp:24-72,
 * src/driver/Orchestrator.cpp:938-1056, 1270-1299). One concentration per surface cell and per
 * soil cell. After each flow step: rain of concentration c_rain mixes into cells wet at the start
 * of the step, then per substep first-order decay, explicit conservative advection by the
 * surface velocities and Darcy fluxes (upwind, or minmod MUSCL), refused above cfl_max (the
 * substep count doubles, up to 4096; beyond that the step leaves the solute unchanged), and
 * implicit diffusion (I - dt D L) C' = C solved by Jacobi-CG. A coupled step carries solute with
 * the SW<->GW exchange water before the groundwater catches up. */
typedef struct {
  double c_rain;      /* rain concentration */
  double k_decay;     /* first-order decay rate [1/s] */
  double D;           /* isotropic diffusion coefficient [m^2/s] */
  double cfl_max;     /* advective Courant ceiling (default 1) */
  double min_depth;   /* surface depth below which a cell takes no rain solute [m] */
  double start_time;  /* no transport in steps that end at or before this time [s] */
  int32_t muscl;      /* 0: first-order upwind, 1: minmod-limited MUSCL */
  int32_t implicit_diffusion; /* 0: none, 1: implicit */
  int32_t substeps;   /* substeps per flow step to start from (>= 1) */
  frehg_solver_config solver; /* the diffusion solve: FP64 CG or BiCGSTAB, Jacobi (reference: CG, rtol 1e-12) */
} frehg_solute_params;

typedef struct {
  double rain_added;      /* solute mass the rain brought [concentration * m^3] */
  double decay_loss;      /* solute mass decay removed */
  double max_courant;     /* largest Courant number of an accepted substep */
  int64_t substeps;       /* accepted substeps */
  int32_t refused_steps;  /* flow steps whose Courant number stayed above cfl_max at 4096 substeps */
  int32_t diffusion_iterations; /* CG iterations of the diffusion solves */
} frehg_solute_stats;

/* After frehg_swe_init and/or frehg_re_init: the solute of whichever solvers the context holds
 * (concentrations start at zero). */
int frehg_solute_init(frehg_ctx_t ctx, const frehg_solute_params* p);
/* Interior initial concentrations (x fastest, then y, then z); either may be NULL. */
int frehg_solute_set_initial(frehg_ctx_t ctx, const double* surface_interior, const double* subsurface_interior);
/* name "surface" ((nx+2)(ny+2)) or "subsurface" ((nx+2)(ny+2)(nz+2)), halo-padded as the other fields. */
int frehg_solute_set_field(frehg_ctx_t ctx, const char* name, const double* src);
int frehg_solute_get_field(frehg_ctx_t ctx, const char* name, double* dst);
/* applySolute for the flow step [t, t + dt] of a surface-only or groundwater-only driver, after
 * its flow step (frehg_coupled_step applies it by itself). */
int frehg_solute_apply(frehg_ctx_t ctx, double t, double dt, double rain);
/* computeSoluteMass (:852-896) and the cumulative totals. */
int frehg_solute_mass(frehg_ctx_t ctx, double* mass, frehg_solute_stats* totals);

/* ---- Checkpoint and restart of the device state: Orchestrator::writeCheckpointNow /
 * loadCheckpointState (src/driver/Orchestrator.cpp:1461-1502, 688-747), Hdf5Writer::
 * writeCheckpoint (src/io/Hdf5Writer.cpp:400-470). HDF5 is not available here: the file is a
 * flat binary (magic "FRHGCKP1", version, which solvers, grid, times, the next groundwater dt,
 * a digest of the solver parameters, then every named field of each solver with its padded
 * (nx+2)(ny+2)[(nz+2)] values), written to path + ".tmp" and renamed over path. A load needs a
 * context set up with the same solvers and grid (else FREHG_ERR_INVALID_ARG); it restores the
 * fields, the coupled-step clocks and the groundwater dt. A context with a solute saves its
 * concentrations too ("sl_surface", "sl_subsurface") and needs them to load (the solute totals
 * start again from zero, as loadCheckpointState's, :722-745); a checkpoint's solute is ignored by
 * a context without one. With the zero or previous-step Krylov
 * start and the same SpMV launch shape (FREHG_SPMV_SHAPE pins it) the continued run is
 * bit-identical to an uninterrupted one; the extrapolated and learned starts begin again from
 * h^n on the first step after a load. */
typedef struct {
  double t, gw_time;    /* coupled-step clocks at the save */
  double gw_dt;         /* groundwater dt of the next step */
  int32_t has_sw, has_gw, nx, ny, nz;
  int32_t params_match; /* 1 if the context's solver parameters equal the saved ones */
  int32_t has_solute;   /* 1 if the checkpoint holds solute concentrations */
} frehg_checkpoint_info;

int frehg_checkpoint_save(frehg_ctx_t ctx, const char* path);
int frehg_checkpoint_load(frehg_ctx_t ctx, const char* path, frehg_checkpoint_info* info);

/* ---- Configuration front end: a Frehg2 v2 YAML file run through this library. Replaces
 * Config::fromFile (src/io/Config.cpp:51-64) + Orchestrator::initialize / run
 * (src/driver/Orchestrator.cpp:192-304, 1595-1615) behind the frehg2 command line
 * (src/driver/main.cpp). The same keys, defaults and gates as the reference (schema v2, the
 * capability gate); options this backend does not implement are refused with
 * FREHG_ERR_INVALID_ARG and a message naming the key: solute transport, polygon sources, probe
 * monitors, formula / polygon initial conditions, HDF5 inputs, wind, subgrid, diffusive wave.
 * Checkpoints (time.dt_checkpoint, time.max_checkpoints) and restarts (initial_conditions.type:
 * restart, file, time) use frehg_checkpoint_save / _load files. Added keys (all optional, defaults = the reference's behaviour):
 * groundwater.solver.precision (fp64 | mixed), .inner_rtol, .max_refine; groundwater.picard.
 * max_iterations (1 = the PCA predictor) and .tolerance; groundwater.initial_guess (zero |
 * previous | extrapolated). `overrides` are "dotted.key=value" strings (value in YAML syntax)
 * applied after loading, like the reference's own tests editing the loaded tree. */
typedef struct {
  int32_t surface_water, groundwater, solute, coupled;
  int32_t nx, ny, nz;
  double dx, dy, dz, botz;
  double dt, t_end, output_interval;
  int64_t max_steps;
  double evaporation_rate;
  int32_t rain_from_file; /* 1: rain_samples (time, rate) pairs interpolated in time */
  int32_t rain_samples;
  double rain_rate;       /* constant rate when rain_from_file = 0 [m/s] */
  frehg_swe_params swe;   /* offset = -min(bathymetry) */
  frehg_re_params re;     /* soil = the first soil class */
  frehg_coupling_params coupling;
  int32_t soil_classes;
  int32_t soil_map;       /* 0: uniform, 2: per column, 3: per cell (soil.map.layers) */
  int32_t recharge;       /* 1: groundwater.recharge sets a top flux field */
  int32_t boundary_regions;
  int32_t gw_initial_head; /* 1: the groundwater starts from heads, 0: from water contents */
  int32_t restart;         /* 1: initial_conditions.type = restart (state from a checkpoint file) */
  double dt_checkpoint;    /* 0: no checkpoints */
  int32_t max_checkpoints;
  frehg_solute_params solute_params; /* the solute: block (when solute = 1, i.e. solute.enabled) */
} frehg_run_setup;

typedef struct {
  double t;                /* simulated time reached [s] */
  int64_t steps;           /* time steps (the groundwater's own steps in a groundwater-only run) */
  int64_t gw_substeps;
  int32_t outputs;         /* field snapshots written (the initial one included) */
  int32_t infos;           /* steps that returned an FREHG_INFO_* code */
  double water_initial, water_final; /* surface + subsurface water volume [m^3] */
  double rain_volume;      /* rain added (the surface step skips the last y row, as the reference) */
  double boundary_outflow; /* net volume the surface boundary regions removed */
  double exchange_volume;  /* net GW -> SW seepage volume */
  double wall_seconds;
  /* the solute ledger of writeSummary (:1666-1675); zero without a solute */
  double solute_initial_mass, solute_final_mass, solute_added_by_rain, solute_decay_loss;
  double solute_relative_mass_error; /* |final - (initial + rain - decay)| / |initial + rain| */
} frehg_run_stats;

/* Loads, gates and maps a configuration without a GPU (no context). err receives the message of
 * a failure (err_len bytes, NUL-terminated). */
int frehg_config_setup(const char* yaml_path, const char* const* overrides, int n_overrides, frehg_run_setup* out,
                       char* err, int err_len);
/* Value at a dotted key of any YAML file (after overrides; no schema check). Returns 0 when the key is missing or
 * null, 1 for a scalar (its text in value), 2 for a sequence and 3 for a mapping (their sizes in
 * value), or FREHG_ERR_INVALID_ARG with the parse error in value. */
int frehg_config_query(const char* yaml_path, const char* const* overrides, int n_overrides, const char* key,
                       char* value, int value_len);
/* A resolved per-cell input of a configuration (no GPU): "bathymetry" (nx*ny, plus the datum
 * offset), "eta0" (nx*ny, physical), "gw_initial" (nx*ny*nz heads or water contents),
 * "top_flux" (nx*ny), "soil_class" (nx*ny or nx*ny*nz), "rain_t" / "rain_v" (the rainfall
 * series), "solute_surface" (nx*ny) / "solute_subsurface" (nx*ny*nz) initial concentrations; x fastest, then y, then z. Copies at most n values into out and returns the count the
 * field holds (0 when the configuration has none), or FREHG_ERR_INVALID_ARG with the message in err. */
int64_t frehg_config_field(const char* yaml_path, const char* const* overrides, int n_overrides, const char* name,
                           double* out, int64_t n, char* err, int err_len);
/* Runs the configuration on ctx (which must hold no solver yet): set-up, the time loop of
 * Orchestrator::runLoopSequential (:1518-1548; async coupling runs the same sequence
 * synchronously), field snapshots and simulation_summary.txt. out_dir (NULL: the directory of
 * output.filename) receives <variable>_<time>.txt files (HDF5 is not available here) with one
 * value per interior cell, x fastest then y then z. */
int frehg_run_config(frehg_ctx_t ctx, const char* yaml_path, const char* const* overrides, int n_overrides,
                     const char* out_dir, frehg_run_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* FREHG_GPU_H */
