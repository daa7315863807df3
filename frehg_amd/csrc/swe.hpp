// Device-resident semi-implicit shallow-water step (replaces src/swe/SweSolver.{hpp,cpp}).
#pragma once

#include <map>
#include <memory>
#include <string>

#include "../../include/frehg_gpu.h"
#include "common.hpp"
#include "krylov.hpp"

namespace frehg {

// One surface boundary region's rule (PolygonBC, src/bc/PolygonRegion.hpp:20-35): kind
// FREHG_BC_*, rate (Q m^3/s or h m), and the region's cell count.
struct SurfaceBcRule {
  int kind;
  int count;
  double rate;
};

class SweState {
 public:
  SweState(int nx, int ny, double dx, double dy, const frehg_swe_params& p, hipStream_t s, Profiler* prof);
  ~SweState();

  void set_params(const frehg_swe_params& p);
  void set_bathymetry(const double* bed_interior);
  void initialize_state(double init_eta);
  // setInitialEtaPhysical + setInitialVelocityConstant(0, 0) + finalizeInitialState
  // (src/swe/SweSolver.cpp:855-924); eta_phys: nx*ny values (i + j*nx), or null for init_eta everywhere.
  void initialize_eta(const double* eta_phys, double init_eta);
  // fields in the reference's packed halo-padded layout ((nx + 2) x (ny + 2))
  void set_field(const std::string& name, const double* src);
  void get_field(const std::string& name, double* dst);
  void step(double rain, double evap, frehg_solve_stats* st);
  double max_cfl();
  // PolygonBC regions (src/bc/PolygonBC.hpp:28-46) built by PolygonIndex::build's rule from
  // polygons; applied after each surface step of a coupled step (and on request).
  void set_boundary_polygons(int nreg, const int32_t* kind, const double* rate, const int32_t* nvert,
                             const double* xy, double x0, double y0);
  bool has_boundaries() const { return nbc_ > 0; }
  // PolygonBC::applySurface over dt, queued on the stream; boundary_outflow() waits and returns
  // the net outflow volume [m^3] of the last application.
  void apply_boundaries_async(double dt);
  double boundary_outflow();

  const Geo& geo() const { return g_; }
  double dx() const { return dx_; }
  double dy() const { return dy_; }
  const frehg_swe_params& params() const { return p_; }
  double* d(const std::string& name) { return field(name); }

 private:
  double* field(const std::string& name);
  // whole device-layout arrays (nsurf() doubles, indexed by sidx)
  void set_raw(const std::string& name, const double* src);
  void get_raw(const std::string& name, double* dst);
  void enq_pre(double evap);
  void enq_post(double rain, double evap);
  void copy_field(const double* src, double* dst);

  Geo g_;  // nz = 1; 2D index = sidx
  double dx_, dy_;
  frehg_swe_params p_;
  hipStream_t s_;
  Profiler* prof_;
  long long n_ = 0;
  std::map<std::string, double*> f_;
  double *D_ = nullptr, *ax_ = nullptr, *ay_ = nullptr, *b_ = nullptr;
  unsigned char* fixed_ = nullptr;
  double* red_ = nullptr;
  double* dres_ = nullptr;  // [0..1] max CFL, [3] boundary outflow
  double* hres_ = nullptr;
  int nbc_ = 0;  // surface boundary regions
  int32_t* bcidx_ = nullptr;
  SurfaceBcRule* bcrules_ = nullptr;
  double* bcpart_ = nullptr;
  unsigned int* dcount_ = nullptr;
  std::unique_ptr<Krylov> kry_;
  // captured step graphs (before / after the surface solve) and what they were captured for
  bool use_graph_ = false;  // FREHG_SW_GRAPH=1: replay the step from captured graphs
  bool solved_ = false;  // the first step runs eagerly and sets the iteration estimate
  int budget_ = 0, shrink_ = 0;
  hipGraphExec_t graph_pre_ = nullptr, graph_post_ = nullptr;
  double graph_rain_ = 0.0, graph_evap_ = 0.0;
  int graph_budget_ = 0;
  frehg_swe_params graph_params_{};
};

}  // namespace frehg
