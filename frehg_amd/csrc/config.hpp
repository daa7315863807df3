// Configuration front end: the Frehg2 v2 YAML schema mapped onto this backend's parameters.
//
// Config mirrors src/io/Config.cpp (dotted-path lookups where a missing or null key takes the
// default, paths resolved against the file's directory, the structural check of
// Config::validateSchema :97-143). build_setup mirrors Orchestrator::initialize
// (src/driver/Orchestrator.cpp:192-304): the v2 schema gate (:85-103), the capability gate
// (:111-173), grid and time (:238-261), rainfall (:313-336), surface water and bathymetry
// (:338-440), groundwater, soil classes and soil map (:442-591, src/soil/SoilConfig.cpp:10-43),
// top recharge (:593-628), boundary polygons (:630-656, src/bc/PolygonConfig.cpp:40-83), initial
// conditions (:658-686, src/ic/ICConfig.cpp:60-143, src/ic/ICApply.cpp:187-233) and the output
// selection (:1077-1175). Options this backend does not implement are refused, as the reference refuses
// the options its solvers do not implement.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "frehg_gpu.h"
#include "yaml.hpp"

namespace frehg {

class ConfigError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Config {
 public:
  // `validate`: apply Config::validateSchema's structural check (off for plain YAML queries).
  static Config from_file(const std::string& path, bool validate = true);
  static Config from_string(const std::string& text, const std::string& dir, const std::string& source = "<string>");

  // Node at a dotted path; null when a key is missing or its value is null.
  const yaml::Node* find(const std::string& dotted) const;
  bool has(const std::string& dotted) const { return find(dotted) != nullptr; }
  double get_double(const std::string& key, double def) const;
  int get_int(const std::string& key, int def) const;
  long long get_long(const std::string& key, long long def) const;
  bool get_bool(const std::string& key, bool def) const;
  std::string get_string(const std::string& key, const std::string& def) const;
  // Paths relative to the configuration file's directory (absolute paths unchanged).
  std::string resolve_path(const std::string& rel) const;
  // "dotted.key=value": the value is parsed as a YAML value and replaces (or adds) the key.
  void apply_override(const std::string& assignment);
  void validate_structure() const;

  const yaml::Node& root() const { return *root_; }
  const std::string& source() const { return source_; }

 private:
  std::shared_ptr<yaml::Node> root_;
  std::string source_, dir_;
  bool validate_ = true;
};

struct SurfaceRegion {
  int kind = FREHG_BC_CRITICAL;  // FREHG_BC_*
  double rate = 0.0;
  std::vector<double> xy;  // x0, y0, x1, y1, ...
};

// Everything a run needs, resolved from the configuration and its data files.
struct RunSetup {
  std::string sim_id, mode, source;
  bool sw = false, gw = false, solute = false, coupled = false;
  int nx = 0, ny = 0, nz = 0;
  double dx = 1, dy = 1, dz = 1, botz = 0, x0 = 0, y0 = 0;
  double dt = 0, t_end = 0, output_interval = 0;
  long long max_steps = 0;
  double evap_rate = 0;
  bool rain_from_file = false;
  double rain_rate = 0;
  std::vector<double> rain_t, rain_v;
  frehg_swe_params swe{};
  std::vector<double> bed;  // interior bathymetry plus the datum offset (nx * ny)
  frehg_re_params re{};
  std::vector<frehg_soil_params> classes;
  std::vector<int32_t> class_idx;  // empty: uniform soil
  bool class_idx_3d = false;
  std::vector<double> top_flux;  // empty: none (nx * ny)
  frehg_coupling_params cp{};
  std::vector<SurfaceRegion> regions;
  std::vector<double> eta0;  // physical initial eta per interior cell (surface water)
  bool gw_init_head = false;
  std::vector<double> gw_init;  // initial head or water content per interior cell
  frehg_solute_params sol{};    // the `solute:` block (used when `solute` is set)
  std::vector<double> sol_surface, sol_subsurface;  // initial concentrations per interior cell
  // "water_depth", "eta", "velocity_u", "velocity_v", "concentration"; "head", "moisture", "qx"...,
  // "concentration"
  std::vector<std::string> sw_out, gw_out;
  std::string out_dir = ".", summary_path = "./simulation_summary.txt";
  std::string stem;  // output.filename without its extension ("" without a filename)
  bool write_fields = false;
  double dt_checkpoint = 0.0;  // time.dt_checkpoint (0: none)
  int max_checkpoints = 2;     // time.max_checkpoints: the newest files kept
  bool restart = false;        // initial_conditions.type = restart
  std::string restart_file;
  bool restart_time_given = false;
  double restart_time = 0.0;
};

RunSetup build_setup(const Config& cfg);

// Rain rate at time t: piecewise-linear series (Orchestrator.cpp:57-71) or the constant rate.
double rain_at(const RunSetup& s, double t);

}  // namespace frehg
