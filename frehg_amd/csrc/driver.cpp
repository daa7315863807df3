// Runs a Frehg2 v2 configuration through the C ABI: the set-up of Orchestrator::initialize and
// the loop of Orchestrator::runLoopSequential (src/driver/Orchestrator.cpp:192-304, 1322-1353,
// 1518-1548), with text snapshots in place of the HDF5 writer and the run summary of
// Orchestrator::writeSummary (:1622-1680).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "config.hpp"
#include "frehg_gpu.h"

namespace frehg {
// api.hip: the context's error message, and whether it already holds a solver
int set_ctx_error(frehg_ctx_t ctx, int rc, const char* msg);
bool ctx_has_solvers(frehg_ctx_t ctx);
}  // namespace frehg

namespace {

using frehg::Config;
using frehg::ConfigError;
using frehg::RunSetup;

void copy_msg(char* dst, int len, const std::string& msg) {
  if (!dst || len <= 0) return;
  const size_t n = std::min((size_t)len - 1, msg.size());
  std::memcpy(dst, msg.data(), n);
  dst[n] = '\0';
}

Config load(const char* path, const char* const* overrides, int n, bool validate = true) {
  if (!path || !*path) throw ConfigError("configuration path is empty");
  Config c = Config::from_file(path, validate);
  for (int i = 0; i < n; ++i) {
    if (!overrides || !overrides[i]) throw ConfigError("null override");
    c.apply_override(overrides[i]);
  }
  return c;
}

// One failed ABI call ends the run with that call's message.
struct RunFailure {
  int rc;
  std::string msg;
};

void check(frehg_ctx_t ctx, int rc, const char* what) {
  if (rc < 0) throw RunFailure{rc, std::string(what) + ": " + frehg_last_error(ctx)};
}

void make_dirs(const std::string& dir) {
  std::string cur;
  for (size_t i = 0; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      if (!cur.empty() && cur != ".") {
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
          throw RunFailure{FREHG_ERR_INVALID_ARG, "cannot create output directory '" + cur + "'"};
      }
    }
    if (i < dir.size()) cur.push_back(dir[i]);
  }
}

// timeKey (src/io/Hdf5Writer.cpp:25-33): an integer when t is within 1e-9 of one, else 12
// significant digits.
std::string time_tag(double t) {
  const long long r = std::llround(t);
  if (std::fabs(t - (double)r) < 1.0e-9) return std::to_string(r);
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.12g", t);
  return buf;
}

class Runner {
 public:
  Runner(frehg_ctx_t ctx, const RunSetup& s, std::string out_dir) : ctx_(ctx), s_(s), dir_(std::move(out_dir)) {}

  void setup() {
    const RunSetup& s = s_;
    if (s.sw) {
      check(ctx_, frehg_swe_init(ctx_, s.nx, s.ny, s.dx, s.dy, &s.swe), "frehg_swe_init");
      check(ctx_, frehg_swe_set_bathymetry(ctx_, s.bed.data()), "frehg_swe_set_bathymetry");
    }
    if (s.gw) {
      check(ctx_, frehg_re_init(ctx_, s.nx, s.ny, s.nz, &s.re), "frehg_re_init");
      if (!s.class_idx.empty())
        check(ctx_,
              frehg_re_set_soil(ctx_, s.classes.data(), (int)s.classes.size(), s.class_idx.data(),
                                s.class_idx_3d ? 1 : 0, nullptr),
              "frehg_re_set_soil");
      if (!s.top_flux.empty()) check(ctx_, frehg_re_set_top_flux(ctx_, s.top_flux.data()), "frehg_re_set_top_flux");
    }
    if (!s.regions.empty()) {
      std::vector<int32_t> kind, nvert;
      std::vector<double> rate, xy;
      for (const auto& r : s.regions) {
        kind.push_back(r.kind);
        rate.push_back(r.rate);
        nvert.push_back((int32_t)(r.xy.size() / 2));
        xy.insert(xy.end(), r.xy.begin(), r.xy.end());
      }
      check(ctx_,
            frehg_swe_set_boundary_polygons(ctx_, (int)kind.size(), kind.data(), rate.data(), nvert.data(), xy.data(),
                                            s.x0, s.y0),
            "frehg_swe_set_boundary_polygons");
    }
    // initial conditions (ICApply.cpp:187-233)
    if (s.sw) check(ctx_, frehg_swe_set_initial_eta(ctx_, s.eta0.data()), "frehg_swe_set_initial_eta");
    if (s.gw)
      check(ctx_,
            frehg_re_set_initial(ctx_, s.gw_init_head ? s.gw_init.data() : nullptr,
                                 s.gw_init_head ? nullptr : s.gw_init.data()),
            "frehg_re_set_initial");
    if (s.solute) {  // buildSolute (:749-783) and the solute initial conditions
      check(ctx_, frehg_solute_init(ctx_, &s.sol), "frehg_solute_init");
      check(ctx_,
            frehg_solute_set_initial(ctx_, s.sol_surface.empty() ? nullptr : s.sol_surface.data(),
                                     s.sol_subsurface.empty() ? nullptr : s.sol_subsurface.data()),
            "frehg_solute_set_initial");
    }
  }

  double solute_mass() {
    double m = 0.0;
    check(ctx_, frehg_solute_mass(ctx_, &m, nullptr), "frehg_solute_mass");
    return m;
  }

  double volume() {
    double sw = 0.0, gw = 0.0;
    check(ctx_, frehg_water_volume(ctx_, &sw, &gw), "frehg_water_volume");
    return sw + gw;
  }

  void write_outputs(double t) {
    ++outputs_;
    last_out_ = t;
    if (!s_.write_fields) return;
    const RunSetup& s = s_;
    const size_t nsx = (size_t)s.nx + 2, nsy = (size_t)s.ny + 2;
    if (s.sw && !s.sw_out.empty()) {
      std::vector<double> eta(nsx * nsy), bot(nsx * nsy), f(nsx * nsy);
      check(ctx_, frehg_swe_get_field(ctx_, "eta", eta.data()), "frehg_swe_get_field");
      check(ctx_, frehg_swe_get_field(ctx_, "bottom", bot.data()), "frehg_swe_get_field");
      for (const std::string& v : s.sw_out) {
        std::vector<double> out((size_t)s.nx * s.ny);
        if (v == "velocity_u" || v == "velocity_v")
          check(ctx_, frehg_swe_get_field(ctx_, v == "velocity_u" ? "uu" : "vv", f.data()), "frehg_swe_get_field");
        if (v == "concentration") check(ctx_, frehg_solute_get_field(ctx_, "surface", f.data()), "frehg_solute_get_field");
        for (int j = 0; j < s.ny; ++j)
          for (int i = 0; i < s.nx; ++i) {
            const size_t c = (size_t)(i + 1) + (size_t)(j + 1) * nsx;
            double x = 0.0;
            if (v == "water_depth") x = std::max(eta[c] - bot[c], 0.0);  // io_pack_depth (:1366-1373)
            else if (v == "eta") x = eta[c];
            else x = f[c];
            out[(size_t)i + (size_t)j * s.nx] = x;
          }
        // the surface and subsurface concentrations sit in separate files (two HDF5 groups in
        // the reference's writer)
        write_file(v == "concentration" ? "surface_concentration" : v, t, out);
      }
    }
    if (s.gw && !s.gw_out.empty()) {
      std::vector<double> f(nsx * nsy * ((size_t)s.nz + 2));
      for (const std::string& v : s.gw_out) {
        const char* field = v == "head" ? "h" : v == "moisture" ? "wc" : v.c_str();
        if (v == "concentration")
          check(ctx_, frehg_solute_get_field(ctx_, "subsurface", f.data()), "frehg_solute_get_field");
        else
          check(ctx_, frehg_re_get_field(ctx_, field, f.data()), "frehg_re_get_field");
        std::vector<double> out((size_t)s.nx * s.ny * s.nz);
        for (int k = 0; k < s.nz; ++k)
          for (int j = 0; j < s.ny; ++j)
            for (int i = 0; i < s.nx; ++i)
              out[(size_t)i + (size_t)j * s.nx + (size_t)k * s.nx * s.ny] =
                  f[(size_t)(i + 1) + (size_t)(j + 1) * nsx + (size_t)(k + 1) * nsx * nsy];
        write_file(v == "concentration" ? "subsurface_concentration" : v, t, out);
      }
    }
  }

  void write_file(const std::string& var, double t, const std::vector<double>& v) {
    const std::string path = dir_ + "/" + var + "_" + time_tag(t) + ".txt";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw RunFailure{FREHG_ERR_INVALID_ARG, "cannot write '" + path + "'"};
    for (double x : v) std::fprintf(f, "%.17g\n", x);
    std::fclose(f);
  }

  // Orchestrator::runLoopSequential with step() (:1322-1353).
  void run(frehg_run_stats* st) {
    const RunSetup& s = s_;
    const auto t0 = std::chrono::steady_clock::now();
    make_dirs(dir_);
    setup();
    const double eps = 1.0e-9 * std::max(1.0, s.t_end);
    auto next_after = [](double interval, double t) {
      if (interval <= 0.0) return std::numeric_limits<double>::max();
      return (std::floor(t / interval + 1.0e-9) + 1.0) * interval;
    };
    double t = 0.0, rain_in = 0.0, outflow = 0.0, exch = 0.0;
    double next_out = s.output_interval > 0.0 ? s.output_interval : s.t_end + 1.0;
    double next_ckpt = s.dt_checkpoint > 0.0 ? s.dt_checkpoint : s.t_end + 1.0;
    double water0 = 0.0;
    if (s.restart) {
      // loadCheckpointState (:688-747): the state, the clocks and the next groundwater dt; no
      // snapshot of the starting state
      frehg_checkpoint_info info;
      check(ctx_, frehg_checkpoint_load(ctx_, s.restart_file.c_str(), &info), "frehg_checkpoint_load");
      if (s.restart_time_given && std::fabs(s.restart_time - info.t) > 1.0e-9 * std::max(1.0, std::fabs(info.t)))
        throw RunFailure{FREHG_ERR_INVALID_ARG, "initial_conditions.time (" + time_tag(s.restart_time) +
                                                    ") differs from the checkpoint's time (" + time_tag(info.t) + ")"};
      t = info.t;
      next_out = next_after(s.output_interval, t);
      next_ckpt = next_after(s.dt_checkpoint, t);
      water0 = volume();
    } else {
      water0 = volume();
      write_outputs(0.0);
    }
    // runLoopSequential (:1521): the solute mass after the initial conditions or the restart
    const double sol0 = s.solute ? solute_mass() : 0.0;
    long long steps = 0, subs = 0;
    int infos = 0;
    const int rain_rows = s.ny - 1;  // the surface step adds rain to every row but the last (:1226-1233)
    while (t < s.t_end - eps && (s.max_steps <= 0 || steps < s.max_steps)) {
      if (s.coupled) {
        const double t_new = t + s.dt, rain = frehg::rain_at(s, t_new);
        frehg_coupled_stats cs;
        const int rc = frehg_coupled_step(ctx_, &s.cp, s.dt, rain, s.evap_rate, s.t_end, &cs);
        check(ctx_, rc, "frehg_coupled_step");
        if (rc > 0) ++infos;
        subs += cs.gw_substeps;
        exch += cs.exchange_volume;
        outflow += cs.boundary_outflow;
        if (rain != 0.0) rain_in += rain * s.dt * s.dx * s.dy * (double)s.nx * rain_rows;
        t = t_new;
      } else if (s.sw) {
        const double t_new = t + s.dt, rain = frehg::rain_at(s, t_new);
        frehg_solve_stats ss;
        check(ctx_, frehg_swe_step(ctx_, rain, s.evap_rate, &ss), "frehg_swe_step");
        if (!s.regions.empty()) {
          double out = 0.0;
          check(ctx_, frehg_swe_apply_boundaries(ctx_, s.dt, &out), "frehg_swe_apply_boundaries");
          outflow += out;
        }
        if (rain != 0.0) rain_in += rain * s.dt * s.dx * s.dy * (double)s.nx * rain_rows;
        if (s.solute) check(ctx_, frehg_solute_apply(ctx_, t, s.dt, rain), "frehg_solute_apply");
        t = t_new;
      } else {
        // groundwater only: the Richards solver adapts its own dt (:1344-1352)
        frehg_re_stats rs;
        const int rc = frehg_re_step(ctx_, &rs);
        check(ctx_, rc, "frehg_re_step");
        if (rc > 0) ++infos;
        if (s.solute) check(ctx_, frehg_solute_apply(ctx_, t, rs.dt_used, 0.0), "frehg_solute_apply");
        t += rs.dt_used;
        ++subs;
      }
      ++steps;
      if (t >= next_out - eps) {
        write_outputs(t);
        next_out = next_after(s.output_interval, t);
      }
      if (s.dt_checkpoint > 0.0 && t >= next_ckpt - eps) {
        write_checkpoint(t);
        next_ckpt = next_after(s.dt_checkpoint, t);
      }
    }
    if (last_out_ < t - eps) write_outputs(t);
    const double water1 = volume();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    frehg_run_stats out{};
    out.t = t;
    out.steps = steps;
    out.gw_substeps = subs;
    out.outputs = outputs_;
    out.infos = infos;
    out.water_initial = water0;
    out.water_final = water1;
    out.rain_volume = rain_in;
    out.boundary_outflow = outflow;
    out.exchange_volume = exch;
    out.wall_seconds = wall;
    if (s.solute) {
      frehg_solute_stats ss;
      double m1 = 0.0;
      check(ctx_, frehg_solute_mass(ctx_, &m1, &ss), "frehg_solute_mass");
      out.solute_initial_mass = sol0;
      out.solute_final_mass = m1;
      out.solute_added_by_rain = ss.rain_added;
      out.solute_decay_loss = ss.decay_loss;
      const double expected = sol0 + ss.rain_added - ss.decay_loss;
      out.solute_relative_mass_error = std::fabs(m1 - expected) / std::max(std::fabs(sol0 + ss.rain_added), 1.0e-300);
    }
    write_summary(out);
    if (st) *st = out;
  }

  // writeCheckpointNow + Hdf5Writer::writeCheckpoint (:1461-1502, Hdf5Writer.cpp:400-480):
  // <stem>.ckpt.<time>.bin, the newest max_checkpoints kept. Like the reference, only a run that
  // writes fields (output.filename set) writes checkpoints.
  void write_checkpoint(double t) {
    if (!s_.write_fields) return;
    std::string stem = s_.stem;
    const size_t slash = stem.find_last_of('/');
    const std::string base = slash == std::string::npos ? stem : stem.substr(slash + 1);
    stem = dir_ + "/" + (base.empty() ? s_.sim_id : base);
    const std::string path = stem + ".ckpt." + time_tag(t) + ".bin";
    check(ctx_, frehg_checkpoint_save(ctx_, path.c_str()), "frehg_checkpoint_save");
    ckpts_.push_back(path);
    while ((int)ckpts_.size() > s_.max_checkpoints) {
      std::remove(ckpts_.front().c_str());
      ckpts_.erase(ckpts_.begin());
    }
  }

  void write_summary(const frehg_run_stats& r) {
    const RunSetup& s = s_;
    const std::string path = dir_ + "/simulation_summary.txt";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw RunFailure{FREHG_ERR_INVALID_ARG, "cannot write '" + path + "'"};
    auto b = [](bool x) { return x ? "true" : "false"; };
    std::fprintf(f, "frehg2_version \"frehg_amd (ABI %d)\"\n", FREHG_ABI_VERSION);
    std::fprintf(f, "simulation_id \"%s\"\n", s.sim_id.c_str());
    std::fprintf(f, "modules {surface_water=%s, groundwater=%s, solute=%s}\n", b(s.sw), b(s.gw), b(s.solute));
    std::fprintf(f, "mpi_ranks 1\n");
    std::fprintf(f, "execution_space \"HIP (one GPU)\"\n");
    std::fprintf(f, "total_runtime_seconds %.9g\n", r.wall_seconds);
    std::fprintf(f, "simulated_time_seconds %.17g\n", r.t);
    std::fprintf(f, "time_steps_completed %lld\n", (long long)r.steps);
    std::fprintf(f, "gw_substeps_completed %lld\n", (long long)r.gw_substeps);
    std::fprintf(f, "output_intervals_completed %d\n", r.outputs);
    std::fprintf(f, "water_initial_volume %.17g\n", r.water_initial);
    std::fprintf(f, "water_final_volume %.17g\n", r.water_final);
    std::fprintf(f, "water_rain_volume_in %.17g\n", r.rain_volume);
    std::fprintf(f, "water_polygon_inflow_volume 0\n");
    std::fprintf(f, "water_polygon_outflow_volume %.17g\n", r.boundary_outflow);
    std::fprintf(f, "water_polygon_well_volume 0\n");
    std::fprintf(f, "coupling_exchange_volume %.17g\n", r.exchange_volume);
    std::fprintf(f, "groundwater_present %s\n", b(s.gw));
    std::fprintf(f, "surface_water_present %s\n", b(s.sw));
    if (s.solute) {
      std::fprintf(f, "solute_initial_mass %.17g\n", r.solute_initial_mass);
      std::fprintf(f, "solute_final_mass %.17g\n", r.solute_final_mass);
      std::fprintf(f, "solute_added_by_rain %.17g\n", r.solute_added_by_rain);
      std::fprintf(f, "solute_decay_loss %.17g\n", r.solute_decay_loss);
      std::fprintf(f, "solute_relative_mass_error %.17g\n", r.solute_relative_mass_error);
    }
    std::fprintf(f, "wall_clock_hours %.9g\n", r.wall_seconds / 3600.0);
    std::fclose(f);
  }

 private:
  frehg_ctx_t ctx_;
  const RunSetup& s_;
  std::string dir_;
  int outputs_ = 0;
  double last_out_ = -1.0;
  std::vector<std::string> ckpts_;
};

void fill_setup(const RunSetup& s, frehg_run_setup* o) {
  std::memset(o, 0, sizeof(*o));
  o->surface_water = s.sw;
  o->groundwater = s.gw;
  o->solute = s.solute;
  o->solute_params = s.sol;
  o->coupled = s.coupled;
  o->nx = s.nx;
  o->ny = s.ny;
  o->nz = s.nz;
  o->dx = s.dx;
  o->dy = s.dy;
  o->dz = s.dz;
  o->botz = s.botz;
  o->dt = s.dt;
  o->t_end = s.t_end;
  o->output_interval = s.output_interval;
  o->max_steps = s.max_steps;
  o->evaporation_rate = s.evap_rate;
  o->rain_from_file = s.rain_from_file;
  o->rain_samples = (int32_t)s.rain_t.size();
  o->rain_rate = s.rain_rate;
  o->swe = s.swe;
  o->re = s.re;
  o->coupling = s.cp;
  o->soil_classes = (int32_t)s.classes.size();
  o->soil_map = s.class_idx.empty() ? 0 : (s.class_idx_3d ? 3 : 2);
  o->recharge = s.top_flux.empty() ? 0 : 1;
  o->boundary_regions = (int32_t)s.regions.size();
  o->gw_initial_head = s.gw_init_head;
  o->restart = s.restart;
  o->dt_checkpoint = s.dt_checkpoint;
  o->max_checkpoints = s.max_checkpoints;
}

}  // namespace

extern "C" {

int frehg_config_setup(const char* yaml_path, const char* const* overrides, int n_overrides, frehg_run_setup* out,
                       char* err, int err_len) {
  try {
    if (!out) throw ConfigError("frehg_config_setup: null output");
    const Config c = load(yaml_path, overrides, n_overrides);
    fill_setup(frehg::build_setup(c), out);
    copy_msg(err, err_len, "");
    return FREHG_OK;
  } catch (const std::exception& e) {
    copy_msg(err, err_len, e.what());
    return FREHG_ERR_INVALID_ARG;
  }
}

int frehg_config_query(const char* yaml_path, const char* const* overrides, int n_overrides, const char* key,
                       char* value, int value_len) {
  try {
    if (!key) throw ConfigError("frehg_config_query: null key");
    const Config c = load(yaml_path, overrides, n_overrides, false);
    const frehg::yaml::Node* n = c.find(key);
    if (!n) {
      copy_msg(value, value_len, "");
      return 0;
    }
    if (n->is_scalar()) {
      copy_msg(value, value_len, n->text);
      return 1;
    }
    copy_msg(value, value_len, std::to_string(n->size()));
    return n->is_seq() ? 2 : 3;
  } catch (const std::exception& e) {
    copy_msg(value, value_len, e.what());
    return FREHG_ERR_INVALID_ARG;
  }
}

int64_t frehg_config_field(const char* yaml_path, const char* const* overrides, int n_overrides, const char* name,
                           double* out, int64_t n, char* err, int err_len) {
  try {
    if (!name) throw ConfigError("frehg_config_field: null name");
    const Config c = load(yaml_path, overrides, n_overrides);
    const RunSetup s = frehg::build_setup(c);
    std::vector<double> v;
    const std::string f = name;
    if (f == "bathymetry") v = s.bed;
    else if (f == "eta0") v = s.eta0;
    else if (f == "gw_initial") v = s.gw_init;
    else if (f == "top_flux") v = s.top_flux;
    else if (f == "soil_class") v.assign(s.class_idx.begin(), s.class_idx.end());
    else if (f == "rain_t") v = s.rain_t;
    else if (f == "rain_v") v = s.rain_v;
    else if (f == "solute_surface") v = s.sol_surface;
    else if (f == "solute_subsurface") v = s.sol_subsurface;
    else throw ConfigError("frehg_config_field: unknown field '" + f + "'");
    if (out && n > 0) std::copy(v.begin(), v.begin() + (long)std::min<int64_t>(n, (int64_t)v.size()), out);
    copy_msg(err, err_len, "");
    return (int64_t)v.size();
  } catch (const std::exception& e) {
    copy_msg(err, err_len, e.what());
    return FREHG_ERR_INVALID_ARG;
  }
}

int frehg_run_config(frehg_ctx_t ctx, const char* yaml_path, const char* const* overrides, int n_overrides,
                     const char* out_dir, frehg_run_stats* stats) {
  if (!ctx) return FREHG_ERR_INVALID_ARG;
  try {
    if (frehg::ctx_has_solvers(ctx))
      return frehg::set_ctx_error(ctx, FREHG_ERR_STATE, "frehg_run_config: the context already holds a solver");
    const Config c = load(yaml_path, overrides, n_overrides);
    const RunSetup s = frehg::build_setup(c);
    RunSetup sd = s;
    if (out_dir && *out_dir) {
      sd.out_dir = out_dir;
      sd.write_fields = true;  // an explicit directory asks for the snapshots
    }
    Runner r(ctx, sd, sd.out_dir);
    r.run(stats);
    return FREHG_OK;
  } catch (const RunFailure& f) {
    return frehg::set_ctx_error(ctx, f.rc, f.msg.c_str());
  } catch (const ConfigError& e) {
    return frehg::set_ctx_error(ctx, FREHG_ERR_INVALID_ARG, e.what());
  } catch (const std::exception& e) {
    return frehg::set_ctx_error(ctx, FREHG_ERR_INVALID_ARG, e.what());
  }
}

}  // extern "C"
