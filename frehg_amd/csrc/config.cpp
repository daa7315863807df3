// Configuration front end (see config.hpp).
#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>

namespace frehg {

namespace {

std::vector<std::string> split_dots(const std::string& path) {
  std::vector<std::string> keys;
  std::string cur;
  for (char c : path) {
    if (c == '.') {
      keys.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  keys.push_back(cur);
  return keys;
}

std::string dir_of(const std::string& path) {
  const auto slash = path.find_last_of('/');
  return slash == std::string::npos ? "" : path.substr(0, slash);
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool ends_with(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

std::vector<double> read_doubles(const std::string& path, const char* what) {
  std::ifstream in(path);
  if (!in) throw ConfigError(std::string(what) + ": cannot open data file '" + path + "'");
  std::vector<double> v;
  double x;
  while (in >> x) v.push_back(x);
  if (!in.eof()) throw ConfigError(std::string(what) + ": '" + path + "' holds a value that is not a number");
  return v;
}

// ESRI ASCII grid (src/io/AsciiRaster.cpp:20-91): header keys in any order and case, the centre
// forms shifted to corners, an optional NODATA_value after the mandatory keys, then nrows rows
// of ncols values, north row first.
struct AsciiGrid {
  int ncols = 0, nrows = 0;
  double nodata = -9999.0;
  std::vector<double> v;
  bool is_nodata(double x) const {
    const double tol = std::max(1e-6, std::fabs(nodata) * 1e-9);
    return std::fabs(x - nodata) <= tol;
  }
};

AsciiGrid read_ascii_grid(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw ConfigError("AsciiRaster: cannot open file '" + path + "'");
  AsciiGrid r;
  bool nc = false, nr = false, hx = false, hy = false, hc = false;
  double dummy = 0.0;
  std::string key;
  while (!(nc && nr && hx && hy && hc)) {
    if (!(in >> key)) throw ConfigError("AsciiRaster: incomplete header in '" + path + "'");
    const std::string k = lower(key);
    if (k == "ncols") {
      in >> r.ncols;
      nc = true;
    } else if (k == "nrows") {
      in >> r.nrows;
      nr = true;
    } else if (k == "xllcorner" || k == "xllcenter") {
      in >> dummy;
      hx = true;
    } else if (k == "yllcorner" || k == "yllcenter") {
      in >> dummy;
      hy = true;
    } else if (k == "cellsize") {
      in >> dummy;
      hc = true;
    } else if (k == "nodata_value") {
      in >> r.nodata;
    } else {
      throw ConfigError("AsciiRaster: unknown header key '" + key + "' in '" + path + "'");
    }
  }
  const std::streampos pos = in.tellg();
  if (in >> key && lower(key) == "nodata_value") {
    in >> r.nodata;
  } else {
    in.clear();
    in.seekg(pos);
  }
  if (r.ncols <= 0 || r.nrows <= 0) throw ConfigError("AsciiRaster: ncols/nrows must be positive in '" + path + "'");
  r.v.resize((size_t)r.ncols * r.nrows);
  for (double& x : r.v)
    if (!(in >> x))
      throw ConfigError("AsciiRaster: not enough data values in '" + path + "' (expected " +
                        std::to_string(r.v.size()) + ")");
  return r;
}

// Model order (i + j nx, j northward) of a raster that has to cover the nx x ny domain.
std::vector<double> grid_to_model(const AsciiGrid& r, int nx, int ny, const std::string& path, const char* what) {
  if (r.ncols != nx || r.nrows != ny)
    throw ConfigError(std::string(what) + ": raster '" + path + "' is " + std::to_string(r.ncols) + "x" +
                      std::to_string(r.nrows) + " but domain is nx=" + std::to_string(nx) + ", ny=" + std::to_string(ny));
  std::vector<double> out((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) out[(size_t)i + (size_t)j * nx] = r.v[(size_t)i + (size_t)(ny - 1 - j) * nx];
  return out;
}

bool raster_format(const std::string& fmt, const std::string& file) {
  if (fmt == "raster" || fmt == "esri" || fmt == "ascii_raster") return true;
  return fmt == "auto" && ends_with(file, ".asc");
}

[[noreturn]] void reject(const std::string& src, const std::string& key, const std::string& why) {
  throw ConfigError("unsupported configuration: '" + key + "' " + why + " (in " + src +
                    "). Disable it or remove the key.");
}

int ksp_of(const std::string& s, const std::string& key, const std::string& src) {
  const std::string k = lower(s);
  if (k == "cg") return FREHG_KSP_CG;
  if (k == "bicgstab" || k == "bcgs") return FREHG_KSP_BICGSTAB;
  reject(src, key, "= '" + s + "': this backend solves with cg or bicgstab");
}

// The preconditioner is diagonal (Jacobi) whatever the configuration names: the solve still
// stops at the configured relative residual, which is what the solution depends on.
void check_pc(const std::string& s, const std::string& key, const std::string& src) {
  const std::string k = lower(s);
  if (k == "jacobi" || k == "sor" || k == "bjacobi" || k == "ilu" || k == "none") return;
  reject(src, key, "= '" + s + "': this backend preconditions with Jacobi (jacobi, sor, bjacobi, ilu accepted)");
}

}  // namespace

// ---- Config ----------------------------------------------------------------------------------
Config Config::from_file(const std::string& path, bool validate) {
  Config c;
  c.source_ = path;
  c.dir_ = dir_of(path);
  c.validate_ = validate;
  try {
    c.root_ = yaml::load_file(path);
  } catch (const yaml::Error& e) {
    throw ConfigError("failed to parse YAML file '" + path + "': " + e.what());
  }
  if (validate) c.validate_structure();
  return c;
}

Config Config::from_string(const std::string& text, const std::string& dir, const std::string& source) {
  Config c;
  c.source_ = source;
  c.dir_ = dir;
  try {
    c.root_ = yaml::parse(text, source);
  } catch (const yaml::Error& e) {
    throw ConfigError(std::string("failed to parse YAML string: ") + e.what());
  }
  c.validate_structure();
  return c;
}

const yaml::Node* Config::find(const std::string& dotted) const {
  const yaml::Node* n = root_.get();
  for (const std::string& k : split_dots(dotted)) {
    if (!n || !n->is_map()) return nullptr;
    n = n->get(k);
    if (!n || n->is_null()) return nullptr;
  }
  return n;
}

namespace {
template <class F>
auto convert(const Config& c, const std::string& key, const char* type, F&& f) -> decltype(f(*c.find(key))) {
  const yaml::Node* n = c.find(key);
  try {
    return f(*n);
  } catch (const yaml::Error&) {
    throw ConfigError("config: '" + key + "' must be " + type + " (in " + c.source() + ")");
  }
}
}  // namespace

double Config::get_double(const std::string& key, double def) const {
  if (!has(key)) return def;
  return convert(*this, key, "a number", [](const yaml::Node& n) { return n.as_double(); });
}
int Config::get_int(const std::string& key, int def) const {
  if (!has(key)) return def;
  return convert(*this, key, "an integer", [](const yaml::Node& n) { return n.as_int(); });
}
long long Config::get_long(const std::string& key, long long def) const {
  if (!has(key)) return def;
  return convert(*this, key, "an integer", [](const yaml::Node& n) { return n.as_long(); });
}
bool Config::get_bool(const std::string& key, bool def) const {
  if (!has(key)) return def;
  return convert(*this, key, "a boolean", [](const yaml::Node& n) { return n.as_bool(); });
}
std::string Config::get_string(const std::string& key, const std::string& def) const {
  if (!has(key)) return def;
  return convert(*this, key, "a string", [](const yaml::Node& n) { return n.as_string(); });
}

std::string Config::resolve_path(const std::string& rel) const {
  if (rel.empty() || rel[0] == '/') return rel;
  if (dir_.empty()) return rel;
  return dir_ + "/" + rel;
}

void Config::apply_override(const std::string& assignment) {
  const size_t eq = assignment.find('=');
  if (eq == std::string::npos || eq == 0) throw ConfigError("override '" + assignment + "' must be key=value");
  const std::string key = assignment.substr(0, eq), value = assignment.substr(eq + 1);
  std::shared_ptr<yaml::Node> v;
  try {
    v = yaml::parse(value.empty() ? "null" : value, "override " + key);
  } catch (const yaml::Error& e) {
    throw ConfigError(std::string("override '") + assignment + "': " + e.what());
  }
  yaml::Node* n = root_.get();
  const std::vector<std::string> keys = split_dots(key);
  for (size_t i = 0; i + 1 < keys.size(); ++i) {
    const yaml::Node* c = n->get(keys[i]);
    if (!c || !c->is_map()) {
      if (c && !c->is_null()) throw ConfigError("override '" + key + "': '" + keys[i] + "' is not a mapping");
      auto m = std::make_shared<yaml::Node>();
      m->kind = yaml::Node::Map;
      n->set(keys[i], m);
      c = m.get();
    }
    n = const_cast<yaml::Node*>(c);
  }
  n->set(keys.back(), v);
  if (validate_) validate_structure();
}

// Config::validateSchema (src/io/Config.cpp:97-143).
void Config::validate_structure() const {
  if (!has("modules")) throw ConfigError("schema error: 'modules' section missing (in " + source_ + ")");
  for (const char* key : {"time.t_end", "time.dt"}) {
    if (!has(key)) throw ConfigError(std::string("schema error: required key '") + key + "' missing (in " + source_ + ")");
    try {
      (void)find(key)->as_double();
    } catch (const yaml::Error&) {
      throw ConfigError(std::string("schema error: '") + key + "' must be numeric (in " + source_ + ")");
    }
  }
  for (const char* key : {"domain.nx", "domain.ny", "domain.nz"}) {
    if (!has(key)) throw ConfigError(std::string("schema error: required key '") + key + "' missing (in " + source_ + ")");
    int v = 0;
    try {
      v = find(key)->as_int();
    } catch (const yaml::Error&) {
      throw ConfigError(std::string("schema error: '") + key + "' must be an integer (in " + source_ + ")");
    }
    if (v <= 0) throw ConfigError(std::string("schema error: '") + key + "' must be positive (in " + source_ + ")");
  }
}

// ---- the run set-up ----------------------------------------------------------------------------
namespace {

// validateSchemaV2 (Orchestrator.cpp:85-103).
void gate_schema_v2(const Config& c) {
  const std::string& src = c.source();
  if (!c.has("schema_version"))
    throw ConfigError("schema error: required top-level 'schema_version' missing (in " + src +
                      "). This is the frozen v2 schema.");
  const std::string ver = c.get_string("schema_version", "");
  if (ver != "2.0")
    throw ConfigError("schema error: schema_version='" + ver + "' is not supported; the production schema is '2.0' (in " +
                      src + ").");
  for (const char* sect : {"simulation", "domain", "time", "modules", "output"})
    if (!c.has(sect))
      throw ConfigError(std::string("schema error: required top-level section '") + sect + "' missing (in " + src + ")");
}

// validateCapabilities (Orchestrator.cpp:111-173), then what this backend adds to it.
void gate_capabilities(const Config& c, bool sw, bool gw, bool solute) {
  const std::string& src = c.source();
  if (sw) {
    if (c.get_bool("surface_water.diffusive_wave", false))
      reject(src, "surface_water.diffusive_wave", "is not implemented (only the dynamic-wave SWE, difuwave==0, is realized)");
    if (c.get_bool("surface_water.wind.enabled", false))
      reject(src, "surface_water.wind.enabled", "wind forcing is not implemented");
    if (c.get_bool("surface_water.subgrid.enabled", false))
      reject(src, "surface_water.subgrid.enabled", "subgrid bathymetry is not implemented");
    const int evap_model = c.get_int("sources.surface.evaporation.model", 0);
    if (evap_model != 0)
      reject(src, "sources.surface.evaporation.model",
             "only constant-rate evaporation (model 0) is implemented; model=" + std::to_string(evap_model) + " is not");
  }
  if (c.get_bool("solute.baroclinic", false))
    reject(src, "solute.baroclinic", "baroclinic (density-driven) coupling is not implemented");
  if (gw) {
    // `groundwater.solver` names the scheme; as a mapping it holds Krylov settings instead
    const yaml::Node* gn = c.find("groundwater.solver");
    if (gn && gn->is_scalar() && gn->text != "pca")
      reject(src, "groundwater.solver",
             "only the PCA predictor-corrector scheme ('pca') is implemented; '" + gn->text +
                 "' (iterative Picard/Newton) is not");
    if (c.get_int("groundwater.iter_solve", 0) != 0)
      reject(src, "groundwater.iter_solve",
             "iterative (Picard/Newton) head solves are not implemented; use the PCA scheme (iter_solve: 0)");
    const yaml::Node* bc = c.find("groundwater.bc_type_gw");
    if (bc && bc->is_seq()) {
      const char* face[6] = {"x-", "x+", "y-", "y+", "z+ (bottom)", "z- (top)"};
      for (int i = 0; i < 6 && i < (int)bc->size(); ++i) {
        int mode = 0;
        try {
          mode = bc->at((size_t)i)->as_int();
        } catch (const yaml::Error&) {
          throw ConfigError("config: 'groundwater.bc_type_gw[" + std::to_string(i) + "]' must be an integer (in " + src + ")");
        }
        const std::string key = "groundwater.bc_type_gw[" + std::to_string(i) + "] (" + face[i] + ")";
        if (i < 4 && mode != 0)
          reject(src, key,
                 "lateral Dirichlet/fixed-flux GW boundaries are not implemented; only no-flux (0) is supported on lateral faces");
        // 3 = free drainage at the bottom: implemented by this backend (not by the reference)
        if (i == 4 && (mode < 0 || mode > 3)) reject(src, key, "must be 0 (no-flux), 1 (Dirichlet), 2 (fixed-flux) or 3 (free drainage)");
        if (i == 5 && (mode < 0 || mode > 2)) reject(src, key, "must be 0 (no-flux), 1 (Dirichlet), or 2 (fixed-flux)");
      }
    }
  }
  (void)solute;  // the reference's gate takes it and checks nothing with it (:172)
  // ---- not implemented by this backend ----
  const yaml::Node* srcs = c.find("sources");
  if (srcs && srcs->is_seq() && srcs->size() > 0)
    reject(src, "sources", "polygon sources (inflow_rate, rainfall_rate, extraction_well) are not implemented by this backend");
  if (c.has("monitors")) reject(src, "monitors", "probe and line monitors are not implemented by this backend");
  const yaml::Node* ic = c.find("initial_conditions");
  if (ic && ic->is_map() && ic->get("regions"))
    reject(src, "initial_conditions.regions", "polygon initial conditions are not implemented by this backend");
  if (c.get_string("output.io_mode", "serial_gather") != "serial_gather")
    reject(src, "output.io_mode", "only serial_gather output is implemented by this backend");
}

frehg_solver_config solver_defaults(double rtol) {
  frehg_solver_config s{};
  s.ksp_type = FREHG_KSP_CG;
  s.pc_type = FREHG_PC_JACOBI;
  s.rtol = rtol;
  s.atol = 1e-50;
  s.max_it = 10000;
  s.precision = FREHG_PREC_FP64;
  s.inner_rtol = 1e-5;
  s.max_refine = 8;
  return s;
}

// A solver section ("surface_water.solver" / "groundwater.solver" when it is a mapping).
frehg_solver_config parse_solver(const Config& c, const std::string& sect, double rtol_default, const std::string& pc_default) {
  frehg_solver_config s = solver_defaults(c.get_double(sect + ".rtol", rtol_default));
  s.ksp_type = ksp_of(c.get_string(sect + ".ksp_type", "cg"), sect + ".ksp_type", c.source());
  check_pc(c.get_string(sect + ".pc_type", pc_default), sect + ".pc_type", c.source());
  s.atol = c.get_double(sect + ".atol", s.atol);
  s.max_it = c.get_int(sect + ".max_it", s.max_it);
  // this backend's additions: FP32 inner solves inside FP64 refinement
  const std::string prec = lower(c.get_string(sect + ".precision", "fp64"));
  if (prec == "fp64") s.precision = FREHG_PREC_FP64;
  else if (prec == "mixed") s.precision = FREHG_PREC_MIXED;
  else reject(c.source(), sect + ".precision", "= '" + prec + "': expected fp64 or mixed");
  s.inner_rtol = c.get_double(sect + ".inner_rtol", s.inner_rtol);
  s.max_refine = c.get_int(sect + ".max_refine", s.max_refine);
  if (!(s.rtol > 0.0)) reject(c.source(), sect + ".rtol", "must be > 0");
  return s;
}

// parseOne (src/soil/SoilConfig.cpp:10-31).
frehg_soil_params parse_soil(const yaml::Node& st, const Config& c, size_t idx) {
  const std::string where = "soil.types[" + std::to_string(idx) + "]";
  if (!st.get("theta_s") || !st.get("theta_r") || !st.get("vg") || !st.get("k_sat"))
    throw ConfigError("parseSoilClasses: a soil.types entry is missing theta_s/theta_r/vg/k_sat (" + where + ")");
  auto num = [&](const yaml::Node* n, const std::string& key) {
    if (!n || n->is_null()) throw ConfigError("parseSoilClasses: " + where + "." + key + " is missing");
    try {
      return n->as_double();
    } catch (const yaml::Error&) {
      throw ConfigError("parseSoilClasses: " + where + "." + key + " must be a number");
    }
  };
  frehg_soil_params s{};
  s.theta_s = num(st.get("theta_s"), "theta_s");
  s.theta_r = num(st.get("theta_r"), "theta_r");
  s.alpha = num(st.get("vg")->get("alpha"), "vg.alpha");
  s.n = num(st.get("vg")->get("n"), "vg.n");
  s.Ks_x = num(st.get("k_sat")->get("x"), "k_sat.x");
  s.Ks_y = num(st.get("k_sat")->get("y"), "k_sat.y");
  s.Ks_z = num(st.get("k_sat")->get("z"), "k_sat.z");
  s.Ss = c.get_double("groundwater.specific_storage", 1.0e-5);
  s.use_vg = c.get_bool("groundwater.use_vg", true) ? 1 : 0;
  s.use_mvg = c.get_bool("groundwater.use_mvg", false) ? 1 : 0;
  s.air_entry = c.get_double("groundwater.air_entry_value", -0.02);
  try {
    if (st.get("specific_storage") && !st.get("specific_storage")->is_null()) s.Ss = st.get("specific_storage")->as_double();
    if (st.get("use_vg") && !st.get("use_vg")->is_null()) s.use_vg = st.get("use_vg")->as_bool() ? 1 : 0;
    if (st.get("use_mvg") && !st.get("use_mvg")->is_null()) s.use_mvg = st.get("use_mvg")->as_bool() ? 1 : 0;
    if (st.get("air_entry_value") && !st.get("air_entry_value")->is_null())
      s.air_entry = st.get("air_entry_value")->as_double();
  } catch (const yaml::Error& e) {
    throw ConfigError("parseSoilClasses: " + where + ": " + e.what());
  }
  return s;
}

// Polygon vertices [[x, y], ...] (src/bc/PolygonConfig.cpp:12-31).
std::vector<double> parse_vertices(const yaml::Node* verts, const std::string& where, const std::string& name) {
  if (!verts || !verts->is_seq() || verts->size() < 3)
    throw ConfigError("PolygonConfig: " + where + " ('" + name + "') must have a 'vertices' list of >= 3 points");
  std::vector<double> xy;
  for (size_t k = 0; k < verts->size(); ++k) {
    const yaml::Node* v = verts->at(k);
    if (!v->is_seq() || v->size() != 2)
      throw ConfigError("PolygonConfig: " + where + " ('" + name + "') vertex " + std::to_string(k) + " must be [x, y]");
    try {
      xy.push_back(v->at(0)->as_double());
      xy.push_back(v->at(1)->as_double());
    } catch (const yaml::Error&) {
      throw ConfigError("PolygonConfig: " + where + " ('" + name + "') vertex " + std::to_string(k) +
                        " has non-numeric coordinates");
    }
  }
  return xy;
}

// Class ids of one 2D layer (Orchestrator.cpp:506-543).
std::vector<int32_t> load_class_layer(const Config& c, const std::string& file, int nx, int ny, int ncls) {
  const std::string fmt = c.get_string("soil.map.format", "auto");
  std::vector<double> raw;
  if (raster_format(fmt, file)) {
    raw = grid_to_model(read_ascii_grid(file), nx, ny, file, "Orchestrator: soil class");
  } else {
    raw = read_doubles(file, "Orchestrator: soil class file");
    if (raw.size() != (size_t)nx * ny)
      throw ConfigError("Orchestrator: soil class file '" + file + "' has " + std::to_string(raw.size()) +
                        " values; expected nx*ny=" + std::to_string(nx * ny));
  }
  std::vector<int32_t> out(raw.size());
  for (size_t i = 0; i < raw.size(); ++i) {
    const long id = std::lround(raw[i]);
    if (id < 0 || id >= ncls)
      throw ConfigError("Orchestrator: soil class id " + std::to_string(id) + " in '" + file + "' is out of range [0," +
                        std::to_string(ncls) + ") (have " + std::to_string(ncls) + " soil.types)");
    out[i] = (int32_t)id;
  }
  return out;
}

// An initial-condition field (ICConfig.cpp:21-49, legacy *_from_file flags :52-62, raster
// loading src/ic/RasterField.cpp:85-119): constant, or values from a list / ESRI grid file.
// Returns the value of every interior cell (nz layers; a 2D field repeats over the layers).
std::vector<double> ic_field(const Config& c, const std::string& key, const std::string& from_file_key,
                             const std::string& file_key, double def, int nx, int ny, int nz, bool* spatial) {
  const yaml::Node* n = c.find(key);
  double value = def;
  std::string kind = "constant", file, fmt = "auto";
  if (n && n->is_scalar()) {
    try {
      value = n->as_double();
    } catch (const yaml::Error&) {
      throw ConfigError("ICConfig: '" + key + "' must be a number");
    }
  } else if (n && n->is_map()) {
    try {
      if (n->get("type") && !n->get("type")->is_null()) kind = n->get("type")->as_string();
      if (n->get("value") && !n->get("value")->is_null()) value = n->get("value")->as_double();
      if (n->get("file") && !n->get("file")->is_null()) file = n->get("file")->as_string();
      if (n->get("format") && !n->get("format")->is_null()) fmt = n->get("format")->as_string();
    } catch (const yaml::Error& e) {
      throw ConfigError("ICConfig: '" + key + "': " + e.what());
    }
    if (kind != "constant" && kind != "raster" && kind != "file") {
      if (kind == "formula" || kind == "polygon" || kind == "restart")
        reject(c.source(), key + ".type", "= " + kind + ": not implemented by this backend (constant, raster and file are)");
      throw ConfigError("ICConfig: unknown initial condition type '" + kind + "'");
    }
  } else if (n) {
    throw ConfigError("ICConfig: '" + key + "' must be a scalar or a mapping");
  }
  if (c.get_bool(from_file_key, false)) {
    kind = "raster";
    file = c.get_string(file_key, "");
    if (file.empty()) throw ConfigError("ICConfig: " + from_file_key + " is true but " + file_key + " is empty");
  }
  const size_t n2 = (size_t)nx * ny, n3 = n2 * (size_t)nz;
  *spatial = kind != "constant";
  if (kind == "constant") return std::vector<double>(n3, value);
  if (fmt == "hdf5" || ends_with(file, ".h5")) reject(c.source(), key, "reads HDF5, which this backend does not");
  const std::string path = c.resolve_path(file);
  std::vector<double> raw;
  if (raster_format(fmt, path)) {
    const AsciiGrid g = read_ascii_grid(path);
    for (double x : g.v)
      if (g.is_nodata(x)) throw ConfigError("RasterField: NODATA in IC raster '" + path + "'");
    raw = grid_to_model(g, nx, ny, path, "RasterField");
  } else {
    raw = read_doubles(path, "RasterField");
  }
  std::vector<double> out(n3);
  if (raw.size() == n3) return raw;
  if (raw.size() == n2 || (nx == 1 && raw.size() == (size_t)ny)) {
    for (int k = 0; k < nz; ++k)
      for (size_t q = 0; q < n2; ++q) out[q + k * n2] = raw[q];
    return out;
  }
  throw ConfigError("RasterField: '" + path + "' has " + std::to_string(raw.size()) + " values; expected " +
                    std::to_string(n2) + " or " + std::to_string(n3));
}

}  // namespace

double rain_at(const RunSetup& s, double tc) {
  if (!s.sw) return 0.0;
  if (!s.rain_from_file) return s.rain_rate;
  // interpSeries (Orchestrator.cpp:57-71)
  const std::vector<double>& t = s.rain_t;
  const std::vector<double>& v = s.rain_v;
  const int n = (int)t.size();
  if (n == 0) return 0.0;
  if (tc == 0.0) return v[0];
  int ind = 1;
  if (ind < n && tc > t[(size_t)ind]) {
    ind += 1;
    while (ind < n && tc > t[(size_t)ind]) ind += 1;
  }
  if (ind >= n) return v[(size_t)n - 1];
  return v[(size_t)ind - 1] + (v[(size_t)ind] - v[(size_t)ind - 1]) * (tc - t[(size_t)ind - 1]) / (t[(size_t)ind] - t[(size_t)ind - 1]);
}

RunSetup build_setup(const Config& c) {
  gate_schema_v2(c);
  RunSetup s;
  const std::string& src = c.source();
  s.source = src;
  s.sim_id = c.get_string("simulation.id", "frehg2");
  s.mode = c.get_string("simulation.mode", "coupled");
  s.sw = c.get_bool("modules.surface_water", false);
  s.gw = c.get_bool("modules.groundwater", false);
  gate_capabilities(c, s.sw, s.gw, c.get_bool("modules.solute", false));
  // solute.enabled is authoritative: modules.solute alone builds no solute (buildSolute, :749-760)
  s.solute = c.get_bool("solute.enabled", false);
  if (s.mode == "surface_water" && !s.sw) throw ConfigError("Orchestrator: mode 'surface_water' requires modules.surface_water");
  if (s.mode == "groundwater" && !s.gw) throw ConfigError("Orchestrator: mode 'groundwater' requires modules.groundwater");
  if (!s.sw && !s.gw) throw ConfigError("Orchestrator: neither modules.surface_water nor modules.groundwater is enabled");
  const std::string cm = c.get_string("coupling.mode", "");
  // async coupling runs the same operation sequence synchronously (Orchestrator::step, :1322-1336)
  s.coupled = s.sw && s.gw && (s.mode == "coupled" || cm == "sync" || cm == "sequential" || cm == "async");
  if (s.sw && s.gw && !s.coupled)
    reject(src, "coupling.mode", "with both modules on and no coupling is not implemented by this backend (set mode: coupled)");

  // grid and time (:238-261)
  s.nx = c.get_int("domain.nx", 0);
  s.ny = c.get_int("domain.ny", 0);
  s.nz = c.get_int("domain.nz", 0);
  s.dx = c.get_double("domain.dx", 1.0);
  s.dy = c.get_double("domain.dy", 1.0);
  s.dz = c.get_double("domain.dz", 1.0);
  s.botz = c.get_double("domain.botz", 0.0);
  s.x0 = c.get_double("domain.x0", 0.0);
  s.y0 = c.get_double("domain.y0", 0.0);
  if (!(s.dx > 0.0) || !(s.dy > 0.0) || !(s.dz > 0.0)) throw ConfigError("Orchestrator: domain.dx/dy/dz must be positive");
  if (c.get_bool("domain.mpi.enabled", false) && (c.get_int("domain.mpi.nx", 1) * c.get_int("domain.mpi.ny", 1) > 1))
    std::fprintf(stderr, "frehg: domain.mpi is set; this backend runs the whole domain on one GPU\n");
  s.dt = c.get_double("time.dt", 0.0);
  s.t_end = c.get_double("time.t_end", 0.0);
  s.max_steps = c.get_long("time.max_steps", 0);
  s.output_interval = c.get_double("time.output_interval", 0.0);
  if (!(s.dt > 0.0)) throw ConfigError("Orchestrator: time.dt must be positive");
  s.evap_rate = c.get_double("sources.surface.evaporation.rate", 0.0);

  // rainfall (:313-336)
  s.rain_from_file = c.get_bool("sources.surface.rainfall.from_file", false);
  s.rain_rate = c.get_double("sources.surface.rainfall.rate", 0.0);
  if (s.rain_from_file) {
    const std::string file = c.resolve_path(c.get_string("sources.surface.rainfall.file", "rain"));
    const std::vector<double> flat = read_doubles(file, "Orchestrator: rainfall");
    if (flat.empty() || flat.size() % 2 != 0)
      throw ConfigError("Orchestrator: rainfall file '" + file + "' must contain an even number of (time, rate) values");
    for (size_t k = 0; k < flat.size() / 2; ++k) {
      s.rain_t.push_back(flat[2 * k]);
      s.rain_v.push_back(flat[2 * k + 1]);
    }
  }

  const size_t n2 = (size_t)s.nx * s.ny;
  // surface water (:338-440)
  if (s.sw) {
    frehg_swe_params& p = s.swe;
    p.gravity = c.get_double("surface_water.gravity", 9.81);
    p.manning = c.get_double("surface_water.manning", 0.0);
    p.min_depth = c.get_double("surface_water.min_depth", 1.0e-8);
    p.visc_x = c.get_double("surface_water.viscosity.x", 0.0);
    p.visc_y = c.get_double("surface_water.viscosity.y", 0.0);
    p.hD = c.get_double("surface_water.h_diffusion_ref", 0.1);
    p.hE = 0.0;
    p.wtfh = c.get_double("surface_water.waterfall_depth", 1.0e-8);
    p.dt = s.dt;
    p.solver = parse_solver(c, "surface_water.solver", 1.0e-12, "jacobi");
    std::vector<double> bath(n2, s.botz);
    if (c.get_bool("domain.bathymetry.from_file", false)) {
      const std::string file = c.resolve_path(c.get_string("domain.bathymetry.file", "bath"));
      const std::string fmt = c.get_string("domain.bathymetry.format", "auto");
      if (raster_format(fmt, file)) {
        const AsciiGrid g = read_ascii_grid(file);
        bath = grid_to_model(g, s.nx, s.ny, file, "Orchestrator: raster DEM");
        // NODATA cells take the lowest valid elevation and stay active (:386-402)
        double vmin = std::numeric_limits<double>::max();
        for (double v : bath)
          if (!g.is_nodata(v)) vmin = std::min(vmin, v);
        if (vmin == std::numeric_limits<double>::max()) vmin = s.botz;
        int nodata = 0;
        for (double& v : bath)
          if (g.is_nodata(v)) {
            v = vmin;
            ++nodata;
          }
        if (nodata > 0)
          std::fprintf(stderr, "frehg: raster DEM '%s' has %d NODATA cell(s), filled with the minimum valid elevation\n",
                       file.c_str(), nodata);
      } else {
        const std::vector<double> raw = read_doubles(file, "Orchestrator: bathymetry");
        if (raw.size() == n2) {
          bath = raw;
        } else if (s.nx == 1 && raw.size() == (size_t)s.ny) {
          bath = raw;
        } else {
          throw ConfigError("Orchestrator: bathymetry file '" + file + "' has " + std::to_string(raw.size()) +
                            " values; expected " + std::to_string(n2) + " (nx*ny) or " + std::to_string(s.ny) +
                            " (column, nx==1)");
        }
      }
    }
    // the legacy datum: offset = -min(bottom) (:424-428)
    double bmin = bath[0];
    for (double b : bath) bmin = std::min(bmin, b);
    p.offset = -bmin;
    s.bed.resize(n2);
    for (size_t q = 0; q < n2; ++q) s.bed[q] = bath[q] + p.offset;
  }

  // groundwater (:442-591)
  if (s.gw) {
    const yaml::Node* types = c.find("soil.types");
    if (!types || !types->is_seq() || types->size() == 0)
      throw ConfigError("Orchestrator: groundwater enabled but soil.types is empty");
    for (size_t k = 0; k < types->size(); ++k) s.classes.push_back(parse_soil(*types->at(k), c, k));
    frehg_re_params& r = s.re;
    r.soil = s.classes[0];
    r.dx = s.dx;
    r.dy = s.dy;
    r.dz = s.dz;
    r.botz = s.botz;
    r.dt = s.dt;
    r.dt_min = c.get_double("groundwater.dt_min", 1.0e-4);
    r.dt_max = c.get_double("groundwater.dt_max", 2.0);
    r.co_max = c.get_double("groundwater.co_max", 2.0);
    r.adaptive_dt = c.get_bool("groundwater.adaptive_dt", true) ? 1 : 0;
    r.use_corrector = c.get_bool("groundwater.use_corrector", true) ? 1 : 0;
    r.use_full3d = c.get_bool("groundwater.full_3d", false) ? 1 : 0;
    r.sim_shallowwater = 0;  // the Orchestrator leaves ReParams::sim_shallowwater at its default
    r.htop = c.get_double("groundwater.htop", 0.0);
    r.hbot = c.get_double("groundwater.hbot", 0.0);
    r.qtop = c.get_double("groundwater.qtop", 0.0);
    r.qbot = c.get_double("groundwater.qbot", 0.0);
    const int bc_default[6] = {0, 0, 0, 0, 0, 1};  // ReParams::bc_type
    for (int i = 0; i < 6; ++i) r.bc_type[i] = bc_default[i];
    const yaml::Node* bc = c.find("groundwater.bc_type_gw");
    if (bc && bc->is_seq())
      for (int i = 0; i < 6 && i < (int)bc->size(); ++i) r.bc_type[i] = bc->at((size_t)i)->as_int();
    // `groundwater.solver` names the scheme ("pca"); Krylov settings sit under it when it is a
    // mapping, as the reference reads them (groundwater.solver.ksp_type, :487-490)
    const yaml::Node* gsn = c.find("groundwater.solver");
    if (gsn && gsn->is_map())
      r.solver = parse_solver(c, "groundwater.solver", 1.0e-10, "sor");
    else
      r.solver = solver_defaults(1.0e-10);
    // this backend's additions: a Picard loop around the PCA step and the Krylov start
    r.picard_max = c.get_int("groundwater.picard.max_iterations", 1);
    r.picard_tol = c.get_double("groundwater.picard.tolerance", 1.0e-8);
    if (r.picard_max < 1) reject(src, "groundwater.picard.max_iterations", "must be >= 1");
    const std::string guess = lower(c.get_string("groundwater.initial_guess", "zero"));
    if (guess == "zero") r.initial_guess = FREHG_GUESS_ZERO;
    else if (guess == "previous") r.initial_guess = FREHG_GUESS_PREVIOUS;
    else if (guess == "extrapolated") r.initial_guess = FREHG_GUESS_EXTRAPOLATED;
    else reject(src, "groundwater.initial_guess", "= '" + guess + "': expected zero, previous or extrapolated");

    // soil map (:495-591): only with soil.map.from_file
    if (c.get_bool("soil.map.from_file", false)) {
      const int ncls = (int)s.classes.size();
      const yaml::Node* layers = c.find("soil.map.layers");
      if (layers && layers->is_seq() && layers->size() > 0) {
        if ((int)layers->size() != s.nz)
          throw ConfigError("Orchestrator: soil.map.layers has " + std::to_string(layers->size()) +
                            " layers but domain nz=" + std::to_string(s.nz) + " (need one class file per vertical layer)");
        s.class_idx.resize(n2 * (size_t)s.nz);
        for (int k = 0; k < s.nz; ++k) {
          const std::string lf = c.resolve_path(layers->at((size_t)k)->as_string());
          const std::vector<int32_t> g = load_class_layer(c, lf, s.nx, s.ny, ncls);
          std::copy(g.begin(), g.end(), s.class_idx.begin() + (long)(n2 * (size_t)k));
        }
        s.class_idx_3d = true;
      } else {
        const std::string file = c.resolve_path(c.get_string("soil.map.file", "soil_class"));
        s.class_idx = load_class_layer(c, file, s.nx, s.ny, ncls);
      }
    }

    // top recharge (:593-628)
    const yaml::Node* rech = c.find("groundwater.recharge");
    if (rech && rech->is_map()) {
      double rate = 0.0;
      try {
        if (rech->get("rate") && !rech->get("rate")->is_null()) rate = rech->get("rate")->as_double();
      } catch (const yaml::Error&) {
        throw ConfigError("config: 'groundwater.recharge.rate' must be a number (in " + src + ")");
      }
      std::vector<double> xy;
      const yaml::Node* verts = rech->get("polygon");
      if (verts && verts->is_seq())
        for (size_t k = 0; k < verts->size(); ++k) {
          const yaml::Node* v = verts->at(k);
          if (v->is_seq() && v->size() >= 2) {
            xy.push_back(v->at(0)->as_double());
            xy.push_back(v->at(1)->as_double());
          }
        }
      const bool whole = xy.size() < 6;  // no polygon: the whole top
      s.top_flux.assign(n2, 0.0);
      for (int j = 0; j < s.ny; ++j)
        for (int i = 0; i < s.nx; ++i) {
          const double xc = s.x0 + (i + 0.5) * s.dx, yc = s.y0 + (j + 0.5) * s.dy;
          bool in = whole;
          if (!whole) {
            // even-odd ray test (src/bc/Polygon.hpp)
            const size_t nv = xy.size() / 2;
            for (size_t a = 0, b = nv - 1; a < nv; b = a++) {
              const double xa = xy[2 * a], ya = xy[2 * a + 1], xb = xy[2 * b], yb = xy[2 * b + 1];
              if (((ya > yc) != (yb > yc)) && (xc < (xb - xa) * (yc - ya) / (yb - ya) + xa)) in = !in;
            }
          }
          if (in) s.top_flux[(size_t)i + (size_t)j * s.nx] = rate;
        }
    }
  }

  if (s.coupled) {
    s.cp.dx = s.dx;
    s.cp.dy = s.dy;
    s.cp.visc = 1.0;
    s.cp.min_depth = c.get_double("surface_water.min_depth", 1.0e-8);
  }

  // solute (SoluteParams::fromConfig, src/solute/SoluteParams.cpp:7-18; buildSolute, :749-783)
  if (s.solute) {
    frehg_solute_params& q = s.sol;
    q.c_rain = c.get_double("solute.c_rain", 0.0);
    q.k_decay = c.get_double("solute.k_decay", 0.0);
    q.D = c.get_double("solute.D", 1.0e-9);
    q.cfl_max = c.get_double("solute.cfl_max", 1.0);
    q.min_depth = c.get_double("solute.min_depth", 1.0e-8);
    q.start_time = c.get_double("solute.start_time", 0.0);
    q.substeps = std::max(1, c.get_int("solute.substeps", 1));
    // the stepper tests the scheme names for equality: anything but "muscl" advects upwind and
    // anything but "implicit" does not diffuse (SoluteStepper.cpp, Advection.cpp:12)
    q.muscl = c.get_string("solute.advection_scheme", "upwind") == "muscl" ? 1 : 0;
    q.implicit_diffusion = c.get_string("solute.diffusion_scheme", "implicit") == "implicit" ? 1 : 0;
    q.solver = parse_solver(c, "solute.solver", 1.0e-12, "jacobi");
    if (q.solver.precision != FREHG_PREC_FP64)
      reject(src, "solute.solver.precision", "must be fp64: the solute diffusion solve runs in FP64");
    if (!(q.cfl_max > 0.0)) throw ConfigError("solute.cfl_max must be > 0");
    if (q.D < 0.0 || q.k_decay < 0.0) throw ConfigError("solute.D and solute.k_decay must be >= 0");
  }

  // boundary polygons (:630-656)
  const yaml::Node* bnds = c.find("boundaries");
  if (bnds && bnds->is_seq()) {
    if (!s.sw && bnds->size() > 0)
      reject(src, "boundaries", "surface boundary regions need modules.surface_water");
    for (size_t i = 0; i < bnds->size(); ++i) {
      const yaml::Node* n = bnds->at(i);
      const std::string where = "boundaries[" + std::to_string(i) + "]";
      if (!n->get("type") || n->get("type")->is_null())
        throw ConfigError("PolygonConfig: " + where + " is missing required key 'type'");
      const std::string type = n->get("type")->as_string();
      std::string name = where;
      if (n->get("name") && !n->get("name")->is_null()) name = n->get("name")->as_string();
      SurfaceRegion reg;
      reg.xy = parse_vertices(n->get("vertices"), where, name);
      // bcKindFromString (src/bc/PolygonRegion.cpp:7-13)
      if (type == "bc_discharge") reg.kind = FREHG_BC_DISCHARGE;
      else if (type == "bc_depth") reg.kind = FREHG_BC_DEPTH;
      else if (type == "bc_critical") reg.kind = FREHG_BC_CRITICAL;
      else
        throw ConfigError("PolygonRegion: unknown boundary type '" + type +
                          "' (expected bc_discharge | bc_depth | bc_critical)");
      if (n->get("rate") && !n->get("rate")->is_null()) {
        try {
          reg.rate = n->get("rate")->as_double();
        } catch (const yaml::Error&) {
          throw ConfigError("PolygonConfig: " + where + ".rate must be a number");
        }
      }
      s.regions.push_back(reg);
    }
  }

  // initial conditions (ICConfig.cpp:60-143, ICApply.cpp:187-233)
  const int nz = s.gw ? s.nz : 1;
  if (s.sw) {
    bool spatial = false;
    std::vector<double> eta = ic_field(c, "initial_conditions.surface_water.eta", "initial_conditions.surface_water.eta_from_file",
                                       "initial_conditions.surface_water.eta_file", c.get_double("domain.botz", 0.0), s.nx, s.ny,
                                       1, &spatial);
    s.eta0.assign(eta.begin(), eta.begin() + (long)n2);
    for (const char* uv : {"initial_conditions.surface_water.u", "initial_conditions.surface_water.v"}) {
      const yaml::Node* n = c.find(uv);
      if (n && !(n->is_scalar() && n->as_double() == 0.0))
        reject(src, uv, "non-zero initial velocities are not implemented by this backend");
    }
    if (c.get_bool("initial_conditions.surface_water.uv_from_file", false))
      reject(src, "initial_conditions.surface_water.uv_from_file", "initial velocities from a file are not implemented by this backend");
  }
  if (s.gw) {
    bool head_spatial = false, wc_spatial = false;
    std::vector<double> h = ic_field(c, "initial_conditions.groundwater.h", "initial_conditions.groundwater.h_from_file",
                                     "initial_conditions.groundwater.h_file", 0.0, s.nx, s.ny, nz, &head_spatial);
    std::vector<double> wc = ic_field(c, "initial_conditions.groundwater.wc", "initial_conditions.groundwater.wc_from_file",
                                      "initial_conditions.groundwater.wc_file", s.re.soil.theta_r, s.nx, s.ny, nz, &wc_spatial);
    // a spatial head wins; otherwise the water content (ICApply.cpp:208-220)
    s.gw_init_head = head_spatial;
    s.gw_init = head_spatial ? h : wc;
  }
  if (s.solute) {  // initial_conditions.solute.surface / .subsurface (ICConfig.cpp:126-129, ICApply.cpp:228-249)
    bool spatial = false;
    if (s.sw) {
      std::vector<double> v = ic_field(c, "initial_conditions.solute.surface", "", "", 0.0, s.nx, s.ny, 1, &spatial);
      s.sol_surface.assign(v.begin(), v.begin() + (long)n2);
    }
    if (s.gw) s.sol_subsurface = ic_field(c, "initial_conditions.solute.subsurface", "", "", 0.0, s.nx, s.ny, nz, &spatial);
  }

  // output selection (:1077-1175)
  const yaml::Node* vars = c.find("output.variables");
  auto add = [](std::vector<std::string>& v, const std::string& n) {
    if (std::find(v.begin(), v.end(), n) == v.end()) v.push_back(n);
  };
  if (!vars || !vars->is_seq() || vars->size() == 0) {
    if (s.sw) s.sw_out = {"water_depth", "eta"};
    if (s.gw) s.gw_out = {"head", "moisture"};
    if (s.solute) {
      if (s.sw) s.sw_out.push_back("concentration");
      if (s.gw) s.gw_out.push_back("concentration");
    }
  } else {
    for (size_t k = 0; k < vars->size(); ++k) {
      const std::string raw = vars->at(k)->as_string();
      const std::string name = lower(raw);
      auto need = [&](bool on, const char* module) {
        if (!on)
          throw ConfigError("output.variables: '" + raw + "' requires modules." + module + " to be enabled");
      };
      if (name == "water_depth" || name == "depth") {
        need(s.sw, "surface_water");
        add(s.sw_out, "water_depth");
      } else if (name == "eta" || name == "wse" || name == "water_surface_elevation") {
        need(s.sw, "surface_water");
        add(s.sw_out, "eta");
      } else if (name == "u" || name == "uu" || name == "velocity_x" || name == "velocity_u") {
        need(s.sw, "surface_water");
        add(s.sw_out, "velocity_u");
      } else if (name == "v" || name == "vv" || name == "velocity_y" || name == "velocity_v") {
        need(s.sw, "surface_water");
        add(s.sw_out, "velocity_v");
      } else if (name == "head" || name == "pressure_head") {
        need(s.gw, "groundwater");
        add(s.gw_out, "head");
      } else if (name == "moisture" || name == "water_content" || name == "wc" || name == "theta") {
        need(s.gw, "groundwater");
        add(s.gw_out, "moisture");
      } else if (name == "qx" || name == "qy" || name == "qz") {
        need(s.gw, "groundwater");
        add(s.gw_out, name);
      } else if (name == "concentration" || name == "conc") {
        if (!s.solute)
          throw ConfigError("output.variables: '" + raw + "' requires modules.solute (or solute.enabled) to be enabled");
        if (s.sw) add(s.sw_out, "concentration");
        if (s.gw) add(s.gw_out, "concentration");
      } else {
        throw ConfigError("output.variables: unknown field '" + raw +
                          "'. Supported: water_depth, eta, u, v (surface); head, moisture, qx, qy, qz (subsurface); "
                          "concentration (solute).");
      }
    }
  }
  const std::string fname = c.get_string("output.filename", "");
  if (fname.empty()) {
    s.out_dir = ".";
    s.write_fields = false;
  } else {
    const std::string resolved = c.resolve_path(fname);
    const std::string d = dir_of(resolved);
    s.out_dir = d.empty() ? "." : d;
    s.write_fields = true;
    const size_t slash = resolved.find_last_of('/'), dot = resolved.find_last_of('.');
    s.stem = (dot != std::string::npos && (slash == std::string::npos || dot > slash)) ? resolved.substr(0, dot) : resolved;
  }
  s.summary_path = s.out_dir + "/simulation_summary.txt";

  // checkpoints (Orchestrator.cpp:253-254, 1461-1502) and restart (ICConfig.cpp:66-74)
  s.dt_checkpoint = c.get_double("time.dt_checkpoint", 0.0);
  s.max_checkpoints = std::max(1, c.get_int("time.max_checkpoints", 2));
  const yaml::Node* icr = c.find("initial_conditions");
  const yaml::Node* ict = (icr && icr->is_map()) ? icr->get("type") : nullptr;
  if (ict && ict->is_scalar() && ict->text == "restart") {
    s.restart = true;
    s.restart_file = c.get_string("initial_conditions.file", "");
    if (s.restart_file.empty()) throw ConfigError("ICConfig: initial_conditions.type=restart requires file");
    s.restart_file = c.resolve_path(s.restart_file);
    s.restart_time_given = c.has("initial_conditions.time");
    s.restart_time = c.get_double("initial_conditions.time", 0.0);
  }
  return s;
}

}  // namespace frehg
