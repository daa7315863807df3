// frehg_run: the command line of this backend, in the place of the frehg2 driver
// (src/driver/main.cpp): runs a v2 YAML configuration on one GPU through the C ABI.
//
//   frehg_run [--set key=value]... [--out DIR] [--device N] [--check] config.yaml
//
// --set edits the loaded configuration (dotted key, YAML value), --out writes the snapshots and
// the summary to DIR, --check loads, gates and maps the configuration without a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "frehg_gpu.h"

namespace {

void usage(FILE* f) {
  std::fprintf(f,
               "Usage: frehg_run [options] config.yaml\n"
               "\n"
               "  Surface water / groundwater coupled model on one GPU (Frehg2 v2 YAML).\n"
               "\n"
               "Options:\n"
               "  -h, --help           Show this help message and exit\n"
               "  -v, --version        Show the library ABI version and exit\n"
               "  --set KEY=VALUE      Replace a configuration value (dotted key, YAML value)\n"
               "  --out DIR            Write snapshots and simulation_summary.txt to DIR\n"
               "  --device N           GPU ordinal (default 0)\n"
               "  --check              Load, gate and map the configuration, then exit (no GPU)\n");
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> sets;
  std::string out, path;
  int device = 0;
  bool check_only = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&](const char* flag) -> std::string {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "frehg_run: %s needs a value\n", flag);
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return 0;
    } else if (a == "-v" || a == "--version") {
      std::printf("frehg_run (frehg_amd ABI %d)\n", frehg_abi_version());
      return 0;
    } else if (a == "--set") {
      sets.push_back(value("--set"));
    } else if (a == "--out") {
      out = value("--out");
    } else if (a == "--device") {
      device = std::atoi(value("--device").c_str());
    } else if (a == "--check") {
      check_only = true;
    } else if (!a.empty() && a[0] == '-') {
      std::fprintf(stderr, "frehg_run: unknown option %s\n", a.c_str());
      usage(stderr);
      return 2;
    } else if (path.empty()) {
      path = a;
    } else {
      std::fprintf(stderr, "frehg_run: more than one configuration file given\n");
      return 2;
    }
  }
  if (path.empty()) {
    usage(stderr);
    return 2;
  }
  std::vector<const char*> ov;
  for (const auto& s : sets) ov.push_back(s.c_str());
  char err[2048];
  frehg_run_setup setup;
  if (frehg_config_setup(path.c_str(), ov.data(), (int)ov.size(), &setup, err, (int)sizeof err) != FREHG_OK) {
    std::fprintf(stderr, "frehg_run: %s\n", err);
    return 1;
  }
  std::printf("frehg_run: %s: %dx%dx%d, surface water %s, groundwater %s, dt %g s, t_end %g s\n", path.c_str(), setup.nx,
              setup.ny, setup.nz, setup.surface_water ? "on" : "off", setup.groundwater ? "on" : "off", setup.dt,
              setup.t_end);
  if (check_only) return 0;
  frehg_ctx_t ctx = nullptr;
  int rc = frehg_create(&ctx, device);
  if (rc != FREHG_OK) {
    std::fprintf(stderr, "frehg_run: no GPU context (code %d)\n", rc);
    return 1;
  }
  frehg_run_stats st;
  rc = frehg_run_config(ctx, path.c_str(), ov.data(), (int)ov.size(), out.empty() ? nullptr : out.c_str(), &st);
  if (rc != FREHG_OK) {
    std::fprintf(stderr, "frehg_run: %s\n", frehg_last_error(ctx));
    frehg_destroy(ctx);
    return 1;
  }
  std::printf("frehg_run: t = %.9g s after %lld steps (%lld groundwater steps), %d snapshots, %.3f s wall\n", st.t,
              (long long)st.steps, (long long)st.gw_substeps, st.outputs, st.wall_seconds);
  std::printf("frehg_run: water %.12g -> %.12g m^3, rain in %.12g, boundary out %.12g, exchange %.12g\n",
              st.water_initial, st.water_final, st.rain_volume, st.boundary_outflow, st.exchange_volume);
  frehg_destroy(ctx);
  return 0;
}
