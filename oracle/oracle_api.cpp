// extern "C" surface of the FP64 oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Loaded by tests/ and by bench.py's cpu_baseline leg through ctypes. Parameter structs are the
// ones of the product ABI (include/frehg_gpu.h) so both sides are driven with identical inputs.
#include <cstring>
#include <map>
#include <omp.h>
#include <exception>
#include <stdexcept>
#include <string>

#include "../include/frehg_gpu.h"
#include "oracle.hpp"

using namespace frehg_oracle;

namespace {
thread_local std::string g_err;

Soil toSoil(const frehg_soil_params& s) {
  Soil o;
  o.alpha = s.alpha;
  o.n = s.n;
  o.theta_s = s.theta_s;
  o.theta_r = s.theta_r;
  o.Ks_x = s.Ks_x;
  o.Ks_y = s.Ks_y;
  o.Ks_z = s.Ks_z;
  o.Ss = s.Ss;
  o.use_vg = s.use_vg;
  o.use_mvg = s.use_mvg;
  o.air_entry = s.air_entry;
  return o;
}

SolverCfg toCfg(const frehg_solver_config& c) {
  SolverCfg o;
  o.rtol = c.rtol;
  o.atol = c.atol;
  o.max_it = c.max_it;
  o.ksp = c.ksp_type == FREHG_KSP_BICGSTAB ? 1 : 0;
  return o;
}

ReParams toRe(const frehg_re_params& p) {
  ReParams o;
  o.soil = toSoil(p.soil);
  o.dx = p.dx;
  o.dy = p.dy;
  o.dz = p.dz;
  o.botz = p.botz;
  o.dt = p.dt;
  o.dt_min = p.dt_min;
  o.dt_max = p.dt_max;
  o.co_max = p.co_max;
  o.adaptive_dt = p.adaptive_dt;
  o.use_corrector = p.use_corrector;
  o.use_full3d = p.use_full3d;
  o.sim_shallowwater = p.sim_shallowwater;
  for (int i = 0; i < 6; ++i) o.bc_type[i] = p.bc_type[i];
  o.htop = p.htop;
  o.hbot = p.hbot;
  o.qtop = p.qtop;
  o.qbot = p.qbot;
  o.solver = toCfg(p.solver);
  o.picard_max = p.picard_max;
  o.picard_tol = p.picard_tol;
  return o;
}

SweParams toSwe(const frehg_swe_params& p) {
  SweParams o;
  o.gravity = p.gravity;
  o.manning = p.manning;
  o.min_depth = p.min_depth;
  o.visc_x = p.visc_x;
  o.visc_y = p.visc_y;
  o.hD = p.hD;
  o.hE = p.hE;
  o.wtfh = p.wtfh;
  o.dt = p.dt;
  o.offset = p.offset;
  o.solver = toCfg(p.solver);
  return o;
}

void fillStats(const SolveStats& s, frehg_solve_stats* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->iterations = s.iterations;
  o->residual_norm = s.residual_norm;
  o->rhs_norm = s.rhs_norm;
  o->converged = s.converged;
}

std::vector<real>* reField(ReSolver* r, const std::string& n) {
  if (n == "h") return &r->h;
  if (n == "hn") return &r->hn;
  if (n == "wc") return &r->wc;
  if (n == "wcn") return &r->wcn;
  if (n == "wch") return &r->wch;
  if (n == "hwc") return &r->hwc;
  if (n == "ch") return &r->ch;
  if (n == "Kx") return &r->Kx;
  if (n == "Ky") return &r->Ky;
  if (n == "Kz") return &r->Kz;
  if (n == "qx") return &r->qx;
  if (n == "qy") return &r->qy;
  if (n == "qz") return &r->qz;
  if (n == "Gxp") return &r->Gxp;
  if (n == "Gxm") return &r->Gxm;
  if (n == "Gyp") return &r->Gyp;
  if (n == "Gym") return &r->Gym;
  if (n == "Gzp") return &r->Gzp;
  if (n == "Gzm") return &r->Gzm;
  if (n == "Gct") return &r->Gct;
  if (n == "Grhs") return &r->Grhs;
  return nullptr;
}

std::vector<real>* sweField(SweSolver* s, const std::string& n) {
#define F(x) \
  if (n == #x) return &s->x;
  F(eta) F(etan) F(bottom) F(bottomXP) F(bottomYP) F(dept) F(deptx) F(depty) F(uu) F(vv) F(un)
  F(vn) F(uy) F(vx) F(Vs) F(Vsn) F(Asx) F(Asy) F(Asz) F(Aszx) F(Aszy) F(Vsx) F(Vsy) F(CDx)
  F(CDy) F(Dx) F(Dy) F(Ex) F(Ey) F(cflx) F(cfly) F(Fu) F(Fv) F(rain) F(evap) F(Sxp) F(Sxm)
  F(Syp) F(Sym) F(Sct) F(Srhs) F(cfl_active) F(wtfx) F(wtfy)
#undef F
  return nullptr;
}

template <class Fn>
int guard(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// OpenMP threads the oracle's parallel loops use.
int orc_num_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

void* orc_re_create(int nx, int ny, int nz, const frehg_re_params* p) {
  try {
    return new ReSolver(nx, ny, nz, toRe(*p));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_re_destroy(void* h) { delete static_cast<ReSolver*>(h); }
int orc_re_set_params(void* h, const frehg_re_params* p) {
  return guard([&] { static_cast<ReSolver*>(h)->setParams(toRe(*p)); });
}
int orc_re_set_soil(void* h, const frehg_soil_params* classes, int ncls, const int32_t* class_idx,
                    int idx3d, const double* ks_cell) {
  return guard([&] {
    auto* r = static_cast<ReSolver*>(h);
    std::vector<Soil> cl;
    for (int c = 0; c < ncls; ++c) cl.push_back(toSoil(classes[c]));
    const size_t n2 = (size_t)r->g.nx * r->g.ny, n3 = n2 * r->g.nz;
    std::vector<int> idx;
    if (class_idx) idx.assign(class_idx, class_idx + (idx3d ? n3 : n2));
    std::vector<real> ks;
    if (ks_cell) ks.assign(ks_cell, ks_cell + n3);
    r->setSoilMap(cl, idx, idx3d != 0, ks);
  });
}
int orc_re_set_top_flux(void* h, const double* q) {
  return guard([&] {
    auto* r = static_cast<ReSolver*>(h);
    r->setTopFluxField(std::vector<real>(q, q + (size_t)r->g.nx * r->g.ny));
  });
}
int orc_re_set_initial(void* h, const double* hv, const double* wv) {
  return guard([&] {
    auto* r = static_cast<ReSolver*>(h);
    const size_t n3 = (size_t)r->g.nx * r->g.ny * r->g.nz;
    if (hv) r->setInitialHead(std::vector<real>(hv, hv + n3));
    else r->setInitialWaterContent(std::vector<real>(wv, wv + n3));
    r->finalizeInitialState();
  });
}
int orc_re_get(void* h, const char* name, double* dst) {
  return guard([&] {
    auto* f = reField(static_cast<ReSolver*>(h), name);
    if (!f) throw std::runtime_error(std::string("unknown GW field ") + name);
    std::memcpy(dst, f->data(), f->size() * sizeof(double));
  });
}
int orc_re_set(void* h, const char* name, const double* src) {
  return guard([&] {
    auto* f = reField(static_cast<ReSolver*>(h), name);
    if (!f) throw std::runtime_error(std::string("unknown GW field ") + name);
    std::memcpy(f->data(), src, f->size() * sizeof(double));
  });
}
int orc_re_reallocate(void* h) {
  return guard([&] { static_cast<ReSolver*>(h)->reallocateAndClamp(); });
}
int orc_re_step(void* h, frehg_re_stats* st) {
  return guard([&] {
    ReStepStats s;
    static_cast<ReSolver*>(h)->advanceStep(&s);
    if (st) {
      std::memset(st, 0, sizeof(*st));
      fillStats(s.solve, &st->solve);
      st->dt_used = s.dt_used;
      st->dt_next = s.dt_next;
      st->picard_iterations = s.picard_iterations;
      st->picard_converged = s.picard_converged ? 1 : 0;
      st->max_dh = s.max_dh;
      st->mass_before = s.mass_before;
      st->mass_after = s.mass_after;
      st->mass_boundary_in = s.mass_boundary_in;
      st->mass_storage_ss = s.mass_storage_ss;
      st->mass_adjust = s.mass_adjust;
      st->mass_residual = s.mass_after - s.mass_before - s.mass_boundary_in - s.mass_storage_ss - s.mass_adjust;
    }
  });
}
double orc_re_dt(void* h) { return static_cast<ReSolver*>(h)->params().dt; }

// Surface boundary regions per oracle SweSolver (PolygonBC + PolygonIndex), applied by
// orc_swe_apply_boundaries and by the coupled step after the SW advance. An entry lives as long
// as its solver: orc_swe_destroy removes it, so a later solver at the same address starts with
// no regions.
static std::map<const void*, PolygonBc>& bc_regions() {
  static std::map<const void*, PolygonBc> m;
  return m;
}

void* orc_swe_create(int nx, int ny, double dx, double dy, const frehg_swe_params* p) {
  try {
    return new SweSolver(nx, ny, dx, dy, toSwe(*p));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_swe_destroy(void* h) {
  bc_regions().erase(h);
  delete static_cast<SweSolver*>(h);
}
int orc_swe_set_bathymetry(void* h, const double* bed) {
  return guard([&] {
    auto* s = static_cast<SweSolver*>(h);
    s->setBathymetry(std::vector<real>(bed, bed + (size_t)s->g.nx * s->g.ny));
  });
}
int orc_swe_initialize_state(void* h, double eta0) {
  return guard([&] { static_cast<SweSolver*>(h)->initializeState(eta0); });
}
// SweSolver::setInitialEtaPhysical + setInitialVelocityConstant(0, 0) + finalizeInitialState
// (src/swe/SweSolver.cpp:855-924): a non-uniform initial surface (nx*ny values, i + j*nx).
int orc_swe_set_initial_eta(void* h, const double* eta_phys) {
  return guard([&] {
    auto* s = static_cast<SweSolver*>(h);
    s->setInitialEtaPhysical(std::vector<real>(eta_phys, eta_phys + (size_t)s->g.nx * s->g.ny));
    s->setInitialVelocityConstant(0.0, 0.0);
    s->finalizeInitialState();
  });
}
int orc_swe_get(void* h, const char* name, double* dst) {
  return guard([&] {
    auto* f = sweField(static_cast<SweSolver*>(h), name);
    if (!f) throw std::runtime_error(std::string("unknown SW field ") + name);
    std::memcpy(dst, f->data(), f->size() * sizeof(double));
  });
}
int orc_swe_set(void* h, const char* name, const double* src) {
  return guard([&] {
    auto* f = sweField(static_cast<SweSolver*>(h), name);
    if (!f) throw std::runtime_error(std::string("unknown SW field ") + name);
    std::memcpy(f->data(), src, f->size() * sizeof(double));
  });
}
int orc_swe_step(void* h, double rain, double evap, frehg_solve_stats* st) {
  return guard([&] {
    SolveStats s;
    static_cast<SweSolver*>(h)->advanceStep(rain, evap, &s);
    fillStats(s, st);
  });
}
double orc_swe_max_cfl(void* h) { return static_cast<SweSolver*>(h)->maxCfl(); }

int orc_couple_exchange(void* swe, void* re, const frehg_coupling_params* cp, double dt,
                        double* volume, double* q_out) {
  return guard([&] {
    CouplingParams c{cp->dx, cp->dy, cp->visc, cp->min_depth};
    std::vector<real> q;
    const real v = couple_exchange(c, *static_cast<SweSolver*>(swe), *static_cast<ReSolver*>(re), dt,
                                   q_out ? &q : nullptr);
    if (volume) *volume = v;
    if (q_out) std::memcpy(q_out, q.data(), q.size() * sizeof(double));
  });
}

// state[5] = {t, gw_time, exchange_volume, gw_substeps, boundary outflow}
int orc_swe_set_boundary_polygons(void* h, int nreg, const int32_t* kind, const double* rate, const int32_t* nvert,
                                  const double* xy, double x0, double y0) {
  return guard([&] {
    auto* s = static_cast<SweSolver*>(h);
    PolygonBc bc;
    size_t off = 0;
    for (int r = 0; r < nreg; ++r) {
      BcPolygonRegion reg;
      reg.kind = kind[r];
      reg.rate = rate[r];
      for (int v = 0; v < nvert[r]; ++v, ++off) reg.vertices.push_back({xy[2 * off], xy[2 * off + 1]});
      bc.regions.push_back(reg);
    }
    bc.build(*s, x0, y0);
    if (nreg == 0) bc_regions().erase(h);
    else bc_regions()[h] = bc;
  });
}
int orc_swe_apply_boundaries(void* h, double dt, double* outflow) {
  return guard([&] {
    auto it = bc_regions().find(h);
    const double v = it == bc_regions().end() ? 0.0 : it->second.apply(*static_cast<SweSolver*>(h), dt);
    if (outflow) *outflow = v;
  });
}
int orc_coupled_step(void* swe, void* re, const frehg_coupling_params* cp, double* state, double dt,
                     double rain, double evap, double t_end) {
  return guard([&] {
    CouplingParams c{cp->dx, cp->dy, cp->visc, cp->min_depth};
    CoupledState cs;
    cs.t = state[0];
    cs.gw_time = state[1];
    cs.exchange_volume = state[2];
    cs.gw_substeps = (int)state[3];
    cs.boundary_outflow = state[4];
    auto it = bc_regions().find(swe);
    cs.bc = it == bc_regions().end() ? nullptr : &it->second;
    coupled_step(c, *static_cast<SweSolver*>(swe), *static_cast<ReSolver*>(re), cs, dt, rain, evap, t_end);
    state[0] = cs.t;
    state[1] = cs.gw_time;
    state[2] = cs.exchange_volume;
    state[3] = cs.gw_substeps;
    state[4] = cs.boundary_outflow;
  });
}

int orc_coupled_step_solute(void* swe, void* re, void* sol, const frehg_coupling_params* cp, double* state,
                            double dt, double rain, double evap, double t_end) {
  return guard([&] {
    CouplingParams c{cp->dx, cp->dy, cp->visc, cp->min_depth};
    CoupledState cs;
    cs.t = state[0];
    cs.gw_time = state[1];
    cs.exchange_volume = state[2];
    cs.gw_substeps = (int)state[3];
    cs.boundary_outflow = state[4];
    auto it = bc_regions().find(swe);
    cs.bc = it == bc_regions().end() ? nullptr : &it->second;
    coupled_step_solute(c, *static_cast<SweSolver*>(swe), *static_cast<ReSolver*>(re), cs,
                        static_cast<SoluteOracle*>(sol), dt, rain, evap, t_end);
    state[0] = cs.t;
    state[1] = cs.gw_time;
    state[2] = cs.exchange_volume;
    state[3] = cs.gw_substeps;
    state[4] = cs.boundary_outflow;
  });
}

// ---- solute ----
void* orc_solute_create(void* swe, void* re, const frehg_solute_params* p) {
  try {
    SoluteParams sp;
    sp.c_rain = p->c_rain;
    sp.k_decay = p->k_decay;
    sp.D = p->D;
    sp.cfl_max = p->cfl_max;
    sp.min_depth = p->min_depth;
    sp.start_time = p->start_time;
    sp.muscl = p->muscl;
    sp.implicit_diffusion = p->implicit_diffusion;
    sp.substeps = p->substeps;
    sp.solver = SolverCfg{p->solver.rtol, p->solver.atol, p->solver.max_it, 0};
    return new SoluteOracle(sp, static_cast<SweSolver*>(swe), static_cast<ReSolver*>(re));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_solute_destroy(void* h) { delete static_cast<SoluteOracle*>(h); }
int orc_solute_get(void* h, const char* name, double* dst) {
  return guard([&] {
    auto* s = static_cast<SoluteOracle*>(h);
    const std::vector<real>& v = std::string(name) == "surface" ? s->cs : s->cg;
    if (v.empty()) throw std::runtime_error("solute field absent");
    std::copy(v.begin(), v.end(), dst);
  });
}
int orc_solute_set(void* h, const char* name, const double* src) {
  return guard([&] {
    auto* s = static_cast<SoluteOracle*>(h);
    std::vector<real>& v = std::string(name) == "surface" ? s->cs : s->cg;
    if (v.empty()) throw std::runtime_error("solute field absent");
    std::copy(src, src + v.size(), v.begin());
  });
}
int orc_solute_apply(void* h, double t, double dt, double rain) {
  return guard([&] { static_cast<SoluteOracle*>(h)->apply(t, dt, rain); });
}
// One stepper step with rain 0; out[0] = Courant number, return value FREHG-style (0) with
// out[1] = 1 if accepted.
int orc_solute_step(void* h, double dt, double* out) {
  return guard([&] {
    real cfl = 0.0;
    const bool ok = static_cast<SoluteOracle*>(h)->step(dt, &cfl);
    out[0] = cfl;
    out[1] = ok ? 1.0 : 0.0;
  });
}
int orc_solute_exchange(void* h, const double* q, int n, double dt, double area) {
  return guard([&] { static_cast<SoluteOracle*>(h)->interfaceExchange(std::vector<real>(q, q + n), dt, area); });
}
int orc_solute_mass(void* h, double* mass, frehg_solute_stats* st) {
  return guard([&] {
    auto* s = static_cast<SoluteOracle*>(h);
    if (mass) *mass = s->mass();
    if (st) {
      st->rain_added = s->tot.rain_added;
      st->decay_loss = s->tot.decay_loss;
      st->max_courant = s->tot.max_courant;
      st->substeps = s->tot.substeps;
      st->refused_steps = s->tot.refused_steps;
      st->diffusion_iterations = s->tot.diffusion_iterations;
    }
  });
}

// Stencil solve in the reference's owned ordering (see frehg_ls_solve in frehg_gpu.h).
int orc_ls_solve(int ndim, int nx, int ny, int nz, const double* diag, const double* const* off,
                 const double* b_owned, double* x_owned, const frehg_solver_config* cfg,
                 frehg_solve_stats* st) {
  return guard([&] {
    StencilOp A;
    A.g.nx = nx;
    A.g.ny = ny;
    A.g.nz = ndim == 3 ? nz : 1;
    A.three_d = ndim == 3;
    A.resize();
    const size_t n = A.diag.size();
    std::vector<real> b(n, 0.0), x;
    std::vector<real>* dirs[6] = {&A.xm, &A.xp, &A.ym, &A.yp, &A.zm, &A.zp};
    const int ndir = ndim == 3 ? 6 : 4;
    for (int k = 0; k < A.g.nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t row = ndim == 3 ? ((size_t)j * nx + i) * nz + k : (size_t)j * nx + i;
          const long long c = ndim == 3 ? A.g.idx(i, j, k) : A.g.sidx(i, j);
          A.diag[c] = diag[row];
          b[c] = b_owned[row];
          for (int d = 0; d < ndir; ++d) (*dirs[d])[c] = off[d][row];
        }
    SolveStats s = pcg_solve(A, b, x, toCfg(*cfg));
    fillStats(s, st);
    for (int k = 0; k < A.g.nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t row = ndim == 3 ? ((size_t)j * nx + i) * nz + k : (size_t)j * nx + i;
          x_owned[row] = x[ndim == 3 ? A.g.idx(i, j, k) : A.g.sidx(i, j)];
        }
  });
}

// Batched Van Genuchten evaluation: theta(h), C(h), K(h, Ks_z), h(theta=wc_in), dKdwc(wc_in).
int orc_vg_eval(const frehg_soil_params* sp, int n, const double* h, const double* wc_in,
                double* theta, double* cap, double* K, double* head, double* dkdwc) {
  return guard([&] {
    const Soil s = toSoil(*sp);
    for (int i = 0; i < n; ++i) {
      if (theta) theta[i] = VG::theta(s, h[i]);
      if (cap) cap[i] = VG::capacity(s, h[i]);
      if (K) K[i] = VG::conductivity(s, h[i], s.Ks_z);
      if (head) head[i] = VG::head(s, wc_in[i]);
      if (dkdwc) dkdwc[i] = VG::dKdwc(s, wc_in[i], s.Ks_z);
    }
  });
}

}  // extern "C"
