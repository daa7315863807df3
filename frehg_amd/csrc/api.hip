// C ABI of the gfx950 backend (include/frehg_gpu.h). Every entry point maps exceptions to the
// return codes of the header; the reference raises std::runtime_error for the same conditions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/frehg_gpu.h"
#include "coupling.hpp"
#include "ensemble.hpp"
#include "mlp.hpp"
#include "noalloc.hpp"
#include "re.hpp"
#include "solute.hpp"
#include "swe.hpp"

struct frehg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  frehg::Profiler prof;
  std::unique_ptr<frehg::ReState> re;
  std::unique_ptr<frehg::SweState> swe;
  std::unique_ptr<frehg::Coupler> cpl;
  std::unique_ptr<frehg::Mlp> mlp;
  std::unique_ptr<frehg::Ensemble> ens;
  std::unique_ptr<frehg::Solute> sol;  // refers to re / swe: reset before either changes
  double t = 0.0, gw_time = 0.0;
};

namespace {

template <class Fn>
int guard(frehg_ctx_t ctx, Fn&& fn) {
  if (!ctx) return FREHG_ERR_INVALID_ARG;
  try {
    FREHG_HIP(hipSetDevice(ctx->device));
    return fn();
  } catch (const frehg::ArgError& e) {
    ctx->err = e.what();
    return FREHG_ERR_INVALID_ARG;
  } catch (const frehg::ConvergenceError& e) {
    ctx->err = e.what();
    return FREHG_ERR_NOT_CONVERGED;
  } catch (const frehg::HipError& e) {
    ctx->err = e.what();
    return FREHG_ERR_HIP;
  } catch (const std::bad_alloc&) {
    ctx->err = "out of host memory";
    return FREHG_ERR_OOM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return FREHG_ERR_STATE;
  }
}

frehg::ReState& need_re(frehg_ctx_t c) {
  if (!c->re) throw std::runtime_error("groundwater solver not initialised (call frehg_re_init)");
  return *c->re;
}
frehg::SweState& need_swe(frehg_ctx_t c) {
  if (!c->swe) throw std::runtime_error("surface solver not initialised (call frehg_swe_init)");
  return *c->swe;
}

// Hand the learned-guess model and the surface depth (when the surface grid is the subsurface's
// top) to the Richards solver.
void bind_mlp(frehg_ctx_t c) {
  if (!c->re) return;
  const double* depth = nullptr;
  if (c->swe && c->swe->geo().nx == c->re->geo().nx && c->swe->geo().ny == c->re->geo().ny)
    depth = c->swe->d("dept");
  c->re->set_mlp(c->mlp.get(), depth);
}

frehg::Mlp& need_mlp(frehg_ctx_t c) {
  if (!c->mlp) {
    c->mlp.reset(new frehg::Mlp(c->stream, &c->prof));
    bind_mlp(c);
  }
  return *c->mlp;
}
frehg::Coupler& need_cpl(frehg_ctx_t c) {
  if (!c->cpl) c->cpl.reset(new frehg::Coupler(c->stream));
  return *c->cpl;
}

}  // namespace

extern "C" {

int frehg_abi_version(void) { return FREHG_ABI_VERSION; }

int frehg_create(frehg_ctx_t* out, int device) {
  if (!out) return FREHG_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FREHG_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return FREHG_ERR_INVALID_ARG;
  frehg_ctx_t c = new (std::nothrow) frehg_ctx_s();
  if (!c) return FREHG_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return FREHG_ERR_HIP;
  }
  // A step waits on the stream several times (Krylov convergence, reallocation sweeps, ledger);
  // spin-waiting wakes the host within a microsecond instead of after an interrupt. The flag can
  // only be set before the device is first used in the process; otherwise the current setting is
  // kept (FREHG_SYNC=yield keeps the default).
  const char* sync = std::getenv("FREHG_SYNC");
  if (!(sync && std::string(sync) == "yield")) (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
  (void)hipGetLastError();
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return FREHG_ERR_HIP;
  }
  c->prof.stream = c->stream;
  // the allocation-free solve's scratch address on this device (once per device)
  try {
    frehg::noalloc_setup(device);
  } catch (const std::exception& e) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FREHG_ERR_HIP;
  }
  *out = c;
  return FREHG_OK;
}

void frehg_destroy(frehg_ctx_t ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->sol.reset();
  ctx->ens.reset();
  ctx->mlp.reset();
  ctx->cpl.reset();
  ctx->re.reset();
  ctx->swe.reset();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* frehg_last_error(frehg_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int frehg_synchronize(frehg_ctx_t ctx) {
  return guard(ctx, [&] {
    FREHG_HIP(hipStreamSynchronize(ctx->stream));
    return FREHG_OK;
  });
}

int frehg_re_init(frehg_ctx_t ctx, int nx, int ny, int nz, const frehg_re_params* p) {
  return guard(ctx, [&] {
    if (!p) throw frehg::ArgError("frehg_re_init: null params");
    ctx->sol.reset();  // a solute set up on the old solver goes with it
    ctx->re.reset();
    ctx->re.reset(new frehg::ReState(nx, ny, nz, *p, ctx->stream, &ctx->prof));
    bind_mlp(ctx);
    return FREHG_OK;
  });
}

int frehg_re_set_params(frehg_ctx_t ctx, const frehg_re_params* p) {
  return guard(ctx, [&] {
    if (!p) throw frehg::ArgError("null params");
    need_re(ctx).set_params(*p);
    return FREHG_OK;
  });
}

int frehg_re_set_soil(frehg_ctx_t ctx, const frehg_soil_params* classes, int ncls, const int32_t* class_idx,
                      int idx3d, const double* ks_cell) {
  return guard(ctx, [&] {
    if (!classes) throw frehg::ArgError("null soil classes");
    need_re(ctx).set_soil(classes, ncls, class_idx, idx3d, ks_cell);
    return FREHG_OK;
  });
}

int frehg_re_set_top_flux(frehg_ctx_t ctx, const double* q) {
  return guard(ctx, [&] {
    if (!q) throw frehg::ArgError("null top flux");
    need_re(ctx).set_top_flux(q);
    return FREHG_OK;
  });
}

int frehg_re_set_initial(frehg_ctx_t ctx, const double* h, const double* wc) {
  return guard(ctx, [&] {
    need_re(ctx).set_initial(h, wc);
    return FREHG_OK;
  });
}

int frehg_re_set_field(frehg_ctx_t ctx, const char* name, const double* src) {
  return guard(ctx, [&] {
    if (!name || !src) throw frehg::ArgError("null argument");
    need_re(ctx).set_field(name, src);
    return FREHG_OK;
  });
}

int frehg_re_get_field(frehg_ctx_t ctx, const char* name, double* dst) {
  return guard(ctx, [&] {
    if (!name || !dst) throw frehg::ArgError("null argument");
    need_re(ctx).get_field(name, dst);
    return FREHG_OK;
  });
}

int frehg_re_step(frehg_ctx_t ctx, frehg_re_stats* st) {
  return guard(ctx, [&] {
    frehg_re_stats local;
    need_re(ctx).step(&local);
    if (st) *st = local;
    if (!local.picard_converged) return FREHG_INFO_PICARD_NOT_CONVERGED;
    return local.guess_fallback ? FREHG_INFO_GUESS_FALLBACK : FREHG_OK;
  });
}

double frehg_re_dt(frehg_ctx_t ctx) { return (ctx && ctx->re) ? ctx->re->dt() : 0.0; }

int frehg_re_prime_dt(frehg_ctx_t ctx, double dt) {
  return guard(ctx, [&] {
    need_re(ctx).prime_dt(dt);
    return FREHG_OK;
  });
}

int frehg_swe_init(frehg_ctx_t ctx, int nx, int ny, double dx, double dy, const frehg_swe_params* p) {
  return guard(ctx, [&] {
    if (!p) throw frehg::ArgError("frehg_swe_init: null params");
    ctx->sol.reset();
    ctx->swe.reset();
    bind_mlp(ctx);
    ctx->swe.reset(new frehg::SweState(nx, ny, dx, dy, *p, ctx->stream, &ctx->prof));
    bind_mlp(ctx);
    return FREHG_OK;
  });
}

int frehg_swe_set_params(frehg_ctx_t ctx, const frehg_swe_params* p) {
  return guard(ctx, [&] {
    if (!p) throw frehg::ArgError("null params");
    need_swe(ctx).set_params(*p);
    return FREHG_OK;
  });
}

int frehg_swe_set_bathymetry(frehg_ctx_t ctx, const double* bed) {
  return guard(ctx, [&] {
    if (!bed) throw frehg::ArgError("null bathymetry");
    need_swe(ctx).set_bathymetry(bed);
    return FREHG_OK;
  });
}

int frehg_swe_initialize_state(frehg_ctx_t ctx, double init_eta) {
  return guard(ctx, [&] {
    need_swe(ctx).initialize_state(init_eta);
    return FREHG_OK;
  });
}

int frehg_swe_set_initial_eta(frehg_ctx_t ctx, const double* eta_interior) {
  return guard(ctx, [&] {
    if (!eta_interior) throw frehg::ArgError("null argument");
    need_swe(ctx).initialize_eta(eta_interior, 0.0);
    return FREHG_OK;
  });
}

int frehg_swe_set_boundary_polygons(frehg_ctx_t ctx, int nregions, const int32_t* kind, const double* rate,
                                    const int32_t* nvert, const double* xy, double x0, double y0) {
  return guard(ctx, [&] {
    need_swe(ctx).set_boundary_polygons(nregions, kind, rate, nvert, xy, x0, y0);
    return FREHG_OK;
  });
}

int frehg_swe_apply_boundaries(frehg_ctx_t ctx, double dt, double* outflow) {
  return guard(ctx, [&] {
    if (!(dt > 0)) throw frehg::ArgError("frehg_swe_apply_boundaries: dt must be > 0");
    auto& swe = need_swe(ctx);
    swe.apply_boundaries_async(dt);
    const double v = swe.boundary_outflow();
    if (outflow) *outflow = v;
    return FREHG_OK;
  });
}

int frehg_swe_set_field(frehg_ctx_t ctx, const char* name, const double* src) {
  return guard(ctx, [&] {
    if (!name || !src) throw frehg::ArgError("null argument");
    need_swe(ctx).set_field(name, src);
    return FREHG_OK;
  });
}

int frehg_swe_get_field(frehg_ctx_t ctx, const char* name, double* dst) {
  return guard(ctx, [&] {
    if (!name || !dst) throw frehg::ArgError("null argument");
    need_swe(ctx).get_field(name, dst);
    return FREHG_OK;
  });
}

int frehg_swe_step(frehg_ctx_t ctx, double rain, double evap, frehg_solve_stats* st) {
  return guard(ctx, [&] {
    frehg_solve_stats local;
    need_swe(ctx).step(rain, evap, &local);
    if (st) *st = local;
    return FREHG_OK;
  });
}

int frehg_swe_max_cfl(frehg_ctx_t ctx, double* out) {
  return guard(ctx, [&] {
    if (!out) throw frehg::ArgError("null output");
    *out = need_swe(ctx).max_cfl();
    return FREHG_OK;
  });
}

int frehg_couple_exchange(frehg_ctx_t ctx, const frehg_coupling_params* cp, double dt, double* volume_out,
                          double* q_out) {
  return guard(ctx, [&] {
    if (!cp) throw frehg::ArgError("null coupling params");
    const double v = need_cpl(ctx).exchange(need_swe(ctx), need_re(ctx), *cp, dt, q_out);
    if (volume_out) *volume_out = v;
    return FREHG_OK;
  });
}

int frehg_coupled_step(frehg_ctx_t ctx, const frehg_coupling_params* cp, double dt, double rain, double evap,
                       double t_end, frehg_coupled_stats* st) {
  return guard(ctx, [&] {
    if (!cp) throw frehg::ArgError("null coupling params");
    frehg_coupled_stats local;
    std::memset(&local, 0, sizeof(local));
    auto& swe = need_swe(ctx);
    auto& re = need_re(ctx);
    const double t_new = ctx->t + dt;  // Orchestrator::step, coupled branch (:1322-1336)
    swe.step(rain, evap, &local.sw);
    // surface boundary regions after the surface step (applySurfacePolygons, Orchestrator.cpp:1240)
    const bool bc = swe.has_boundaries();
    if (bc) swe.apply_boundaries_async(dt);
    auto& cpl = need_cpl(ctx);
    const double t0 = ctx->t;
    cpl.set_capture_rates(ctx->sol != nullptr);
    cpl.exchange_async(swe, re, *cp, dt);  // its volume is read after the groundwater step
    // the exchange carries solute with its water (applyCouplingExchange, :1260-1268)
    if (ctx->sol) ctx->sol->interface_exchange(cpl.rates(), dt, cp->dx * cp->dy);
    const double eps = 1.0e-9 * std::max(1.0, t_end);  // gwCatchUpTo (:1243-1252)
    int fallback = 0, iters = 0, unconverged = 0;
    while (ctx->gw_time < t_new - eps) {
      re.step(&local.gw);
      iters += local.gw.solve.iterations;
      fallback |= local.gw.guess_fallback;
      unconverged |= !local.gw.picard_converged;
      ctx->gw_time += re.dt_used();
      ++local.gw_substeps;
    }
    local.gw.solve.iterations = iters;
    if (ctx->sol) ctx->sol->apply(t0, dt, rain);  // applySolute after the flow step (:1333)
    local.exchange_volume = cpl.exchanged_volume();
    local.boundary_outflow = bc ? swe.boundary_outflow() : 0.0;
    ctx->t = t_new;
    local.t = ctx->t;
    local.gw_time = ctx->gw_time;
    if (st) *st = local;
    if (unconverged) return FREHG_INFO_PICARD_NOT_CONVERGED;
    return fallback ? FREHG_INFO_GUESS_FALLBACK : FREHG_OK;
  });
}

int frehg_water_volume(frehg_ctx_t ctx, double* surface, double* subsurface) {
  return guard(ctx, [&] {
    if (!ctx->swe && !ctx->re) throw std::runtime_error("no solver initialised");
    need_cpl(ctx).volume(ctx->swe.get(), ctx->re.get(), surface, subsurface);
    return FREHG_OK;
  });
}

int frehg_ls_solve(frehg_ctx_t ctx, int ndim, int nx, int ny, int nz, const double* diag,
                   const double* const* offdiag, const double* b_owned, double* x_owned,
                   const frehg_solver_config* cfg, frehg_solve_stats* st) {
  return guard(ctx, [&] {
    if (ndim != 2 && ndim != 3) throw frehg::ArgError("frehg_ls_solve: ndim must be 2 or 3");
    if (nx < 1 || ny < 1 || (ndim == 3 && nz < 1)) throw frehg::ArgError("frehg_ls_solve: bad sizes");
    if (!diag || !offdiag || !b_owned || !x_owned || !cfg) throw frehg::ArgError("frehg_ls_solve: null argument");
    const bool three = ndim == 3;
    const int NZ = three ? nz : 1;
    const frehg::Geo g = frehg::make_geo(nx, ny, NZ);
    const long long n = three ? g.ncell() : g.nsurf();
    const int ndir = three ? 6 : 4;
    for (int d = 0; d < ndir; ++d)
      if (!offdiag[d]) throw frehg::ArgError("frehg_ls_solve: null off-diagonal");
    std::vector<double> D((size_t)n, 0.0), ax((size_t)n, 0.0), ay((size_t)n, 0.0), az((size_t)n, 0.0),
        b((size_t)n, 0.0), bn((size_t)n, 0.0), x((size_t)n, 0.0);
    std::vector<unsigned char> fixed((size_t)n, 0);
    std::vector<double> off[6];
    for (int d = 0; d < 6; ++d) off[d].assign((size_t)n, 0.0);
    auto row = [&](int i, int j, int k) -> size_t {
      return three ? ((size_t)j * nx + i) * nz + k : (size_t)j * nx + i;
    };
    auto pidx = [&](int i, int j, int k) -> long long { return three ? g.idx(i, j, k) : g.sidx(i, j); };
    const long long step[6] = {-1, 1, -g.sy, g.sy, -g.sz, g.sz};
    for (int k = 0; k < NZ; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t r = row(i, j, k);
          const long long c = pidx(i, j, k);
          D[(size_t)c] = diag[r];
          b[(size_t)c] = bn[(size_t)c] = b_owned[r];
          const bool inside[6] = {i > 0, i < nx - 1, j > 0, j < ny - 1, three && k > 0, three && k < nz - 1};
          bool all_zero = true;
          for (int d = 0; d < ndir; ++d) {
            off[d][(size_t)c] = inside[d] ? offdiag[d][r] : 0.0;
            if (off[d][(size_t)c] != 0.0) all_zero = false;
          }
          if (!(D[(size_t)c] != 0.0)) throw frehg::ArgError("frehg_ls_solve: zero diagonal (Jacobi needs D != 0)");
          if (!three && all_zero) {
            fixed[(size_t)c] = 1;
            x[(size_t)c] = b[(size_t)c] / D[(size_t)c];
          }
        }
    // Fold decoupled rows into their neighbours. CG needs the remaining couplings symmetric;
    // BiCGSTAB keeps each row's own -x/-y/-z couplings.
    const frehg::KrylovConfig kc = frehg::krylov_config(*cfg, "frehg_ls_solve");
    const bool nonsym = kc.ksp == FREHG_KSP_BICGSTAB;
    std::vector<double> axm, aym, azm;
    if (nonsym) {
      axm.assign((size_t)n, 0.0);
      aym.assign((size_t)n, 0.0);
      azm.assign((size_t)n, 0.0);
    }
    std::vector<double>* dst[6] = {&axm, &ax, &aym, &ay, &azm, &az};
    for (int k = 0; k < NZ; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long long c = pidx(i, j, k);
          if (fixed[(size_t)c]) continue;
          for (int d = 0; d < ndir; ++d) {
            const double a = off[d][(size_t)c];
            if (a == 0.0) continue;
            const long long nb = c + step[d];
            if (fixed[(size_t)nb]) {
              b[(size_t)c] -= a * x[(size_t)nb];
              continue;
            }
            if (nonsym) {
              (*dst[d])[(size_t)c] = a;
            } else if (d & 1) {
              const double back = off[d - 1][(size_t)nb];
              if (std::fabs(a - back) > 1e-12 * (std::fabs(a) + std::fabs(back)))
                throw frehg::ArgError("frehg_ls_solve: operator is not symmetric (CG requires A = A^T)");
              (*dst[d])[(size_t)c] = a;
            }
          }
        }
    auto& s = ctx->stream;
    double *dD, *dax, *day, *daz = nullptr, *db, *dbn, *dx, *daxm = nullptr, *daym = nullptr, *dazm = nullptr;
    unsigned char* dfixed = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    std::vector<void*> allocs;
    auto alloc = [&](double** p) {
      FREHG_HIP(hipMalloc(p, bytes));
      allocs.push_back(*p);
    };
    struct Free {
      std::vector<void*>& a;
      ~Free() {
        for (void* p : a) (void)hipFree(p);
      }
    } freer{allocs};
    alloc(&dD);
    alloc(&dax);
    alloc(&day);
    if (three) alloc(&daz);
    alloc(&db);
    alloc(&dbn);
    alloc(&dx);
    FREHG_HIP(hipMemcpy(dD, D.data(), bytes, hipMemcpyHostToDevice));
    FREHG_HIP(hipMemcpy(dax, ax.data(), bytes, hipMemcpyHostToDevice));
    FREHG_HIP(hipMemcpy(day, ay.data(), bytes, hipMemcpyHostToDevice));
    if (three) FREHG_HIP(hipMemcpy(daz, az.data(), bytes, hipMemcpyHostToDevice));
    FREHG_HIP(hipMemcpy(db, b.data(), bytes, hipMemcpyHostToDevice));
    FREHG_HIP(hipMemcpy(dbn, bn.data(), bytes, hipMemcpyHostToDevice));
    FREHG_HIP(hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice));
    if (nonsym) {
      alloc(&daxm);
      alloc(&daym);
      FREHG_HIP(hipMemcpy(daxm, axm.data(), bytes, hipMemcpyHostToDevice));
      FREHG_HIP(hipMemcpy(daym, aym.data(), bytes, hipMemcpyHostToDevice));
      if (three) {
        alloc(&dazm);
        FREHG_HIP(hipMemcpy(dazm, azm.data(), bytes, hipMemcpyHostToDevice));
      }
    }
    if (!three) {
      FREHG_HIP(hipMalloc(&dfixed, (size_t)n));
      allocs.push_back(dfixed);
      FREHG_HIP(hipMemcpy(dfixed, fixed.data(), (size_t)n, hipMemcpyHostToDevice));
    }
    frehg::Krylov kry(g, three, s, &ctx->prof);
    frehg::StencilView<double> A{dD, dax, day, daz, daxm, daym, dazm};
    const frehg::SolveResult sr = kry.solve(A, nullptr, db, dbn, dx, dfixed, true, false, kc);
    FREHG_HIP(hipMemcpy(x.data(), dx, bytes, hipMemcpyDeviceToHost));
    for (int k = 0; k < NZ; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) x_owned[row(i, j, k)] = x[(size_t)pidx(i, j, k)];
    if (st) {
      std::memset(st, 0, sizeof(*st));
      st->iterations = sr.iterations;
      st->residual_norm = sr.residual_norm;
      st->rhs_norm = sr.rhs_norm;
      st->converged = sr.converged ? 1 : 0;
    }
    return FREHG_OK;
  });
}

int frehg_ls_solve_noalloc(frehg_ctx_t ctx, int ndim, int nx, int ny, int nz, const double* diag,
                           const double* const* offdiag, const double* b_owned, double* x_owned,
                           const frehg_solver_config* cfg, frehg_solve_stats* st) {
  return guard(ctx, [&] {
    if (!cfg) throw frehg::ArgError("frehg_ls_solve_noalloc: null argument");
    frehg::ls_solve_noalloc(ctx->device, ctx->stream, ndim, nx, ny, nz, diag, offdiag, b_owned, x_owned, *cfg, st);
    return FREHG_OK;
  });
}

int frehg_profile_enable(frehg_ctx_t ctx, int enable) {
  return guard(ctx, [&] {
    ctx->prof.enabled = enable != 0;
    return FREHG_OK;
  });
}

int frehg_profile_select(frehg_ctx_t ctx, uint32_t kernel_mask) {
  return guard(ctx, [&] {
    ctx->prof.mask = kernel_mask;
    return FREHG_OK;
  });
}

int frehg_profile_read(frehg_ctx_t ctx, int kt, double* total_ms, int64_t* launches) {
  return guard(ctx, [&] {
    if (kt < 0 || kt >= FREHG_KT_COUNT) throw frehg::ArgError("bad kernel type");
    long long n = 0;
    double ms = 0.0;
    ctx->prof.read(kt, &ms, &n);
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
    return FREHG_OK;
  });
}

int frehg_profile_reset(frehg_ctx_t ctx) {
  return guard(ctx, [&] {
    ctx->prof.reset();
    return FREHG_OK;
  });
}

int frehg_spmv_shape(frehg_ctx_t ctx, int precision, int* cells_per_lane, int* kchunk) {
  return guard(ctx, [&] {
    frehg::ReState& re = need_re(ctx);
    if (precision != FREHG_PREC_FP64 && precision != FREHG_PREC_MIXED)
      throw frehg::ArgError("frehg_spmv_shape: precision must be FREHG_PREC_FP64 or FREHG_PREC_MIXED");
    const frehg::Krylov::SpmvShape& sh = re.krylov().spmv_shape(precision == FREHG_PREC_MIXED ? 1 : 0);
    if (cells_per_lane) *cells_per_lane = sh.cpl;
    if (kchunk) *kchunk = sh.kc;
    return FREHG_OK;
  });
}

frehg::Ensemble& need_ens(frehg_ctx_t c) {
  if (!c->ens) throw std::runtime_error("ensemble not initialised (call frehg_ens_init)");
  return *c->ens;
}

int frehg_ens_init(frehg_ctx_t ctx, int nboxes, const int32_t* dims, const frehg_re_params* params) {
  return guard(ctx, [&] {
    ctx->ens.reset();
    ctx->ens.reset(new frehg::Ensemble(nboxes, dims, params, ctx->stream, &ctx->prof));
    return FREHG_OK;
  });
}

int frehg_ens_set_soil(frehg_ctx_t ctx, int box, const frehg_soil_params* classes, int ncls, const int32_t* class_idx,
                       int idx3d, const double* ks_cell) {
  return guard(ctx, [&] {
    need_ens(ctx).set_soil(box, classes, ncls, class_idx, idx3d, ks_cell);
    return FREHG_OK;
  });
}

int frehg_ens_set_initial(frehg_ctx_t ctx, int box, const double* h, const double* wc) {
  return guard(ctx, [&] {
    need_ens(ctx).set_initial(box, h, wc);
    return FREHG_OK;
  });
}

int frehg_ens_get_field(frehg_ctx_t ctx, int box, const char* name, double* dst) {
  return guard(ctx, [&] {
    if (!name || !dst) throw frehg::ArgError("null argument");
    need_ens(ctx).get_field(box, name, dst);
    return FREHG_OK;
  });
}

int frehg_ens_step(frehg_ctx_t ctx, int use_graph, frehg_ens_stats* stats, frehg_re_stats* per_box) {
  return guard(ctx, [&] {
    need_ens(ctx).step(use_graph, stats, per_box);
    return FREHG_OK;
  });
}

int frehg_mlp_init(frehg_ctx_t ctx, uint64_t seed) {
  return guard(ctx, [&] {
    need_mlp(ctx).init(seed);
    return FREHG_OK;
  });
}

int frehg_mlp_log(frehg_ctx_t ctx, int stride, int64_t capacity) {
  return guard(ctx, [&] {
    need_mlp(ctx).set_logging(stride, capacity);
    return FREHG_OK;
  });
}

int frehg_mlp_logged(frehg_ctx_t ctx, int64_t* n) {
  return guard(ctx, [&] {
    if (!n) throw frehg::ArgError("frehg_mlp_logged: null output");
    *n = need_mlp(ctx).logged();
    return FREHG_OK;
  });
}

int frehg_mlp_train(frehg_ctx_t ctx, int steps, int batch, double lr, uint64_t seed, double* loss_first,
                    double* loss_last) {
  return guard(ctx, [&] {
    const double l = need_mlp(ctx).train(steps, batch, lr, seed, loss_first);
    if (loss_last) *loss_last = l;
    return FREHG_OK;
  });
}

int frehg_mlp_get_params(frehg_ctx_t ctx, float* params, float* norm) {
  return guard(ctx, [&] {
    need_mlp(ctx).get_params(params, norm);
    return FREHG_OK;
  });
}

int frehg_mlp_set_params(frehg_ctx_t ctx, const float* params, const float* norm) {
  return guard(ctx, [&] {
    need_mlp(ctx).set_params(params, norm);
    return FREHG_OK;
  });
}

int frehg_mlp_loss_grad(frehg_ctx_t ctx, const float* x, const float* y, int n, float* grad, double* loss) {
  return guard(ctx, [&] {
    const double l = need_mlp(ctx).loss_grad(x, y, n, grad);
    if (loss) *loss = l;
    return FREHG_OK;
  });
}

int frehg_mlp_gate(frehg_ctx_t ctx, double gate) {
  return guard(ctx, [&] {
    if (!(gate >= 0.0)) throw frehg::ArgError("frehg_mlp_gate: gate must be >= 0");
    need_mlp(ctx).set_gate(gate);
    return FREHG_OK;
  });
}

int frehg_mlp_inferred_cells(frehg_ctx_t ctx, int64_t* cells) {
  return guard(ctx, [&] {
    if (!cells) throw frehg::ArgError("frehg_mlp_inferred_cells: null output");
    *cells = need_mlp(ctx).inferred_cells();
    return FREHG_OK;
  });
}

int frehg_mlp_increment(frehg_ctx_t ctx, double* dst) {
  return guard(ctx, [&] {
    if (!dst) throw frehg::ArgError("frehg_mlp_increment: null output");
    need_mlp(ctx);
    need_re(ctx).learned_increment(dst);
    return FREHG_OK;
  });
}

int frehg_mlp_start(frehg_ctx_t ctx, double* dst) {
  return guard(ctx, [&] {
    if (!dst) throw frehg::ArgError("frehg_mlp_start: null output");
    need_mlp(ctx);
    need_re(ctx).learned_start(dst);
    return FREHG_OK;
  });
}

int frehg_mlp_predict(frehg_ctx_t ctx, const float* x, int n, float* out) {
  return guard(ctx, [&] {
    need_mlp(ctx).predict_batch(x, n, out);
    return FREHG_OK;
  });
}

}  // extern "C"

// ---- solute ----------------------------------------------------------------------------------
extern "C" {

int frehg_solute_init(frehg_ctx_t ctx, const frehg_solute_params* p) {
  return guard(ctx, [&] {
    if (!p) throw frehg::ArgError("frehg_solute_init: null params");
    ctx->sol.reset();
    ctx->sol.reset(new frehg::Solute(*p, ctx->swe.get(), ctx->re.get(), ctx->stream, &ctx->prof));
    return FREHG_OK;
  });
}

int frehg_solute_set_initial(frehg_ctx_t ctx, const double* surface_interior, const double* subsurface_interior) {
  return guard(ctx, [&] {
    if (!ctx->sol) throw std::runtime_error("frehg_solute_set_initial: call frehg_solute_init first");
    ctx->sol->set_initial(surface_interior, subsurface_interior);
    return FREHG_OK;
  });
}

int frehg_solute_set_field(frehg_ctx_t ctx, const char* name, const double* src) {
  return guard(ctx, [&] {
    if (!ctx->sol) throw std::runtime_error("frehg_solute_set_field: call frehg_solute_init first");
    if (!name || !src) throw frehg::ArgError("null argument");
    ctx->sol->set_field(name, src);
    return FREHG_OK;
  });
}

int frehg_solute_get_field(frehg_ctx_t ctx, const char* name, double* dst) {
  return guard(ctx, [&] {
    if (!ctx->sol) throw std::runtime_error("frehg_solute_get_field: call frehg_solute_init first");
    if (!name || !dst) throw frehg::ArgError("null argument");
    ctx->sol->get_field(name, dst);
    return FREHG_OK;
  });
}

int frehg_solute_apply(frehg_ctx_t ctx, double t, double dt, double rain) {
  return guard(ctx, [&] {
    if (!ctx->sol) throw std::runtime_error("frehg_solute_apply: call frehg_solute_init first");
    ctx->sol->apply(t, dt, rain);
    return FREHG_OK;
  });
}

int frehg_solute_mass(frehg_ctx_t ctx, double* mass, frehg_solute_stats* totals) {
  return guard(ctx, [&] {
    if (!ctx->sol) throw std::runtime_error("frehg_solute_mass: call frehg_solute_init first");
    if (mass) *mass = ctx->sol->mass();
    if (totals) *totals = ctx->sol->totals();
    return FREHG_OK;
  });
}

}  // extern "C"

// ---- checkpoint / restart ------------------------------------------------------------------
namespace {

const char* const kSweStateNames[] = {"eta", "etan", "bottom", "bottomXP", "bottomYP", "dept", "deptx", "depty",
                                      "uu", "vv", "un", "vn", "uy", "vx", "Vs", "Vsn", "Asx", "Asy", "Asz",
                                      "Aszx", "Aszy", "Vsx", "Vsy", "CDx", "CDy", "Dx", "Dy", "Ex", "Ey",
                                      "cflx", "cfly", "Fu", "Fv", "Sxp", "Sxm", "Syp", "Sym", "Sct", "Srhs",
                                      "cfl_active", "wtfx", "wtfy"};
const char* const kReStateNames[] = {"h", "hn", "wc", "wcn", "wch", "hwc", "ch", "Kx", "Ky", "Kz",
                                     "qx", "qy", "qz", "Gct", "Grhs"};
constexpr char kCkptMagic[8] = {'F', 'R', 'H', 'G', 'C', 'K', 'P', '1'};
constexpr int32_t kCkptVersion = 1;

uint64_t fnv1a(const void* p, size_t n, uint64_t h) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

uint64_t params_digest(frehg_ctx_t ctx) {
  uint64_t h = 1469598103934665603ull;
  if (ctx->re) {
    frehg_re_params p = ctx->re->params();
    p.dt = 0.0;  // the adaptive dt is state, not a parameter
    h = fnv1a(&p, sizeof p, h);
  }
  if (ctx->swe) h = fnv1a(&ctx->swe->params(), sizeof(frehg_swe_params), h);
  return h;
}

struct File {
  FILE* f;
  explicit File(FILE* ff) : f(ff) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

void put(FILE* f, const void* p, size_t n, const std::string& path) {
  if (n && std::fwrite(p, 1, n, f) != n) throw frehg::ArgError("checkpoint: write failed on '" + path + "'");
}
void get(FILE* f, void* p, size_t n, const std::string& path) {
  if (n && std::fread(p, 1, n, f) != n) throw frehg::ArgError("checkpoint: '" + path + "' is truncated");
}

}  // namespace

extern "C" {

int frehg_checkpoint_save(frehg_ctx_t ctx, const char* path) {
  return guard(ctx, [&] {
    if (!path || !*path) throw frehg::ArgError("frehg_checkpoint_save: empty path");
    if (!ctx->re && !ctx->swe) throw frehg::ArgError("frehg_checkpoint_save: no solver to save");
    const std::string final_path = path, tmp = final_path + ".tmp";
    {
      File f(std::fopen(tmp.c_str(), "wb"));
      if (!f.f) throw frehg::ArgError("frehg_checkpoint_save: cannot create '" + tmp + "'");
      const frehg::Geo& g = ctx->re ? ctx->re->geo() : ctx->swe->geo();
      const int32_t head[6] = {kCkptVersion, ctx->swe ? 1 : 0, ctx->re ? 1 : 0, g.nx, g.ny, ctx->re ? g.nz : 1};
      const double times[3] = {ctx->t, ctx->gw_time, ctx->re ? ctx->re->dt() : 0.0};
      const uint64_t digest = params_digest(ctx);
      put(f.f, kCkptMagic, sizeof kCkptMagic, tmp);
      put(f.f, head, sizeof head, tmp);
      put(f.f, times, sizeof times, tmp);
      put(f.f, &digest, sizeof digest, tmp);
      const bool sol_s = ctx->sol && ctx->swe, sol_g = ctx->sol && ctx->re;
      const int32_t nfields = (ctx->swe ? (int32_t)(sizeof kSweStateNames / sizeof(char*)) : 0) +
                              (ctx->re ? (int32_t)(sizeof kReStateNames / sizeof(char*)) : 0) + (sol_s ? 1 : 0) +
                              (sol_g ? 1 : 0);
      put(f.f, &nfields, sizeof nfields, tmp);
      auto field = [&](const char* prefix, const char* name, size_t count, auto&& read) {
        const std::string key = std::string(prefix) + name;
        const int32_t len = (int32_t)key.size();
        const int64_t cnt = (int64_t)count;
        std::vector<double> buf(count);
        read(name, buf.data());
        put(f.f, &len, sizeof len, tmp);
        put(f.f, key.data(), key.size(), tmp);
        put(f.f, &cnt, sizeof cnt, tmp);
        put(f.f, buf.data(), count * sizeof(double), tmp);
      };
      const size_t n2 = (size_t)(g.nx + 2) * (g.ny + 2);
      if (ctx->swe)
        for (const char* nm : kSweStateNames)
          field("sw_", nm, n2, [&](const char* n, double* d) { ctx->swe->get_field(n, d); });
      if (ctx->re)
        for (const char* nm : kReStateNames)
          field("gw_", nm, n2 * (size_t)(g.nz + 2), [&](const char* n, double* d) { ctx->re->get_field(n, d); });
      // solute_sw_conc / solute_gw_conc of writeCheckpointNow (:1461-1502)
      if (sol_s) field("sl_", "surface", n2, [&](const char* n, double* d) { ctx->sol->get_field(n, d); });
      if (sol_g)
        field("sl_", "subsurface", n2 * (size_t)(g.nz + 2), [&](const char* n, double* d) { ctx->sol->get_field(n, d); });
      if (std::fflush(f.f) != 0) throw frehg::ArgError("frehg_checkpoint_save: write failed on '" + tmp + "'");
    }
    if (std::rename(tmp.c_str(), final_path.c_str()) != 0)
      throw frehg::ArgError("frehg_checkpoint_save: cannot rename '" + tmp + "' to '" + final_path + "'");
    return FREHG_OK;
  });
}

int frehg_checkpoint_load(frehg_ctx_t ctx, const char* path, frehg_checkpoint_info* info) {
  return guard(ctx, [&] {
    if (!path || !*path) throw frehg::ArgError("frehg_checkpoint_load: empty path");
    if (!ctx->re && !ctx->swe) throw frehg::ArgError("frehg_checkpoint_load: set up the solvers first");
    const std::string p = path;
    File f(std::fopen(path, "rb"));
    if (!f.f) throw frehg::ArgError("frehg_checkpoint_load: cannot open '" + p + "'");
    char magic[8];
    int32_t head[6];
    double times[3];
    uint64_t digest = 0;
    int32_t nfields = 0;
    get(f.f, magic, sizeof magic, p);
    if (std::memcmp(magic, kCkptMagic, sizeof magic) != 0)
      throw frehg::ArgError("frehg_checkpoint_load: '" + p + "' is not a checkpoint of this library");
    get(f.f, head, sizeof head, p);
    if (head[0] != kCkptVersion) throw frehg::ArgError("frehg_checkpoint_load: unsupported checkpoint version");
    get(f.f, times, sizeof times, p);
    get(f.f, &digest, sizeof digest, p);
    get(f.f, &nfields, sizeof nfields, p);
    if ((head[1] != 0) != (ctx->swe != nullptr))
      throw frehg::ArgError(head[1] ? "frehg_checkpoint_load: the checkpoint has surface-water state, the context none"
                                    : "frehg_checkpoint_load: checkpoint has no SW state");
    if ((head[2] != 0) != (ctx->re != nullptr))
      throw frehg::ArgError(head[2] ? "frehg_checkpoint_load: the checkpoint has groundwater state, the context none"
                                    : "frehg_checkpoint_load: checkpoint has no GW state");
    const frehg::Geo& g = ctx->re ? ctx->re->geo() : ctx->swe->geo();
    if (head[3] != g.nx || head[4] != g.ny || head[5] != (ctx->re ? g.nz : 1))
      throw frehg::ArgError("frehg_checkpoint_load: grid differs from the checkpoint's (" + std::to_string(head[3]) + "x" +
                            std::to_string(head[4]) + "x" + std::to_string(head[5]) + ")");
    // read everything first: a bad file leaves the context unchanged
    std::vector<std::pair<std::string, std::vector<double>>> fields;
    for (int32_t q = 0; q < nfields; ++q) {
      int32_t len = 0;
      int64_t cnt = 0;
      get(f.f, &len, sizeof len, p);
      if (len <= 0 || len > 64) throw frehg::ArgError("frehg_checkpoint_load: '" + p + "' is corrupt");
      std::string key((size_t)len, '\0');
      get(f.f, &key[0], (size_t)len, p);
      get(f.f, &cnt, sizeof cnt, p);
      const size_t n2 = (size_t)(g.nx + 2) * (g.ny + 2);
      auto known = [&](const char* prefix, const char* const* names, size_t n) {
        if (key.rfind(prefix, 0) != 0) return false;
        for (size_t q2 = 0; q2 < n; ++q2)
          if (key.compare(3, std::string::npos, names[q2]) == 0) return true;
        return false;
      };
      const char* const kSolNames[] = {"surface", "subsurface"};
      if (!known("sw_", kSweStateNames, sizeof kSweStateNames / sizeof(char*)) &&
          !known("gw_", kReStateNames, sizeof kReStateNames / sizeof(char*)) && !known("sl_", kSolNames, 2))
        throw frehg::ArgError("frehg_checkpoint_load: unknown field '" + key + "' in '" + p + "'");
      const bool three = key.rfind("gw_", 0) == 0 || key == "sl_subsurface";
      const size_t want = three ? n2 * (size_t)(g.nz + 2) : n2;
      if (cnt != (int64_t)want) throw frehg::ArgError("frehg_checkpoint_load: field '" + key + "' size mismatch");
      std::vector<double> buf((size_t)cnt);
      get(f.f, buf.data(), buf.size() * sizeof(double), p);
      fields.emplace_back(std::move(key), std::move(buf));
    }
    auto has = [&](const std::string& k) {
      for (auto& kv : fields)
        if (kv.first == k) return true;
      return false;
    };
    if (ctx->swe)
      for (const char* nm : kSweStateNames)
        if (!has(std::string("sw_") + nm))
          throw frehg::ArgError(std::string("frehg_checkpoint_load: missing field 'sw_") + nm + "'");
    if (ctx->re)
      for (const char* nm : kReStateNames)
        if (!has(std::string("gw_") + nm))
          throw frehg::ArgError(std::string("frehg_checkpoint_load: missing field 'gw_") + nm + "'");
    const bool has_solute = has("sl_surface") || has("sl_subsurface");
    if (ctx->sol) {  // loadCheckpointState (:722-737)
      if (!has_solute) throw frehg::ArgError("frehg_checkpoint_load: solute set up but the checkpoint has no solute state");
      if (ctx->swe && !has("sl_surface")) throw frehg::ArgError("frehg_checkpoint_load: missing field 'sl_surface'");
      if (ctx->re && !has("sl_subsurface")) throw frehg::ArgError("frehg_checkpoint_load: missing field 'sl_subsurface'");
    }
    for (auto& kv : fields) {
      if (kv.first.rfind("sw_", 0) == 0) ctx->swe->set_field(kv.first.substr(3), kv.second.data());
      else if (kv.first.rfind("gw_", 0) == 0) ctx->re->set_field(kv.first.substr(3), kv.second.data());
      else if (ctx->sol) ctx->sol->set_field(kv.first.substr(3), kv.second.data());
    }
    if (ctx->sol) ctx->sol->reset_totals();
    ctx->t = times[0];
    ctx->gw_time = times[1];
    if (ctx->re) ctx->re->prime_dt(times[2]);
    if (info) {
      info->t = times[0];
      info->gw_time = times[1];
      info->gw_dt = times[2];
      info->has_sw = head[1];
      info->has_gw = head[2];
      info->nx = head[3];
      info->ny = head[4];
      info->nz = head[5];
      info->params_match = digest == params_digest(ctx) ? 1 : 0;
      info->has_solute = has_solute ? 1 : 0;
    }
    return FREHG_OK;
  });
}

}  // extern "C"

// Used by the configuration driver (driver.cpp), which runs above the C ABI.
namespace frehg {
int set_ctx_error(frehg_ctx_t ctx, int rc, const char* msg) {
  if (ctx) ctx->err = msg ? msg : "";
  return rc;
}
bool ctx_has_solvers(frehg_ctx_t ctx) { return ctx && (ctx->re || ctx->swe || ctx->ens); }
}  // namespace frehg
