// FP64 CPU oracle of the passive solute (test infrastructure: the checker of the HIP solute and
// part of the CPU baseline). Restates src/solute/Advection.cpp, Diffusion.cpp, SourceSink.cpp,
// SoluteStepper.cpp and the solute steps of src/driver/Orchestrator.cpp on the oracle's own
// SweSolver / ReSolver fields.
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace frehg_oracle {

namespace {

real minmod(real a, real b) {  // Advection.cpp:10-13
  if (a * b <= 0.0) return 0.0;
  return (std::fabs(a) < std::fabs(b)) ? a : b;
}

real faceValue(bool muscl, real uf, real cLL, real cL, real cR, real cRR) {  // Advection.cpp:16-27
  if (uf > 0.0) {
    if (!muscl) return cL;
    return cL + 0.5 * minmod(cL - cLL, cR - cL);
  }
  if (uf < 0.0) {
    if (!muscl) return cR;
    return cR - 0.5 * minmod(cRR - cR, cR - cL);
  }
  return 0.0;
}

// advectSurface (Advection.cpp:31-101): returns the max Courant number; refuses above cfl_max.
real advectSurface(std::vector<real>& conc, const SweSolver& swe, const SoluteParams& p, real dt) {
  const int nx = swe.g.nx, ny = swe.g.ny;
  const real dx = swe.dx, dy = swe.dy;
  const Grid& g = swe.g;
  const std::vector<real>& u = swe.uu;
  const std::vector<real>& v = swe.vv;
  real cfl = 0.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const real uE = (i + 1 < nx) ? u[g.sidx(i, j)] : 0.0;
      const real uW = (i - 1 >= 0) ? u[g.sidx(i - 1, j)] : 0.0;
      const real vN = (j + 1 < ny) ? v[g.sidx(i, j)] : 0.0;
      const real vS = (j - 1 >= 0) ? v[g.sidx(i, j - 1)] : 0.0;
      const real out_x = std::max(uE, 0.0) + std::max(-uW, 0.0);
      const real out_y = std::max(vN, 0.0) + std::max(-vS, 0.0);
      cfl = std::max(cfl, out_x * dt / dx + out_y * dt / dy);
    }
  if (cfl > p.cfl_max) return cfl;
  const bool muscl = p.muscl != 0;
  std::vector<real> cold((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) cold[(size_t)(i + j * nx)] = conc[g.sidx(i, j)];
  auto C = [&](int ii, int jj) { return cold[(size_t)(ii + jj * nx)]; };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      real Fe = 0.0, Fw = 0.0, Fn = 0.0, Fs = 0.0;
      if (i + 1 < nx) {
        const real uf = u[g.sidx(i, j)];
        Fe = uf * faceValue(muscl, uf, (i - 1 >= 0) ? C(i - 1, j) : C(i, j), C(i, j), C(i + 1, j),
                            (i + 2 < nx) ? C(i + 2, j) : C(i + 1, j));
      }
      if (i - 1 >= 0) {
        const real uf = u[g.sidx(i - 1, j)];
        Fw = uf * faceValue(muscl, uf, (i - 2 >= 0) ? C(i - 2, j) : C(i - 1, j), C(i - 1, j), C(i, j),
                            (i + 1 < nx) ? C(i + 1, j) : C(i, j));
      }
      if (j + 1 < ny) {
        const real uf = v[g.sidx(i, j)];
        Fn = uf * faceValue(muscl, uf, (j - 1 >= 0) ? C(i, j - 1) : C(i, j), C(i, j), C(i, j + 1),
                            (j + 2 < ny) ? C(i, j + 2) : C(i, j + 1));
      }
      if (j - 1 >= 0) {
        const real uf = v[g.sidx(i, j - 1)];
        Fs = uf * faceValue(muscl, uf, (j - 2 >= 0) ? C(i, j - 2) : C(i, j - 1), C(i, j - 1), C(i, j),
                            (j + 1 < ny) ? C(i, j + 1) : C(i, j));
      }
      conc[g.sidx(i, j)] = C(i, j) - dt / dx * (Fe - Fw) - dt / dy * (Fn - Fs);
    }
  return cfl;
}

// advectSubsurface (Advection.cpp:103-195) with the Darcy fluxes qx, qy, qz of the RE step.
real advectSubsurface(std::vector<real>& conc, const ReSolver& re, const SoluteParams& p, real dt) {
  const Grid& g = re.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real dx = re.params().dx, dy = re.params().dy, dz = re.params().dz;
  real cfl = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const real qE = (i + 1 < nx) ? re.qx[g.idx(i, j, k)] : 0.0;
        const real qW = (i - 1 >= 0) ? re.qx[g.idx(i - 1, j, k)] : 0.0;
        const real qN = (j + 1 < ny) ? re.qy[g.idx(i, j, k)] : 0.0;
        const real qS = (j - 1 >= 0) ? re.qy[g.idx(i, j - 1, k)] : 0.0;
        const real qD = (k + 1 < nz) ? re.qz[g.idx(i, j, k)] : 0.0;
        const real qU = (k - 1 >= 0) ? re.qz[g.idx(i, j, k - 1)] : 0.0;
        const real ox = std::max(qE, 0.0) + std::max(-qW, 0.0);
        const real oy = std::max(qN, 0.0) + std::max(-qS, 0.0);
        const real oz = std::max(qD, 0.0) + std::max(-qU, 0.0);
        cfl = std::max(cfl, ox * dt / dx + oy * dt / dy + oz * dt / dz);
      }
  if (cfl > p.cfl_max) return cfl;
  const bool muscl = p.muscl != 0;
  std::vector<real> cold((size_t)nx * ny * nz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) cold[(size_t)(i + j * nx + k * nx * ny)] = conc[g.idx(i, j, k)];
  auto C = [&](int ii, int jj, int kk) { return cold[(size_t)(ii + jj * nx + kk * nx * ny)]; };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        real Fe = 0.0, Fw = 0.0, Fn = 0.0, Fs = 0.0, Fd = 0.0, Fu = 0.0;
        if (i + 1 < nx) {
          const real uf = re.qx[g.idx(i, j, k)];
          Fe = uf * faceValue(muscl, uf, (i - 1 >= 0) ? C(i - 1, j, k) : C(i, j, k), C(i, j, k), C(i + 1, j, k),
                              (i + 2 < nx) ? C(i + 2, j, k) : C(i + 1, j, k));
        }
        if (i - 1 >= 0) {
          const real uf = re.qx[g.idx(i - 1, j, k)];
          Fw = uf * faceValue(muscl, uf, (i - 2 >= 0) ? C(i - 2, j, k) : C(i - 1, j, k), C(i - 1, j, k), C(i, j, k),
                              (i + 1 < nx) ? C(i + 1, j, k) : C(i, j, k));
        }
        if (j + 1 < ny) {
          const real uf = re.qy[g.idx(i, j, k)];
          Fn = uf * faceValue(muscl, uf, (j - 1 >= 0) ? C(i, j - 1, k) : C(i, j, k), C(i, j, k), C(i, j + 1, k),
                              (j + 2 < ny) ? C(i, j + 2, k) : C(i, j + 1, k));
        }
        if (j - 1 >= 0) {
          const real uf = re.qy[g.idx(i, j - 1, k)];
          Fs = uf * faceValue(muscl, uf, (j - 2 >= 0) ? C(i, j - 2, k) : C(i, j - 1, k), C(i, j - 1, k), C(i, j, k),
                              (j + 1 < ny) ? C(i, j + 1, k) : C(i, j, k));
        }
        if (k + 1 < nz) {
          const real uf = re.qz[g.idx(i, j, k)];
          Fd = uf * faceValue(muscl, uf, (k - 1 >= 0) ? C(i, j, k - 1) : C(i, j, k), C(i, j, k), C(i, j, k + 1),
                              (k + 2 < nz) ? C(i, j, k + 2) : C(i, j, k + 1));
        }
        if (k - 1 >= 0) {
          const real uf = re.qz[g.idx(i, j, k - 1)];
          Fu = uf * faceValue(muscl, uf, (k - 2 >= 0) ? C(i, j, k - 2) : C(i, j, k - 1), C(i, j, k - 1), C(i, j, k),
                              (k + 1 < nz) ? C(i, j, k + 1) : C(i, j, k));
        }
        conc[g.idx(i, j, k)] =
            C(i, j, k) - dt / dx * (Fe - Fw) - dt / dy * (Fn - Fs) - dt / dz * (Fd - Fu);
      }
  return cfl;
}

// DiffusionSolver::solve2D / solve3D (Diffusion.cpp:30-179): (I - dt D L) C' = C, Neumann walls,
// Jacobi-CG from x0 = 0.
int diffuse(std::vector<real>& conc, const Grid& grid, bool three, real dx, real dy, real dz, real D, real dt,
            const SolverCfg& cfg) {
  if (D <= 0.0 || dt <= 0.0) return 0;
  StencilOp A;
  A.g = grid;
  A.three_d = three;
  A.resize();
  const real cx = D * dt / (dx * dx), cy = D * dt / (dy * dy), cz = D * dt / (dz * dz);
  const int nx = grid.nx, ny = grid.ny, nz = three ? grid.nz : 1;
  std::vector<real> b(conc.size(), 0.0), x(conc.size(), 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = three ? grid.idx(i, j, k) : grid.sidx(i, j);
        real diag = 1.0;
        if (i - 1 >= 0) diag += cx;
        if (i + 1 < nx) diag += cx;
        if (j - 1 >= 0) diag += cy;
        if (j + 1 < ny) diag += cy;
        if (three) {
          if (k - 1 >= 0) diag += cz;
          if (k + 1 < nz) diag += cz;
        }
        A.diag[c] = diag;
        A.xm[c] = (i - 1 >= 0) ? -cx : 0.0;
        A.xp[c] = (i + 1 < nx) ? -cx : 0.0;
        A.ym[c] = (j - 1 >= 0) ? -cy : 0.0;
        A.yp[c] = (j + 1 < ny) ? -cy : 0.0;
        if (three) {
          A.zm[c] = (k - 1 >= 0) ? -cz : 0.0;
          A.zp[c] = (k + 1 < nz) ? -cz : 0.0;
        }
        b[c] = conc[c];
      }
  const SolveStats st = pcg_solve(A, b, x, cfg);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = three ? grid.idx(i, j, k) : grid.sidx(i, j);
        conc[c] = x[c];
      }
  return st.iterations;
}

}  // namespace

SoluteOracle::SoluteOracle(const SoluteParams& p, SweSolver* swe, ReSolver* re) : p_(p), swe_(swe), re_(re) {
  if (swe) {
    cs.assign((size_t)swe->g.nsurf(), 0.0);
    dx_ = swe->dx;
    dy_ = swe->dy;
  }
  if (re) {
    cg.assign((size_t)re->g.ncell(), 0.0);
    dx_ = re->params().dx;
    dy_ = re->params().dy;
    dz_ = re->params().dz;
  }
}

// computeSoluteMass (Orchestrator.cpp:852-896).
real SoluteOracle::mass() const {
  real total = 0.0;
  if (swe_) {
    real s = 0.0;
    const Grid& g = swe_->g;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const long long c = g.sidx(i, j);
        s += cs[c] * std::max(0.0, swe_->eta[c] - swe_->bottom[c]) * (dx_ * dy_);
      }
    total += s;
  }
  if (re_) {
    real s = 0.0;
    const Grid& g = re_->g;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const long long c = g.idx(i, j, k);
          s += cg[c] * re_->wc[c] * (dx_ * dy_ * dz_);
        }
    total += s;
  }
  return total;
}

// applyCouplingSoluteExchange (Orchestrator.cpp:1270-1299) + applyInterfaceExchange
// (SourceSink.cpp:22-53).
void SoluteOracle::interfaceExchange(const std::vector<real>& q, real dt, real area) {
  if (!swe_ || !re_) return;
  const Grid& gs = swe_->g;
  for (int j = 0; j < gs.ny; ++j)
    for (int i = 0; i < gs.nx; ++i) {
      const long long c = gs.sidx(i, j), t = re_->g.idx(i, j, 0);
      const real v = -q[(size_t)(i + j * gs.nx)] * dt / area;
      if (v > 0.0) {
        const real h = re_->wc[t] * dz_;
        if (h > 0.0) {
          real h0 = h - v;
          if (h0 < 0.0) h0 = 0.0;
          cg[t] = (cg[t] * h0 + cs[c] * v) / h;
        }
      } else if (v < 0.0) {
        const real s = -v;
        const real d = std::max(0.0, swe_->eta[c] - swe_->bottom[c]);
        if (d > 0.0) {
          real d0 = d - s;
          if (d0 < 0.0) d0 = 0.0;
          cs[c] = (cs[c] * d0 + cg[t] * s) / d;
        }
      }
    }
}

// SoluteStepper::step with rain = 0 (SoluteStepper.cpp:24-72): decay, advection of both domains
// on copies (a refusal discards them), implicit diffusion, commit.
bool SoluteOracle::step(real dt, real* cfl_out) {
  std::vector<real> ws = cs, wg = cg;
  if (p_.k_decay > 0.0 && dt > 0.0) {
    const real fac = std::exp(-p_.k_decay * dt);
    if (swe_)
      for (int j = 0; j < swe_->g.ny; ++j)
        for (int i = 0; i < swe_->g.nx; ++i) ws[swe_->g.sidx(i, j)] *= fac;
    if (re_)
      for (int k = 0; k < re_->g.nz; ++k)
        for (int j = 0; j < re_->g.ny; ++j)
          for (int i = 0; i < re_->g.nx; ++i) wg[re_->g.idx(i, j, k)] *= fac;
  }
  real cfl = 0.0;
  if (swe_) cfl = std::max(cfl, advectSurface(ws, *swe_, p_, dt));
  if (re_) cfl = std::max(cfl, advectSubsurface(wg, *re_, p_, dt));
  *cfl_out = cfl;
  if (cfl > p_.cfl_max) return false;
  if (p_.implicit_diffusion && p_.D > 0.0) {
    if (swe_) tot.diffusion_iterations += diffuse(ws, swe_->g, false, dx_, dy_, dz_, p_.D, dt, p_.solver);
    if (re_) tot.diffusion_iterations += diffuse(wg, re_->g, true, dx_, dy_, dz_, p_.D, dt, p_.solver);
  }
  cs.swap(ws);
  cg.swap(wg);
  return true;
}

// Orchestrator::applySolute (:938-1056).
void SoluteOracle::apply(real t, real dt, real rain) {
  if (dt <= 0.0) return;
  if (t + dt <= p_.start_time) return;
  const real area = dx_ * dy_;
  if (swe_ && rain > 0.0 && p_.c_rain != 0.0) {
    const Grid& g = swe_->g;
    const real dw = rain * dt;
    real added = 0.0;
    for (int j = 0; j < g.ny; ++j) {
      if (j == g.ny - 1) continue;  // the surface step gives the last row no rain
      for (int i = 0; i < g.nx; ++i) {
        const long long c = g.sidx(i, j);
        const real post = std::max(0.0, swe_->eta[c] - swe_->bottom[c]);
        const real d0 = std::max(0.0, post - dw);
        if (d0 <= p_.min_depth) continue;
        cs[c] = (d0 * cs[c] + dw * p_.c_rain) / (d0 + dw);
        added += dw * p_.c_rain * area;
      }
    }
    tot.rain_added += added;
  }
  const real mass_before = p_.k_decay > 0.0 ? mass() : 0.0;
  int nsub = p_.substeps;
  const int cap = 4096;
  for (;;) {
    const real sub = dt / (real)nsub;
    bool refused = false;
    real cmax = 0.0;
    for (int s = 0; s < nsub; ++s) {
      real cfl = 0.0;
      if (!step(sub, &cfl)) {
        refused = true;
        break;
      }
      cmax = std::max(cmax, cfl);
      ++tot.substeps;
    }
    if (!refused) {
      tot.max_courant = std::max(tot.max_courant, cmax);
      break;
    }
    if (nsub < cap) {
      nsub *= 2;
    } else {
      tot.refused_steps += 1;
      break;
    }
  }
  if (p_.k_decay > 0.0) tot.decay_loss += std::max(0.0, mass_before - mass());
}

// Orchestrator::step, coupled branch, with the solute steps (:1260-1268, 1322-1336).
void coupled_step_solute(const CouplingParams& cp, SweSolver& swe, ReSolver& re, CoupledState& cs,
                         SoluteOracle* sol, real dt, real rain, real evap, real t_end_hint) {
  const real t0 = cs.t, t_new = cs.t + dt;
  swe.advanceStep(rain, evap);
  if (cs.bc) cs.boundary_outflow += cs.bc->apply(swe, dt);
  std::vector<real> q;
  cs.exchange_volume += couple_exchange(cp, swe, re, dt, sol ? &q : nullptr);
  if (sol) sol->interfaceExchange(q, dt, cp.dx * cp.dy);
  const real eps = 1.0e-9 * std::max<real>(1.0, t_end_hint);
  while (cs.gw_time < t_new - eps) {
    re.advanceStep();
    cs.gw_time += re.dt_used;
    ++cs.gw_substeps;
  }
  if (sol) sol->apply(t0, dt, rain);
  cs.t = t_new;
}

}  // namespace frehg_oracle
