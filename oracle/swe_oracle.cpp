// Oracle semi-implicit shallow-water step -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Follows src/swe/SweSolver.cpp, serial path (parallel() == false: every ghost is a domain
// edge, filled zero-gradient); line numbers cite that file.
#include <algorithm>
#include <stdexcept>

#include "oracle.hpp"

namespace frehg_oracle {

SweSolver::SweSolver(int nx, int ny, real dx_, real dy_, const SweParams& p) : dx(dx_), dy(dy_), p_(p) {
  g.nx = nx;
  g.ny = ny;
  g.nz = 1;
  const size_t n = (size_t)g.nsurf();
  for (auto* a : {&eta, &etan, &bottom, &bottomXP, &bottomYP, &dept, &deptx, &depty, &uu, &vv, &un,
                  &vn, &uy, &vx, &Vs, &Vsn, &Asx, &Asy, &Asz, &Aszx, &Aszy, &Vsx, &Vsy, &CDx, &CDy,
                  &Dx, &Dy, &Ex, &Ey, &cflx, &cfly, &Fu, &Fv, &rain, &evap, &Sxp, &Sxm, &Syp, &Sym,
                  &Sct, &Srhs, &cfl_active, &wtfx, &wtfy})
    a->assign(n, 0.0);
}

void SweSolver::ghostsZeroGradient(std::vector<real>& f) {  // :25-36
  const int nx = g.nx, ny = g.ny;
  for (int i = 0; i < nx; ++i) {
    f[s(i, -1)] = f[s(i, 0)];
    f[s(i, ny)] = f[s(i, ny - 1)];
  }
  for (int j = 0; j < ny; ++j) {
    f[s(-1, j)] = f[s(0, j)];
    f[s(nx, j)] = f[s(nx - 1, j)];
  }
}

void SweSolver::setBathymetry(const std::vector<real>& physical) {  // :144-157
  if ((int)physical.size() != g.nx * g.ny) throw std::runtime_error("setBathymetry: expected nx*ny values");
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) bottom[s(i, j)] = physical[(size_t)(i + j * g.nx)];
  updateBottomFaces();
}

void SweSolver::updateBottomFaces() {  // :159-205
  const int nx = g.nx, ny = g.ny;
  ghostsZeroGradient(bottom);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j);
      bottomXP[c] = bottom[c];
      if (bottom[s(i + 1, j)] > bottom[c]) bottomXP[c] = bottom[s(i + 1, j)];
      bottomYP[c] = bottom[c];
      if (bottom[s(i, j + 1)] > bottom[c]) bottomYP[c] = bottom[s(i, j + 1)];
    }
  for (int i = 0; i < nx; ++i) {
    bottomYP[s(i, -1)] = bottom[s(i, 0)];
    if (bottom[s(i, -1)] > bottom[s(i, 0)]) bottomYP[s(i, -1)] = bottom[s(i, -1)];
  }
  for (int i = 0; i < nx; ++i) bottomYP[s(i, ny)] = bottomYP[s(i, ny - 1)];
  for (int j = 0; j < ny; ++j) {
    bottomXP[s(-1, j)] = bottom[s(0, j)];
    if (bottom[s(-1, j)] > bottom[s(0, j)]) bottomXP[s(-1, j)] = bottom[s(-1, j)];
  }
  for (int j = 0; j < ny; ++j) bottomXP[s(nx, j)] = bottomXP[s(nx - 1, j)];
}

void SweSolver::updateDepth() {  // :207-304
  const int nx = g.nx, ny = g.ny;
  const real md = p_.min_depth;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j);
      const real diff = eta[c] - bottom[c];
      if (diff > 0.0 && diff < md) eta[c] = bottom[c];
    }
  const long long n = g.nsurf();
  for (long long c = 0; c < n; ++c) {
    dept[c] = eta[c] - bottom[c];
    if (dept[c] <= md) dept[c] = 0.0;
    deptx[c] = 0.0;
    depty[c] = 0.0;
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j), ip = s(i + 1, j), jp = s(i, j + 1);
      real eh = eta[c], bh = bottom[c];
      if (eta[ip] > eh) eh = eta[ip];
      if (bottom[ip] > bh) bh = bottom[ip];
      deptx[c] = eh - bh;
      eh = eta[c];
      bh = bottom[c];
      if (eta[jp] > eh) eh = eta[jp];
      if (bottom[jp] > bh) bh = bottom[jp];
      depty[c] = eh - bh;
    }
  auto face = [&](long long a, long long b) {
    real eh = eta[a], bh = bottom[a];
    if (eta[b] > eh) eh = eta[b];
    if (bottom[b] > bh) bh = bottom[b];
    return eh - bh;
  };
  for (int i = 0; i < nx; ++i) depty[s(i, -1)] = face(s(i, 0), s(i, -1));
  for (int i = 0; i < nx; ++i) depty[s(i, ny)] = face(s(i, ny - 1), s(i, ny));
  for (int j = 0; j < ny; ++j) deptx[s(-1, j)] = face(s(0, j), s(-1, j));
  for (int j = 0; j < ny; ++j) deptx[s(nx, j)] = face(s(nx - 1, j), s(nx, j));
  for (long long c = 0; c < n; ++c) {
    if (deptx[c] < 0.0) deptx[c] = 0.0;
    if (depty[c] < 0.0) depty[c] = 0.0;
  }
}

void SweSolver::updateGeometry() {  // :306-369
  const int nx = g.nx, ny = g.ny;
  const long long n = g.nsurf();
  for (long long c = 0; c < n; ++c) {
    Vsn[c] = Vs[c];
    Vs[c] = dept[c] * dx * dy;
    Asz[c] = dept[c] > 0.0 ? dx * dy : 0.0;
    Asx[c] = deptx[c] * dy;
    Asy[c] = depty[c] * dx;
    if (nx == 1) Asx[c] = 0.0;
    if (ny == 1) Asy[c] = 0.0;
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j), ip = s(i + 1, j), jp = s(i, j + 1);
      Vsx[c] = 0.5 * (Vs[c] + Vs[ip]);
      Vsy[c] = 0.5 * (Vs[c] + Vs[jp]);
      Aszx[c] = 0.5 * (Asz[c] + Asz[ip]);
      Aszy[c] = 0.5 * (Asz[c] + Asz[jp]);
    }
  ghostsZeroGradient(Vs);
  ghostsZeroGradient(Asx);
  ghostsZeroGradient(Asy);
  ghostsZeroGradient(Asz);
  for (int i = 0; i < nx; ++i) {
    Vsy[s(i, -1)] = Vs[s(i, 0)];
    Aszy[s(i, -1)] = Asz[s(i, 0)];
  }
  for (int j = 0; j < ny; ++j) {
    Vsx[s(-1, j)] = Vs[s(0, j)];
    Aszx[s(-1, j)] = Asz[s(0, j)];
  }
}

void SweSolver::enforceSurfBc() {  // :371-384
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      if (eta[c] < bottom[c]) eta[c] = bottom[c];
    }
  ghostsZeroGradient(eta);
}

void SweSolver::momentumSource() {  // :386-444
  const real dt = p_.dt, viscx = p_.visc_x, viscy = p_.visc_y;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j), iM = s(i - 1, j), iP = s(i + 1, j), jM = s(i, j - 1), jP = s(i, j + 1);
      const real u = uu[c], v = vv[c], vxc = vx[c], uyc = uy[c];
      real advX = (0.5 / dx) * ((u + std::fabs(u)) * (u - uu[iM]) + (u - std::fabs(u)) * (uu[iP] - u)) +
                  (0.5 / dy) * ((vxc + std::fabs(vxc)) * (u - uu[jM]) + (vxc - std::fabs(vxc)) * (uu[jP] - u));
      real advY = (0.5 / dx) * ((uyc + std::fabs(uyc)) * (v - vv[iM]) + (uyc - std::fabs(uyc)) * (vv[iP] - v)) +
                  (0.5 / dy) * ((v + std::fabs(v)) * (v - vv[jM]) + (v - std::fabs(v)) * (vv[jP] - v));
      if (u == 0.0 || cflx[c] > 0.7) advX = 0.0;
      else if (cflx[c] > 0.5) advX = advX * (0.7 - cflx[c]) / (0.7 - 0.5);
      if (v == 0.0 || cfly[c] > 0.7) advY = 0.0;
      else if (cfly[c] > 0.5) advY = advY * (0.7 - cfly[c]) / (0.7 - 0.5);
      real difX = 0.0, difY = 0.0;
      if (Vsx[c] > 0.0)
        difX = (viscx / Vsx[c] / dx) * (Asx[c] * (uu[iP] - u) - Asx[c] * (u - uu[iM])) +
               (viscy / Vsx[c] / dy) * (Asy[c] * (uu[jP] - u) - Asy[jM] * (u - uu[jM]));
      if (Vsy[c] > 0.0)
        difY = (viscx / Vsy[c] / dx) * (Asx[c] * (vv[iP] - v) - Asx[iM] * (v - vv[iM])) +
               (viscy / Vsy[c] / dy) * (Asy[c] * (vv[jP] - v) - Asy[c] * (v - vv[jM]));
      const real velx = std::sqrt(u * u + vxc * vxc);
      const real vely = std::sqrt(uyc * uyc + v * v);
      real facdx = 0.0, facdy = 0.0;
      if (Vsx[c] > 0.0) facdx = Aszx[c] / Vsx[c];
      if (Vsy[c] > 0.0) facdy = Aszy[c] / Vsy[c];
      Dx[c] = 1.0 / (0.5 * dt * CDx[c] * velx * facdx + 1.0);
      Dy[c] = 1.0 / (0.5 * dt * CDy[c] * vely * facdy + 1.0);
      Ex[c] = (u + dt * (difX - advX)) * Dx[c];
      Ey[c] = (v + dt * (difY - advY)) * Dy[c];
    }
}

void SweSolver::shallowwaterRhs() {  // :446-462
  const real dt = p_.dt;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j), iM = s(i - 1, j), jM = s(i, j - 1);
      Srhs[c] = eta[c] * Asz[c] - dt * (Asx[c] * Ex[c] - Asx[iM] * Ex[iM] + Asy[c] * Ey[c] - Asy[jM] * Ey[jM]);
      if (eta[c] - bottom[c] > p_.hE) Srhs[c] -= evap[c] * Asz[c] * dt;
    }
}

void SweSolver::shallowwaterMatCoeff() {  // :464-526 (serial)
  const int nx = g.nx, ny = g.ny;
  const real dt = p_.dt;
  const real coef = p_.gravity * dt * dt;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j), iM = s(i - 1, j), jM = s(i, j - 1);
      Sxp[c] = Sxm[c] = Syp[c] = Sym[c] = 0.0;
      if (Vsx[c] > 0.0) Sxp[c] = coef * Asx[c] * Asx[c] * Dx[c] / Vsx[c];
      if (Vsx[iM] > 0.0) Sxm[c] = coef * Asx[iM] * Asx[iM] * Dx[iM] / Vsx[iM];
      if (Vsy[c] > 0.0) Syp[c] = coef * Asy[c] * Asy[c] * Dy[c] / Vsy[c];
      if (Vsy[jM] > 0.0) Sym[c] = coef * Asy[jM] * Asy[jM] * Dy[jM] / Vsy[jM];
      Sct[c] = Asz[c] + Sxp[c] + Sxm[c] + Syp[c] + Sym[c];
      if (dept[c] == 0.0) {
        Sct[c] = dx * dy;
        Srhs[c] = eta[c] * dx * dy;
        if (uu[c] == 0.0 && uu[iM] == 0.0 && vv[c] == 0.0 && vv[jM] == 0.0) Sxp[c] = Sxm[c] = Syp[c] = Sym[c] = 0.0;
      } else if (Sct[c] == 0.0) {
        Sct[c] = dx * dy;
        Srhs[c] = eta[c] * dx * dy;
      }
      if (i == 0) {
        Sct[c] -= Sxm[c];
        if (nx == 1) Sct[c] -= Sxp[c];
      } else if (i == nx - 1) {
        Sct[c] -= Sxp[c];
      }
      if (j == 0) {
        Sct[c] -= Sym[c];
        if (ny == 1) Sct[c] -= Syp[c];
      } else if (j == ny - 1) {
        Sct[c] -= Syp[c];
      }
    }
}

void SweSolver::assembleAndSolve(SolveStats* st) {  // :528-595
  StencilOp A;
  A.g = g;
  A.three_d = false;
  A.resize();
  const int nx = g.nx, ny = g.ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const long long c = s(i, j);
      A.diag[c] = Sct[c];
      A.xm[c] = -Sxm[c];
      A.xp[c] = -Sxp[c];
      A.ym[c] = -Sym[c];
      A.yp[c] = -Syp[c];
      const bool zw = (i == 0) || A.xm[c] == 0.0, ze = (i == nx - 1) || A.xp[c] == 0.0;
      const bool zs = (j == 0) || A.ym[c] == 0.0, zn = (j == ny - 1) || A.yp[c] == 0.0;
      A.fixed[c] = (zw && ze && zs && zn) ? 1 : 0;
    }
  std::vector<real> x;
  SolveStats ss = pcg_solve(A, Srhs, x, p_.solver);
  if (st) *st = ss;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) eta[s(i, j)] = x[s(i, j)];
}

void SweSolver::cflLimiter() {  // :597-624
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      const real diff = eta[c] - bottom[c];
      if (dept[c] <= 0.0 && diff > 0.0) {
        int wet = 0;
        if (dept[s(i + 1, j)] > 0.0) wet = 1;
        else if (dept[s(i - 1, j)] > 0.0) wet = 1;
        else if (dept[s(i, j + 1)] > 0.0) wet = 1;
        else if (dept[s(i, j - 1)] > 0.0) wet = 1;
        if (wet == 0) {
          eta[c] = bottom[c];
          cfl_active[c] = 1.0;
        }
      }
    }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      const real diff = eta[c] - bottom[c];
      if (diff > 0.0 && diff < p_.min_depth) eta[c] = bottom[c];
    }
}

void SweSolver::evaprain(real rain_rate, real evap_rate) {  // :626-652
  const real dt = p_.dt;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i)
      if (j != g.ny - 1) eta[s(i, j)] += rain_rate * dt;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      if (eta[c] - bottom[c] > p_.hE) eta[c] -= evap_rate * dt;
    }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      const real diff = eta[c] - bottom[c];
      if (diff < p_.min_depth) eta[c] = bottom[c];
      dept[c] = eta[c] - bottom[c];
    }
}

void SweSolver::updateDragCoef() {  // :654-672
  const real coef = p_.gravity * p_.manning * p_.manning;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      if (Vs[c] > 0.0) {
        const real effh = Vs[c] / (dx * dy);
        const real expo = (effh < p_.hD) ? (2.0 / 3.0) : (1.0 / 3.0);
        CDx[c] = coef / std::pow(effh, expo);
        CDy[c] = coef / std::pow(effh, expo);
      }
    }
}

void SweSolver::waterfallLocation() {  // :674-701
  const real wtfh = p_.wtfh;
  std::fill(wtfx.begin(), wtfx.end(), 0.0);
  std::fill(wtfy.begin(), wtfy.end(), 0.0);
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      real facd = eta[s(i + 1, j)] - bottomXP[c];
      if (eta[c] < bottomXP[c] && facd > wtfh) wtfx[c] = -1.0;
      facd = eta[s(i, j + 1)] - bottomYP[c];
      if (eta[c] < bottomYP[c] && facd > wtfh) wtfy[c] = -1.0;
    }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j), iM = s(i - 1, j), jM = s(i, j - 1);
      real facd = eta[iM] - bottomXP[iM];
      if (eta[c] < bottomXP[iM] && facd > wtfh) wtfx[c] = 1.0;
      facd = eta[jM] - bottomYP[jM];
      if (eta[c] < bottomYP[jM] && facd > wtfh) wtfy[c] = 1.0;
    }
}

void SweSolver::updateVelocity() {  // :703-770
  const real dt = p_.dt, wtfh = p_.wtfh, coef = p_.gravity * dt;
  un = uu;
  vn = vv;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      const real effhx = (Vsx[c] > 0.0) ? Asx[c] / Vsx[c] : 0.0;
      const real effhy = (Vsy[c] > 0.0) ? Asy[c] / Vsy[c] : 0.0;
      uu[c] = (Ex[c] - coef * effhx * (eta[s(i + 1, j)] - eta[c])) * Dx[c];
      vv[c] = (Ey[c] - coef * effhy * (eta[s(i, j + 1)] - eta[c])) * Dy[c];
    }
  for (int j = 0; j < g.ny; ++j)  // sequential limiter pass :736-758
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j), iM = s(i - 1, j), jM = s(i, j - 1);
      if (Asx[c] < wtfh * dy) uu[c] = 0.0;
      if (Asy[c] < wtfh * dx) vv[c] = 0.0;
      if (dept[c] < wtfh) {
        if (uu[c] > 0.0) uu[c] = 0.0;
        if (uu[iM] < 0.0) uu[iM] = 0.0;
        if (vv[c] > 0.0) vv[c] = 0.0;
        if (vv[jM] < 0.0) vv[jM] = 0.0;
      }
      if (cfl_active[c] == 1.0) {
        uu[c] = 0.0;
        uu[iM] = 0.0;
        vv[c] = 0.0;
        vv[jM] = 0.0;
        cfl_active[c] = 0.0;
      }
    }
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      Fu[c] = uu[c] * Asx[c];
      Fv[c] = vv[c] * Asy[c];
      cflx[c] = std::fabs(uu[c] * dt / dx);
      cfly[c] = std::fabs(vv[c] * dt / dy);
    }
}

void SweSolver::enforceVeloBc() {  // :772-778
  ghostsZeroGradient(uu);
  ghostsZeroGradient(vv);
}

void SweSolver::interpVelocity() {  // :780-792
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      uy[c] = 0.25 * (uu[c] + uu[s(i - 1, j)] + uu[s(i, j + 1)] + uu[s(i - 1, j + 1)]);
      vx[c] = 0.25 * (vv[c] + vv[s(i, j - 1)] + vv[s(i + 1, j)] + vv[s(i + 1, j - 1)]);
    }
}

void SweSolver::advanceStep(real rain_rate, real evap_rate, SolveStats* st) {  // :803-853
  const long long n = g.nsurf();
  for (long long c = 0; c < n; ++c) {
    rain[c] = rain_rate;
    evap[c] = evap_rate;
    etan[c] = eta[c];
  }
  enforceSurfBc();
  momentumSource();
  shallowwaterRhs();
  shallowwaterMatCoeff();
  assembleAndSolve(st);
  enforceSurfBc();
  cflLimiter();
  evaprain(rain_rate, evap_rate);
  updateDepth();
  updateDepth();
  updateDragCoef();
  waterfallLocation();
  updateVelocity();
  enforceVeloBc();
  interpVelocity();
  updateGeometry();
}

void SweSolver::setInitialEtaPhysical(const std::vector<real>& e) {  // :855-870
  if ((int)e.size() != g.nx * g.ny) throw std::runtime_error("setInitialEtaPhysical: expected nx*ny values");
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      eta[c] = e[(size_t)(i + j * g.nx)] + p_.offset;
      if (eta[c] < bottom[c]) eta[c] = bottom[c];
    }
}

void SweSolver::setInitialVelocityConstant(real u, real v) {  // :872-884
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = s(i, j);
      uu[c] = u;
      vv[c] = v;
      CDx[c] = 0.0;
      CDy[c] = 0.0;
    }
}

void SweSolver::finalizeInitialState() {  // :904-924
  enforceSurfBc();
  ghostsZeroGradient(uu);
  ghostsZeroGradient(vv);
  updateDepth();
  updateGeometry();
  etan = eta;
  un = uu;
  vn = vv;
  Vsn = Vs;
}

void SweSolver::initializeState(real init_eta) {  // :926-934
  setInitialEtaPhysical(std::vector<real>((size_t)(g.nx * g.ny), init_eta));
  setInitialVelocityConstant(0.0, 0.0);
  finalizeInitialState();
}

real SweSolver::maxCfl() const {  // :954-975
  const real dmin = std::min(dx, dy);
  real cfl = 0.0;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const long long c = g.sidx(i, j);
      const real hh = dept[c];
      if (hh <= 0.0) continue;
      const real speed = std::sqrt(uu[c] * uu[c] + vv[c] * vv[c]);
      const real local = (speed + std::sqrt(p_.gravity * hh)) * p_.dt / dmin;
      if (local > cfl) cfl = local;
    }
  return cfl;
}

// ---------------------------------------------------------------------------------------------
// Polygon boundary regions.

// src/bc/Polygon.hpp: even-odd ray casting (+x ray), bounding-box reject, ring closed.
static bool polygon_contains(const std::vector<std::array<double, 2>>& v, double x, double y) {
  const size_t n = v.size();
  if (n < 3) return false;
  double xmin = v[0][0], xmax = v[0][0], ymin = v[0][1], ymax = v[0][1];
  for (const auto& p : v) {
    xmin = std::min(xmin, p[0]);
    xmax = std::max(xmax, p[0]);
    ymin = std::min(ymin, p[1]);
    ymax = std::max(ymax, p[1]);
  }
  if (x < xmin || x > xmax || y < ymin || y > ymax) return false;
  bool inside = false;
  for (size_t i = 0, j = n - 1; i < n; j = i++) {
    const double xi = v[i][0], yi = v[i][1], xj = v[j][0], yj = v[j][1];
    if ((yi > y) != (yj > y)) {
      const double xc = xj + (y - yj) * (xi - xj) / (yi - yj);
      if (x < xc) inside = !inside;
    }
  }
  return inside;
}

void PolygonBc::build(const SweSolver& swe, double x0, double y0) {  // PolygonIndex.cpp:9-33
  index.assign((size_t)swe.g.nsurf(), -1);
  count.assign(regions.size(), 0);
  for (int j = 0; j < swe.g.ny; ++j)
    for (int i = 0; i < swe.g.nx; ++i) {
      const double x = x0 + (i + 0.5) * swe.dx, y = y0 + (j + 0.5) * swe.dy;
      for (size_t p = 0; p < regions.size(); ++p)
        if (polygon_contains(regions[p].vertices, x, y)) {
          index[(size_t)swe.g.sidx(i, j)] = (int)p;
          ++count[p];  // globalColumnCounts (PolygonIndex.cpp:45-56), one rank
          break;
        }
    }
}

real PolygonBc::apply(SweSolver& swe, real dt) const {  // PolygonBC.cpp:16-78
  if (regions.empty()) return 0.0;
  const real area = swe.dx * swe.dy;
  const real gravity = swe.params().gravity;
  real net = 0.0;
  for (int j = 0; j < swe.g.ny; ++j)
    for (int i = 0; i < swe.g.nx; ++i) {
      const long long c = swe.g.sidx(i, j);
      const int p = index[(size_t)c];
      if (p < 0) continue;
      const BcPolygonRegion& r = regions[(size_t)p];
      const real depth = std::max(swe.eta[c] - swe.bottom[c], 0.0);
      if (r.kind == 1) {  // depth: prescribed ponded depth
        const real hbc = std::max(r.rate, 0.0);
        const real old = swe.eta[c];
        swe.eta[c] = swe.bottom[c] + hbc;
        net += (old - swe.eta[c]) * area;
      } else if (r.kind == 0) {  // discharge Q spread over the region's cells, outflow limited
        const int gc = count[(size_t)p];
        if (gc <= 0) continue;
        real vol = r.rate * dt / (real)gc;
        if (vol > depth * area) vol = depth * area;
        swe.eta[c] -= vol / area;
        net += vol;
      } else {  // critical-depth weir: q = sqrt(g) h^1.5 min(dx, dy), limited
        const real q = std::sqrt(gravity) * std::pow(depth, 1.5) * std::min(swe.dx, swe.dy);
        real vol = q * dt;
        if (vol > depth * area) vol = depth * area;
        swe.eta[c] -= vol / area;
        net += vol;
      }
    }
  swe.updateDepth();
  swe.updateGeometry();
  return net;
}

// ---------------------------------------------------------------------------------------------
// src/coupling/Coupling.cpp:25-164. Soil is class 0 (re.params().soil), as in the reference.
real couple_exchange(const CouplingParams& cp, SweSolver& swe, ReSolver& re, real dt,
                     std::vector<real>* q_out) {
  const int nx = swe.g.nx, ny = swe.g.ny;
  const Soil soil = re.params().soil;
  const real Az = cp.dx * cp.dy;
  const real dz = re.params().dz;
  std::vector<real> q((size_t)(nx * ny), 0.0);
  for (int j = 0; j < ny; ++j)  // computeExchangeRates :49-57
    for (int i = 0; i < nx; ++i) {
      const long long si = swe.g.sidx(i, j), ti = re.g.idx(i, j, 0);
      const real dept = std::max(swe.eta[si] - swe.bottom[si], 0.0);
      real qv = Az * soil.Ks_z * cp.visc * (re.h[ti] - dept) / (0.5 * dz);
      if (dept <= cp.min_depth && qv < 0.0) qv = 0.0;
      q[(size_t)(i + j * nx)] = qv;
    }
  for (int j = 0; j < ny; ++j)  // limitExchange :78-98
    for (int i = 0; i < nx; ++i) {
      const size_t oi = (size_t)(i + j * nx);
      real qv = q[oi];
      if (qv == 0.0) continue;
      const long long si = swe.g.sidx(i, j), ti = re.g.idx(i, j, 0);
      const real V = Az * dz;
      const real wc_top = re.wc[ti];
      if (qv < 0.0) {
        const real dept = std::max(swe.eta[si] - swe.bottom[si], 0.0);
        const real q_min_surf = -dept * Az / dt;
        const real q_min_gw = -(soil.theta_s - wc_top) * V / dt;
        qv = std::max(qv, std::max(q_min_surf, q_min_gw));
      } else {
        const real q_max_gw = (wc_top - soil.theta_r) * V / dt;
        qv = std::min(qv, q_max_gw);
      }
      q[oi] = qv;
    }
  if (q_out) *q_out = q;
  for (int j = 0; j < ny; ++j)  // applyExchangeToSurface :113-118
    for (int i = 0; i < nx; ++i) {
      const long long si = swe.g.sidx(i, j);
      const real dEta = q[(size_t)(i + j * nx)] * dt / Az;
      swe.eta[si] += dEta;
      swe.dept[si] = std::max(swe.eta[si] - swe.bottom[si], 0.0);
    }
  for (int j = 0; j < ny; ++j)  // applyExchangeToGroundwater :134-142
    for (int i = 0; i < nx; ++i) {
      const long long ti = re.g.idx(i, j, 0);
      const real V = Az * dz;
      real wc_new = re.wc[ti] - q[(size_t)(i + j * nx)] * dt / V;
      wc_new = std::min(std::max(wc_new, soil.theta_r), soil.theta_s);
      re.wc[ti] = wc_new;
      re.h[ti] = VG::head(soil, wc_new);
    }
  real total = 0.0;
  for (size_t k = 0; k < q.size(); ++k) total += q[k] * dt;
  return total;
}

// Orchestrator.cpp:1322-1336: the SW advance, the boundary regions (applySurfacePolygons, :1240,
// :1301-1308), the exchange, the GW catch-up (gwCatchUpTo, :1243-1252); without solute.
void coupled_step(const CouplingParams& cp, SweSolver& swe, ReSolver& re, CoupledState& cs,
                  real dt, real rain, real evap, real t_end_hint) {
  coupled_step_solute(cp, swe, re, cs, nullptr, dt, rain, evap, t_end_hint);
}

}  // namespace frehg_oracle
