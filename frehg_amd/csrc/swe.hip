// Semi-implicit shallow-water step on the GPU, following SweSolver::advanceStep
// (src/swe/SweSolver.cpp:803-853), serial path (every ghost is a domain edge). Fusions:
//   surfbc     enforceSurfBc (:371-384) incl. zero-gradient ghosts
//   momentum   momentumSource (:386-444)
//   assemble   shallowwaterRhs + shallowwaterMatCoeff (:446-526)
//   fixrows    decoupled rows (all off-diagonals 0: dry cells at rest) are solved exactly
//   symm       symmetric couplings of the remaining rows, their rhs with decoupled rows folded in
//   [Krylov]   Jacobi-PCG on the 5-point system (:528-595)
//   post       enforceSurfBc + cflLimiter + evaprain (:371-384, :597-652); the second
//              updateDepth of the step is a repeat of the first on the serial path and the
//              first one's snap finds nothing after evaprain's small-depth removal
//   depth      updateDepth (:207-304)
//   dragwtf    updateDragCoef + waterfallLocation (:654-701)
//   vel        updateVelocity with its limiter pass (:703-770) + enforceVeloBc (:772-778).
//              The limiter only ever sets a face velocity to 0 under conditions on the sign of
//              that velocity or on its two cells, so its sequential sweep has an order-free
//              closed form evaluated per face here.
//   interp     interpVelocity (:780-792), clears cfl_active (:755)
//   geom       updateGeometry (:306-369)
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "swe.hpp"

namespace frehg {

namespace {

constexpr int NT = 256;

struct SwK {
  Geo g;
  double dx, dy, dt, gravity, manning, min_depth, visc_x, visc_y, hD, hE, wtfh;
};

// src/bc/Polygon.hpp: even-odd ray casting with a +x ray, bounding-box reject, ring closed.
bool polygon_contains(const double* xy, int n, double x, double y) {
  if (n < 3) return false;
  double xmin = xy[0], xmax = xy[0], ymin = xy[1], ymax = xy[1];
  for (int v = 1; v < n; ++v) {
    xmin = std::min(xmin, xy[2 * v]);
    xmax = std::max(xmax, xy[2 * v]);
    ymin = std::min(ymin, xy[2 * v + 1]);
    ymax = std::max(ymax, xy[2 * v + 1]);
  }
  if (x < xmin || x > xmax || y < ymin || y > ymax) return false;
  bool inside = false;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    const double xi = xy[2 * i], yi = xy[2 * i + 1], xj = xy[2 * j], yj = xy[2 * j + 1];
    if ((yi > y) != (yj > y) && x < xj + (y - yj) * (xi - xj) / (yi - yj)) inside = !inside;
  }
  return inside;
}

struct C2 {
  int i, j;
  long long c;
  bool valid;
};

__device__ inline C2 cell2(const Geo& g) {
  C2 t;
  t.i = blockIdx.x * 64 + (threadIdx.x & 63);
  t.j = blockIdx.y * 4 + (threadIdx.x >> 6);
  t.valid = t.i < g.nx && t.j < g.ny;
  t.c = t.valid ? g.sidx(t.i, t.j) : 0;
  return t;
}

// Zero-gradient ghost copy by the owning edge thread (SweSolver.cpp:25-36).
__device__ inline void ghost_copy(const Geo& g, const C2& t, double* f, double v) {
  if (t.i == 0) f[t.c - 1] = v;
  if (t.i == g.nx - 1) f[t.c + 1] = v;
  if (t.j == 0) f[t.c - g.sy] = v;
  if (t.j == g.ny - 1) f[t.c + g.sy] = v;
}

__global__ __launch_bounds__(NT) void k_swe_surfbc(SwK P, double* eta, const double* bottom) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  double e = eta[t.c];
  if (e < bottom[t.c]) e = bottom[t.c];
  eta[t.c] = e;
  ghost_copy(P.g, t, eta, e);
}

struct MomIn {
  const double *uu, *vv, *vx, *uy, *cflx, *cfly, *Vsx, *Vsy, *Asx, *Asy, *Aszx, *Aszy, *CDx, *CDy;
};

__global__ __launch_bounds__(NT) void k_swe_momentum(SwK P, MomIn m, double* Dx, double* Dy, double* Ex,
                                                      double* Ey) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c, iM = c - 1, iP = c + 1, jM = c - g.sy, jP = c + g.sy;
  const double dx = P.dx, dy = P.dy, dt = P.dt;
  const double u = m.uu[c], v = m.vv[c], vxc = m.vx[c], uyc = m.uy[c];
  double advX = (0.5 / dx) * ((u + fabs(u)) * (u - m.uu[iM]) + (u - fabs(u)) * (m.uu[iP] - u)) +
                (0.5 / dy) * ((vxc + fabs(vxc)) * (u - m.uu[jM]) + (vxc - fabs(vxc)) * (m.uu[jP] - u));
  double advY = (0.5 / dx) * ((uyc + fabs(uyc)) * (v - m.vv[iM]) + (uyc - fabs(uyc)) * (m.vv[iP] - v)) +
                (0.5 / dy) * ((v + fabs(v)) * (v - m.vv[jM]) + (v - fabs(v)) * (m.vv[jP] - v));
  const double cx = m.cflx[c], cy = m.cfly[c];
  if (u == 0.0 || cx > 0.7) advX = 0.0;
  else if (cx > 0.5) advX = advX * (0.7 - cx) / (0.7 - 0.5);
  if (v == 0.0 || cy > 0.7) advY = 0.0;
  else if (cy > 0.5) advY = advY * (0.7 - cy) / (0.7 - 0.5);
  double difX = 0.0, difY = 0.0;
  const double vsx = m.Vsx[c], vsy = m.Vsy[c];
  if (vsx > 0.0)
    difX = (P.visc_x / vsx / dx) * (m.Asx[c] * (m.uu[iP] - u) - m.Asx[c] * (u - m.uu[iM])) +
           (P.visc_y / vsx / dy) * (m.Asy[c] * (m.uu[jP] - u) - m.Asy[jM] * (u - m.uu[jM]));
  if (vsy > 0.0)
    difY = (P.visc_x / vsy / dx) * (m.Asx[c] * (m.vv[iP] - v) - m.Asx[iM] * (v - m.vv[iM])) +
           (P.visc_y / vsy / dy) * (m.Asy[c] * (m.vv[jP] - v) - m.Asy[c] * (v - m.vv[jM]));
  const double velx = sqrt(u * u + vxc * vxc);
  const double vely = sqrt(uyc * uyc + v * v);
  double facdx = 0.0, facdy = 0.0;
  if (vsx > 0.0) facdx = m.Aszx[c] / vsx;
  if (vsy > 0.0) facdy = m.Aszy[c] / vsy;
  const double dxv = 1.0 / (0.5 * dt * m.CDx[c] * velx * facdx + 1.0);
  const double dyv = 1.0 / (0.5 * dt * m.CDy[c] * vely * facdy + 1.0);
  Dx[c] = dxv;
  Dy[c] = dyv;
  Ex[c] = (u + dt * (difX - advX)) * dxv;
  Ey[c] = (v + dt * (difY - advY)) * dyv;
}

struct AsmIn {
  const double *eta, *bottom, *Asz, *Asx, *Asy, *Ex, *Ey, *Vsx, *Vsy, *Dx, *Dy, *dept, *uu, *vv;
};
struct AsmOut {
  double *Sxp, *Sxm, *Syp, *Sym, *Sct, *Srhs;
};

__global__ __launch_bounds__(NT) void k_swe_assemble(SwK P, AsmIn a, AsmOut o, double evap) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c, iM = c - 1, jM = c - g.sy;
  const double dt = P.dt, dxdy_x = P.dx, dxdy_y = P.dy;
  const double e = a.eta[c];
  double rhs = e * a.Asz[c] - dt * (a.Asx[c] * a.Ex[c] - a.Asx[iM] * a.Ex[iM] + a.Asy[c] * a.Ey[c] - a.Asy[jM] * a.Ey[jM]);
  if (e - a.bottom[c] > P.hE) rhs -= evap * a.Asz[c] * dt;
  const double coef = P.gravity * dt * dt;
  double sxp = 0.0, sxm = 0.0, syp = 0.0, sym = 0.0;
  if (a.Vsx[c] > 0.0) sxp = coef * a.Asx[c] * a.Asx[c] * a.Dx[c] / a.Vsx[c];
  if (a.Vsx[iM] > 0.0) sxm = coef * a.Asx[iM] * a.Asx[iM] * a.Dx[iM] / a.Vsx[iM];
  if (a.Vsy[c] > 0.0) syp = coef * a.Asy[c] * a.Asy[c] * a.Dy[c] / a.Vsy[c];
  if (a.Vsy[jM] > 0.0) sym = coef * a.Asy[jM] * a.Asy[jM] * a.Dy[jM] / a.Vsy[jM];
  double sct = a.Asz[c] + sxp + sxm + syp + sym;
  if (a.dept[c] == 0.0) {
    sct = dxdy_x * dxdy_y;
    rhs = e * dxdy_x * dxdy_y;
    if (a.uu[c] == 0.0 && a.uu[iM] == 0.0 && a.vv[c] == 0.0 && a.vv[jM] == 0.0) sxp = sxm = syp = sym = 0.0;
  } else if (sct == 0.0) {
    sct = dxdy_x * dxdy_y;
    rhs = e * dxdy_x * dxdy_y;
  }
  if (t.i == 0) {
    sct -= sxm;
    if (g.nx == 1) sct -= sxp;
  } else if (t.i == g.nx - 1) {
    sct -= sxp;
  }
  if (t.j == 0) {
    sct -= sym;
    if (g.ny == 1) sct -= syp;
  } else if (t.j == g.ny - 1) {
    sct -= syp;
  }
  o.Sxp[c] = sxp;
  o.Sxm[c] = sxm;
  o.Syp[c] = syp;
  o.Sym[c] = sym;
  o.Sct[c] = sct;
  o.Srhs[c] = rhs;
}

// Rows whose in-domain off-diagonals are all zero are decoupled: x = b / diag exactly.
__global__ __launch_bounds__(NT) void k_swe_fixrows(SwK P, AsmOut o, unsigned char* fixed, double* eta) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  const bool fx = (t.i == 0 || o.Sxm[c] == 0.0) && (t.i == g.nx - 1 || o.Sxp[c] == 0.0) &&
                  (t.j == 0 || o.Sym[c] == 0.0) && (t.j == g.ny - 1 || o.Syp[c] == 0.0);
  fixed[c] = fx ? 1 : 0;
  if (fx) eta[c] = o.Srhs[c] / o.Sct[c];
}

// Symmetric couplings between non-decoupled rows; decoupled neighbours folded into the rhs.
// Matrix row (SweSolver.cpp:554-568): W -Sxm, E -Sxp, S -Sym, N -Syp.
__global__ __launch_bounds__(NT) void k_swe_symm(SwK P, AsmOut o, const unsigned char* fixed, const double* eta,
                                                  double* ax, double* ay, double* b) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  const bool fc = fixed[c] != 0;
  ax[c] = (!fc && t.i < g.nx - 1 && !fixed[c + 1]) ? -o.Sxp[c] : 0.0;
  ay[c] = (!fc && t.j < g.ny - 1 && !fixed[c + g.sy]) ? -o.Syp[c] : 0.0;
  double acc = o.Srhs[c];
  if (!fc) {
    if (t.i > 0 && fixed[c - 1]) acc -= (-o.Sxm[c]) * eta[c - 1];
    if (t.i < g.nx - 1 && fixed[c + 1]) acc -= (-o.Sxp[c]) * eta[c + 1];
    if (t.j > 0 && fixed[c - g.sy]) acc -= (-o.Sym[c]) * eta[c - g.sy];
    if (t.j < g.ny - 1 && fixed[c + g.sy]) acc -= (-o.Syp[c]) * eta[c + g.sy];
  }
  b[c] = acc;
}

// enforceSurfBc + cflLimiter + evaprain on the solved surface (SweSolver.cpp:829-831).
__global__ __launch_bounds__(NT) void k_swe_post(SwK P, double* eta, const double* bottom, const double* dept,
                                                  double* cfl_active, double rain, double evap) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  const double b = bottom[c];
  double e = eta[c];
  if (e < b) e = b;
  ghost_copy(g, t, eta, e);
  if (dept[c] <= 0.0 && e - b > 0.0) {
    const bool wet = dept[c + 1] > 0.0 || dept[c - 1] > 0.0 || dept[c + g.sy] > 0.0 || dept[c - g.sy] > 0.0;
    if (!wet) {
      e = b;
      cfl_active[c] = 1.0;
    }
  }
  {
    const double diff = e - b;
    if (diff > 0.0 && diff < P.min_depth) e = b;
  }
  if (t.j != g.ny - 1) e += rain * P.dt;
  if (e - b > P.hE) e -= evap * P.dt;
  if (e - b < P.min_depth) e = b;
  eta[c] = e;
}

// updateDepth (SweSolver.cpp:207-304) without the interior snap (a no-op after evaprain).
__device__ inline double face_depth(const double* eta, const double* bottom, long long a, long long b) {
  double eh = eta[a], bh = bottom[a];
  if (eta[b] > eh) eh = eta[b];
  if (bottom[b] > bh) bh = bottom[b];
  return eh - bh;
}

__global__ __launch_bounds__(NT) void k_swe_depth(SwK P, const double* eta, const double* bottom, double* dept,
                                                   double* deptx, double* depty) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  auto dp = [&](long long q) {
    double d = eta[q] - bottom[q];
    if (d <= P.min_depth) d = 0.0;
    dept[q] = d;
  };
  dp(c);
  if (t.i == 0) dp(c - 1);
  if (t.i == g.nx - 1) dp(c + 1);
  if (t.j == 0) dp(c - g.sy);
  if (t.j == g.ny - 1) dp(c + g.sy);
  deptx[c] = fmax(face_depth(eta, bottom, c, c + 1), 0.0);
  depty[c] = fmax(face_depth(eta, bottom, c, c + g.sy), 0.0);
  if (t.j == 0) depty[c - g.sy] = fmax(face_depth(eta, bottom, c, c - g.sy), 0.0);
  if (t.j == g.ny - 1) depty[c + g.sy] = fmax(face_depth(eta, bottom, c, c + g.sy), 0.0);
  if (t.i == 0) deptx[c - 1] = fmax(face_depth(eta, bottom, c, c - 1), 0.0);
  if (t.i == g.nx - 1) deptx[c + 1] = fmax(face_depth(eta, bottom, c, c + 1), 0.0);
}

// updateDragCoef + waterfallLocation (SweSolver.cpp:654-701).
__global__ __launch_bounds__(NT) void k_swe_dragwtf(SwK P, const double* Vs, const double* eta, const double* bXP,
                                                     const double* bYP, double* CDx, double* CDy, double* wtfx,
                                                     double* wtfy) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  if (Vs[c] > 0.0) {
    const double coef = P.gravity * P.manning * P.manning;
    const double effh = Vs[c] / (P.dx * P.dy);
    const double expo = (effh < P.hD) ? (2.0 / 3.0) : (1.0 / 3.0);
    const double cd = coef / pow(effh, expo);
    CDx[c] = cd;
    CDy[c] = cd;
  }
  double wx = 0.0, wy = 0.0;
  const double e = eta[c];
  if (e < bXP[c] && eta[c + 1] - bXP[c] > P.wtfh) wx = -1.0;
  if (e < bYP[c] && eta[c + g.sy] - bYP[c] > P.wtfh) wy = -1.0;
  if (e < bXP[c - 1] && eta[c - 1] - bXP[c - 1] > P.wtfh) wx = 1.0;
  if (e < bYP[c - g.sy] && eta[c - g.sy] - bYP[c - g.sy] > P.wtfh) wy = 1.0;
  wtfx[c] = wx;
  wtfy[c] = wy;
}

struct VelIn {
  const double *Ex, *Ey, *Dx, *Dy, *Vsx, *Vsy, *Asx, *Asy, *eta, *dept, *cfl_active;
};
struct VelOut {
  double *uu, *vv, *Fu, *Fv, *cflx, *cfly;
};

__device__ inline double vel_x(const SwK& P, const VelIn& a, long long c) {
  const double coef = P.gravity * P.dt;
  const double effhx = (a.Vsx[c] > 0.0) ? a.Asx[c] / a.Vsx[c] : 0.0;
  return (a.Ex[c] - coef * effhx * (a.eta[c + 1] - a.eta[c])) * a.Dx[c];
}
__device__ inline double vel_y(const SwK& P, const VelIn& a, long long c, long long sy) {
  const double coef = P.gravity * P.dt;
  const double effhy = (a.Vsy[c] > 0.0) ? a.Asy[c] / a.Vsy[c] : 0.0;
  return (a.Ey[c] - coef * effhy * (a.eta[c + sy] - a.eta[c])) * a.Dy[c];
}

__global__ __launch_bounds__(NT) void k_swe_vel(SwK P, VelIn a, VelOut o) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  const double wt = P.wtfh;
  // x face between c and c+1
  double u = vel_x(P, a, c);
  {
    bool z = a.Asx[c] < wt * P.dy || a.cfl_active[c] == 1.0 || (u > 0.0 && a.dept[c] < wt);
    if (t.i + 1 < g.nx) z = z || a.cfl_active[c + 1] == 1.0 || (u < 0.0 && a.dept[c + 1] < wt);
    if (z) u = 0.0;
  }
  double v = vel_y(P, a, c, g.sy);
  {
    bool z = a.Asy[c] < wt * P.dx || a.cfl_active[c] == 1.0 || (v > 0.0 && a.dept[c] < wt);
    if (t.j + 1 < g.ny) z = z || a.cfl_active[c + g.sy] == 1.0 || (v < 0.0 && a.dept[c + g.sy] < wt);
    if (z) v = 0.0;
  }
  o.uu[c] = u;
  o.vv[c] = v;
  o.Fu[c] = u * a.Asx[c];
  o.Fv[c] = v * a.Asy[c];
  o.cflx[c] = fabs(u * P.dt / P.dx);
  o.cfly[c] = fabs(v * P.dt / P.dy);
  ghost_copy(g, t, o.uu, u);
  ghost_copy(g, t, o.vv, v);
}

__global__ __launch_bounds__(NT) void k_swe_interp(SwK P, const double* uu, const double* vv, double* uy, double* vx,
                                                    double* cfl_active) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  uy[c] = 0.25 * (uu[c] + uu[c - 1] + uu[c + g.sy] + uu[c - 1 + g.sy]);
  vx[c] = 0.25 * (vv[c] + vv[c - g.sy] + vv[c + 1] + vv[c + 1 - g.sy]);
  cfl_active[c] = 0.0;
}

struct GeoOut {
  double *Vs, *Asx, *Asy, *Asz, *Aszx, *Aszy, *Vsx, *Vsy;
};

// updateGeometry (SweSolver.cpp:306-369): face means read the neighbour's freshly computed
// value (ghost neighbours: from their own depth, before the ghost fill), then edge ghosts.
__global__ __launch_bounds__(NT) void k_swe_geom(SwK P, const double* dept, const double* deptx, const double* depty,
                                                  GeoOut o) {
  const C2 t = cell2(P.g);
  if (!t.valid) return;
  const Geo& g = P.g;
  const long long c = t.c;
  const double dxdy = P.dx * P.dy;
  auto vs = [&](long long q) { return dept[q] * P.dx * P.dy; };
  auto asz = [&](long long q) { return dept[q] > 0.0 ? dxdy : 0.0; };
  const double Vs = vs(c), Asz = asz(c);
  double Asx = deptx[c] * P.dy, Asy = depty[c] * P.dx;
  if (g.nx == 1) Asx = 0.0;
  if (g.ny == 1) Asy = 0.0;
  o.Vs[c] = Vs;
  o.Asz[c] = Asz;
  o.Asx[c] = Asx;
  o.Asy[c] = Asy;
  o.Vsx[c] = 0.5 * (Vs + vs(c + 1));
  o.Vsy[c] = 0.5 * (Vs + vs(c + g.sy));
  o.Aszx[c] = 0.5 * (Asz + asz(c + 1));
  o.Aszy[c] = 0.5 * (Asz + asz(c + g.sy));
  ghost_copy(g, t, o.Vs, Vs);
  ghost_copy(g, t, o.Asx, Asx);
  ghost_copy(g, t, o.Asy, Asy);
  ghost_copy(g, t, o.Asz, Asz);
  if (t.j == 0) {
    o.Vsy[c - g.sy] = Vs;
    o.Aszy[c - g.sy] = Asz;
  }
  if (t.i == 0) {
    o.Vsx[c - 1] = Vs;
    o.Aszx[c - 1] = Asz;
  }
}

// PolygonBC::applySurface (src/bc/PolygonBC.cpp:16-78): each interior cell tagged with a region
// takes its rule (prescribed depth; discharge Q dt spread over the region's cells; critical-depth
// weir q = sqrt(g) h^1.5 min(dx, dy)), outflow limited by the ponded volume, then updateDepth's
// interior snap (SweSolver.cpp:219-227) for the cells the rule changed (every other cell was
// snapped at the end of the step). Per-block partial of the net outflow volume.
__global__ __launch_bounds__(NT) void k_swe_boundary(SwK P, const int32_t* __restrict__ idx,
                                                      const SurfaceBcRule* __restrict__ rules, double dt,
                                                      double* __restrict__ eta, const double* __restrict__ bottom,
                                                      double* __restrict__ part) {
  __shared__ double lds[NT / kWave];
  const C2 t = cell2(P.g);
  double out = 0.0;
  if (t.valid) {
    const int r = idx[t.c];
    if (r >= 0) {
      const SurfaceBcRule R = rules[r];
      const double area = P.dx * P.dy;
      const double e0 = eta[t.c], b = bottom[t.c];
      const double depth = fmax(e0 - b, 0.0);
      double e = e0;
      if (R.kind == FREHG_BC_DEPTH) {
        e = b + fmax(R.rate, 0.0);
        out = (e0 - e) * area;
      } else if (R.kind == FREHG_BC_DISCHARGE) {
        if (R.count > 0) {
          double vol = R.rate * dt / (double)R.count;
          if (vol > depth * area) vol = depth * area;
          e = e0 - vol / area;
          out = vol;
        }
      } else {
        const double q = sqrt(P.gravity) * pow(depth, 1.5) * fmin(P.dx, P.dy);
        double vol = q * dt;
        if (vol > depth * area) vol = depth * area;
        e = e0 - vol / area;
        out = vol;
      }
      const double diff = e - b;
      if (diff > 0.0 && diff < P.min_depth) e = b;
      eta[t.c] = e;
    }
  }
  const double s = block_sum<NT>(out, lds);
  if (threadIdx.x == 0) part[blockIdx.x + gridDim.x * blockIdx.y] = s;
}

// maxCfl (SweSolver.cpp:954-975).
__global__ __launch_bounds__(NT) void k_swe_maxcfl(SwK P, const double* dept, const double* uu, const double* vv,
                                                    double* part, unsigned int nblocks) {
  __shared__ double lm[NT / kWave];
  const C2 t = cell2(P.g);
  double v = 0.0;
  if (t.valid) {
    const double h = dept[t.c];
    if (h > 0.0) {
      const double speed = sqrt(uu[t.c] * uu[t.c] + vv[t.c] * vv[t.c]);
      v = (speed + sqrt(P.gravity * h)) * P.dt / fmin(P.dx, P.dy);
    }
  }
  v = wave_max(v);
  if (threadIdx.x % kWave == 0) lm[threadIdx.x / kWave] = v;
  __syncthreads();
  const unsigned int bid = blockIdx.x + gridDim.x * blockIdx.y;
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int q = 0; q < NT / kWave; ++q) a = fmax(a, lm[q]);
    part[bid] = a;
    part[nblocks + bid] = 0.0;
  }
}

const char* kNames[] = {"eta",  "etan", "bottom", "bottomXP", "bottomYP", "dept", "deptx", "depty", "uu",
                        "vv",   "un",   "vn",     "uy",       "vx",       "Vs",   "Vsn",   "Asx",   "Asy",
                        "Asz",  "Aszx", "Aszy",   "Vsx",      "Vsy",      "CDx",  "CDy",   "Dx",    "Dy",
                        "Ex",   "Ey",   "cflx",   "cfly",     "Fu",       "Fv",   "Sxp",   "Sxm",   "Syp",
                        "Sym",  "Sct",  "Srhs",   "cfl_active", "wtfx",   "wtfy"};

__global__ __launch_bounds__(NT) void k_copy_field(const double* __restrict__ src, double* __restrict__ dst,
                                                   long long n) {
  const long long c = (long long)blockIdx.x * NT + threadIdx.x;
  if (c < n) dst[c] = src[c];
}

}  // namespace

SweState::SweState(int nx, int ny, double dx, double dy, const frehg_swe_params& p, hipStream_t s, Profiler* prof)
    : g_(make_geo(nx, ny, 1)), dx_(dx), dy_(dy), p_(p), s_(s), prof_(prof) {
  if (nx < 1 || ny < 1) throw ArgError("frehg_swe_init: grid sizes must be >= 1");
  if (!(dx > 0 && dy > 0)) throw ArgError("frehg_swe_init: dx, dy must be > 0");
  n_ = g_.nsurf();
  const size_t bytes = (size_t)n_ * sizeof(double);
  for (const char* nm : kNames) {
    double* ptr = nullptr;
    FREHG_HIP(hipMalloc(&ptr, bytes));
    FREHG_HIP(hipMemsetAsync(ptr, 0, bytes, s_));
    f_[nm] = ptr;
  }
  for (double** q : {&D_, &ax_, &ay_, &b_}) {
    FREHG_HIP(hipMalloc(q, bytes));
    FREHG_HIP(hipMemsetAsync(*q, 0, bytes, s_));
  }
  FREHG_HIP(hipMalloc(&fixed_, (size_t)n_));
  FREHG_HIP(hipMemsetAsync(fixed_, 0, (size_t)n_, s_));
  const unsigned int nb = ceil_div(nx, 64) * ceil_div(ny, 4);
  FREHG_HIP(hipMalloc(&red_, 2 * (size_t)nb * sizeof(double)));
  FREHG_HIP(hipMalloc(&dres_, 4 * sizeof(double)));
  FREHG_HIP(hipHostMalloc(&hres_, 4 * sizeof(double), hipHostMallocDefault));
  FREHG_HIP(hipMalloc(&dcount_, 16));
  FREHG_HIP(hipMemsetAsync(dcount_, 0, 16, s_));
  kry_.reset(new Krylov(g_, false, s_, prof_));
  if (const char* env = std::getenv("FREHG_SW_GRAPH")) use_graph_ = std::atoi(env) != 0;
  set_params(p);
  FREHG_HIP(hipStreamSynchronize(s_));
}

SweState::~SweState() {
  for (hipGraphExec_t ge : {graph_pre_, graph_post_})
    if (ge) (void)hipGraphExecDestroy(ge);
  for (auto& kv : f_) (void)hipFree(kv.second);
  for (void* q : {(void*)D_, (void*)ax_, (void*)ay_, (void*)b_, (void*)fixed_, (void*)red_, (void*)dres_,
                  (void*)dcount_, (void*)bcidx_, (void*)bcrules_, (void*)bcpart_})
    if (q) (void)hipFree(q);
  if (hres_) (void)hipHostFree(hres_);
}

void SweState::set_params(const frehg_swe_params& p) {
  if (!(p.dt > 0)) throw ArgError("frehg_swe: dt must be > 0");
  (void)krylov_config(p.solver, "frehg_swe");
  p_ = p;
}

double* SweState::field(const std::string& name) {
  auto it = f_.find(name);
  if (it == f_.end()) throw ArgError("unknown SW field '" + name + "'");
  return it->second;
}

void SweState::set_field(const std::string& name, const double* src) {
  copy_field_in(g_, 1, field(name), src, s_);
  FREHG_HIP(hipStreamSynchronize(s_));
}

void SweState::get_field(const std::string& name, double* dst) {
  copy_field_out(g_, 1, dst, field(name), s_);
  FREHG_HIP(hipStreamSynchronize(s_));
}

void SweState::set_raw(const std::string& name, const double* src) {
  FREHG_HIP(hipMemcpyAsync(field(name), src, (size_t)n_ * sizeof(double), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

void SweState::get_raw(const std::string& name, double* dst) {
  FREHG_HIP(hipMemcpyAsync(dst, field(name), (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

static SwK make_swk(const Geo& g, double dx, double dy, const frehg_swe_params& p) {
  SwK P;
  P.g = g;
  P.dx = dx;
  P.dy = dy;
  P.dt = p.dt;
  P.gravity = p.gravity;
  P.manning = p.manning;
  P.min_depth = p.min_depth;
  P.visc_x = p.visc_x;
  P.visc_y = p.visc_y;
  P.hD = p.hD;
  P.hE = p.hE;
  P.wtfh = p.wtfh;
  return P;
}

// setBathymetry + updateBottomFaces (SweSolver.cpp:144-205), evaluated once on the host.
void SweState::set_bathymetry(const double* bed) {
  const int nx = g_.nx, ny = g_.ny;
  std::vector<double> b((size_t)n_, 0.0), bxp((size_t)n_, 0.0), byp((size_t)n_, 0.0);
  get_raw("bottom", b.data());
  get_raw("bottomXP", bxp.data());
  get_raw("bottomYP", byp.data());
  auto s = [&](int i, int j) { return (size_t)g_.sidx(i, j); };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) b[s(i, j)] = bed[(size_t)i + (size_t)j * nx];
  for (int i = 0; i < nx; ++i) {
    b[s(i, -1)] = b[s(i, 0)];
    b[s(i, ny)] = b[s(i, ny - 1)];
  }
  for (int j = 0; j < ny; ++j) {
    b[s(-1, j)] = b[s(0, j)];
    b[s(nx, j)] = b[s(nx - 1, j)];
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      bxp[s(i, j)] = b[s(i + 1, j)] > b[s(i, j)] ? b[s(i + 1, j)] : b[s(i, j)];
      byp[s(i, j)] = b[s(i, j + 1)] > b[s(i, j)] ? b[s(i, j + 1)] : b[s(i, j)];
    }
  for (int i = 0; i < nx; ++i) byp[s(i, -1)] = b[s(i, -1)] > b[s(i, 0)] ? b[s(i, -1)] : b[s(i, 0)];
  for (int i = 0; i < nx; ++i) byp[s(i, ny)] = byp[s(i, ny - 1)];
  for (int j = 0; j < ny; ++j) bxp[s(-1, j)] = b[s(-1, j)] > b[s(0, j)] ? b[s(-1, j)] : b[s(0, j)];
  for (int j = 0; j < ny; ++j) bxp[s(nx, j)] = bxp[s(nx - 1, j)];
  set_raw("bottom", b.data());
  set_raw("bottomXP", bxp.data());
  set_raw("bottomYP", byp.data());
}

// initializeState (SweSolver.cpp:926-934) with finalizeInitialState (:904-924).
void SweState::initialize_state(double init_eta) { initialize_eta(nullptr, init_eta); }

void SweState::initialize_eta(const double* eta_phys, double init_eta) {
  const int nx = g_.nx, ny = g_.ny;
  std::vector<double> eta((size_t)n_), bot((size_t)n_);
  get_raw("eta", eta.data());
  get_raw("bottom", bot.data());
  auto s = [&](int i, int j) { return (size_t)g_.sidx(i, j); };
  // setInitialEtaPhysical (:855-870) + enforceSurfBc (:371-384)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      eta[s(i, j)] = (eta_phys ? eta_phys[(size_t)i + (size_t)j * nx] : init_eta) + p_.offset;
      if (eta[s(i, j)] < bot[s(i, j)]) eta[s(i, j)] = bot[s(i, j)];
    }
  for (int i = 0; i < nx; ++i) {
    eta[s(i, -1)] = eta[s(i, 0)];
    eta[s(i, ny)] = eta[s(i, ny - 1)];
  }
  for (int j = 0; j < ny; ++j) {
    eta[s(-1, j)] = eta[s(0, j)];
    eta[s(nx, j)] = eta[s(nx - 1, j)];
  }
  // updateDepth's interior snap (:219-227); ghosts keep the pre-snap copy
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double diff = eta[s(i, j)] - bot[s(i, j)];
      if (diff > 0.0 && diff < p_.min_depth) eta[s(i, j)] = bot[s(i, j)];
    }
  set_raw("eta", eta.data());
  // setInitialVelocityConstant(0, 0) (:872-884) + zero-gradient velocity ghosts (:906-907)
  for (const char* nm : {"uu", "vv", "CDx", "CDy"}) {
    std::vector<double> a((size_t)n_);
    get_raw(nm, a.data());
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) a[s(i, j)] = 0.0;
    if (std::string(nm) == "uu" || std::string(nm) == "vv") {
      for (int i = 0; i < nx; ++i) {
        a[s(i, -1)] = a[s(i, 0)];
        a[s(i, ny)] = a[s(i, ny - 1)];
      }
      for (int j = 0; j < ny; ++j) {
        a[s(-1, j)] = a[s(0, j)];
        a[s(nx, j)] = a[s(nx - 1, j)];
      }
    }
    set_raw(nm, a.data());
  }
  const SwK P = make_swk(g_, dx_, dy_, p_);
  const dim3 grid(ceil_div(nx, 64), ceil_div(ny, 4)), blk(NT);
  hipLaunchKernelGGL(k_swe_depth, grid, blk, 0, s_, P, f_["eta"], f_["bottom"], f_["dept"], f_["deptx"], f_["depty"]);
  FREHG_LAUNCH_CHECK();
  GeoOut go{f_["Vs"], f_["Asx"], f_["Asy"], f_["Asz"], f_["Aszx"], f_["Aszy"], f_["Vsx"], f_["Vsy"]};
  hipLaunchKernelGGL(k_swe_geom, grid, blk, 0, s_, P, f_["dept"], f_["deptx"], f_["depty"], go);
  FREHG_LAUNCH_CHECK();
  const size_t bytes = (size_t)n_ * sizeof(double);
  FREHG_HIP(hipMemcpyAsync(f_["etan"], f_["eta"], bytes, hipMemcpyDeviceToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(f_["un"], f_["uu"], bytes, hipMemcpyDeviceToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(f_["vn"], f_["vv"], bytes, hipMemcpyDeviceToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(f_["Vsn"], f_["Vs"], bytes, hipMemcpyDeviceToDevice, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

// Kernels of the step before the surface solve: savePrevious (eta), surface BCs, momentum
// coefficients, assembly and the symmetric form of the system.
void SweState::enq_pre(double evap) {
  const SwK P = make_swk(g_, dx_, dy_, p_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4)), blk(NT);
  auto F = [&](const char* nm) { return f_[nm]; };
  copy_field(F("eta"), F("etan"));
  hipLaunchKernelGGL(k_swe_surfbc, grid, blk, 0, s_, P, F("eta"), F("bottom"));
  FREHG_LAUNCH_CHECK();
  MomIn m{F("uu"), F("vv"), F("vx"), F("uy"), F("cflx"), F("cfly"), F("Vsx"), F("Vsy"),
          F("Asx"), F("Asy"), F("Aszx"), F("Aszy"), F("CDx"), F("CDy")};
  hipLaunchKernelGGL(k_swe_momentum, grid, blk, 0, s_, P, m, F("Dx"), F("Dy"), F("Ex"), F("Ey"));
  FREHG_LAUNCH_CHECK();
  AsmIn a{F("eta"), F("bottom"), F("Asz"), F("Asx"), F("Asy"), F("Ex"), F("Ey"), F("Vsx"),
          F("Vsy"), F("Dx"),     F("Dy"),  F("dept"), F("uu"), F("vv")};
  AsmOut o{F("Sxp"), F("Sxm"), F("Syp"), F("Sym"), F("Sct"), F("Srhs")};
  hipLaunchKernelGGL(k_swe_assemble, grid, blk, 0, s_, P, a, o, evap);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_swe_fixrows, grid, blk, 0, s_, P, o, fixed_, F("eta"));
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_swe_symm, grid, blk, 0, s_, P, o, fixed_, F("eta"), ax_, ay_, b_);
  FREHG_LAUNCH_CHECK();
}

// Kernels of the step after the surface solve: depth, drag and wet/dry faces, velocities,
// their interpolation and the surface geometry for the next step.
void SweState::enq_post(double rain, double evap) {
  const SwK P = make_swk(g_, dx_, dy_, p_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4)), blk(NT);
  auto F = [&](const char* nm) { return f_[nm]; };
  hipLaunchKernelGGL(k_swe_post, grid, blk, 0, s_, P, F("eta"), F("bottom"), F("dept"), F("cfl_active"), rain, evap);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_swe_depth, grid, blk, 0, s_, P, F("eta"), F("bottom"), F("dept"), F("deptx"), F("depty"));
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_swe_dragwtf, grid, blk, 0, s_, P, F("Vs"), F("eta"), F("bottomXP"), F("bottomYP"), F("CDx"),
                     F("CDy"), F("wtfx"), F("wtfy"));
  FREHG_LAUNCH_CHECK();
  copy_field(F("uu"), F("un"));
  copy_field(F("vv"), F("vn"));
  VelIn vi{F("Ex"), F("Ey"), F("Dx"), F("Dy"), F("Vsx"), F("Vsy"), F("Asx"), F("Asy"), F("eta"), F("dept"),
           F("cfl_active")};
  VelOut vo{F("uu"), F("vv"), F("Fu"), F("Fv"), F("cflx"), F("cfly")};
  hipLaunchKernelGGL(k_swe_vel, grid, blk, 0, s_, P, vi, vo);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_swe_interp, grid, blk, 0, s_, P, F("uu"), F("vv"), F("uy"), F("vx"), F("cfl_active"));
  FREHG_LAUNCH_CHECK();
  copy_field(F("Vs"), F("Vsn"));
  GeoOut go{F("Vs"), F("Asx"), F("Asy"), F("Asz"), F("Aszx"), F("Aszy"), F("Vsx"), F("Vsy")};
  hipLaunchKernelGGL(k_swe_geom, grid, blk, 0, s_, P, F("dept"), F("deptx"), F("depty"), go);
  FREHG_LAUNCH_CHECK();
}

void SweState::copy_field(const double* src, double* dst) {
  hipLaunchKernelGGL(k_copy_field, dim3(ceil_div(n_, NT)), dim3(NT), 0, s_, src, dst, n_);
  FREHG_LAUNCH_CHECK();
}

// With FREHG_SW_GRAPH=1 the step's kernels are replayed from two captured graphs around the
// surface solve (its start and a budget of CG iterations are in the first): a graph launch
// replaces a host launch per kernel of this launch-bound 2D step. The graphs are captured again
// when the parameters, the forcing or the budget change; iterations past the budget run eagerly
// (Krylov::finish_cg). Off by default: on the 512 x 512 coupled workload the replayed kernels
// run back to back but the graph is dispatched in segments separated by 10-40 us, and the step
// is no faster than with eager launches.
void SweState::step(double rain, double evap, frehg_solve_stats* st) {
  KrylovConfig kc = krylov_config(p_.solver, "frehg_swe_step");
  kc.precision = 0;  // the surface system is small; it stays FP64
  StencilView<double> A{f_["Sct"], ax_, ay_, nullptr};
  SolveResult sr;
  if (use_graph_ && kc.ksp == FREHG_KSP_CG && solved_) {
    // budget: one iteration more than the last solve took; grown at once, shrunk only after it
    // has been too large for 16 steps (each change costs a capture)
    const int want = std::max(2, kry_->last_iterations() + 1);
    if (budget_ == 0 || want > budget_) {
      budget_ = want;
      shrink_ = 0;
    } else if (want < budget_ && ++shrink_ >= 16) {
      budget_ = want;
      shrink_ = 0;
    } else if (want == budget_) {
      shrink_ = 0;
    }
    const bool same = graph_pre_ && graph_rain_ == rain && graph_evap_ == evap && graph_budget_ == budget_ &&
                      std::memcmp(&graph_params_, &p_, sizeof(p_)) == 0;
    if (!same) {
      for (hipGraphExec_t* ge : {&graph_pre_, &graph_post_})
        if (*ge) {
          FREHG_HIP(hipGraphExecDestroy(*ge));
          *ge = nullptr;
        }
      auto capture = [&](hipGraphExec_t* out, auto&& enqueue) {
        hipGraph_t g;
        FREHG_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
          enqueue();
        } catch (...) {
          (void)hipStreamEndCapture(s_, &g);
          throw;
        }
        FREHG_HIP(hipStreamEndCapture(s_, &g));
        FREHG_HIP(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
        FREHG_HIP(hipGraphDestroy(g));
      };
      capture(&graph_pre_, [&] {
        enq_pre(evap);
        kry_->enqueue_cg(A, b_, f_["Srhs"], f_["eta"], fixed_, true, kc, budget_);
      });
      capture(&graph_post_, [&] { enq_post(rain, evap); });
      graph_rain_ = rain;
      graph_evap_ = evap;
      graph_budget_ = budget_;
      graph_params_ = p_;
    }
    FREHG_HIP(hipGraphLaunch(graph_pre_, s_));
    sr = kry_->finish_cg(A, f_["eta"], kc, budget_);
    FREHG_HIP(hipGraphLaunch(graph_post_, s_));
  } else {
    enq_pre(evap);
    sr = kry_->solve(A, nullptr, b_, f_["Srhs"], f_["eta"], fixed_, true, false, kc);
    enq_post(rain, evap);
    solved_ = true;
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = sr.iterations;
    st->residual_norm = sr.residual_norm;
    st->rhs_norm = sr.rhs_norm;
    st->converged = sr.converged ? 1 : 0;
  }
}

void SweState::set_boundary_polygons(int nreg, const int32_t* kind, const double* rate, const int32_t* nvert,
                                     const double* xy, double x0, double y0) {
  if (nreg < 0) throw ArgError("frehg_swe_set_boundary_polygons: nregions must be >= 0");
  if (nreg > 0 && (!kind || !rate || !nvert || !xy)) throw ArgError("frehg_swe_set_boundary_polygons: null argument");
  // PolygonIndex::build (src/bc/PolygonIndex.cpp:9-33): cell centres, first match wins; the
  // per-region cell counts of globalColumnCounts (:45-56)
  std::vector<SurfaceBcRule> rules((size_t)nreg);
  std::vector<size_t> off((size_t)nreg + 1, 0);
  for (int r = 0; r < nreg; ++r) {
    if (kind[r] != FREHG_BC_DISCHARGE && kind[r] != FREHG_BC_DEPTH && kind[r] != FREHG_BC_CRITICAL)
      throw ArgError("frehg_swe_set_boundary_polygons: unknown region kind");
    if (nvert[r] < 0) throw ArgError("frehg_swe_set_boundary_polygons: negative vertex count");
    rules[(size_t)r].kind = kind[r];
    rules[(size_t)r].rate = rate[r];
    rules[(size_t)r].count = 0;
    off[(size_t)r + 1] = off[(size_t)r] + (size_t)nvert[r];
  }
  std::vector<int32_t> idx((size_t)n_, -1);
  for (int j = 0; j < g_.ny; ++j)
    for (int i = 0; i < g_.nx; ++i) {
      const double x = x0 + (i + 0.5) * dx_, y = y0 + (j + 0.5) * dy_;
      for (int r = 0; r < nreg; ++r)
        if (polygon_contains(xy + 2 * off[(size_t)r], nvert[r], x, y)) {
          idx[(size_t)g_.sidx(i, j)] = r;
          ++rules[(size_t)r].count;
          break;
        }
    }
  for (void* q : {(void*)bcidx_, (void*)bcrules_})
    if (q) FREHG_HIP(hipFree(q));
  bcidx_ = nullptr;
  bcrules_ = nullptr;
  nbc_ = nreg;
  if (nreg == 0) return;
  FREHG_HIP(hipMalloc(&bcidx_, (size_t)n_ * sizeof(int32_t)));
  FREHG_HIP(hipMalloc(&bcrules_, (size_t)nreg * sizeof(SurfaceBcRule)));
  if (!bcpart_) {
    const unsigned int nb = ceil_div(g_.nx, 64) * ceil_div(g_.ny, 4);
    FREHG_HIP(hipMalloc(&bcpart_, (size_t)nb * sizeof(double)));
  }
  FREHG_HIP(hipMemcpy(bcidx_, idx.data(), (size_t)n_ * sizeof(int32_t), hipMemcpyHostToDevice));
  FREHG_HIP(hipMemcpy(bcrules_, rules.data(), (size_t)nreg * sizeof(SurfaceBcRule), hipMemcpyHostToDevice));
}

void SweState::apply_boundaries_async(double dt) {
  if (nbc_ == 0) {
    hres_[3] = 0.0;
    return;
  }
  const SwK P = make_swk(g_, dx_, dy_, p_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4)), blk(NT);
  hipLaunchKernelGGL(k_swe_boundary, grid, blk, 0, s_, P, (const int32_t*)bcidx_, (const SurfaceBcRule*)bcrules_, dt,
                     f_["eta"], (const double*)f_["bottom"], bcpart_);
  FREHG_LAUNCH_CHECK();
  // depth and geometry refresh (PolygonBC.cpp:74-77); updateGeometry first keeps Vs as Vsn
  hipLaunchKernelGGL(k_swe_depth, grid, blk, 0, s_, P, f_["eta"], f_["bottom"], f_["dept"], f_["deptx"], f_["depty"]);
  FREHG_LAUNCH_CHECK();
  copy_field(f_["Vs"], f_["Vsn"]);
  GeoOut go{f_["Vs"], f_["Asx"], f_["Asy"], f_["Asz"], f_["Aszx"], f_["Aszy"], f_["Vsx"], f_["Vsy"]};
  hipLaunchKernelGGL(k_swe_geom, grid, blk, 0, s_, P, f_["dept"], f_["deptx"], f_["depty"], go);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)bcpart_, grid.x * grid.y, 1,
                     dres_ + 3);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(hres_ + 3, dres_ + 3, sizeof(double), hipMemcpyDeviceToHost, s_));
}

double SweState::boundary_outflow() {
  FREHG_HIP(hipStreamSynchronize(s_));  // returns at once when the stream has already drained
  return hres_[3];
}

double SweState::max_cfl() {
  const SwK P = make_swk(g_, dx_, dy_, p_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4)), blk(NT);
  hipLaunchKernelGGL(k_swe_maxcfl, grid, blk, 0, s_, P, f_["dept"], f_["uu"], f_["vv"], red_, grid.x * grid.y);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_maxmin, dim3(1), dim3(kRedThreads), 0, s_, (const double*)red_, grid.x * grid.y, dres_);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(hres_, dres_, sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  return hres_[0];
}

}  // namespace frehg
