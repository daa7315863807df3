// Oracle Richards-equation PCA step -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Follows src/re/ReSolver.cpp (serial, non-baroclinic path); line numbers cite that file.
#include <algorithm>
#include <stdexcept>

#include "oracle.hpp"

namespace frehg_oracle {

ReSolver::ReSolver(int nx, int ny, int nz, const ReParams& p) : p_(p) {
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  const size_t n = (size_t)g.ncell();
  for (auto* a : {&h, &hn, &wc, &wcn, &wch, &hwc, &ch, &Kx, &Ky, &Kz, &qx, &qy, &qz, &Gxp, &Gxm,
                  &Gyp, &Gym, &Gzp, &Gzm, &Gct, &Grhs})
    a->assign(n, 0.0);
  qtop_field.assign((size_t)g.nsurf(), 0.0);
}

void ReSolver::setSoilMap(const std::vector<Soil>& classes, const std::vector<int>& class_idx,
                          bool idx3d, const std::vector<real>& ks_cell) {
  classes_ = classes;
  class_idx_ = class_idx;
  idx3d_ = idx3d;
  ks_cell_ = ks_cell;
}

// ReSolver.cpp:87-96: class of the clamped (i,j,k); uniform params_.soil when no map.
Soil ReSolver::soilAt(int i, int j, int k) const {
  const int ci = i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i);
  const int cj = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
  const int ck = k < 0 ? 0 : (k >= g.nz ? g.nz - 1 : k);
  Soil s = p_.soil;
  if (!classes_.empty()) {
    int cls = 0;
    if (!class_idx_.empty()) {
      const size_t o = idx3d_ ? (size_t)ci + (size_t)cj * g.nx + (size_t)ck * g.nx * g.ny
                              : (size_t)ci + (size_t)cj * g.nx;
      cls = class_idx_[o];
    }
    s = classes_[(size_t)cls];
  }
  if (!ks_cell_.empty()) {
    const real ks = ks_cell_[(size_t)ci + (size_t)cj * g.nx + (size_t)ck * g.nx * g.ny];
    s.Ks_x = s.Ks_y = s.Ks_z = ks;
  }
  return s;
}

void ReSolver::setTopFluxField(const std::vector<real>& q) {
  if ((int)q.size() != g.nx * g.ny) throw std::runtime_error("setTopFluxField: expected nx*ny values");
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) qtop_field[(size_t)g.sidx(i, j)] = q[(size_t)(i + j * g.nx)];
  has_qtop_field = true;
}

void ReSolver::initializeGeometry() {}  // uniform dz, actv = 1 (ReSolver.cpp:104-131)

void ReSolver::setInitialWaterContent(const std::vector<real>& w) {  // :135-153
  if ((long long)w.size() != (long long)g.nx * g.ny * g.nz)
    throw std::runtime_error("setInitialWaterContent: expected nx*ny*nz values");
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        const real v = w[(size_t)i + (size_t)j * g.nx + (size_t)k * g.nx * g.ny];
        wc[c] = v;
        h[c] = VG::head(s, v);
        wcn[c] = v;
        hn[c] = h[c];
      }
}

void ReSolver::setInitialHead(const std::vector<real>& hv) {  // :155-173
  if ((long long)hv.size() != (long long)g.nx * g.ny * g.nz)
    throw std::runtime_error("setInitialHead: expected nx*ny*nz values");
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        const real v = hv[(size_t)i + (size_t)j * g.nx + (size_t)k * g.nx * g.ny];
        h[c] = v;
        wc[c] = VG::theta(s, v);
        wcn[c] = wc[c];
        hn[c] = v;
      }
}

void ReSolver::finalizeInitialState() {  // :175-178
  updateDiagnostics();
  dtn = p_.dt;
}

real ReSolver::soilKs(int i, int j, int k, int dir) const {
  const Soil s = soilAt(i, j, k);
  return dir == 0 ? s.Ks_x : (dir == 1 ? s.Ks_y : s.Ks_z);
}

void ReSolver::updateDiagnostics() {  // :199-215
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)g.nx * g.ny * g.nz))
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        hwc[c] = VG::head(s, wc[c]);
        wch[c] = VG::theta(s, h[c]);
        ch[c] = VG::capacity(s, h[c]);
      }
}

void ReSolver::computeKFace() {  // :217-380
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long sy = nx + 2, sz = g.plane();
  const int bc2 = p_.bc_type[2];
  const bool full3d = p_.use_full3d != 0;
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        const Soil sC = soilAt(i, j, k);
        {  // Kx :243-250
          const Soil sXp = soilAt(i + 1, j, k);
          const real Kp = VG::conductivity(sXp, h[c + 1], sXp.Ks_x);
          const real Km = VG::conductivity(sC, h[c], sC.Ks_x);
          Kx[c] = 0.5 * (Kp + Km);
        }
        {  // Ky :251-265 (serial: this rank owns the y+ wall)
          const Soil sYp = soilAt(i, j + 1, k);
          const real Kp = VG::conductivity(sYp, h[c + sy], sYp.Ks_y);
          const real Km = VG::conductivity(sC, h[c], sC.Ks_y);
          Ky[c] = 0.5 * (Kp + Km);
          if (j == ny - 1) {
            if (bc2 == 0) Ky[c] = 0.0;
            else if (bc2 == 1) Ky[c] = VG::conductivity(sC, h[c], sC.Ks_y);
          }
        }
        {  // Kz :272-284 (k+1 is never a top cell on a flat column; all cells active)
          const Soil sZp = soilAt(i, j, k + 1);
          const real Kp = VG::conductivity(sZp, h[c + sz], sZp.Ks_z);
          const real Km = VG::conductivity(sC, h[c], sC.Ks_z);
          if (k + 1 >= nz) Kz[c] = Km;
          else Kz[c] = 0.5 * (Kp + Km);
        }
      }
  if (!full3d) {  // :287-310
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) {
        const long long im = g.idx(-1, j, k);
        const Soil s0 = soilAt(0, j, k);
        const real Kp = VG::conductivity(s0, h[im], s0.Ks_x);
        Kx[g.idx(0, j, k)] = 0.5 * (Kp + Kp);
        if (p_.bc_type[0] == 0) Kx[g.idx(-1, j, k)] = 0.0;
        if (p_.bc_type[1] == 0) Kx[g.idx(nx, j, k)] = 0.0;
      }
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i) {
        const long long jm = g.idx(i, -1, k);
        const Soil s0 = soilAt(i, 0, k);
        const real Kp = VG::conductivity(s0, h[jm], s0.Ks_y);
        Ky[g.idx(i, 0, k)] = 0.5 * (Kp + Kp);
        if (p_.bc_type[3] == 0) Ky[g.idx(i, -1, k)] = 0.0;
      }
  } else {  // :311-337 (serial: every global wall is owned)
    for (int k = 0; k < nz; ++k) {
      for (int j = 0; j < ny; ++j) {
        if (p_.bc_type[0] == 0) Kx[g.idx(-1, j, k)] = 0.0;
        if (p_.bc_type[1] == 0) Kx[g.idx(nx - 1, j, k)] = 0.0;
      }
      for (int i = 0; i < nx; ++i) {
        if (p_.bc_type[3] == 0) Ky[g.idx(i, -1, k)] = 0.0;
        if (p_.bc_type[2] == 0) Ky[g.idx(i, ny - 1, k)] = 0.0;
      }
    }
  }
  for (int j = 0; j < ny; ++j)  // bottom :338-344
    for (int i = 0; i < nx; ++i) {
      Kz[g.idx(i, j, nz)] = 0.0;
      if (p_.bc_type[4] == 0) Kz[g.idx(i, j, nz - 1)] = 0.0;
    }
  for (int j = 0; j < ny; ++j)  // top :345-363
    for (int i = 0; i < nx; ++i) {
      const long long c = g.idx(i, j, 0), km = g.idx(i, j, -1);
      const Soil s = soilAt(i, j, 0);
      const real Kp = VG::conductivity(s, h[c], s.Ks_z);
      if (p_.sim_shallowwater) {
      } else if (p_.bc_type[5] == 1) {
        Kz[km] = s.Ks_z;
      } else if (p_.bc_type[5] == 0) {
        Kz[km] = 0.0;
      } else {
        Kz[km] = Kp;
      }
    }
  if (!full3d) {  // :365-379 (sequential scatter; every write is a zero)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long long c = g.idx(i, j, k);
          if (wc[c] < soilAt(i, j, k).theta_s) {
            Kx[c] = 0.0;
            Kx[c - 1] = 0.0;
            Ky[c] = 0.0;
            Ky[c - sy] = 0.0;
          }
        }
  }
}

void ReSolver::matCoeff() {  // :397-470
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real dx = p_.dx, dy = p_.dy, dtg = p_.dt;
  const long long sy = nx + 2, sz = g.plane();
  const int bc2 = p_.bc_type[2], bc3 = p_.bc_type[3], bc4 = p_.bc_type[4], bc5 = p_.bc_type[5];
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        const Soil s = soilAt(i, j, k);
        const long long im = c - 1, jm = c - sy, km = c - sz;
        const real dzk = dzCell();
        const real Ax_k = dy * dzk;
        const real Ay_k = dx * dzk;
        Gxp[c] = -Kx[c] * dtg * Ax_k / dx;
        Gxm[c] = -Kx[im] * dtg * Ax_k / dx;
        Gyp[c] = -Ky[c] * dtg * Ay_k / dy;
        Gym[c] = -Ky[jm] * dtg * Ay_k / dy;
        if (j == ny - 1 && bc2 == 1) Gyp[c] = 2.0 * Gyp[c];
        if (j == 0 && bc3 == 1) Gym[c] = 2.0 * Gym[c];
        const real dz_down = dzk;  // uniform dz (both branches of :440)
        const real dzf_p = 0.5 * (dzk + dz_down);
        Gzp[c] = -Kz[c] * dtg / (dzk * dzf_p);
        Gzp[c] *= V();
        const real dzf_m = (k == 0) ? 0.5 * dzk : 0.5 * (dzk + dzk);
        Gzm[c] = -Kz[km] * dtg / (dzk * dzf_m);
        Gzm[c] *= V();
        Gct[c] = (ch[c] + s.Ss * wcn[c] / s.theta_s);
        Gct[c] *= V();
        Gct[c] -= (Gxp[c] + Gxm[c] + Gyp[c] + Gym[c]);
        if (!(k == nz - 1 && bc4 != 1)) Gct[c] -= Gzp[c];
        if (k == 0) {
          if (p_.sim_shallowwater) {
          } else if (bc5 == 1) {
            Gct[c] -= Gzm[c];
          }
        } else {
          Gct[c] -= Gzm[c];
        }
      }
}

void ReSolver::rhs() {  // :472-549 (serial: global == local)
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real dtg = p_.dt;
  const long long sy = nx + 2, sz = g.plane();
  const int bc4 = p_.bc_type[4], bc5 = p_.bc_type[5];
  const bool sim_sw = p_.sim_shallowwater != 0;
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        const Soil s = soilAt(i, j, k);
        const long long im = c - 1, jm = c - sy, km = c - sz, ip = c + 1, jp = c + sy, kp = c + sz;
        const real Vc = V();
        const real dzk = dzCell();
        real r = (ch[c] + s.Ss * wcn[c] / s.theta_s) * hn[c] * Vc;
        r -= dtg * Vc * Kz[c] / dzk;
        r += dtg * Vc * Kz[km] / dzk;
        if (i == nx - 1) r -= Gxp[c] * hn[ip];
        if (i == 0) r -= Gxm[c] * hn[im];
        if (j == ny - 1) r -= Gyp[c] * hn[jp];
        if (j == 0) r -= Gym[c] * hn[jm];
        if (k == nz - 1) {
          if (bc4 == 1) r += Gzp[c] * hn[kp];
          if (bc4 == 2) r += dtg * Vc * (p_.qbot + Kz[c]) / dzk;
        }
        if (k == 0 && !sim_sw && bc5 == 1) r -= Gzm[c] * p_.htop;
        if (k == 0 && !sim_sw && bc5 == 2) {
          const real qcol = has_qtop_field ? qtop_field[(size_t)g.sidx(i, j)] : p_.qtop;
          r += -dtg * Vc * (qcol + Kz[g.idx(i, j, -1)]) / dzk;
        }
        Grhs[c] = r;
      }
}

void ReSolver::assembleAndSolve(SolveStats* st) {  // :560-636 (COO rows -> stencil rows)
  StencilOp A;
  A.g = g;
  A.three_d = true;
  A.resize();
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const long long c = g.idx(i, j, k);
        A.diag[c] = Gct[c];
        A.xm[c] = Gxm[c];
        A.xp[c] = Gxp[c];
        A.ym[c] = Gym[c];
        A.yp[c] = Gyp[c];
        A.zm[c] = Gzm[c];
        A.zp[c] = Gzp[c];
      }
  std::vector<real> x;
  SolveStats s = pcg_solve(A, Grhs, x, p_.solver);
  if (st) *st = s;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) h[g.idx(i, j, k)] = x[g.idx(i, j, k)];
}

void ReSolver::enforceHeadBc() {  // :638-669
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      h[g.idx(-1, j, k)] = h[g.idx(0, j, k)];
      h[g.idx(nx, j, k)] = h[g.idx(nx - 1, j, k)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      h[g.idx(i, -1, k)] = h[g.idx(i, 0, k)];
      h[g.idx(i, ny, k)] = h[g.idx(i, ny - 1, k)];
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      h[g.idx(i, j, nz)] = (p_.bc_type[4] == 1) ? p_.hbot : h[g.idx(i, j, nz - 1)];
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      h[g.idx(i, j, -1)] =
          (!p_.sim_shallowwater && p_.bc_type[5] == 1) ? p_.htop : h[g.idx(i, j, 0)];
}

void ReSolver::groundwaterFlux() {  // :671-743 with darcyFluxZ :24-58
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real dx = p_.dx, dy = p_.dy;
  const long long sy = nx + 2, sz = g.plane();
  const real faceA = dx * dy;
  const int bc4 = p_.bc_type[4], bc5 = p_.bc_type[5];
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        if (p_.use_full3d) {
          const real dzk = dzCell();
          qx[c] = (dy * dzk) * Kx[c] * (h[c + 1] - h[c]) / dx;
          qy[c] = (dx * dzk) * Ky[c] * (h[c + sy] - h[c]) / dy;
        } else {
          qx[c] = 0.0;
          qy[c] = 0.0;
        }
        // darcyFluxZ(back = false)
        const long long kp = c + sz;
        real delta = 0.5 * (dzCell() + dzCell());
        real q;
        if (k == nz - 1 && bc4 == 0) {
          q = 0.0;
        } else {
          if (k == nz - 1) delta = 0.5 * dzCell();
          const Soil sC = soilAt(i, j, k), sZp = soilAt(i, j, k + 1);
          const real kc = VG::conductivity(sC, h[c], sC.Ks_z);
          const real ks = VG::conductivity(sZp, h[kp], sZp.Ks_z);
          const real kface = 0.5 * (kc + ks);
          q = faceA * kface * ((h[kp] - h[c]) / delta - 1.0);
        }
        qz[c] = q;
      }
  for (int j = 0; j < ny; ++j)  // re_flux_top :711-725
    for (int i = 0; i < nx; ++i) {
      const long long km = g.idx(i, j, -1), c = g.idx(i, j, 0);
      if (bc5 == 2) {
        const real qcol = has_qtop_field ? qtop_field[(size_t)g.sidx(i, j)] : p_.qtop;
        qz[km] = qcol * faceA;
      } else {
        const Soil st = soilAt(i, j, 0);
        const real delta = 0.5 * dzCell();
        real kface = st.Ks_z;
        if (bc5 == 0) kface = 0.0;
        qz[km] = faceA * kface * ((h[c] - h[km]) / delta - 1.0);
      }
    }
  if (bc4 == 2)  // re_flux_bot :729-734
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) qz[g.idx(i, j, nz - 1)] = p_.qbot * faceA;
}

void ReSolver::updateWaterContent() {  // :762-786
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real dtg = p_.dt;
  const long long sy = nx + 2, sz = g.plane();
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        const real Vc = V();
        const real coeff = Vc * (1.0 + s.Ss * (h[c] - hn[c]) / s.theta_s);
        const real dqx = dtg * (qx[c] - qx[c - 1]);
        const real dqy = dtg * (qy[c] - qy[c - sy]);
        const real dqz = dtg * (qz[c] - qz[c - sz]);
        wc[c] = (wcn[c] * Vc + dqx + dqy + dqz) / coeff;
      }
}

void ReSolver::reallocateWaterContent() {  // :788-837
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long sy = nx + 2, sz = g.plane();
  // wch/hwc of a cell depend only on its own h and its own pre-visit wc, so they are evaluated
  // in parallel first; the order-dependent part below is the reference's sequential sweep.
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)nx * ny * nz))
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        wch[c] = VG::theta(s, h[c]);
        hwc[c] = VG::head(s, wc[c]);
      }
  auto adjacentToSat = [&](int i, int j, int k) {  // :802-817
    const long long c = g.idx(i, j, k);
    if (wc[c + sz] >= soilAt(i, j, k + 1).theta_s) return true;
    if (Kx[c] != 0.0 && wc[c + 1] >= soilAt(i + 1, j, k).theta_s) return true;
    if (Kx[c - 1] != 0.0 && wc[c - 1] >= soilAt(i - 1, j, k).theta_s) return true;
    if (Ky[c] != 0.0 && wc[c + sy] >= soilAt(i, j + 1, k).theta_s) return true;
    if (Ky[c - sy] != 0.0 && wc[c - sy] >= soilAt(i, j - 1, k).theta_s) return true;
    if (k == 0) {
      if (!p_.sim_shallowwater && p_.bc_type[5] == 1 && p_.htop >= 0.0) return true;
    } else if (wc[c - sz] >= soilAt(i, j, k - 1).theta_s) {
      return true;
    }
    return false;
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        if (wc[c] >= s.theta_s) {
          wc[c] = s.theta_s;
        } else if (!adjacentToSat(i, j, k)) {
          if (wc[c] < 0.9999 * s.theta_s) h[c] = hwc[c];
        } else {
          wc[c] = wch[c];
        }
      }
}

void ReSolver::clampWaterContent() {  // :839-853
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)g.nx * g.ny * g.nz))
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        if (wc[c] > s.theta_s) wc[c] = s.theta_s;
        else if (wc[c] < s.theta_r) wc[c] = s.theta_r;
      }
}

void ReSolver::enforceMoistureBc() {  // :855-878
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      wc[g.idx(-1, j, k)] = wc[g.idx(0, j, k)];
      wc[g.idx(nx, j, k)] = wc[g.idx(nx - 1, j, k)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      wc[g.idx(i, -1, k)] = wc[g.idx(i, 0, k)];
      wc[g.idx(i, ny, k)] = wc[g.idx(i, ny - 1, k)];
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      wc[g.idx(i, j, nz)] = wc[g.idx(i, j, nz - 1)];
      wc[g.idx(i, j, -1)] = wc[g.idx(i, j, 0)];
    }
}

void ReSolver::adaptiveTimeStep() {  // :880-929 (max/min are exact, so order-free)
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long sy = nx + 2, sz = g.plane();
  real dq_max = 0.0, dt_Comin = 1.0e8;
  p_.dt = dtn;
  const real dtv = p_.dt;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : dq_max) reduction(min : dt_Comin) if(omp_wide((long long)nx * ny * nz))
  for (int k = 1; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const Soil s = soilAt(i, j, k);
        const long long c = g.idx(i, j, k);
        const real Ax = p_.dy * dzCell(), Ay = p_.dx * dzCell(), Az = p_.dx * p_.dy;
        const real qin = qx[c] / Ax + qy[c] / Ay + qz[c] / Az;
        const real qou = qx[c - 1] / Ax + qy[c - sy] / Ay + qz[c - sz] / Az;
        const real dq = std::fabs(qin - qou) * dtv / dzCell();
        dq_max = std::max(dq_max, dq);
        if (wc[c] < s.theta_s) {
          const real d = VG::dKdwc(s, wc[c], s.Ks_z);
          const real dt_Co = p_.co_max * dzCell() / d;
          dt_Comin = std::min(dt_Comin, dt_Co);
        }
      }
  if (dq_max > 0.02) p_.dt *= 0.75;
  else if (dq_max < 0.01) p_.dt *= 1.25;
  if (p_.dt > dt_Comin) p_.dt = dt_Comin;
  if (p_.dt > p_.dt_max) p_.dt = p_.dt_max;
  if (p_.dt < p_.dt_min) p_.dt = p_.dt_min;
  dtn = p_.dt;
}

static bool pmax_is_one(const ReParams& p) { return p.picard_max <= 1; }

void ReSolver::advanceStep(ReStepStats* st) {  // :931-969
  dt_used = p_.dt;
  const size_t n = h.size();
#pragma omp parallel for schedule(static) if(omp_wide((long long)n))
  for (long long c = 0; c < (long long)n; ++c) {  // savePreviousStepFields :190-197
    hn[c] = h[c];
    wcn[c] = wc[c];
  }
  // Picard outer loop (a new method; picard_max = 1 is the reference's predictor): K and C at the
  // iterate h^m (h^0 = h^n, ghost heads from the head BCs), the storage and rhs terms at time level
  // n as in the predictor, solve from x0 = 0 for h^{m+1}; stop when max |h^{m+1} - h^m| <= tol.
  SolveStats ss;
  int picard = 0, iters = 0;
  bool converged = pmax_is_one(p_);
  std::vector<real> hm;
  const int pmax = std::max(1, p_.picard_max);
  for (;;) {
    updateDiagnostics();
    computeKFace();
    matCoeff();
    rhs();
    if (pmax > 1) hm = h;
    assembleAndSolve(&ss);
    iters += ss.iterations;
    ++picard;
    enforceHeadBc();
    if (pmax == 1) break;
    real delta = 0.0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const long long c = g.idx(i, j, k);
          delta = std::max(delta, std::fabs(h[c] - hm[c]));
        }
    if (delta <= p_.picard_tol) {
      converged = true;
      break;
    }
    if (picard >= pmax) break;  // the cap: the step continues from the last iterate
  }
  ss.iterations = iters;
  computeKFace();
  groundwaterFlux();
  if (p_.use_corrector) {
    updateWaterContent();
  } else {
#pragma omp parallel for collapse(2) schedule(static) if(omp_wide((long long)g.nx * g.ny * g.nz))
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wc[g.idx(i, j, k)] = VG::theta(soilAt(i, j, k), h[g.idx(i, j, k)]);
  }
  // Ledger terms around the corrector (not part of the reference; the checker of the GPU ledger).
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const real Vc = V();
  real before = 0.0, ss_term = 0.0, bnd = 0.0;
  std::vector<real> wcorr((size_t)nx * ny * nz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        const Soil s = soilAt(i, j, k);
        wcorr[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny] = wc[c];
        before += wcn[c] * Vc;
        if (p_.use_corrector) ss_term -= wc[c] * Vc * s.Ss * (h[c] - hn[c]) / s.theta_s;
        if (k == 0) bnd -= dt_used * qz[c - g.plane()];
        if (k == nz - 1) bnd += dt_used * qz[c];
      }
  updateDiagnostics();
  reallocateWaterContent();
  clampWaterContent();
  enforceMoistureBc();
  real after = 0.0, adjust = 0.0, dhmax = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long c = g.idx(i, j, k);
        after += wc[c] * Vc;
        adjust += (wc[c] - wcorr[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny]) * Vc;
        dhmax = std::max(dhmax, std::fabs(h[c] - hn[c]));
      }
  if (p_.adaptive_dt) adaptiveTimeStep();
  if (st) {
    st->solve = ss;
    st->picard_iterations = picard;
    st->picard_converged = converged;
    st->dt_used = dt_used;
    st->dt_next = p_.dt;
    st->max_dh = dhmax;
    st->mass_before = before;
    st->mass_after = after;
    st->mass_boundary_in = bnd;
    st->mass_storage_ss = ss_term;
    st->mass_adjust = adjust;
  }
}

}  // namespace frehg_oracle
