// Per-cell device code of the Richards PCA step, shared by the single-box kernels (re.hip) and
// the batched ensemble kernels (ensemble.hip). Sequence and arithmetic follow
// ReSolver::advanceStep (src/re/ReSolver.cpp:931-969); each function names the reference code it
// restates. Every function works on one interior cell `t` of the box described by `P`, with field
// pointers at the box's own halo-padded storage.
#pragma once

#include "../../include/frehg_gpu.h"
#include "common.hpp"

namespace frehg {
namespace redev {

constexpr int kReNT = 256;  // threads per tile: 64 (i) x 4 (j), one layer

struct ReK {
  Geo g;
  SoilView soil;
  double dx, dy, dz, dt, htop, hbot, qtop, qbot, co_max;
  int full3d, sim_sw, bc4, bc5, use_corrector;
  const double* qtop_field;  // surface padded or null
};

struct Cell {
  int i, j, k;
  long long c;
  bool valid;
};

// Cell of this thread in tile (bx, by, bz) of 64 x 4 columns, one layer (bz = k).
__device__ inline Cell cell_at(const Geo& g, unsigned int bx, unsigned int by, unsigned int bz) {
  Cell t;
  t.i = bx * 64 + (threadIdx.x & 63);
  t.j = by * 4 + (threadIdx.x >> 6);
  t.k = bz;
  t.valid = t.i < g.nx && t.j < g.ny;
  t.c = t.valid ? g.idx(t.i, t.j, t.k) : 0;
  return t;
}

// Cell of this thread in a single-box launch over (ceil(nx/64), ceil(ny/4), nz) tiles.
__device__ inline Cell cell_of(const Geo& g) {
  const TileId b = xcd_tile();
  return cell_at(g, b.x, b.y, b.z);
}

// Conductivity on the +x face of cell (i,j,k) at flat index c (i may be -1: the x- wall).
// ReSolver::computeKFace, ReSolver.cpp:243-250 (main loop), :287-299 (vertical-only x- quirk),
// :329-330 (3D walls), :365-379 (vertical-only lateral zeroing in unsaturated cells).
__device__ inline double face_x(const ReK& P, const double* kr, const double* wref, long long c, int i) {
  const Geo& g = P.g;
  if (i < 0) return 0.0;
  if (P.full3d) {
    if (i >= g.nx - 1) return 0.0;
    return 0.5 * (P.soil.Ksx(c + 1) * kr[c + 1] + P.soil.Ksx(c) * kr[c]);
  }
  double K;
  if (i == 0) {
    const double Kp = P.soil.Ksx(c) * kr[c - 1];
    K = 0.5 * (Kp + Kp);
  } else {
    K = 0.5 * (P.soil.Ksx(c + 1) * kr[c + 1] + P.soil.Ksx(c) * kr[c]);
  }
  if (wref[c] < P.soil.at(c).theta_s) K = 0.0;
  if (i + 1 <= g.nx - 1 && wref[c + 1] < P.soil.at(c + 1).theta_s) K = 0.0;
  return K;
}

// +y face of cell (i,j,k): ReSolver.cpp:251-265, :301-310, :333-334, :365-379.
__device__ inline double face_y(const ReK& P, const double* kr, const double* wref, long long c, int j) {
  const Geo& g = P.g;
  if (j < 0) return 0.0;
  if (P.full3d) {
    if (j >= g.ny - 1) return 0.0;
    return 0.5 * (P.soil.Ksy(c + g.sy) * kr[c + g.sy] + P.soil.Ksy(c) * kr[c]);
  }
  double K;
  if (j == 0) {
    const double Kp = P.soil.Ksy(c) * kr[c - g.sy];
    K = 0.5 * (Kp + Kp);
  } else if (j == g.ny - 1) {
    K = 0.0;
  } else {
    K = 0.5 * (P.soil.Ksy(c + g.sy) * kr[c + g.sy] + P.soil.Ksy(c) * kr[c]);
  }
  if (wref[c] < P.soil.at(c).theta_s) K = 0.0;
  if (j + 1 <= g.ny - 1 && wref[c + g.sy] < P.soil.at(c + g.sy).theta_s) K = 0.0;
  return K;
}

// +z (downward) face of cell k (k = -1: the top boundary face). ReSolver.cpp:272-284 (main),
// :338-344 (bottom), :345-363 (top: Ks_z for Dirichlet even when htop < 0).
__device__ inline double face_z(const ReK& P, const double* kr, long long c, int k) {
  const Geo& g = P.g;
  if (k < 0) {
    if (P.sim_sw) return 0.0;
    const long long ct = c + g.sz;
    if (P.bc5 == 1) return P.soil.Ksz(ct);
    if (P.bc5 == 0) return 0.0;
    return P.soil.Ksz(ct) * kr[ct];
  }
  if (k < g.nz - 1) return 0.5 * (P.soil.Ksz(c + g.sz) * kr[c + g.sz] + P.soil.Ksz(c) * kr[c]);
  return P.bc4 == 0 ? 0.0 : P.soil.Ksz(c) * kr[c];
}

// K/Ks and C at the interior cell and kr at its face-adjacent ghost cells (whose soil is the
// cell's own: ghost indices are clamped). ReSolver::updateDiagnostics (ReSolver.cpp:199-215).
__device__ inline void constitutive_with_ghosts(const ReK& P, const Cell& t, const double* h,
                                                 double* kr, double* ch, double* wch) {
  const Geo& g = P.g;
  const SoilClass& s = P.soil.at(t.c);
  const VGState v = vg_state(s, h[t.c]);
  kr[t.c] = v.kr;
  ch[t.c] = v.cap;
  if (wch) wch[t.c] = v.theta;
  const long long gh[6] = {t.i == 0 ? t.c - 1 : -1,         t.i == g.nx - 1 ? t.c + 1 : -1,
                           t.j == 0 ? t.c - g.sy : -1,      t.j == g.ny - 1 ? t.c + g.sy : -1,
                           t.k == 0 ? t.c - g.sz : -1,      t.k == g.nz - 1 ? t.c + g.sz : -1};
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if (gh[q] >= 0) kr[gh[q]] = vg_state(s, h[gh[q]]).kr;
}

__device__ __forceinline__ void re_pre(const ReK& P, const Cell& t, const double* hn, double* kr, double* ch) {
  constitutive_with_ghosts(P, t, hn, kr, ch, nullptr);
}

// groundwaterMatCoeff + groundwaterRhs (ReSolver.cpp:397-549) as a symmetric stencil:
// D = Gct, ax/ay/az = Gxp/Gyp/Gzp where the +neighbour is interior (Gxm(c) == Gxp(c-1) etc.).
__device__ __forceinline__ void re_assemble(const ReK& P, const Cell& t, const double* kr, const double* ch, const double* wcn, const double* hn, double* D, double* ax, double* ay, double* az, double* rhs, float* D32, float* ax32, float* ay32, float* az32) {
  const Geo& g = P.g;
  const long long c = t.c;
  const double Kxp = face_x(P, kr, wcn, c, t.i), Kxm = face_x(P, kr, wcn, c - 1, t.i - 1);
  const double Kyp = face_y(P, kr, wcn, c, t.j), Kym = face_y(P, kr, wcn, c - g.sy, t.j - 1);
  const double Kzp = face_z(P, kr, c, t.k), Kzm = face_z(P, kr, c - g.sz, t.k - 1);
  const SoilClass& s = P.soil.at(c);
  const double dt = P.dt, dzk = P.dz;
  const double Ax = P.dy * dzk, Ay = P.dx * dzk, V = P.dx * P.dy * dzk;
  const double Gxp = -Kxp * dt * Ax / P.dx;
  const double Gxm = -Kxm * dt * Ax / P.dx;
  const double Gyp = -Kyp * dt * Ay / P.dy;
  const double Gym = -Kym * dt * Ay / P.dy;
  const double dzf_p = 0.5 * (dzk + dzk);
  double Gzp = -Kzp * dt / (dzk * dzf_p);
  Gzp *= V;
  const double dzf_m = (t.k == 0) ? 0.5 * dzk : 0.5 * (dzk + dzk);
  double Gzm = -Kzm * dt / (dzk * dzf_m);
  Gzm *= V;
  const double st = ch[c] + s.Ss * wcn[c] / s.theta_s;
  double Gct = st;
  Gct *= V;
  Gct -= (Gxp + Gxm + Gyp + Gym);
  if (!(t.k == g.nz - 1 && P.bc4 != 1)) Gct -= Gzp;
  if (t.k == 0) {
    if (!P.sim_sw && P.bc5 == 1) Gct -= Gzm;
  } else {
    Gct -= Gzm;
  }
  double r = st * hn[c] * V;
  r -= dt * V * Kzp / dzk;
  r += dt * V * Kzm / dzk;
  if (t.i == g.nx - 1) r -= Gxp * hn[c + 1];
  if (t.i == 0) r -= Gxm * hn[c - 1];
  if (t.j == g.ny - 1) r -= Gyp * hn[c + g.sy];
  if (t.j == 0) r -= Gym * hn[c - g.sy];
  if (t.k == g.nz - 1) {
    if (P.bc4 == 1) r += Gzp * hn[c + g.sz];
    if (P.bc4 == 2) r += dt * V * (P.qbot + Kzp) / dzk;
  }
  if (t.k == 0 && !P.sim_sw && P.bc5 == 1) r -= Gzm * P.htop;
  if (t.k == 0 && !P.sim_sw && P.bc5 == 2) {
    const double qcol = P.qtop_field ? P.qtop_field[g.sidx(t.i, t.j)] : P.qtop;
    r += -dt * V * (qcol + Kzm) / dzk;
  }
  const double axv = (t.i < g.nx - 1) ? Gxp : 0.0;
  const double ayv = (t.j < g.ny - 1) ? Gyp : 0.0;
  const double azv = (t.k < g.nz - 1) ? Gzp : 0.0;
  D[c] = Gct;
  ax[c] = axv;
  ay[c] = ayv;
  az[c] = azv;
  rhs[c] = r;
  if (D32) {
    D32[c] = (float)Gct;
    ax32[c] = (float)axv;
    ay32[c] = (float)ayv;
    az32[c] = (float)azv;
  }
}

// enforceHeadBc (ReSolver.cpp:638-669) + kr, C and theta at the solved head (the second
// updateDiagnostics' ch and wch, ReSolver.cpp:963, :825).
__device__ __forceinline__ void re_post(const ReK& P, const Cell& t, double* h, double* kr, double* ch, double* wch) {
  const Geo& g = P.g;
  const double hc = h[t.c];
  if (t.i == 0) h[t.c - 1] = hc;
  if (t.i == g.nx - 1) h[t.c + 1] = hc;
  if (t.j == 0) h[t.c - g.sy] = hc;
  if (t.j == g.ny - 1) h[t.c + g.sy] = hc;
  if (t.k == g.nz - 1) h[t.c + g.sz] = (P.bc4 == 1) ? P.hbot : hc;
  if (t.k == 0) h[t.c - g.sz] = (!P.sim_sw && P.bc5 == 1) ? P.htop : hc;
  // ghost values written above are re-read by this same thread only
  constitutive_with_ghosts(P, t, h, kr, ch, wch);
}

// Second computeKFace + groundwaterFlux with darcyFluxZ (ReSolver.cpp:947-948, 24-58, 671-743).
__device__ __forceinline__ void re_flux(const ReK& P, const Cell& t, const double* h, const double* kr, const double* wcn, double* Kx, double* Ky, double* Kz, double* qx, double* qy, double* qz) {
  const Geo& g = P.g;
  const long long c = t.c;
  const double Kxp = face_x(P, kr, wcn, c, t.i);
  const double Kyp = face_y(P, kr, wcn, c, t.j);
  const double Kzp = face_z(P, kr, c, t.k);
  Kx[c] = Kxp;
  Ky[c] = Kyp;
  Kz[c] = Kzp;
  if (t.i == 0) Kx[c - 1] = 0.0;
  if (t.i == g.nx - 1 && !P.full3d) Kx[c + 1] = 0.0;
  if (t.j == 0) Ky[c - g.sy] = 0.0;
  if (t.k == g.nz - 1) Kz[c + g.sz] = 0.0;
  const double dzk = P.dz;
  const double faceA = P.dx * P.dy;
  if (P.full3d) {
    qx[c] = (P.dy * dzk) * Kxp * (h[c + 1] - h[c]) / P.dx;
    qy[c] = (P.dx * dzk) * Kyp * (h[c + g.sy] - h[c]) / P.dy;
  } else {
    qx[c] = 0.0;
    qy[c] = 0.0;
  }
  double q;
  if (t.k == g.nz - 1 && P.bc4 == 0) {
    q = 0.0;
  } else {
    const double delta = (t.k == g.nz - 1) ? 0.5 * dzk : 0.5 * (dzk + dzk);
    const double kc = P.soil.Ksz(c) * kr[c];
    const double ks = P.soil.Ksz(c + g.sz) * kr[c + g.sz];
    const double kface = 0.5 * (kc + ks);
    q = faceA * kface * ((h[c + g.sz] - h[c]) / delta - 1.0);
  }
  if (t.k == g.nz - 1 && P.bc4 == 2) q = P.qbot * faceA;
  qz[c] = q;
  if (t.k == 0) {
    Kz[c - g.sz] = face_z(P, kr, c - g.sz, -1);
    double qt;
    if (P.bc5 == 2) {
      const double qcol = P.qtop_field ? P.qtop_field[g.sidx(t.i, t.j)] : P.qtop;
      qt = qcol * faceA;
    } else {
      const double kface = (P.bc5 == 0) ? 0.0 : P.soil.Ksz(c);
      qt = faceA * kface * ((h[c] - h[c - g.sz]) / (0.5 * dzk) - 1.0);
    }
    qz[c - g.sz] = qt;
  }
}

// Stores the corrected theta of cell t, h(theta) and the saturation flags; ghost neighbours keep
// the moisture ghosts of the previous step (ReSolver.cpp:855-878).
__device__ inline void correct_cell(const ReK& P, const Cell& t, double w, double th, const double* wcn, double* wc,
                                    double* hwc, unsigned char* sat0, unsigned char* cand) {
  const Geo& g = P.g;
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  wc[c] = w;
  hwc[c] = vg_head(s, w);
  const unsigned char s0 = (w >= s.theta_s) ? 1 : 0;
  sat0[c] = s0;
  cand[c] = (!s0 && th >= s.theta_s) ? 1 : 0;
  const long long gh[6] = {t.i == 0 ? c - 1 : -1,         t.i == g.nx - 1 ? c + 1 : -1,
                           t.j == 0 ? c - g.sy : -1,      t.j == g.ny - 1 ? c + g.sy : -1,
                           t.k == 0 ? c - g.sz : -1,      t.k == g.nz - 1 ? c + g.sz : -1};
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if (gh[q] >= 0) {
      sat0[gh[q]] = (wcn[gh[q]] >= s.theta_s) ? 1 : 0;
      cand[gh[q]] = 0;
    }
}

// updateWaterContent (ReSolver.cpp:762-786) or theta(h) without the corrector (:953-961) for one
// cell, plus its ledger terms: theta^n V, the specific-storage term and the top/bottom inflow.
__device__ __forceinline__ void re_correct(const ReK& P, const Cell& t, const double* h, const double* hn,
                                           const double* wcn, const double* qx, const double* qy, const double* qz,
                                           const double* wch, double* wc, double* hwc, unsigned char* sat0,
                                           unsigned char* cand, double& l_before, double& l_ss, double& l_bnd) {
  const Geo& g = P.g;
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  const double V = P.dx * P.dy * P.dz;
  double w;
  const double th = wch[c];
  if (P.use_corrector) {
    const double coeff = V * (1.0 + s.Ss * (h[c] - hn[c]) / s.theta_s);
    const double dqx = P.dt * (qx[c] - qx[c - 1]);
    const double dqy = P.dt * (qy[c] - qy[c - g.sy]);
    const double dqz = P.dt * (qz[c] - qz[c - g.sz]);
    w = (wcn[c] * V + dqx + dqy + dqz) / coeff;
    l_ss = -w * V * s.Ss * (h[c] - hn[c]) / s.theta_s;
  } else {
    w = th;
  }
  l_before = wcn[c] * V;
  if (t.k == 0) l_bnd -= P.dt * qz[c - g.sz];
  if (t.k == g.nz - 1) l_bnd += P.dt * qz[c];
  correct_cell(P, t, w, th, wcn, wc, hwc, sat0, cand);
}

// adjacentToSat (ReSolver.cpp:802-817) with `prev` holding the post-visit saturation state of
// the cells the i->j->k sweep has already visited and `sat0` the pre-visit state of the rest.
__device__ inline bool adjacent_sat(const ReK& P, const Cell& t, const double* Kx, const double* Ky,
                                    const unsigned char* sat0, const unsigned char* prev) {
  const Geo& g = P.g;
  const long long c = t.c;
  if (sat0[c + g.sz]) return true;
  if (Kx[c] != 0.0 && sat0[c + 1]) return true;
  if (Kx[c - 1] != 0.0 && prev[c - 1]) return true;
  if (Ky[c] != 0.0 && sat0[c + g.sy]) return true;
  if (Ky[c - g.sy] != 0.0 && prev[c - g.sy]) return true;
  if (t.k == 0) return (!P.sim_sw && P.bc5 == 1 && P.htop >= 0.0);
  return prev[c - g.sz] != 0;
}

// One Jacobi sweep of the monotone map whose unique fixed point is the sequential sweep's
// saturation state: a candidate (unsaturated theta, saturated theta(h)) becomes saturated iff
// it is adjacent to saturation, reading visited neighbours from `prev`.
__device__ __forceinline__ void re_realloc(const ReK& P, const Cell& t, const double* Kx, const double* Ky, const unsigned char* sat0, const unsigned char* cand, const unsigned char* prev, unsigned char* next, int* changed) {
  unsigned char v = sat0[t.c];
  if (cand[t.c]) v = adjacent_sat(P, t, Kx, Ky, sat0, prev) ? 1 : 0;
  next[t.c] = v;
  if (v != prev[t.c]) atomicOr(changed, 1);
}

// reallocateWaterContent branches with the converged saturation state + clampWaterContent +
// enforceMoistureBc (ReSolver.cpp:819-878) for one cell, plus its ledger terms.
__device__ __forceinline__ void re_finalize(const ReK& P, const Cell& t, const double* Kx, const double* Ky,
                                            const unsigned char* sat0, const unsigned char* fin, const double* wch,
                                            const double* hwc, const double* hn, double* wc, double* h,
                                            double& l_after, double& l_adj, double& dh) {
  const Geo& g = P.g;
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  const double wcorr = wc[c];
  double w = wcorr;
  if (sat0[c]) {
    w = s.theta_s;
  } else if (!adjacent_sat(P, t, Kx, Ky, sat0, fin)) {
    if (w < 0.9999 * s.theta_s) h[c] = hwc[c];
  } else {
    w = wch[c];
  }
  if (w > s.theta_s) w = s.theta_s;
  else if (w < s.theta_r) w = s.theta_r;
  wc[c] = w;
  if (t.i == 0) wc[c - 1] = w;
  if (t.i == g.nx - 1) wc[c + 1] = w;
  if (t.j == 0) wc[c - g.sy] = w;
  if (t.j == g.ny - 1) wc[c + g.sy] = w;
  if (t.k == g.nz - 1) wc[c + g.sz] = w;
  if (t.k == 0) wc[c - g.sz] = w;
  const double V = P.dx * P.dy * P.dz;
  l_after = w * V;
  l_adj = (w - wcorr) * V;
  dh = fabs(h[c] - hn[c]);
}

// setInitialHead / setInitialWaterContent (ReSolver.cpp:135-173) + finalizeInitialState
// diagnostics (:175-178) on interior cells.
__device__ __forceinline__ void re_initial(const ReK& P, const Cell& t, int from_head, double* h, double* wc, double* hn, double* wcn, double* wch, double* hwc, double* ch) {
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  if (from_head) {
    wc[c] = vg_theta_ref(s, h[c]);
  } else {
    h[c] = vg_head(s, wc[c]);
  }
  hn[c] = h[c];
  wcn[c] = wc[c];
  const VGState v = vg_state(s, h[c]);
  hwc[c] = vg_head(s, wc[c]);
  wch[c] = v.theta;
  ch[c] = v.cap;
}

inline ReK make_rek(const Geo& g, const SoilView& sv, const frehg_re_params& p, const double* qtop) {
  ReK P;
  P.g = g;
  P.soil = sv;
  P.dx = p.dx;
  P.dy = p.dy;
  P.dz = p.dz;
  P.dt = p.dt;
  P.htop = p.htop;
  P.hbot = p.hbot;
  P.qtop = p.qtop;
  P.qbot = p.qbot;
  P.co_max = p.co_max;
  P.full3d = p.use_full3d ? 1 : 0;
  P.sim_sw = p.sim_shallowwater ? 1 : 0;
  P.bc4 = p.bc_type[4];
  P.bc5 = p.bc_type[5];
  P.use_corrector = p.use_corrector ? 1 : 0;
  P.qtop_field = qtop;
  return P;
}

}  // namespace redev
}  // namespace frehg
