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

// Parameters of one box's step. The template arguments select the soil lookup variant (see
// SoilViewT) and whether lateral flow is on (F3D: 0 off, 1 on, 2 read from full3d at run time);
// ReK, the runtime-checked variant, is what host code builds and stores.
template <int CLS, int KS, int F3D = 2>
struct ReKT {
  Geo g;
  SoilViewT<CLS, KS> soil;
  double dx, dy, dz, dt, htop, hbot, qtop, qbot, co_max;
  int full3d, sim_sw, bc4, bc5, use_corrector;
  const double* qtop_field;  // surface padded or null
  // Products of the grid spacing and dt that the per-cell code multiplies by instead of dividing
  // by a spacing at every cell (set by make_rek): the coupling factors of the matrix (dt A / d per
  // face, with the interior and half-cell top distances in z), dt V / dz of the gravity terms, and
  // the area-over-distance factors of the Darcy fluxes.
  double gx, gy, gzi, gzt, gdz, fx, fy, izi, izh;
  __device__ inline bool full() const {
    if constexpr (F3D == 2) return full3d != 0;
    else return F3D == 1;
  }
};
using ReK = ReKT<kRuntime, kRuntime, 2>;

// The same parameters with the soil lookup fixed to the given variant.
template <int CLS, int KS, int F3D>
__host__ __device__ inline ReKT<CLS, KS, F3D> rek_as(const ReK& P) {
  ReKT<CLS, KS, F3D> Q;
  Q.g = P.g;
  Q.soil.classes = P.soil.classes;
  Q.soil.cls = P.soil.cls;
  Q.soil.ks = P.soil.ks;
  Q.soil.c0 = P.soil.c0;
  Q.dx = P.dx;
  Q.dy = P.dy;
  Q.dz = P.dz;
  Q.dt = P.dt;
  Q.htop = P.htop;
  Q.hbot = P.hbot;
  Q.qtop = P.qtop;
  Q.qbot = P.qbot;
  Q.co_max = P.co_max;
  Q.full3d = P.full3d;
  Q.sim_sw = P.sim_sw;
  Q.bc4 = P.bc4;
  Q.bc5 = P.bc5;
  Q.use_corrector = P.use_corrector;
  Q.qtop_field = P.qtop_field;
  Q.gx = P.gx;
  Q.gy = P.gy;
  Q.gzi = P.gzi;
  Q.gzt = P.gzt;
  Q.gdz = P.gdz;
  Q.fx = P.fx;
  Q.fy = P.fy;
  Q.izi = P.izi;
  Q.izh = P.izh;
  return Q;
}

// rek_as with the variant given as the target type.
template <class K>
struct RekTraits;
template <int CLS, int KS, int F3D>
struct RekTraits<ReKT<CLS, KS, F3D>> {
  static constexpr int cls = CLS, ks = KS, f3d = F3D;
};
template <class K>
__host__ __device__ inline K rek_to(const ReK& P) {
  return rek_as<RekTraits<K>::cls, RekTraits<K>::ks, RekTraits<K>::f3d>(P);
}

// Variant code of P for with_soil_variant_code: 1..6 as in with_soil_variant, 0 = run time.
inline int soil_variant_code(const ReK& P) {
  const bool c = P.soil.cls != nullptr, k = P.soil.ks != nullptr, f3 = P.full3d != 0;
  if (!c) return k ? (f3 ? 1 : 2) : (f3 ? 3 : 4);
  return k ? 5 : 6;
}

// Calls f(Q) with Q of the variant `code` (0: the runtime-checked ReK) built from P.
template <typename F>
inline void with_soil_variant_code(int code, F&& f) {
  switch (code) {
    case 1: f(ReKT<kAbsent, kPresent, 1>{}); break;
    case 2: f(ReKT<kAbsent, kPresent, 0>{}); break;
    case 3: f(ReKT<kAbsent, kAbsent, 1>{}); break;
    case 4: f(ReKT<kAbsent, kAbsent, 0>{}); break;
    case 5: f(ReKT<kPresent, kPresent, 2>{}); break;
    case 6: f(ReKT<kPresent, kAbsent, 2>{}); break;
    default: f(ReK{});
  }
}

// Calls f(Q) with Q = P fixed to the variant matching which soil arrays P has (and, for one soil
// class, whether lateral flow is on).
template <typename F>
inline void with_soil_variant(const ReK& P, F&& f) {
  const bool c = P.soil.cls != nullptr, k = P.soil.ks != nullptr, f3 = P.full3d != 0;
  if (!c && k) {
    if (f3) f(rek_as<kAbsent, kPresent, 1>(P));
    else f(rek_as<kAbsent, kPresent, 0>(P));
  } else if (!c) {
    if (f3) f(rek_as<kAbsent, kAbsent, 1>(P));
    else f(rek_as<kAbsent, kAbsent, 0>(P));
  } else if (k) {
    f(rek_as<kPresent, kPresent, 2>(P));
  } else {
    f(rek_as<kPresent, kAbsent, 2>(P));
  }
}

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

// Face conductivities. Every operand is loaded whichever branch applies (the stencil points lie
// in the halo-padded rows), so that the loads of all faces of a cell issue together instead of
// one dependent round trip per branch; the arithmetic is that of the reference branch.
//
// Conductivity on the +x face of cell (i,j,k) at flat index c (i may be -1: the x- wall).
// ReSolver::computeKFace, ReSolver.cpp:243-250 (main loop), :287-299 (vertical-only x- quirk),
// :329-330 (3D walls), :365-379 (vertical-only lateral zeroing in unsaturated cells).
template <class RK>
__device__ inline double face_x(const RK& P, const double* __restrict__ kr, const double* __restrict__ wref,
                                long long c, int i) {
#pragma clang fp contract(off)
  const Geo& g = P.g;
  const double ks0 = P.soil.Ksx(c), ks1 = P.soil.Ksx(c + 1);
  const double kr0 = kr[c], kr1 = kr[c + 1];
  const double Kmid = 0.5 * (ks1 * kr1 + ks0 * kr0);
  if (P.full()) return (i < 0 || i >= g.nx - 1) ? 0.0 : Kmid;
  const double krm = kr[i == 0 ? c - 1 : c];
  const double w0 = wref[c], w1 = wref[c + 1];
  const double ts0 = P.soil.at(c).theta_s, ts1 = P.soil.at(c + 1).theta_s;
  const double Kp = ks0 * krm;
  double K = (i == 0) ? 0.5 * (Kp + Kp) : Kmid;
  if (w0 < ts0) K = 0.0;
  if (i + 1 <= g.nx - 1 && w1 < ts1) K = 0.0;
  return i < 0 ? 0.0 : K;
}

// +y face of cell (i,j,k): ReSolver.cpp:251-265, :301-310, :333-334, :365-379.
template <class RK>
__device__ inline double face_y(const RK& P, const double* __restrict__ kr, const double* __restrict__ wref,
                                long long c, int j) {
#pragma clang fp contract(off)
  const Geo& g = P.g;
  const double ks0 = P.soil.Ksy(c), ks1 = P.soil.Ksy(c + g.sy);
  const double kr0 = kr[c], kr1 = kr[c + g.sy];
  const double Kmid = 0.5 * (ks1 * kr1 + ks0 * kr0);
  if (P.full()) return (j < 0 || j >= g.ny - 1) ? 0.0 : Kmid;
  const double krm = kr[j == 0 ? c - g.sy : c];
  const double w0 = wref[c], w1 = wref[c + g.sy];
  const double ts0 = P.soil.at(c).theta_s, ts1 = P.soil.at(c + g.sy).theta_s;
  const double Kp = ks0 * krm;
  double K = (j == 0) ? 0.5 * (Kp + Kp) : (j == g.ny - 1 ? 0.0 : Kmid);
  if (w0 < ts0) K = 0.0;
  if (j + 1 <= g.ny - 1 && w1 < ts1) K = 0.0;
  return j < 0 ? 0.0 : K;
}

// +z (downward) face of cell k (k = -1: the top boundary face). ReSolver.cpp:272-284 (main),
// :338-344 (bottom), :345-363 (top: Ks_z for Dirichlet even when htop < 0).
template <class RK>
__device__ inline double face_z(const RK& P, const double* __restrict__ kr, long long c, int k) {
#pragma clang fp contract(off)
  const Geo& g = P.g;
  const double ks0 = P.soil.Ksz(c), ks1 = P.soil.Ksz(c + g.sz);
  const double kr0 = kr[c], kr1 = kr[c + g.sz];
  if (k < 0) return P.sim_sw ? 0.0 : (P.bc5 == 1 ? ks1 : (P.bc5 == 0 ? 0.0 : ks1 * kr1));
  if (k < g.nz - 1) return 0.5 * (ks1 * kr1 + ks0 * kr0);
  return P.bc4 == 0 ? 0.0 : ks0 * kr0;
}

// K/Ks and C at the interior cell and kr at its face-adjacent ghost cells (whose soil is the
// cell's own: ghost indices are clamped). ReSolver::updateDiagnostics (ReSolver.cpp:199-215).
// A ghost whose head equals the cell's (every no-flux face) takes the cell's kr, the same
// function of the same operands; only a ghost with a head of its own (Dirichlet faces, or a
// boundary cell whose head the reallocation replaced) is evaluated, so that a wave holding one
// boundary cell no longer runs the constitutive law again for the whole wave.
template <class RK>
__device__ inline void constitutive_with_ghosts(const RK& P, const Cell& t, const double* h,
                                                 double* kr, double* ch, double* wch) {
  const Geo& g = P.g;
  const SoilClass& s = P.soil.at(t.c);
  const double hc = h[t.c];
  const VGState v = vg_state(s, hc);
  kr[t.c] = v.kr;
  ch[t.c] = v.cap;
  if (wch) wch[t.c] = v.theta;
  const long long gh[6] = {t.i == 0 ? t.c - 1 : -1,         t.i == g.nx - 1 ? t.c + 1 : -1,
                           t.j == 0 ? t.c - g.sy : -1,      t.j == g.ny - 1 ? t.c + g.sy : -1,
                           t.k == 0 ? t.c - g.sz : -1,      t.k == g.nz - 1 ? t.c + g.sz : -1};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    if (gh[q] < 0) continue;
    const double hg = h[gh[q]];
    kr[gh[q]] = (hg == hc) ? v.kr : vg_state(s, hg).kr;
  }
}

template <class RK>
__device__ __forceinline__ void re_pre(const RK& P, const Cell& t, const double* hn, double* kr, double* ch) {
  constitutive_with_ghosts(P, t, hn, kr, ch, nullptr);
}

// re_pre when kr and C already hold their values at the head of every cell except those flagged
// in hmod (the post-solve evaluation of the previous step, ReSolver.cpp:963, at the same heads):
// only flagged interior cells are re-evaluated. Ghost heads, and so ghost kr, are unchanged
// since that evaluation.
template <class RK>
__device__ __forceinline__ void re_pre_stale(const RK& P, const Cell& t, const unsigned char* hmod, const double* hn,
                                             double* kr, double* ch) {
  if (!hmod[t.c]) return;
  const VGState v = vg_state(P.soil.at(t.c), hn[t.c]);
  kr[t.c] = v.kr;
  ch[t.c] = v.cap;
}

// One row of the assembled system: D = Gct, the couplings to the six neighbours as stored (zero
// where the neighbour is a ghost), rhs = Grhs.
struct ReRow {
  double D, xp, xm, yp, ym, zp, zm, rhs;
};

// groundwaterMatCoeff + groundwaterRhs (ReSolver.cpp:397-549) as a symmetric stencil:
// D = Gct, ax/ay/az = Gxp/Gyp/Gzp where the +neighbour is interior (Gxm(c) == Gxp(c-1) etc.).
template <class RK>
__device__ __forceinline__ ReRow re_assemble_row(const RK& P, const Cell& t, const double* kr, const double* ch,
                                                 const double* wcn, const double* hn) {
  const Geo& g = P.g;
  const long long c = t.c;
  const double Kxp = face_x(P, kr, wcn, c, t.i), Kxm = face_x(P, kr, wcn, c - 1, t.i - 1);
  const double Kyp = face_y(P, kr, wcn, c, t.j), Kym = face_y(P, kr, wcn, c - g.sy, t.j - 1);
  const double Kzp = face_z(P, kr, c, t.k), Kzm = face_z(P, kr, c - g.sz, t.k - 1);
  const SoilClass& s = P.soil.at(c);
  const double V = P.dx * P.dy * P.dz;
  // the reference's -K dt A / d per face, with dt A / d formed once per launch (make_rek)
  const double Gxp = -Kxp * P.gx;
  const double Gxm = -Kxm * P.gx;
  const double Gyp = -Kyp * P.gy;
  const double Gym = -Kym * P.gy;
  const double Gzp = -Kzp * P.gzi;
  const double Gzm = -Kzm * (t.k == 0 ? P.gzt : P.gzi);  // half-cell distance to the top face
  const double st = ch[c] + wcn[c] * s.ss_ts;
  double Gct = st;
  Gct *= V;
  Gct -= (Gxp + Gxm + Gyp + Gym);
  if (!(t.k == g.nz - 1 && P.bc4 != 1)) Gct -= Gzp;
  if (t.k == 0) {
    if (!P.sim_sw && P.bc5 == 1) Gct -= Gzm;
  } else {
    Gct -= Gzm;
  }
  const double hxp = hn[c + 1], hxm = hn[c - 1], hyp = hn[c + g.sy], hym = hn[c - g.sy], hzp = hn[c + g.sz];
  double r = st * hn[c] * V;
  r -= Kzp * P.gdz;
  r += Kzm * P.gdz;
  if (t.i == g.nx - 1) r -= Gxp * hxp;
  if (t.i == 0) r -= Gxm * hxm;
  if (t.j == g.ny - 1) r -= Gyp * hyp;
  if (t.j == 0) r -= Gym * hym;
  if (t.k == g.nz - 1) {
    if (P.bc4 == 1) r += Gzp * hzp;
    if (P.bc4 == 2) r += (P.qbot + Kzp) * P.gdz;
  }
  if (t.k == 0 && !P.sim_sw && P.bc5 == 1) r -= Gzm * P.htop;
  if (t.k == 0 && !P.sim_sw && P.bc5 == 2) {
    const double qcol = P.qtop_field ? P.qtop_field[g.sidx(t.i, t.j)] : P.qtop;
    r -= (qcol + Kzm) * P.gdz;
  }
  ReRow row;
  row.D = Gct;
  row.xp = (t.i < g.nx - 1) ? Gxp : 0.0;
  row.xm = (t.i > 0) ? Gxm : 0.0;
  row.yp = (t.j < g.ny - 1) ? Gyp : 0.0;
  row.ym = (t.j > 0) ? Gym : 0.0;
  row.zp = (t.k < g.nz - 1) ? Gzp : 0.0;
  row.zm = (t.k > 0) ? Gzm : 0.0;
  row.rhs = r;
  return row;
}

// Stores row c: D, ax/ay/az (the + couplings), rhs, and the FP32 copies when D32 is set.
__device__ __forceinline__ void re_store_row(long long c, const ReRow& row, double* D, double* ax, double* ay,
                                             double* az, double* rhs, float* D32, float* ax32, float* ay32,
                                             float* az32) {
  D[c] = row.D;
  ax[c] = row.xp;
  ay[c] = row.yp;
  az[c] = row.zp;
  rhs[c] = row.rhs;
  if (D32) {
    D32[c] = (float)row.D;
    ax32[c] = (float)row.xp;
    ay32[c] = (float)row.yp;
    az32[c] = (float)row.zp;
  }
}

template <class RK>
__device__ __forceinline__ void re_assemble(const RK& P, const Cell& t, const double* kr, const double* ch,
                                            const double* wcn, const double* hn, double* D, double* ax, double* ay,
                                            double* az, double* rhs, float* D32, float* ax32, float* ay32,
                                            float* az32) {
  re_store_row(t.c, re_assemble_row(P, t, kr, ch, wcn, hn), D, ax, ay, az, rhs, D32, ax32, ay32, az32);
}

// enforceHeadBc (ReSolver.cpp:638-669) + kr, C and theta at the solved head (the second
// updateDiagnostics' ch and wch, ReSolver.cpp:963, :825).
template <class RK>
__device__ __forceinline__ void re_post(const RK& P, const Cell& t, double* h, double* kr, double* ch, double* wch) {
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

// Darcy fluxes through the +x / +y / +z face of interior cell c (groundwaterFlux with
// darcyFluxZ, ReSolver.cpp:24-58, 671-743), given the face conductivity. Evaluated without
// contraction so that a face's flux has the same bits whichever of its two cells computes it.
template <class RK>
__device__ inline double flux_x(const RK& P, const double* h, double Kxp, long long c) {
#pragma clang fp contract(off)
  return P.fx * Kxp * (h[c + 1] - h[c]);
}
template <class RK>
__device__ inline double flux_y(const RK& P, const double* h, double Kyp, long long c) {
#pragma clang fp contract(off)
  return P.fy * Kyp * (h[c + P.g.sy] - h[c]);
}
template <class RK>
__device__ inline double flux_z(const RK& P, const double* h, const double* kr, long long c, int k) {
#pragma clang fp contract(off)
  const Geo& g = P.g;
  const double faceA = P.dx * P.dy;
  if (k == g.nz - 1 && P.bc4 == 2) return P.qbot * faceA;
  if (k == g.nz - 1 && P.bc4 == 0) return 0.0;
  const double idelta = (k == g.nz - 1) ? P.izh : P.izi;  // 1 / distance between the heads
  const double kc = P.soil.Ksz(c) * kr[c];
  const double ks = P.soil.Ksz(c + g.sz) * kr[c + g.sz];
  const double kface = 0.5 * (kc + ks);
  return faceA * kface * ((h[c + g.sz] - h[c]) * idelta - 1.0);
}
// Top-boundary flux of the column of top cell c (darcyFluxZ's k = -1 face, ReSolver.cpp:24-58).
template <class RK>
__device__ inline double flux_top(const RK& P, const Cell& t, const double* h) {
#pragma clang fp contract(off)
  const Geo& g = P.g;
  const long long c = t.c;
  const double faceA = P.dx * P.dy;
  if (P.bc5 == 2) {
    const double qcol = P.qtop_field ? P.qtop_field[g.sidx(t.i, t.j)] : P.qtop;
    return qcol * faceA;
  }
  const double kface = (P.bc5 == 0) ? 0.0 : P.soil.Ksz(c);
  return faceA * kface * ((h[c] - h[c - g.sz]) * P.izh - 1.0);
}

// Faces and fluxes of cell t at h^{n+1} (second computeKFace + groundwaterFlux,
// ReSolver.cpp:947-948): the cell's + faces and, for a top cell, the top face of the column.
struct CellFlux {
  double Kxp, Kyp, Kzp, Kzt, qxp, qyp, qzp, qt;
};
template <class RK>
__device__ __forceinline__ CellFlux flux_values(const RK& P, const Cell& t, const double* __restrict__ h,
                                                const double* __restrict__ kr, const double* __restrict__ wcn) {
  const Geo& g = P.g;
  const long long c = t.c;
  CellFlux f;
  f.Kxp = face_x(P, kr, wcn, c, t.i);
  f.Kyp = face_y(P, kr, wcn, c, t.j);
  f.Kzp = face_z(P, kr, c, t.k);
  f.qxp = P.full() ? flux_x(P, h, f.Kxp, c) : 0.0;
  f.qyp = P.full() ? flux_y(P, h, f.Kyp, c) : 0.0;
  f.qzp = flux_z(P, h, kr, c, t.k);
  f.Kzt = 0.0;
  f.qt = 0.0;
  if (t.k == 0) {
    f.Kzt = face_z(P, kr, c - g.sz, -1);
    f.qt = flux_top(P, t, h);
  }
  return f;
}
// Stores Kx/Ky/Kz and qx/qy/qz of the cell's + faces, the top face of a top cell, and the zero
// ghost-face conductivities of boundary cells.
template <class RK>
__device__ __forceinline__ void store_flux(const RK& P, const Cell& t, const CellFlux& f, double* __restrict__ Kx,
                                           double* __restrict__ Ky, double* __restrict__ Kz, double* __restrict__ qx,
                                           double* __restrict__ qy, double* __restrict__ qz) {
  const Geo& g = P.g;
  const long long c = t.c;
  Kx[c] = f.Kxp;
  Ky[c] = f.Kyp;
  Kz[c] = f.Kzp;
  if (t.i == 0) Kx[c - 1] = 0.0;
  if (t.i == g.nx - 1 && !P.full()) Kx[c + 1] = 0.0;
  if (t.j == 0) Ky[c - g.sy] = 0.0;
  if (t.k == g.nz - 1) Kz[c + g.sz] = 0.0;
  qx[c] = f.qxp;
  qy[c] = f.qyp;
  qz[c] = f.qzp;
  if (t.k == 0) {
    Kz[c - g.sz] = f.Kzt;
    qz[c - g.sz] = f.qt;
  }
}
template <class RK>
__device__ __forceinline__ CellFlux re_flux(const RK& P, const Cell& t, const double* h, const double* kr,
                                            const double* wcn, double* Kx, double* Ky, double* Kz, double* qx,
                                            double* qy, double* qz) {
  const CellFlux f = flux_values(P, t, h, kr, wcn);
  store_flux(P, t, f, Kx, Ky, Kz, qx, qy, qz);
  return f;
}

// Stores the corrected theta of cell t, h(theta) and the saturation flags; ghost neighbours keep
// the moisture ghosts of the previous step (ReSolver.cpp:855-878).
template <class RK>
__device__ inline void correct_cell(const RK& P, const Cell& t, double w, double th, const double* wcn, double* wc,
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
// cell from the fluxes through its + faces (qxp, qyp, qzp) and - faces (qxm, qym, qzm), plus its
// ledger terms: theta^n V, the specific-storage term and the top/bottom inflow.
template <class RK>
__device__ __forceinline__ void correct_fluxes(const RK& P, const Cell& t, const double* h, const double* hn,
                                               const double* wcn, double qxp, double qxm, double qyp, double qym,
                                               double qzp, double qzm, const double* wch, double* wc, double* hwc,
                                               unsigned char* sat0, unsigned char* cand, double& l_before,
                                               double& l_ss, double& l_bnd) {
  const Geo& g = P.g;
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  const double V = P.dx * P.dy * P.dz;
  double w;
  const double th = wch[c];
  if (P.use_corrector) {
    const double coeff = V * (1.0 + (h[c] - hn[c]) * s.ss_ts);
    const double dqx = P.dt * (qxp - qxm);
    const double dqy = P.dt * (qyp - qym);
    const double dqz = P.dt * (qzp - qzm);
    w = (wcn[c] * V + dqx + dqy + dqz) / coeff;
    l_ss = -w * V * (h[c] - hn[c]) * s.ss_ts;
  } else {
    w = th;
  }
  l_before = wcn[c] * V;
  if (t.k == 0) l_bnd -= P.dt * qzm;
  if (t.k == g.nz - 1) l_bnd += P.dt * qzp;
  correct_cell(P, t, w, th, wcn, wc, hwc, sat0, cand);
}

// The corrector with the face fluxes read from the stored qx / qy / qz.
template <class RK>
__device__ __forceinline__ void re_correct(const RK& P, const Cell& t, const double* h, const double* hn,
                                           const double* wcn, const double* qx, const double* qy, const double* qz,
                                           const double* wch, double* wc, double* hwc, unsigned char* sat0,
                                           unsigned char* cand, double& l_before, double& l_ss, double& l_bnd) {
  const Geo& g = P.g;
  const long long c = t.c;
  correct_fluxes(P, t, h, hn, wcn, qx[c], qx[c - 1], qy[c], qy[c - g.sy], qz[c], qz[c - g.sz], wch, wc, hwc, sat0,
                 cand, l_before, l_ss, l_bnd);
}

// Fluxes + corrector of one cell in one pass: the cell stores its + face fluxes (re_flux) and
// evaluates the fluxes through its - faces with the same functions of the same operands as the
// neighbour that stores them (identical bits), instead of re-reading them. Ghost-face fluxes
// (i = 0 / j = 0 x- and y- faces), which no cell writes, are read. All loads precede the stores.
// With kf != nullptr the face conductivities and fluxes are not stored (re_flux rebuilds them from
// the post-solve head, kr and theta^n; the ghost-face entries read here are never written by
// either); kf receives the nonzero-face bits that the reallocation needs (bit 0: Kx != 0, bit 1:
// Ky != 0), zero at the x- and y- ghost faces.
template <class RK>
__device__ __forceinline__ void re_flux_correct(const RK& P, const Cell& t, const double* __restrict__ h,
                                                const double* __restrict__ kr, const double* __restrict__ hn,
                                                const double* __restrict__ wcn, const double* __restrict__ wch,
                                                double* __restrict__ Kx, double* __restrict__ Ky,
                                                double* __restrict__ Kz, double* __restrict__ qx,
                                                double* __restrict__ qy, double* __restrict__ qz,
                                                double* __restrict__ wc, double* __restrict__ hwc,
                                                unsigned char* __restrict__ sat0, unsigned char* __restrict__ cand,
                                                unsigned char* __restrict__ kf, double& l_before, double& l_ss,
                                                double& l_bnd) {
  const Geo& g = P.g;
  const long long c = t.c;
  const CellFlux f = flux_values(P, t, h, kr, wcn);
  const double qx_ghost = qx[c - 1], qy_ghost = qy[c - g.sy];
  const double qxm = t.i > 0 ? (P.full() ? flux_x(P, h, face_x(P, kr, wcn, c - 1, t.i - 1), c - 1) : 0.0) : qx_ghost;
  const double qym =
      t.j > 0 ? (P.full() ? flux_y(P, h, face_y(P, kr, wcn, c - g.sy, t.j - 1), c - g.sy) : 0.0) : qy_ghost;
  const double qzm = (t.k > 0) ? flux_z(P, h, kr, c - g.sz, t.k - 1) : f.qt;
  correct_fluxes(P, t, h, hn, wcn, f.qxp, qxm, f.qyp, qym, f.qzp, qzm, wch, wc, hwc, sat0, cand, l_before, l_ss,
                 l_bnd);
  if (kf) {
    kf[c] = (unsigned char)((f.Kxp != 0.0 ? 1 : 0) | (f.Kyp != 0.0 ? 2 : 0));
    if (t.i == 0) kf[c - 1] = 0;
    if (t.j == 0) kf[c - g.sy] = 0;
  } else {
    store_flux(P, t, f, Kx, Ky, Kz, qx, qy, qz);
  }
}

// Nonzero-face tests of the reallocation, from the stored face conductivities or from the face
// bits of re_flux_correct.
struct FacesK {
  const double* __restrict__ Kx;
  const double* __restrict__ Ky;
  __device__ bool x(long long c) const { return Kx[c] != 0.0; }
  __device__ bool y(long long c) const { return Ky[c] != 0.0; }
};
struct FacesF {
  const unsigned char* __restrict__ f;
  __device__ bool x(long long c) const { return (f[c] & 1) != 0; }
  __device__ bool y(long long c) const { return (f[c] & 2) != 0; }
};

// adjacentToSat (ReSolver.cpp:802-817) with `prev` holding the post-visit saturation state of
// the cells the i->j->k sweep has already visited and `sat0` the pre-visit state of the rest.
// The reference's early-exit chain as one boolean expression over operands loaded together.
template <class RK, class FA>
__device__ inline bool adjacent_sat(const RK& P, const Cell& t, const FA& fa, const unsigned char* __restrict__ sat0,
                                    const unsigned char* __restrict__ prev) {
  const Geo& g = P.g;
  const long long c = t.c;
  const unsigned char below = sat0[c + g.sz], xp = sat0[c + 1], xm = prev[c - 1], yp = sat0[c + g.sy],
                      ym = prev[c - g.sy], above = prev[c - g.sz];
  const bool kx0 = fa.x(c), kxm = fa.x(c - 1), ky0 = fa.y(c), kym = fa.y(c - g.sy);
  const bool top = (t.k == 0) ? (!P.sim_sw && P.bc5 == 1 && P.htop >= 0.0) : (above != 0);
  return below || (kx0 && xp) || (kxm && xm) || (ky0 && yp) || (kym && ym) || top;
}

// One Jacobi sweep of the monotone map whose unique fixed point is the sequential sweep's
// saturation state: a candidate (unsaturated theta, saturated theta(h)) becomes saturated iff
// it is adjacent to saturation, reading visited neighbours from `prev`.
template <class RK, class FA>
__device__ __forceinline__ void re_realloc(const RK& P, const Cell& t, const FA& fa,
                                           const unsigned char* __restrict__ sat0,
                                           const unsigned char* __restrict__ cand,
                                           const unsigned char* __restrict__ prev, unsigned char* __restrict__ next,
                                           int* changed) {
  unsigned char v = sat0[t.c];
  if (cand[t.c]) v = adjacent_sat(P, t, fa, sat0, prev) ? 1 : 0;  // candidates are few
  next[t.c] = v;
  if (v != prev[t.c]) atomicOr(changed, 1);
}

// reallocateWaterContent branches with the converged saturation state + clampWaterContent +
// enforceMoistureBc (ReSolver.cpp:819-878) for one cell, plus its ledger terms. hmod (optional)
// receives 1 where the head was replaced by h(theta), else 0.
// hout (optional): the final head goes to hout (every interior cell, and the ghost heads of
// boundary cells copied from h) and h keeps the post-solve head; without it h is updated in place.
template <class RK, class FA>
__device__ __forceinline__ void re_finalize(const RK& P, const Cell& t, const FA& fa,
                                            const unsigned char* __restrict__ sat0,
                                            const unsigned char* __restrict__ fin, const double* __restrict__ wch,
                                            const double* __restrict__ hwc, const double* __restrict__ hn,
                                            double* __restrict__ wc, double* __restrict__ h,
                                            double* __restrict__ hout, unsigned char* __restrict__ hmod,
                                            double& l_after, double& l_adj, double& dh) {
  const Geo& g = P.g;
  const long long c = t.c;
  const SoilClass& s = P.soil.at(c);
  const double wcorr = wc[c], hw = hwc[c], hc = h[c], whc = wch[c], hnc = hn[c];
  const unsigned char s0 = sat0[c];
  const bool adj = adjacent_sat(P, t, fa, sat0, fin);
  double w = wcorr;
  bool replace = false;
  if (s0) {
    w = s.theta_s;
  } else if (!adj) {
    replace = w < 0.9999 * s.theta_s;
  } else {
    w = whc;
  }
  const double hnew = replace ? hw : hc;
  if (w > s.theta_s) w = s.theta_s;
  else if (w < s.theta_r) w = s.theta_r;
  if (hout) {
    hout[c] = hnew;
    const long long gh[6] = {t.i == 0 ? c - 1 : -1,         t.i == g.nx - 1 ? c + 1 : -1,
                             t.j == 0 ? c - g.sy : -1,      t.j == g.ny - 1 ? c + g.sy : -1,
                             t.k == 0 ? c - g.sz : -1,      t.k == g.nz - 1 ? c + g.sz : -1};
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (gh[q] >= 0) hout[gh[q]] = h[gh[q]];
  } else if (replace) {
    h[c] = hnew;
  }
  wc[c] = w;
  if (t.i == 0) wc[c - 1] = w;
  if (t.i == g.nx - 1) wc[c + 1] = w;
  if (t.j == 0) wc[c - g.sy] = w;
  if (t.j == g.ny - 1) wc[c + g.sy] = w;
  if (t.k == g.nz - 1) wc[c + g.sz] = w;
  if (t.k == 0) wc[c - g.sz] = w;
  if (hmod) hmod[c] = (replace && hnew != hc) ? 1 : 0;
  const double V = P.dx * P.dy * P.dz;
  l_after = w * V;
  l_adj = (w - wcorr) * V;
  dh = fabs(hnew - hnc);
}

// setInitialHead / setInitialWaterContent (ReSolver.cpp:135-173) + finalizeInitialState
// diagnostics (:175-178) on interior cells.
template <class RK>
__device__ __forceinline__ void re_initial(const RK& P, const Cell& t, int from_head, double* h, double* wc, double* hn, double* wcn, double* wch, double* hwc, double* ch) {
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
  const double V = p.dx * p.dy * p.dz;
  P.gx = p.dt * (p.dy * p.dz) / p.dx;
  P.gy = p.dt * (p.dx * p.dz) / p.dy;
  P.gzi = p.dt / (p.dz * p.dz) * V;
  P.gzt = p.dt / (p.dz * (0.5 * p.dz)) * V;
  P.gdz = p.dt * V / p.dz;
  P.fx = (p.dy * p.dz) / p.dx;
  P.fy = (p.dx * p.dz) / p.dy;
  P.izi = 1.0 / p.dz;
  P.izh = 1.0 / (0.5 * p.dz);
  return P;
}

}  // namespace redev
}  // namespace frehg
