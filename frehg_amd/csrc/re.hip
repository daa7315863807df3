// Richards-equation PCA step on the GPU. Sequence and arithmetic follow
// ReSolver::advanceStep (src/re/ReSolver.cpp:931-969); each kernel names the reference
// functions it fuses. Lateral boundaries are closed (the capability gate rejects anything
// else, src/driver/Orchestrator.cpp:155-163).
//
// Kernels per step (one thread per interior cell unless stated):
//   pre        kr = K/Ks and C at h^n, incl. the face-adjacent ghost cells          (:199-215)
//              (after a step: only the cells whose head changed since post, see hmod)
//   assemble   face conductivities -> 7-point coefficients, Gct and Grhs          (:217-549)
//   [Krylov]   Jacobi-PCG, x0 = 0 / h^n / learned                                   (:560-636)
//   post       ghost heads, kr and C at h^{n+1}                                     (:638-669)
//   flux+corr  faces at h^{n+1}, Darcy fluxes qx qy qz, mass-conservative theta,
//              wch/hwc, saturation flags                                 (:671-743, :762-786)
//              (by default Kx..qz are not stored: face bits for the reallocation instead, and
//              the fields are rebuilt from the post-solve state when read, see flux_pending_)
//   realloc    monotone fixed point of the sequential saturation sweep              (:788-837)
//   finalize   reallocation, clamp, moisture ghosts                                 (:788-878)
//   dtadapt    flux-imbalance max and Courant min (adaptive dt only)                (:880-929)
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <vector>

#include "mlp.hpp"
#include "re.hpp"
#include "re_device.hpp"

namespace frehg {

using redev::make_rek;

namespace {

using namespace redev;
constexpr int NT = kReNT;

// Cell of this thread; ZF selects the tile order (0: x fastest, 1: layer fastest). Layer-fastest
// order lets the kernels with vertical stencils (assemble, flux + corrector) find the layers above
// and below in L2; measured faster for all per-cell kernels except finalize.
template <int ZF>
__device__ inline Cell tile_cell(const Geo& g, TileId& b) {
  b = ZF ? zfast_tile() : xcd_tile();
  return cell_at(g, b.x, b.y, b.z);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_pre(K P, const double* __restrict__ hn, double* __restrict__ kr,
                                                double* __restrict__ ch) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (t.valid) re_pre(P, t, hn, kr, ch);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_pre_stale(K P, const unsigned char* __restrict__ hmod,
                                                      const double* __restrict__ hn, double* __restrict__ kr,
                                                      double* __restrict__ ch) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (t.valid) re_pre_stale(P, t, hmod, hn, kr, ch);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_assemble(K P, const double* __restrict__ kr, const double* __restrict__ ch,
                                                     const double* __restrict__ wcn, const double* __restrict__ hn,
                                                     double* __restrict__ D, double* __restrict__ ax,
                                                     double* __restrict__ ay, double* __restrict__ az,
                                                     double* __restrict__ rhs, float* __restrict__ D32,
                                                     float* __restrict__ ax32, float* __restrict__ ay32,
                                                     float* __restrict__ az32) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (t.valid) re_assemble(P, t, kr, ch, wcn, hn, D, ax, ay, az, rhs, D32, ax32, ay32, az32);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_post(K P, double* __restrict__ h, double* __restrict__ kr,
                                                 double* __restrict__ ch, double* __restrict__ wch) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (t.valid) re_post(P, t, h, kr, ch, wch);
}

// Fluxes + corrector (re_flux_correct); ledger partials (sum theta^n V, specific-storage term,
// boundary inflow) per block to part[0..2][nblocks].
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_flux_correct(K P, const double* __restrict__ h, const double* __restrict__ kr,
                                                         const double* __restrict__ hn, const double* __restrict__ wcn,
                                                         const double* __restrict__ wch, double* __restrict__ Kx,
                                                         double* __restrict__ Ky, double* __restrict__ Kz,
                                                         double* __restrict__ qx, double* __restrict__ qy,
                                                         double* __restrict__ qz, double* __restrict__ wc,
                                                         double* __restrict__ hwc, unsigned char* __restrict__ sat0,
                                                         unsigned char* __restrict__ cand,
                                                         unsigned char* __restrict__ kf, double* __restrict__ part,
                                                         unsigned int nblocks, int* __restrict__ flags, int nflags) {
  __shared__ double lds3[3 * (NT / kWave)];
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (b.linear == 0 && (int)threadIdx.x < nflags) flags[threadIdx.x] = 0;  // reallocation sweep flags
  double l_before = 0.0, l_ss = 0.0, l_bnd = 0.0;
  if (t.valid)
    re_flux_correct(P, t, h, kr, hn, wcn, wch, Kx, Ky, Kz, qx, qy, qz, wc, hwc, sat0, cand, kf, l_before, l_ss,
                    l_bnd);
  double v[3] = {l_before, l_ss, l_bnd};
  block_sums<NT, 3>(v, lds3);
  if (threadIdx.x == 0) {
    part[b.linear] = v[0];
    part[nblocks + b.linear] = v[1];
    part[2 * nblocks + b.linear] = v[2];
  }
}

// Kx..qz at the post-solve head (the stores the flux + corrector kernel leaves out by default).
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_flux(K P, const double* __restrict__ h, const double* __restrict__ kr,
                                                 const double* __restrict__ wcn, double* __restrict__ Kx,
                                                 double* __restrict__ Ky, double* __restrict__ Kz,
                                                 double* __restrict__ qx, double* __restrict__ qy,
                                                 double* __restrict__ qz) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (t.valid) (void)re_flux(P, t, h, kr, wcn, Kx, Ky, Kz, qx, qy, qz);
}

// One Jacobi sweep of the monotone map whose unique fixed point is the sequential sweep's
// saturation state (see re_realloc).
template <class K, class FA>
__global__ __launch_bounds__(NT) void k_re_realloc_iter(K P, FA fa, const unsigned char* __restrict__ sat0,
                                                         const unsigned char* __restrict__ cand,
                                                         const unsigned char* __restrict__ prev,
                                                         unsigned char* __restrict__ next, int* changed) {
  const Cell t = cell_of(P.g);
  if (t.valid) re_realloc(P, t, fa, sat0, cand, prev, next, changed);
}

// Reallocation + clamp + moisture BCs; ledger partials (sum theta^{n+1} V, sum (theta^{n+1} -
// theta_c) V) to part[3..4][nblocks] and max |h^{n+1} - h^n| to mx[nblocks].
template <class K, int ZF, class FA>
__global__ __launch_bounds__(NT) void k_re_finalize(K P, FA fa, const unsigned char* __restrict__ sat0,
                                                     const unsigned char* __restrict__ fin,
                                                     const double* __restrict__ wch, const double* __restrict__ hwc,
                                                     const double* __restrict__ hn, double* __restrict__ wc,
                                                     double* __restrict__ h, double* __restrict__ hout,
                                                     unsigned char* __restrict__ hmod,
                                                     double* __restrict__ part, double* __restrict__ mx,
                                                     unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  __shared__ double lmax[NT / kWave];
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  double l_after = 0.0, l_adj = 0.0, dh = 0.0;
  if (t.valid) re_finalize(P, t, fa, sat0, fin, wch, hwc, hn, wc, h, hout, hmod, l_after, l_adj, dh);
  const double a = block_sum<NT>(l_after, lds);
  const double bs = block_sum<NT>(l_adj, lds);
  dh = wave_max(dh);
  if ((threadIdx.x % kWave) == 0) lmax[threadIdx.x / kWave] = dh;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lmax[0];
    for (int q = 1; q < NT / kWave; ++q) m = fmax(m, lmax[q]);
    part[3 * nblocks + b.linear] = a;
    part[4 * nblocks + b.linear] = bs;
    mx[b.linear] = m;
    mx[nblocks + b.linear] = 0.0;
  }
}

// adaptiveTimeStep reductions (ReSolver.cpp:891-913): max flux imbalance and min Courant dt
// over non-top cells. max/min are exact, so the result is order-independent.
__global__ __launch_bounds__(NT) void k_re_dtadapt(ReK P, const double* qx, const double* qy,
                                                    const double* qz, const double* wc, double* part,
                                                    unsigned int nblocks) {
  const Cell t = cell_of(P.g);
  double dqm = 0.0, dtc = 1.0e8;
  if (t.valid && t.k > 0) {
    const Geo& g = P.g;
    const long long c = t.c;
    const SoilClass& s = P.soil.at(c);
    const double Ax = P.dy * P.dz, Ay = P.dx * P.dz, Az = P.dx * P.dy;
    const double qin = qx[c] / Ax + qy[c] / Ay + qz[c] / Az;
    const double qou = qx[c - 1] / Ax + qy[c - g.sy] / Ay + qz[c - g.sz] / Az;
    dqm = fabs(qin - qou) * P.dt / P.dz;
    if (wc[c] < s.theta_s) dtc = P.co_max * P.dz / vg_dKdwc(s, wc[c], P.soil.Ksz(c));
  }
  dqm = wave_max(dqm);
  dtc = wave_min(dtc);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __shared__ double lmax[NT / kWave], lmin[NT / kWave];
  if (lane == 0) {
    lmax[w] = dqm;
    lmin[w] = dtc;
  }
  __syncthreads();
  const unsigned int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (threadIdx.x == 0) {
    double a = lmax[0], b = lmin[0];
    for (int q = 1; q < NT / kWave; ++q) {
      a = fmax(a, lmax[q]);
      b = fmin(b, lmin[q]);
    }
    part[bid] = a;
    part[nblocks + bid] = b;
  }
}

// k_re_post at the Picard iterate h, with the Picard test: per-block max |h - ref| over interior
// cells to part[0..nblocks) (part[nblocks..) = 0 for k_reduce_maxmin), and hm = h (the iterate
// the next test compares against; hm may be ref itself).
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_post_delta(K P, double* __restrict__ h, double* __restrict__ kr,
                                                       double* __restrict__ ch, double* __restrict__ wch,
                                                       const double* ref, double* hm, double* part,
                                                       unsigned int nblocks) {
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  double d = 0.0;
  if (t.valid) {
    const double hv = h[t.c];
    d = fabs(hv - ref[t.c]);
    hm[t.c] = hv;
    re_post(P, t, h, kr, ch, wch);
  }
  d = wave_max(d);
  __shared__ double lmax[NT / kWave];
  if ((threadIdx.x % kWave) == 0) lmax[threadIdx.x / kWave] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lmax[0];
    for (int q = 1; q < NT / kWave; ++q) m = fmax(m, lmax[q]);
    part[b.linear] = m;
    part[nblocks + b.linear] = 0.0;
  }
}

// The first test of a solve from the Picard iterate x = h^m, taken before its system is stored:
// each cell forms its row as k_re_assemble does (the -x/-y/-z couplings of a row are the stored
// +couplings of the neighbour rows, the same products of the same face conductivities) and
// r = b - A x in the solvers' order of terms. Per block: sum (r/D)^2 and sum (b/D)^2 over the
// interior cells to part[0..nblocks) and part[nblocks..2 nblocks); the solvers stop before
// their first update when ||D^-1 r|| <= max(rtol ||D^-1 b||, atol).
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_check(K P, const double* __restrict__ kr, const double* __restrict__ ch,
                                                  const double* __restrict__ wcn, const double* __restrict__ hn,
                                                  const double* __restrict__ x, double* __restrict__ part,
                                                  unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  double srr = 0.0, sbb = 0.0;
  if (t.valid) {
    const ReRow row = re_assemble_row(P, t, kr, ch, wcn, hn);
    const Geo& g = P.g;
    const long long c = t.c;
    double y = row.D * x[c] + row.xp * x[c + 1] + row.xm * x[c - 1] + row.yp * x[c + g.sy] + row.ym * x[c - g.sy];
    y += row.zp * x[c + g.sz] + row.zm * x[c - g.sz];
    const double r = row.rhs - y;
    const double z = r / row.D, bz = row.rhs / row.D;
    srr = z * z;
    sbb = bz * bz;
  }
  const double a = block_sum<NT>(srr, lds);
  const double bs = block_sum<NT>(sbb, lds);
  if (threadIdx.x == 0) {
    part[b.linear] = a;
    part[nblocks + b.linear] = bs;
  }
}

// Cell `lane` of 64-cell row chunk ch (chunks numbered along x, then rows j, then layers k: the
// chunks of the learned start's inference).
__device__ inline Cell chunk_cell(const Geo& g, long long ch, int lane) {
  const int iblocks = (g.nx + 63) / 64;
  const long long row = ch / iblocks;
  Cell t;
  t.i = (int)(ch - row * iblocks) * 64 + lane;
  t.j = (int)(row % g.ny);
  t.k = (int)(row / g.ny);
  t.valid = t.i < g.nx;
  t.c = t.valid ? g.idx(t.i, t.j, t.k) : 0;
  return t;
}

// The Picard loop's first iterate h^0 = h^n + (hp - h) (h^n plus the previous step's solve
// increment: hp holds that step's post-solve head and h its starting head) fused with the
// post-solve pass at h^0 that follows it (k_re_post_delta's work, without its unused test):
// head-BC ghosts, K, C and theta at h^0, and hp = h^0 for the first Picard test. One wave per
// 64-cell row chunk. GATED (the learned start): the cells whose |hp - h| exceeds the gate are
// marked in masks[chunk]; a chunk with a marked cell gets h^0 only, its post-solve pass waits for
// the inference (k_re_post_chunks). hp and h are read and written at the thread's own cell only.
template <class K, bool GATED>
__global__ __launch_bounds__(NT) void k_re_start_post(K P, const double* __restrict__ hn, double* hp, double* h,
                                                       double* __restrict__ kr, double* __restrict__ ch,
                                                       double* __restrict__ wch, double gate,
                                                       unsigned long long* __restrict__ masks, long long nchunks) {
  const long long q = (long long)blockIdx.x * (NT / kWave) + (threadIdx.x >> 6);
  if (q >= nchunks) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  const Cell t = chunk_cell(P.g, q, lane);
  double hv = 0.0;
  bool need = false;
  if (t.valid) {
    const double inc = hp[t.c] - h[t.c];
    hv = hn[t.c] + inc;
    h[t.c] = hv;
    need = GATED && fabs(inc) > gate;
  }
  if (GATED) {
    const unsigned long long m = __ballot(need);
    if (lane == 0) masks[q] = m;
    if (m != 0ull) return;  // wave-uniform
  }
  if (t.valid) {
    hp[t.c] = hv;
    re_post(P, t, h, kr, ch, wch);
  }
}

// The post-solve pass at h^0 (as in k_re_start_post) for the chunks with a marked cell, after the
// inference has written their heads.
template <class K>
__global__ __launch_bounds__(NT) void k_re_post_chunks(K P, const unsigned long long* __restrict__ masks,
                                                        long long nchunks, double* h, double* __restrict__ kr,
                                                        double* __restrict__ ch, double* __restrict__ wch,
                                                        double* __restrict__ hm) {
  const long long q = (long long)blockIdx.x * (NT / kWave) + (threadIdx.x >> 6);
  if (q >= nchunks || masks[q] == 0ull) return;  // wave-uniform
  const Cell t = chunk_cell(P.g, q, threadIdx.x & 63);
  if (t.valid) {
    hm[t.c] = h[t.c];
    re_post(P, t, h, kr, ch, wch);
  }
}

// k_re_check fused with k_re_flux_correct at the same iterate h (both form the same face
// conductivities from kr and theta^n): the test's partials to cpart[0..2 nblocks), the flux and
// corrector outputs and ledger partials as k_re_flux_correct writes them. When the test passes
// the loop ends at h and these are the step's fluxes and corrector; when it fails the loop goes
// on (nothing it reads is written here) and the flux kernel runs again at its last iterate.
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_check_flux_correct(
    K P, const double* __restrict__ h, const double* __restrict__ kr, const double* __restrict__ ch,
    const double* __restrict__ hn, const double* __restrict__ wcn, const double* __restrict__ wch,
    double* __restrict__ Kx, double* __restrict__ Ky, double* __restrict__ Kz, double* __restrict__ qx,
    double* __restrict__ qy, double* __restrict__ qz, double* __restrict__ wc, double* __restrict__ hwc,
    unsigned char* __restrict__ sat0, unsigned char* __restrict__ cand, unsigned char* __restrict__ kf,
    double* __restrict__ part, double* __restrict__ cpart, unsigned int nblocks, int* __restrict__ flags,
    int nflags) {
  __shared__ double lds5[5 * (NT / kWave)];
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  if (b.linear == 0 && (int)threadIdx.x < nflags) flags[threadIdx.x] = 0;  // reallocation sweep flags
  double srr = 0.0, sbb = 0.0, l_before = 0.0, l_ss = 0.0, l_bnd = 0.0;
  if (t.valid) {
    const ReRow row = re_assemble_row(P, t, kr, ch, wcn, hn);
    const Geo& g = P.g;
    const long long c = t.c;
    double y = row.D * h[c] + row.xp * h[c + 1] + row.xm * h[c - 1] + row.yp * h[c + g.sy] + row.ym * h[c - g.sy];
    y += row.zp * h[c + g.sz] + row.zm * h[c - g.sz];
    const double r = row.rhs - y;
    const double z = r / row.D, bz = row.rhs / row.D;
    srr = z * z;
    sbb = bz * bz;
    re_flux_correct(P, t, h, kr, hn, wcn, wch, Kx, Ky, Kz, qx, qy, qz, wc, hwc, sat0, cand, kf, l_before, l_ss,
                    l_bnd);
  }
  double v[5] = {l_before, l_ss, l_bnd, srr, sbb};
  block_sums<NT, 5>(v, lds5);
  if (threadIdx.x == 0) {
    part[b.linear] = v[0];
    part[nblocks + b.linear] = v[1];
    part[2 * nblocks + b.linear] = v[2];
    cpart[b.linear] = v[3];
    cpart[nblocks + b.linear] = v[4];
  }
}

// k_re_assemble with the FP32 copies, fused with the first sweep of the mixed-precision solve
// from x that follows it (the paired refinement sweep with the start of the inner FP32 CG): from
// the row in registers, r = b - A x in FP64 (the sweep's order of terms), r32 = (float) r,
// z32 = r32 / D32, xw = x (the copy that makes one refinement end in x), and per block the five
// partials of that sweep at stride nblocks: ||D^-1 r||^2, ||D^-1 b||^2, r32.z32, z32.z32 and
// ||D32^-1 r32||^2.
template <class K, int ZF>
__global__ __launch_bounds__(NT) void k_re_assemble_sweep(K P, const double* __restrict__ kr,
                                                           const double* __restrict__ ch,
                                                           const double* __restrict__ wcn,
                                                           const double* __restrict__ hn,
                                                           const double* __restrict__ x, double* __restrict__ D,
                                                           double* __restrict__ ax, double* __restrict__ ay,
                                                           double* __restrict__ az, double* __restrict__ rhs,
                                                           float* __restrict__ D32, float* __restrict__ ax32,
                                                           float* __restrict__ ay32, float* __restrict__ az32,
                                                           double* __restrict__ xw, float* __restrict__ r32,
                                                           float* __restrict__ z32, double* __restrict__ part,
                                                           unsigned int nblocks) {
  __shared__ double lds5[5 * (NT / kWave)];
  TileId b;
  const Cell t = tile_cell<ZF>(P.g, b);
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (t.valid) {
    const ReRow row = re_assemble_row(P, t, kr, ch, wcn, hn);
    const long long c = t.c;
    re_store_row(c, row, D, ax, ay, az, rhs, D32, ax32, ay32, az32);
    const Geo& g = P.g;
    const double xc = x[c];
    double y = row.D * xc + row.xp * x[c + 1] + row.xm * x[c - 1] + row.yp * x[c + g.sy] + row.ym * x[c - g.sy];
    y += row.zp * x[c + g.sz] + row.zm * x[c - g.sz];
    const double r = row.rhs - y;
    const double z = r / row.D, bz = row.rhs / row.D;
    const float rT = (float)r, dd = (float)row.D;
    const float zT = (float)((double)rT / (double)dd);
    const double ib = (double)rT / (double)dd;
    xw[c] = xc;
    r32[c] = rT;
    z32[c] = zT;
    s[0] = z * z;
    s[1] = bz * bz;
    s[2] = (double)rT * (double)zT;
    s[3] = (double)zT * (double)zT;
    s[4] = ib * ib;
  }
  block_sums<NT, 5>(s, lds5);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 5; ++q) part[q * nblocks + b.linear] = s[q];
  }
}

__global__ __launch_bounds__(NT) void k_re_initial(ReK P, int from_head, double* h, double* wc, double* hn,
                                                    double* wcn, double* wch, double* hwc, double* ch) {
  const Cell t = cell_of(P.g);
  if (t.valid) re_initial(P, t, from_head, h, wc, hn, wcn, wch, hwc, ch);
}

// x0 = h^n on interior cells (previous-step initial guess).
__global__ __launch_bounds__(NT) void k_copy_interior(Geo g, const double* src, double* dst) {
  const Cell t = cell_of(g);
  if (!t.valid) return;
  dst[t.c] = src[t.c];
}

// h <- h^n + (hp - h): h^n plus the previous step's solve increment, hp holding that step's
// post-solve head and h its starting head on entry.
__global__ __launch_bounds__(NT) void k_extrap_interior(Geo g, const double* hn, const double* hp, double* h) {
  const Cell t = cell_of(g);
  if (!t.valid) return;
  h[t.c] = hn[t.c] + (hp[t.c] - h[t.c]);
}

__global__ __launch_bounds__(NT) void k_sub_interior(Geo g, const double* a, const double* b, double* dst) {
  const Cell t = cell_of(g);
  if (!t.valid) return;
  dst[t.c] = a[t.c] - b[t.c];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
ReState::ReState(int nx, int ny, int nz, const frehg_re_params& p, hipStream_t s, Profiler* prof)
    : g_(make_geo(nx, ny, nz)), p_(p), s_(s), prof_(prof) {
  if (nx < 1 || ny < 1 || nz < 1) throw ArgError("frehg_re_init: grid sizes must be >= 1");
  try {
    build(p);
  } catch (...) {
    release();  // a failed set-up frees what it allocated
    throw;
  }
}

void ReState::build(const frehg_re_params& p) {
  const int nx = g_.nx, ny = g_.ny, nz = g_.nz;
  n_ = g_.ncell();
  const size_t bytes = (size_t)n_ * sizeof(double);
  const char* lz = std::getenv("FREHG_LAZY_FLUX");
  lazy_flux_ = !(lz && lz[0] == '0');
  const char* pc = std::getenv("FREHG_PICARD_CHECK");
  picard_check_ = !(pc && pc[0] == '0');
  const char* as = std::getenv("FREHG_ASSEMBLE_SWEEP");
  assemble_sweep_ = !(as && as[0] == '0');
  const char* cf = std::getenv("FREHG_CHECK_FLUX");
  check_flux_ = !(cf && cf[0] == '0');
  const char* fs = std::getenv("FREHG_FUSED_START");
  fused_start_ = !(fs && fs[0] == '0');
  for (double** f : {&h_, &hn_, &hp_, &wc_, &wcn_, &wch_, &hwc_, &ch_, &Kx_, &Ky_, &Kz_, &qx_, &qy_, &qz_, &kr_,
                     &D_, &ax_, &ay_, &az_, &rhs_}) {
    FREHG_HIP(hipMalloc(f, bytes));
    FREHG_HIP(hipMemsetAsync(*f, 0, bytes, s_));
  }
  for (unsigned char** f : {&sat0_, &cand_, &satA_, &satB_, &hmod_, &kf_}) {
    FREHG_HIP(hipMalloc(f, (size_t)n_));
    FREHG_HIP(hipMemsetAsync(*f, 0, (size_t)n_, s_));
  }
  FREHG_HIP(hipMalloc(&dflag_, kSweepFlags * sizeof(int)));
  FREHG_HIP(hipMemsetAsync(dflag_, 0, kSweepFlags * sizeof(int), s_));
  FREHG_HIP(hipHostMalloc(&hflag_, 16, hipHostMallocDefault));
  const unsigned int nb = ceil_div(nx, 64) * ceil_div(ny, 4) * nz;
  FREHG_HIP(hipMalloc(&red_, 2 * (size_t)nb * sizeof(double)));
  FREHG_HIP(hipMalloc(&led_, 7 * (size_t)nb * sizeof(double)));
  FREHG_HIP(hipMalloc(&dres_, 16 * sizeof(double)));
  FREHG_HIP(hipMalloc(&red2_, 5 * kRedStage * sizeof(double)));
  FREHG_HIP(hipHostMalloc(&hres_, 16 * sizeof(double), hipHostMallocDefault));
  FREHG_HIP(hipMalloc(&dcount_, 16));
  FREHG_HIP(hipMemsetAsync(dcount_, 0, 16, s_));
  frehg_soil_params sp = p.soil;
  set_soil(&sp, 1, nullptr, 0, nullptr);
  kry_.reset(new Krylov(g_, true, s_, prof_));
  set_params(p);
  FREHG_HIP(hipStreamSynchronize(s_));
}

ReState::~ReState() { release(); }

void ReState::release() {
  for (void* f : {(void*)h_, (void*)hn_, (void*)hp_, (void*)wc_, (void*)wcn_, (void*)wch_, (void*)hwc_, (void*)ch_,
                  (void*)Kx_, (void*)Ky_, (void*)Kz_, (void*)qx_, (void*)qy_, (void*)qz_, (void*)kr_,
                  (void*)D_, (void*)ax_, (void*)ay_, (void*)az_, (void*)rhs_, (void*)D32_, (void*)ax32_,
                  (void*)ay32_, (void*)az32_, (void*)sat0_, (void*)cand_, (void*)satA_, (void*)satB_, (void*)hmod_,
                  (void*)kf_, (void*)dflag_, (void*)red_, (void*)red2_, (void*)led_, (void*)dres_, (void*)dcount_, (void*)classes_, (void*)cls_,
                  (void*)ks_, (void*)qtop_, (void*)dhp_})
    if (f) (void)hipFree(f);
  if (hflag_) (void)hipHostFree(hflag_);
  if (hres_) (void)hipHostFree(hres_);
}

void ReState::ensure_dhp() {
  if (dhp_) return;
  FREHG_HIP(hipMalloc(&dhp_, (size_t)n_ * sizeof(double)));
  FREHG_HIP(hipMemsetAsync(dhp_, 0, (size_t)n_ * sizeof(double), s_));
}

void ReState::set_params(const frehg_re_params& p) {
  for (int b = 0; b < 4; ++b)
    if (p.bc_type[b] != 0)
      throw ArgError("frehg_re: lateral GW boundaries must be no-flux (0); Dirichlet/fixed-flux "
                     "lateral faces are not implemented (src/driver/Orchestrator.cpp:155-163)");
  if (p.bc_type[4] < 0 || p.bc_type[4] > 3) throw ArgError("frehg_re: bottom bc must be 0, 1, 2 or 3");
  if (p.bc_type[5] < 0 || p.bc_type[5] > 2) throw ArgError("frehg_re: top bc must be 0, 1 or 2");
  if (!(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.dt > 0)) throw ArgError("frehg_re: dx, dy, dz, dt must be > 0");
  if (p.picard_max < 1) throw ArgError("frehg_re: picard_max must be >= 1 (1 = the reference's predictor)");
  if (p.picard_max > 1 && !(p.picard_tol > 0.0)) throw ArgError("frehg_re: picard_tol must be > 0");
  if (p.initial_guess < FREHG_GUESS_ZERO || p.initial_guess > FREHG_GUESS_EXTRAPOLATED)
    throw ArgError("frehg_re: unknown initial_guess");
  (void)krylov_config(p.solver, "frehg_re");
  const bool mixed = p.solver.precision == FREHG_PREC_MIXED;
  if (mixed && !D32_) {
    const size_t b32 = (size_t)n_ * sizeof(float);
    for (float** f : {&D32_, &ax32_, &ay32_, &az32_}) {
      FREHG_HIP(hipMalloc(f, b32));
      FREHG_HIP(hipMemsetAsync(*f, 0, b32, s_));
    }
  }
  flush_lazy();
  const bool first = (dtn_ == 0.0);
  p_ = p;
  kc_fresh_ = false;
  if (first) dtn_ = p.dt;
}

void ReState::set_soil(const frehg_soil_params* classes, int ncls, const int32_t* class_idx, int idx3d,
                       const double* ks_cell) {
  if (ncls < 1) throw ArgError("frehg_re_set_soil: need at least one soil class");
  flush_lazy();
  kc_fresh_ = false;
  host_classes_.clear();
  for (int c = 0; c < ncls; ++c) {
    const frehg_soil_params& s = classes[c];
    if (!(s.n > 1.0) || !(s.theta_s > s.theta_r) || !(s.alpha > 0))
      throw ArgError("frehg_re_set_soil: invalid class (need n > 1, theta_s > theta_r, alpha > 0)");
    host_classes_.push_back(make_soil_class(s.alpha, s.n, s.theta_s, s.theta_r, s.Ks_x, s.Ks_y, s.Ks_z, s.Ss,
                                            s.use_vg, s.use_mvg, s.air_entry));
  }
  if (classes_) FREHG_HIP(hipFree(classes_));
  FREHG_HIP(hipMalloc(&classes_, host_classes_.size() * sizeof(SoilClass)));
  FREHG_HIP(hipMemcpy(classes_, host_classes_.data(), host_classes_.size() * sizeof(SoilClass),
                      hipMemcpyHostToDevice));
  const int nx = g_.nx, ny = g_.ny, nz = g_.nz;
  auto clamp = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  if (cls_) {
    FREHG_HIP(hipFree(cls_));
    cls_ = nullptr;
  }
  if (class_idx) {
    std::vector<int32_t> pad((size_t)n_, 0);
    for (int k = -1; k <= nz; ++k)
      for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i) {
          const int ci = clamp(i, nx), cj = clamp(j, ny), ck = clamp(k, nz);
          const size_t o = idx3d ? (size_t)ci + (size_t)cj * nx + (size_t)ck * nx * ny : (size_t)ci + (size_t)cj * nx;
          const int32_t v = class_idx[o];
          if (v < 0 || v >= ncls) throw ArgError("frehg_re_set_soil: class index out of range");
          pad[(size_t)g_.idx(i, j, k)] = v;
        }
    FREHG_HIP(hipMalloc(&cls_, (size_t)n_ * sizeof(int32_t)));
    FREHG_HIP(hipMemcpy(cls_, pad.data(), (size_t)n_ * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  if (ks_) {
    FREHG_HIP(hipFree(ks_));
    ks_ = nullptr;
  }
  if (ks_cell) {
    std::vector<double> pad((size_t)n_, 0.0);
    for (int k = -1; k <= nz; ++k)
      for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i)
          pad[(size_t)g_.idx(i, j, k)] =
              ks_cell[(size_t)clamp(i, nx) + (size_t)clamp(j, ny) * nx + (size_t)clamp(k, nz) * nx * ny];
    FREHG_HIP(hipMalloc(&ks_, (size_t)n_ * sizeof(double)));
    FREHG_HIP(hipMemcpy(ks_, pad.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice));
  }
}

void ReState::set_top_flux(const double* q) {
  flush_lazy();
  std::vector<double> pad((size_t)g_.nsurf(), 0.0);
  for (int j = 0; j < g_.ny; ++j)
    for (int i = 0; i < g_.nx; ++i) pad[(size_t)g_.sidx(i, j)] = q[(size_t)i + (size_t)j * g_.nx];
  if (!qtop_) FREHG_HIP(hipMalloc(&qtop_, pad.size() * sizeof(double)));
  FREHG_HIP(hipMemcpy(qtop_, pad.data(), pad.size() * sizeof(double), hipMemcpyHostToDevice));
}

double* ReState::field_ptr(const std::string& n) {
  if (n == "h") return h_;
  if (n == "hn") return hn_;
  if (n == "wc") return wc_;
  if (n == "wcn") return wcn_;
  if (n == "wch") return wch_;
  if (n == "hwc") return hwc_;
  if (n == "ch") return ch_;
  if (n == "Kx") return Kx_;
  if (n == "Ky") return Ky_;
  if (n == "Kz") return Kz_;
  if (n == "qx") return qx_;
  if (n == "qy") return qy_;
  if (n == "qz") return qz_;
  if (n == "Gct") return D_;
  if (n == "Grhs") return rhs_;
  throw ArgError("unknown GW field '" + n + "'");
}

void ReState::set_field(const std::string& name, const double* src) {
  double* dst = field_ptr(name);
  flush_lazy();
  kc_fresh_ = false;
  inc_valid_ = false;
  copy_field_in(g_, g_.nz + 2, dst, src, s_);
  FREHG_HIP(hipStreamSynchronize(s_));
}

void ReState::get_field(const std::string& name, double* dst) {
  const double* src = field_ptr(name);
  if (src == Kx_ || src == Ky_ || src == Kz_ || src == qx_ || src == qy_ || src == qz_) materialize_fluxes();
  if (src == D_ || src == rhs_) materialize_operator();
  copy_field_out(g_, g_.nz + 2, dst, src, s_);
  FREHG_HIP(hipStreamSynchronize(s_));
}


void ReState::set_initial(const double* h_int, const double* wc_int) {
  if ((h_int == nullptr) == (wc_int == nullptr))
    throw ArgError("frehg_re_set_initial: pass exactly one of h / wc");
  const double* src = h_int ? h_int : wc_int;
  flush_lazy();
  kc_fresh_ = false;
  std::vector<double> pad((size_t)n_, 0.0);
  FREHG_HIP(hipMemcpy(pad.data(), h_int ? h_ : wc_, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < g_.nz; ++k)
    for (int j = 0; j < g_.ny; ++j)
      for (int i = 0; i < g_.nx; ++i)
        pad[(size_t)g_.idx(i, j, k)] = src[(size_t)i + (size_t)j * g_.nx + (size_t)k * g_.nx * g_.ny];
  FREHG_HIP(hipMemcpy(h_int ? h_ : wc_, pad.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice));
  const ReK P = make_rek(g_, soil_view(), p_, qtop_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4), g_.nz);
  hipLaunchKernelGGL(k_re_initial, grid, dim3(NT), 0, s_, P, h_int ? 1 : 0, h_, wc_, hn_, wcn_, wch_, hwc_, ch_);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipStreamSynchronize(s_));
  dtn_ = p_.dt;
  inc_valid_ = false;
}

void ReState::learned_increment(double* dst) {
  if (!mlp_ || !mlp_->ready()) throw ArgError("frehg_mlp_increment: no trained model");
  DeviceBuffer<double> tmp((size_t)n_);
  FREHG_HIP(hipMemsetAsync(tmp.get(), 0, (size_t)n_ * sizeof(double), s_));
  mlp_->predict_guess(*this, tmp.get(), depth_, nullptr, nullptr, h_, wc_);
  std::vector<double> pad((size_t)n_), hh((size_t)n_);
  FREHG_HIP(hipMemcpyAsync(pad.data(), tmp.get(), (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipMemcpyAsync(hh.data(), h_, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  for (int k = 0; k < g_.nz; ++k)
    for (int j = 0; j < g_.ny; ++j)
      for (int i = 0; i < g_.nx; ++i) {
        const size_t c = (size_t)g_.idx(i, j, k);
        dst[(size_t)i + (size_t)j * g_.nx + (size_t)k * g_.nx * g_.ny] = pad[c] - hh[c];
      }
}

void ReState::learned_start(double* dst) {
  if (!mlp_ || !mlp_->ready()) throw ArgError("frehg_mlp_start: no trained model");
  // The next step swaps h and h^n first: its h^n is h_ now, and its previous solve increment is
  // the post-solve head hp_ minus the last step's starting head, now in hn_.
  DeviceBuffer<double> tmp((size_t)n_);
  FREHG_HIP(hipMemsetAsync(tmp.get(), 0, (size_t)n_ * sizeof(double), s_));
  mlp_->predict_guess(*this, tmp.get(), depth_, inc_valid_ ? hp_ : nullptr, inc_valid_ ? hn_ : nullptr, h_, wc_);
  std::vector<double> pad((size_t)n_);
  FREHG_HIP(hipMemcpyAsync(pad.data(), tmp.get(), (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  for (int k = 0; k < g_.nz; ++k)
    for (int j = 0; j < g_.ny; ++j)
      for (int i = 0; i < g_.nx; ++i)
        dst[(size_t)i + (size_t)j * g_.nx + (size_t)k * g_.nx * g_.ny] = pad[(size_t)g_.idx(i, j, k)];
}

void ReState::materialize_operator() {
  if (!op_pending_) return;
  op_pending_ = false;
  frehg_re_params q = p_;
  q.dt = dt_used_;  // the step's; an adaptive step has already set p_.dt to the next one
  const ReK P = make_rek(g_, soil_view(), q, qtop_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4), g_.nz);
  with_soil_variant(P, [&](const auto& PK) {
    using K = std::decay_t<decltype(PK)>;
    hipLaunchKernelGGL((k_re_assemble<K, 1>), grid, dim3(NT), 0, s_, PK, (const double*)kr_, (const double*)ch_,
                       (const double*)wcn_, (const double*)hn_, D_, ax_, ay_, az_, rhs_, nullptr, nullptr, nullptr,
                       nullptr);
    FREHG_LAUNCH_CHECK();
  });
}

void ReState::materialize_fluxes() {
  if (!flux_pending_) return;
  const ReK P = make_rek(g_, soil_view(), p_, qtop_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4), g_.nz);
  with_soil_variant(P, [&](const auto& PK) {
    using K = std::decay_t<decltype(PK)>;
    hipLaunchKernelGGL((k_re_flux<K, 1>), grid, dim3(NT), 0, s_, PK, (const double*)hp_, (const double*)kr_,
                       (const double*)wcn_, Kx_, Ky_, Kz_, qx_, qy_, qz_);
    FREHG_LAUNCH_CHECK();
  });
  flux_pending_ = false;
}

void ReState::step(frehg_re_stats* st) {
  // Reject a request that cannot be served before any state changes.
  if (p_.initial_guess == FREHG_GUESS_LEARNED && (!mlp_ || !mlp_->ready()))
    throw ArgError("frehg_re_step: learned guess requested but no trained model loaded");
  frehg_re_stats local;
  std::memset(&local, 0, sizeof(local));
  dt_used_ = p_.dt;
  const bool mixed = p_.solver.precision == FREHG_PREC_MIXED;
  const ReK P = make_rek(g_, soil_view(), p_, qtop_);
  const dim3 grid(ceil_div(g_.nx, 64), ceil_div(g_.ny, 4), g_.nz);
  const dim3 blk(NT);
  const unsigned int nbl = grid.x * grid.y * grid.z;
  // savePreviousStepFields (ReSolver.cpp:190-197): hn <- h, wcn <- wc. The arrays are swapped
  // instead of copied; the new h and wc buffers are fully rewritten below (corner ghosts, which
  // no kernel reads or writes, hold the same values in both buffers).
  std::swap(h_, hn_);
  std::swap(wc_, wcn_);
  const bool kc_fresh = kc_fresh_;
  kc_fresh_ = false;  // set again once this step has completed
  flux_pending_ = false;  // this step's fluxes supersede the previous step's
  op_pending_ = false;    // and its system the previous step's
  // With the Picard loop the learned head is the loop's starting iterate h^0 (below): the system
  // is assembled at it, and K and C at h^n are not needed (the network's features evaluate their
  // own from h^n).
  const bool learned_h0 = p_.initial_guess == FREHG_GUESS_LEARNED && p_.picard_max > 1;
  // The extrapolated head as h^0 likewise.
  const bool extrap_h0 = p_.initial_guess == FREHG_GUESS_EXTRAPOLATED && p_.picard_max > 1;
  const bool start_h0 = learned_h0 || extrap_h0;
  // The previous step's solve increment d^{n-1} = hp_ - h_ here (its post-solve head minus its
  // starting head, both kept since): the extrapolated start adds it to h^n, the learned start
  // adds it and the network's prediction, and pair logging subtracts it from the logged
  // increment (a copy, taken before the guess overwrites h_). A step without one (the first
  // after set-up) logs no pairs.
  const bool logging = mlp_ && mlp_->logging();
  const bool track_inc = p_.initial_guess == FREHG_GUESS_EXTRAPOLATED ||
                         p_.initial_guess == FREHG_GUESS_LEARNED || logging;
  const bool have_inc = inc_valid_;
  inc_valid_ = false;
  const double* inc0 = nullptr;  // d^{n-1} for the logged pairs
  if (logging && have_inc) {
    ensure_dhp();
    hipLaunchKernelGGL(k_sub_interior, grid, blk, 0, s_, g_, (const double*)hp_, (const double*)h_, dhp_);
    FREHG_LAUNCH_CHECK();
    inc0 = dhp_;
  }
  if (!start_h0) {
    hipEvent_t ev0 = prof_->start(FREHG_KT_GW_ASSEMBLE);
    hipEvent_t ev1 = prof_->stop();
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      if (kc_fresh)
        hipExtLaunchKernelGGL((k_re_pre_stale<K, 1>), grid, blk, 0, s_, ev0, nullptr, 0, PK,
                              (const unsigned char*)hmod_, (const double*)hn_, kr_, ch_);
      else
        hipExtLaunchKernelGGL((k_re_pre<K, 1>), grid, blk, 0, s_, ev0, nullptr, 0, PK, (const double*)hn_, kr_, ch_);
      FREHG_LAUNCH_CHECK();
      hipExtLaunchKernelGGL((k_re_assemble<K, 1>), grid, blk, 0, s_, nullptr, ev1, 0, PK, (const double*)kr_,
                            (const double*)ch_, (const double*)wcn_, (const double*)hn_, D_, ax_, ay_, az_, rhs_,
                            mixed ? D32_ : nullptr, ax32_, ay32_, az32_);
      FREHG_LAUNCH_CHECK();
    });
  }
  // Initial guess in h (the solve's x): zero (PETSc default), h^n, h^n + the previous step's
  // solve increment, or h^n + MLP increment.
  bool x0_zero = true, guard = false;
  bool post0 = false;  // the start kernels took the post-solve pass at h^0
  const long long nchunks = (long long)ceil_div(g_.nx, 64) * g_.ny * g_.nz;  // 64-cell row chunks
  if (p_.initial_guess == FREHG_GUESS_PREVIOUS || (p_.initial_guess == FREHG_GUESS_EXTRAPOLATED && !have_inc)) {
    hipLaunchKernelGGL(k_copy_interior, grid, blk, 0, s_, g_, hn_, h_);
    FREHG_LAUNCH_CHECK();
    x0_zero = false;
  } else if (p_.initial_guess == FREHG_GUESS_EXTRAPOLATED) {
    if (extrap_h0 && fused_start_) {  // with the post-solve pass at h^0
      with_soil_variant(P, [&](const auto& PK) {
        using K = std::decay_t<decltype(PK)>;
        hipLaunchKernelGGL((k_re_start_post<K, false>), dim3(ceil_div(nchunks, NT / kWave)), blk, 0, s_, PK,
                           (const double*)hn_, hp_, h_, kr_, ch_, wch_, 0.0, nullptr, nchunks);
        FREHG_LAUNCH_CHECK();
      });
      post0 = true;
    } else {
      hipLaunchKernelGGL(k_extrap_interior, grid, blk, 0, s_, g_, (const double*)hn_, (const double*)hp_, h_);
      FREHG_LAUNCH_CHECK();
    }
    x0_zero = false;
  } else if (p_.initial_guess == FREHG_GUESS_LEARNED) {
    if (!mlp_ || !mlp_->ready()) throw ArgError("frehg_re_step: learned guess requested but no trained model loaded");
    FREHG_HIP(hipMemsetAsync(dres_ + 10, 0, sizeof(double), s_));
    if (learned_h0 && have_inc && mlp_->gate() > 0.0 && fused_start_) {
      // the gated start's base pass with the post-solve pass at h^0 of the chunks the network
      // leaves alone; the network's chunks get theirs after the inference
      unsigned long long* masks = mlp_->chunk_masks(nchunks);
      with_soil_variant(P, [&](const auto& PK) {
        using K = std::decay_t<decltype(PK)>;
        hipLaunchKernelGGL((k_re_start_post<K, true>), dim3(ceil_div(nchunks, NT / kWave)), blk, 0, s_, PK,
                           (const double*)hn_, hp_, h_, kr_, ch_, wch_, mlp_->gate(), masks, nchunks);
        FREHG_LAUNCH_CHECK();
      });
      mlp_->infer_masked(*this, h_, depth_, dres_ + 10);
      with_soil_variant(P, [&](const auto& PK) {
        using K = std::decay_t<decltype(PK)>;
        hipLaunchKernelGGL((k_re_post_chunks<K>), dim3(ceil_div(nchunks, NT / kWave)), blk, 0, s_, PK,
                           (const unsigned long long*)masks, nchunks, h_, kr_, ch_, wch_, hp_);
        FREHG_LAUNCH_CHECK();
      });
      post0 = true;
    } else {
      mlp_->predict_guess(*this, h_, depth_, have_inc ? hp_ : nullptr, have_inc ? h_ : nullptr, nullptr, nullptr,
                          dres_ + 10);
    }
    x0_zero = false;
    guard = true;
  }
  // The learned (or extrapolated) head as the Picard loop's h^0: K and C at it (with its head-BC
  // ghosts), the assembly with the time-level-n terms, and the first stopping test against it
  // (kept in hp_).
  const KrylovConfig kc = krylov_config(p_.solver, "frehg_re_step");
  // The system at kr_, ch_ (FP32 copies when fp32), for a solve from x = h_ when from_h. A
  // mixed-precision solve from h_ with fresh FP32 copies and CG inside takes its first sweep in
  // the assembly kernel (the row in registers instead of re-read); returns the number of blocks
  // of that sweep's partials for the solve (0: the solve takes the sweep itself).
  auto assemble = [&](bool fp32, bool from_h) -> unsigned int {
    const bool sweep = assemble_sweep_ && mixed && fp32 && from_h && kc.ksp == FREHG_KSP_CG;
    const Krylov::SweepTarget st = sweep ? kry_->first_sweep_target() : Krylov::SweepTarget{};
    if (sweep && nbl > st.cap) throw ArgError("frehg_re_step: sweep partials exceed the solver's capacity");
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      if (sweep)
        hipLaunchKernelGGL((k_re_assemble_sweep<K, 1>), grid, blk, 0, s_, PK, (const double*)kr_, (const double*)ch_,
                           (const double*)wcn_, (const double*)hn_, (const double*)h_, D_, ax_, ay_, az_, rhs_,
                           D32_, ax32_, ay32_, az32_, st.xw, st.r32, st.z32, st.part, nbl);
      else
        hipLaunchKernelGGL((k_re_assemble<K, 1>), grid, blk, 0, s_, PK, (const double*)kr_, (const double*)ch_,
                           (const double*)wcn_, (const double*)hn_, D_, ax_, ay_, az_, rhs_, fp32 ? D32_ : nullptr,
                           ax32_, ay32_, az32_);
      FREHG_LAUNCH_CHECK();
    });
    return sweep ? nbl : 0u;
  };
  unsigned int swept = 0;
  if (start_h0) {
    if (!post0)
      with_soil_variant(P, [&](const auto& PK) {
        using K = std::decay_t<decltype(PK)>;
        // post-solve pass at h^0, which also keeps h^0 in hp_ (its test output is not used)
        hipLaunchKernelGGL((k_re_post_delta<K, 1>), grid, blk, 0, s_, PK, h_, kr_, ch_, wch_, (const double*)hn_,
                           hp_, red_, nbl);
        FREHG_LAUNCH_CHECK();
      });
    swept = assemble(mixed, true);
  }
  StencilView<double> A{D_, ax_, ay_, az_};
  StencilView<float> A32{D32_, ax32_, ay32_, az32_};
  SolveResult sr = kry_->solve(A, mixed ? &A32 : nullptr, rhs_, nullptr, h_, nullptr, x0_zero, guard, kc, swept);
  // A learned Picard start whose residual did not drop below the zero guess's: the solve above
  // started from zero, but on the operator assembled at the rejected head. The loop restarts as
  // the plain method: h^0 = h^n, K and C at h^n, x0 = 0.
  bool from_hn = !start_h0;
  if (start_h0 && sr.guess_rejected) {
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      hipLaunchKernelGGL((k_re_pre<K, 1>), grid, blk, 0, s_, PK, (const double*)hn_, kr_, ch_);
      FREHG_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_re_assemble<K, 1>), grid, blk, 0, s_, PK, (const double*)kr_, (const double*)ch_,
                         (const double*)wcn_, (const double*)hn_, D_, ax_, ay_, az_, rhs_, mixed ? D32_ : nullptr,
                         ax32_, ay32_, az32_);
      FREHG_LAUNCH_CHECK();
    });
    const SolveResult s0 = kry_->solve(A, mixed ? &A32 : nullptr, rhs_, nullptr, h_, nullptr, true, false, kc);
    sr.iterations += s0.iterations;
    sr.refinements += s0.refinements;
    sr.residual_norm = s0.residual_norm;
    sr.rhs_norm = s0.rhs_norm;
    sr.converged = s0.converged;
    sr.fallback_fp64 = sr.fallback_fp64 || s0.fallback_fp64;
    from_hn = true;
  }
  local.solve.iterations = sr.iterations;
  local.solve.refinements = sr.refinements;
  local.solve.residual_norm = sr.residual_norm;
  local.solve.rhs_norm = sr.rhs_norm;
  local.solve.converged = sr.converged ? 1 : 0;
  local.solve.fallback = sr.fallback_fp64 ? 1 : 0;
  local.guess_fallback = sr.guess_rejected ? 1 : 0;
  local.picard_iterations = 1;
  local.picard_converged = p_.picard_max > 1 ? 0 : 1;
  bool refined = sr.refinements > 0;  // the last solve took a refinement (mixed precision)
  // Picard outer loop (picard_max > 1; a new method, the reference's step is its first iterate):
  // K and C at the iterate h^m with its head-BC ghosts, the storage and rhs terms at time level n
  // as in the predictor, solve for h^{m+1} (from x0 = 0 with the zero guess, else from h^m); stop
  // when max |h^{m+1} - h^m| <= picard_tol (h^0 = h^n, or the learned head).
  // The post-solve kernel (head BCs, K, C and theta at the iterate) runs at every iterate with
  // the test fused in; at the last one it is the step's own post-solve evaluation.
  bool post_done = false, flux_done = false;
  for (int m = 1; p_.picard_max > 1; ++m) {
    const double* ref = (m == 1 && from_hn) ? hn_ : hp_;
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      hipLaunchKernelGGL((k_re_post_delta<K, 1>), grid, blk, 0, s_, PK, h_, kr_, ch_, wch_, ref, hp_, red_, nbl);
      FREHG_LAUNCH_CHECK();
    });
    post_done = true;
    reduce_maxmin(red_, nbl, red2_, dres_, s_);
    FREHG_LAUNCH_CHECK();
    const bool zero = p_.initial_guess == FREHG_GUESS_ZERO;
    // A solve from h^m first tests the residual of h^m on the system assembled at h^m and returns
    // h^m unchanged when the test passes, which ends the loop (below). That test is taken here
    // on rows formed in registers, without storing the system; only when it fails is the system
    // stored and solved. The stored system then lags one iterate behind; it is assembled when
    // read or before its inputs change (op_pending_). The residual test is queued behind the
    // Picard test without waiting for its outcome, so one host round trip serves both; with the
    // fluxes and corrector at h^m taken in the same pass (the default), those are the step's when
    // either test ends the loop at h^m.
    const bool spec = !zero && picard_check_ && m < p_.picard_max;
    if (spec) {
      hipEvent_t ev0 = check_flux_ ? prof_->start(FREHG_KT_GW_CORRECT) : nullptr;
      hipEvent_t ev1 = check_flux_ ? prof_->stop() : nullptr;
      with_soil_variant(P, [&](const auto& PK) {
        using K = std::decay_t<decltype(PK)>;
        if (check_flux_)
          hipExtLaunchKernelGGL((k_re_check_flux_correct<K, 1>), grid, blk, 0, s_, ev0, ev1, 0, PK,
                                (const double*)h_, (const double*)kr_, (const double*)ch_, (const double*)hn_,
                                (const double*)wcn_, (const double*)wch_, Kx_, Ky_, Kz_, qx_, qy_, qz_, wc_, hwc_,
                                sat0_, cand_, lazy_flux_ ? kf_ : nullptr, led_, red_, nbl, dflag_, kSweepFlags);
        else
          hipLaunchKernelGGL((k_re_check<K, 1>), grid, blk, 0, s_, PK, (const double*)kr_, (const double*)ch_,
                             (const double*)wcn_, (const double*)hn_, (const double*)h_, red_, nbl);
        FREHG_LAUNCH_CHECK();
      });
      reduce_sums(red_, nbl, 2, red2_, dres_ + 11, s_);
      FREHG_LAUNCH_CHECK();
    }
    // [0] the Picard test's max |dh|, [11..12] the residual test's sums (the slots between are
    // rewritten by the ledger copy at the end of the step)
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, (spec ? 13 : 1) * sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    if (hres_[0] <= p_.picard_tol) {
      local.picard_converged = 1;
      flux_done = spec && check_flux_;
      break;
    }
    if (m >= p_.picard_max) break;  // the cap: the step continues from the last iterate
    if (spec) {
      const double rr = hres_[11], bb = hres_[12];
      const double ttol = std::fmax(kc.rtol * std::sqrt(bb), kc.atol);
      if (rr <= ttol * ttol) {
        local.solve.residual_norm = std::sqrt(rr);
        local.solve.rhs_norm = std::sqrt(bb);
        local.picard_iterations = m + 1;
        local.picard_converged = 1;
        post_done = true;
        flux_done = check_flux_;
        op_pending_ = true;
        break;
      }
    }
    // The FP32 copy of the operator is refreshed when the coming solve will need inner solves:
    // from x0 = 0, or after a solve that took a refinement. A solve from h^m that converged at
    // its first residual check leaves the previous copy in place (its inner solves, if any,
    // would only serve refinements whose outer residuals use this iterate's FP64 operator).
    const bool fresh32 = mixed && (zero || refined);
    const unsigned int swept_m = assemble(fresh32, !zero);
    post_done = false;
    const SolveResult sm = kry_->solve(A, mixed ? &A32 : nullptr, rhs_, nullptr, h_, nullptr, zero, false, kc, swept_m);
    local.solve.iterations += sm.iterations;
    local.solve.refinements += sm.refinements;
    local.solve.residual_norm = sm.residual_norm;
    local.solve.rhs_norm = sm.rhs_norm;
    local.solve.converged = (local.solve.converged && sm.converged) ? 1 : 0;
    local.solve.fallback |= sm.fallback_fp64 ? 1 : 0;
    local.picard_iterations = m + 1;
    refined = sm.refinements > 0;
    // A solve from h^m that returned it unchanged (converged at its first residual check, no
    // update): h^{m+1} = h^m bit for bit, so the test passes (max |dh| = 0), and the post-solve
    // pass at h^{m+1} would repeat the one just done at h^m, whose results are still in place.
    if (!zero && sm.iterations == 0 && sm.refinements == 0 && !sm.fallback_fp64) {
      post_done = true;
      local.picard_converged = 1;
      break;
    }
  }
  if (logging && inc0) mlp_->log_pair(*this, hn_, h_, depth_, inc0);
  // Without lazy fluxes finalize writes the final head over the post-solve head: keep the latter
  // in hp_ for the next step's increment.
  if (track_inc && !lazy_flux_) {
    hipLaunchKernelGGL(k_copy_interior, grid, blk, 0, s_, g_, (const double*)h_, hp_);
    FREHG_LAUNCH_CHECK();
  }
  if (!flux_done) {
    hipEvent_t ev0 = prof_->start(FREHG_KT_GW_CORRECT);
    hipEvent_t ev1 = prof_->stop();
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      if (!post_done) {
        hipExtLaunchKernelGGL((k_re_post<K, 1>), grid, blk, 0, s_, ev0, nullptr, 0, PK, h_, kr_, ch_, wch_);
        FREHG_LAUNCH_CHECK();
      }
      hipExtLaunchKernelGGL((k_re_flux_correct<K, 1>), grid, blk, 0, s_, post_done ? ev0 : nullptr, ev1, 0, PK,
                            (const double*)h_,
                            (const double*)kr_, (const double*)hn_, (const double*)wcn_, (const double*)wch_, Kx_, Ky_,
                            Kz_, qx_, qy_, qz_, wc_, hwc_, sat0_, cand_, lazy_flux_ ? kf_ : nullptr, led_, nbl,
                            dflag_, kSweepFlags);
      FREHG_LAUNCH_CHECK();
    });
  }
  // Reallocation fixed point: sweep until the saturation state stops changing. Sweep 0 reads
  // sat0 itself; later sweeps alternate between satA and satB. Sweep s reports a change in
  // dflag_[s] (cleared by the flux kernel above).
  const unsigned char* prev = sat0_;
  for (int it = 0; it < g_.nx * g_.ny * g_.nz + 1; ++it) {
    unsigned char* next = (it & 1) ? satB_ : satA_;
    const int f = it < kSweepFlags ? it : kSweepFlags - 1;
    if (it >= kSweepFlags) FREHG_HIP(hipMemsetAsync(dflag_ + f, 0, sizeof(int), s_));
    with_soil_variant(P, [&](const auto& PK) {
      using K = std::decay_t<decltype(PK)>;
      if (lazy_flux_)
        hipLaunchKernelGGL((k_re_realloc_iter<K, FacesF>), grid, blk, 0, s_, PK, FacesF{kf_},
                           (const unsigned char*)sat0_, (const unsigned char*)cand_, prev, next, dflag_ + f);
      else
        hipLaunchKernelGGL((k_re_realloc_iter<K, FacesK>), grid, blk, 0, s_, PK, FacesK{Kx_, Ky_},
                           (const unsigned char*)sat0_, (const unsigned char*)cand_, prev, next, dflag_ + f);
      FREHG_LAUNCH_CHECK();
    });
    FREHG_HIP(hipMemcpyAsync(hflag_, dflag_ + f, sizeof(int), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    prev = next;
    if (hflag_[0] == 0) break;
  }
  // Lazy fluxes: the final head goes to hp_ and h_ keeps the post-solve head; the swap below
  // leaves the final head in h_ and the post-solve head (a flux rebuild input) in hp_.
  with_soil_variant(P, [&](const auto& PK) {
    using K = std::decay_t<decltype(PK)>;
    if (lazy_flux_)
      hipLaunchKernelGGL((k_re_finalize<K, 0, FacesF>), grid, blk, 0, s_, PK, FacesF{kf_},
                         (const unsigned char*)sat0_, prev, (const double*)wch_, (const double*)hwc_,
                         (const double*)hn_, wc_, h_, hp_, hmod_, led_, led_ + 5 * (size_t)nbl, nbl);
    else
      hipLaunchKernelGGL((k_re_finalize<K, 0, FacesK>), grid, blk, 0, s_, PK, FacesK{Kx_, Ky_},
                         (const unsigned char*)sat0_, prev, (const double*)wch_, (const double*)hwc_,
                         (const double*)hn_, wc_, h_, nullptr, hmod_, led_, led_ + 5 * (size_t)nbl, nbl);
    FREHG_LAUNCH_CHECK();
  });
  if (lazy_flux_) {
    std::swap(h_, hp_);
    flux_pending_ = true;
  }
  // ledger: [before, storage_ss, boundary_in, after, adjust] sums, then max |dh|
  reduce_sums(led_, nbl, 5, red2_, dres_ + 2, s_);
  FREHG_LAUNCH_CHECK();
  reduce_maxmin(led_ + 5 * (size_t)nbl, nbl, red2_, dres_ + 7, s_);
  FREHG_LAUNCH_CHECK();
  if (p_.adaptive_dt) {
    materialize_fluxes();
    const unsigned int nb = grid.x * grid.y * grid.z;
    hipLaunchKernelGGL(k_re_dtadapt, grid, blk, 0, s_, P, qx_, qy_, qz_, wc_, red_, nb);
    FREHG_LAUNCH_CHECK();
    reduce_maxmin(red_, nb, red2_, dres_, s_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    // ReSolver::adaptiveTimeStep (ReSolver.cpp:890-928)
    const double dq_max = hres_[0], dt_Comin = hres_[1];
    p_.dt = dtn_;
    if (dq_max > 0.02) p_.dt *= 0.75;
    else if (dq_max < 0.01) p_.dt *= 1.25;
    if (p_.dt > dt_Comin) p_.dt = dt_Comin;
    if (p_.dt > p_.dt_max) p_.dt = p_.dt_max;
    if (p_.dt < p_.dt_min) p_.dt = p_.dt_min;
    dtn_ = p_.dt;
  }
  FREHG_HIP(hipMemcpyAsync(hres_ + 2, dres_ + 2, 9 * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  // a network output that was not finite: those cells started from h^n + d^{n-1} instead
  if (p_.initial_guess == FREHG_GUESS_LEARNED && hres_[10] != 0.0) local.guess_fallback = 1;
  local.mass_before = hres_[2];
  local.mass_storage_ss = hres_[3];
  local.mass_boundary_in = hres_[4];
  local.mass_after = hres_[5];
  local.mass_adjust = hres_[6];
  local.max_dh = hres_[7];
  local.mass_residual =
      local.mass_after - local.mass_before - local.mass_boundary_in - local.mass_storage_ss - local.mass_adjust;
  local.dt_used = dt_used_;
  local.dt_next = p_.dt;
  kc_fresh_ = true;
  inc_valid_ = track_inc || lazy_flux_;
  if (st) *st = local;
}

}  // namespace frehg
