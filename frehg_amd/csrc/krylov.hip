// Jacobi-PCG and Jacobi-BiCGSTAB kernels for the 7-/5-point stencil systems of the Richards and
// shallow-water steps. One CG iteration is two kernels, with no reduction kernel in between:
//   spmv:   every block first reduces the previous update's partials r.z, z.z (convergence test,
//           beta = r.z / r.z_old), then p = z + beta p_prev (recomputed at each stencil point),
//           q = A p, partial p.q
//   update: every block first reduces the spmv's partials p.q (alpha = r.z / p.q), then
//           x += alpha p, r -= alpha q, z = r / D, partials r.z and z.z
// Each thread owns one (i, j) column of a 64 x 4 tile and marches down k over a chunk of
// layers, keeping the k-1 / k / k+1 values of p and the az coupling of the layer above in
// registers, so every array is streamed from HBM once per iteration.
// One BiCGSTAB iteration (right Jacobi preconditioning, y = D^-1 p, z = D^-1 s) is three kernels:
//   spmv1:  p = r + beta (p_prev - omega v_prev) and y = p / D at each stencil point, v = A y,
//           partial rh.v;                        fin: alpha = rho / rh.v
//   spmv2:  s = r - alpha v and z = s / D at each stencil point, t = A z, partials t.s, t.t,
//           ||D^-1 s||^2;                        fin: omega = t.s / t.t (0 once s has converged)
//   update: x += alpha y + omega z, r = s - omega t, partials rh.r, ||D^-1 r||^2;
//                                                fin: convergence test, beta, rho
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "krylov.hpp"

namespace frehg {

namespace {

constexpr int NT = 256;
constexpr int TX = 64;
constexpr int TY = 4;

struct Tile {
  int i, j, k0, k1;
  bool valid;
};

__device__ inline Tile tile_of(const Geo& g, int kchunk) {
  const TileId b = xcd_tile();
  Tile t;
  t.i = b.x * TX + (threadIdx.x & (TX - 1));
  t.j = b.y * TY + (threadIdx.x / TX);
  t.k0 = b.z * kchunk;
  t.k1 = min(g.nz, t.k0 + kchunk);
  t.valid = t.i < g.nx && t.j < g.ny;
  return t;
}

__device__ inline unsigned int block_id() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

template <bool THREE_D>
__device__ inline long long base_index(const Geo& g, const Tile& t) {
  return THREE_D ? g.idx(t.i, t.j, t.k0) : g.sidx(t.i, t.j);
}

template <typename T, bool THREE_D>
__device__ inline double apply_at(const Geo& g, const StencilView<T>& A, const T* x, long long c) {
  double y = (double)A.D[c] * (double)x[c] + (double)A.ax[c] * (double)x[c + 1] +
             (double)A.ax[c - 1] * (double)x[c - 1] + (double)A.ay[c] * (double)x[c + g.sy] +
             (double)A.ay[c - g.sy] * (double)x[c - g.sy];
  if (THREE_D) y += (double)A.az[c] * (double)x[c + g.sz] + (double)A.az[c - g.sz] * (double)x[c - g.sz];
  return y;
}

// -x/-y/-z couplings of row c: the row's own (nonsymmetric operators) or the neighbour's +x/+y/+z.
template <typename T>
__device__ inline double cxm(const StencilView<T>& A, long long c) {
  return A.axm ? (double)A.axm[c] : (double)A.ax[c - 1];
}
template <typename T>
__device__ inline double cym(const StencilView<T>& A, long long c, long long sy) {
  return A.aym ? (double)A.aym[c] : (double)A.ay[c - sy];
}
template <typename T>
__device__ inline double czm(const StencilView<T>& A, long long c, long long sz) {
  return A.azm ? (double)A.azm[c] : (double)A.az[c - sz];
}

// (A v)(c) for cell (i, j, k) of the domain; V(cc) is read only at in-domain stencil points, so
// it may divide by D without touching ghost cells.
template <typename T, bool THREE_D, typename F>
__device__ inline double apply_in_domain(const Geo& g, const StencilView<T>& A, long long c, int i, int j, int k,
                                         double vc, F&& V) {
  double y = (double)A.D[c] * vc;
  if (i + 1 < g.nx) y += (double)A.ax[c] * V(c + 1);
  if (i > 0) y += cxm(A, c) * V(c - 1);
  if (j + 1 < g.ny) y += (double)A.ay[c] * V(c + g.sy);
  if (j > 0) y += cym(A, c, g.sy) * V(c - g.sy);
  if (THREE_D) {
    if (k + 1 < g.nz) y += (double)A.az[c] * V(c + g.sz);
    if (k > 0) y += czm(A, c, g.sz) * V(c - g.sz);
  }
  return y;
}

// r = b - A x (or b when x0 = 0), z = r / D; partials r.z, z.z, (bn/D)^2. Iteration 0 reads z as
// p_prev (beta = 0), so no p_prev is cleared.
template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_init(Geo g, int kchunk, StencilView<T> A, const T* b,
                                              const T* bn, T* x, T* r, T* z,
                                              const unsigned char* fixed, int x0_zero,
                                              double* part, unsigned int nblocks) {
  __shared__ double lds[3 * (NT / kWave)];
  const Tile t = tile_of(g, kchunk);
  double srz = 0.0, szz = 0.0, sbb = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const bool fx = fixed != nullptr && fixed[c] != 0;
      double ax = 0.0;
      if (!x0_zero) ax = apply_at<T, THREE_D>(g, A, x, c);
      const double rv = fx ? 0.0 : (double)b[c] - ax;
      const T rT = (T)rv;
      const T zT = (T)((double)rT / (double)A.D[c]);
      r[c] = rT;
      z[c] = zT;
      srz += (double)rT * (double)zT;
      szz += (double)zT * (double)zT;
      const double bz = (double)(bn ? bn[c] : b[c]) / (double)A.D[c];
      sbb += bz * bz;
      // x0 = 0 (only when x0_zero, in which case x is not read at all). Without fixed rows x is
      // left as it is: the first update writes x = alpha p (x_fresh).
      if (x0_zero && fixed != nullptr && !fx) x[c] = (T)0;
    }
  }
  double v[3] = {srz, szz, sbb};
  block_sums<NT, 3>(v, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = v[0];
    part[nblocks + bid] = v[1];
    part[2 * nblocks + bid] = v[2];
  }
}

__global__ __launch_bounds__(kRedThreads) void k_fin_init(const double* part, unsigned int nblocks, KrylovCtl* ctl,
                                                          double rtol, double atol, int max_it, int guard,
                                                          int x0_zero) {
  __shared__ double lds[kRedThreads / kWave];
  const double rz = block_sum_partials<kRedThreads>(part, nblocks, lds);
  const double zz = block_sum_partials<kRedThreads>(part + nblocks, nblocks, lds);
  const double bb = block_sum_partials<kRedThreads>(part + 2 * nblocks, nblocks, lds);
  if (threadIdx.x == 0) {
    const double ttol = fmax(rtol * sqrt(bb), atol);
    ctl->rz = rz;
    ctl->rzb[0] = rz;
    ctl->rzb[1] = rz;
    ctl->zz = zz;
    ctl->bb = bb;
    ctl->ttol2 = ttol * ttol;
    ctl->iters = 0;
    ctl->beta = 0.0;
    ctl->alpha = 0.0;
    ctl->breakdown = 0;
    ctl->max_it = max_it;
    ctl->rr_guess = zz;
    ctl->guess_rejected = (guard && !x0_zero && !(zz < bb)) ? 1 : 0;
    ctl->converged = (zz <= ctl->ttol2) ? 1 : 0;
    ctl->done = (ctl->converged || ctl->guess_rejected) ? 1 : 0;
  }
}

// Start of CG iteration `it` in the SpMV. For it > 0, PETSc's convergence test on the residual
// of update it - 1 and beta = r.z / r.z_old, from that update's partials, which every block reduces
// itself in the same order (so all blocks agree bit for bit); the lead block records the outcome
// in ctl. False when the solve has ended (the block then does nothing). The done flag is read by
// one thread and shared through LDS, so a block's threads agree even while the lead block of the
// same launch is setting it (both reach the same decision).
template <int NTB>
__device__ inline bool cg_spmv_begin(KrylovCtl* ctl, const double* part_u, unsigned int nb_u, int it, bool lead,
                                     double* lds, int* sflag, double* beta) {
  if (threadIdx.x == 0) *sflag = ctl->done;
  __syncthreads();
  if (*sflag) return false;
  if (it == 0) {
    *beta = 0.0;
    return true;
  }
  const double rz = block_sum_partials<NTB>(part_u, nb_u, lds);
  const double zz = block_sum_partials<NTB>(part_u + nb_u, nb_u, lds);
  const bool conv = zz <= ctl->ttol2;
  const bool stop = conv || it >= ctl->max_it;
  if (lead && threadIdx.x == 0) {
    ctl->iters = it;
    ctl->zz = zz;
    ctl->rz = rz;
    ctl->rzb[it & 1] = rz;
    if (stop) {
      ctl->converged = conv ? 1 : 0;
      ctl->done = 1;
    }
  }
  *beta = rz / ctl->rzb[(it - 1) & 1];
  return !stop;
}

// Start of the CG update of iteration `it`: alpha = r.z / p.q from the SpMV's partials (reduced by
// every block); p.q <= 0 is a breakdown (recorded by the lead block; every block stops).
template <int NTB>
__device__ inline bool cg_update_begin(KrylovCtl* ctl, const double* part_s, unsigned int nb_s, int it, bool lead,
                                       double* lds, int* sflag, double* alpha) {
  if (threadIdx.x == 0) *sflag = ctl->done;
  __syncthreads();
  if (*sflag) return false;
  const double pq = block_sum_partials<NTB>(part_s, nb_s, lds);
  const bool brk = !(pq > 0.0) || !isfinite(pq);
  if (lead && threadIdx.x == 0) {
    ctl->pq = pq;
    if (brk) {
      ctl->breakdown = 1;
      ctl->done = 1;
    }
  }
  *alpha = ctl->rzb[it & 1] / pq;
  return !brk;
}

// Operands of one layer of the z-march: z and p_prev at the four lateral neighbours and the
// cell below, and the row's coefficients. The next layer's set is loaded before the current
// layer is computed, so each wave keeps a full layer of loads in flight.
template <typename T>
struct SpmvLayer {
  T zxm, pxm, zxp, pxp, zym, pym, zyp, pyp, zzp, pzp, d, axc, axm, ayc, aym, azc;
};

template <typename T, bool THREE_D>
__device__ __forceinline__ void spmv_load(const Geo& g, const StencilView<T>& A, const T* __restrict__ z,
                                          const T* __restrict__ pold, long long c, SpmvLayer<T>& v) {
  v.zxm = z[c - 1];
  v.pxm = pold[c - 1];
  v.zxp = z[c + 1];
  v.pxp = pold[c + 1];
  v.zym = z[c - g.sy];
  v.pym = pold[c - g.sy];
  v.zyp = z[c + g.sy];
  v.pyp = pold[c + g.sy];
  v.d = A.D[c];
  v.axc = A.ax[c];
  v.axm = A.ax[c - 1];
  v.ayc = A.ay[c];
  v.aym = A.ay[c - g.sy];
  if (THREE_D) {
    v.zzp = z[c + g.sz];
    v.pzp = pold[c + g.sz];
    v.azc = A.az[c];
  }
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_spmv(Geo g, int kchunk, StencilView<T> A, const T* __restrict__ z,
                                              const T* __restrict__ pold, T* __restrict__ pnew, T* __restrict__ q,
                                              double* part, KrylovCtl* ctl, const double* part_u, unsigned int nb_u,
                                              int it) {
  __shared__ double lds[NT / kWave];
  __shared__ int sflag;
  double beta;
  if (!cg_spmv_begin<NT>(ctl, part_u, nb_u, it, block_id() == 0, lds, &sflag, &beta)) return;
  const Tile t = tile_of(g, kchunk);
  double acc = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    auto P = [&](T zz, T pp) -> double { return (double)(T)((double)zz + beta * (double)pp); };
    double pm = 0.0, azm = 0.0;
    if (THREE_D) {
      pm = P(z[c - g.sz], pold[c - g.sz]);
      azm = (double)A.az[c - g.sz];
    }
    double pc = P(z[c], pold[c]);
    SpmvLayer<T> cur, nxt;
    spmv_load<T, THREE_D>(g, A, z, pold, c, cur);
    for (int k = t.k0; k < t.k1; ++k) {
      const bool more = THREE_D && k + 1 < t.k1;
      if (more) spmv_load<T, THREE_D>(g, A, z, pold, c + g.sz, nxt);
      const double pp = THREE_D ? P(cur.zzp, cur.pzp) : 0.0;
      double y = (double)cur.d * pc + (double)cur.axc * P(cur.zxp, cur.pxp) + (double)cur.axm * P(cur.zxm, cur.pxm) +
                 (double)cur.ayc * P(cur.zyp, cur.pyp) + (double)cur.aym * P(cur.zym, cur.pym);
      if (THREE_D) y += (double)cur.azc * pp + azm * pm;
      const T yT = (T)y;
      pnew[c] = (T)pc;
      q[c] = yT;
      acc += pc * (double)yT;
      pm = pc;
      pc = pp;
      azm = (double)cur.azc;
      c += g.sz;
      if (more) cur = nxt;
    }
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part[block_id()] = s;
}

// Wave-wide lane shifts (DPP wave_shr:1 / wave_shl:1): lane l receives the value of lane l - 1
// (shr) or l + 1 (shl); the lane without a source receives 0.
__device__ __forceinline__ double lane_from_left(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_right(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Two horizontally adjacent cells (i even), loaded as one aligned 2 x sizeof(T) access.
template <typename T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};
template <typename T>
__device__ __forceinline__ Pair<T> ld_pair(const T* __restrict__ v, long long c) {
  return *reinterpret_cast<const Pair<T>*>(v + c);
}

// k_init for x0 = 0 on a 3D operator without fixed rows (r = b, z = b / D, the same partials),
// on 128 x 4 tiles of cell pairs with kchunk layers per block: an elementwise pass with aligned
// paired loads and few partials. A pair straddling the row end handles its first cell only.
template <typename T>
__global__ __launch_bounds__(NT) void k_init_zero_pair(Geo g, int kchunk, const T* __restrict__ D,
                                                        const T* __restrict__ b, const T* __restrict__ bn,
                                                        T* __restrict__ r, T* __restrict__ z, double* part) {
  __shared__ double lds[NT / kWave];
  const TileId tb = xcd_tile();
  const int i = tb.x * 128 + 2 * (threadIdx.x & 63), j = tb.y * TY + (threadIdx.x >> 6);
  const int k0 = tb.z * kchunk, k1 = min(g.nz, k0 + kchunk);
  double srz = 0.0, szz = 0.0, sbb = 0.0;
  auto cell = [&](T bc, T bnc, T dc, T& ro, T& zo) {
    const T rT = (T)(double)bc;
    const T zT = (T)((double)rT / (double)dc);
    ro = rT;
    zo = zT;
    srz += (double)rT * (double)zT;
    szz += (double)zT * (double)zT;
    const double bz = (double)bnc / (double)dc;
    sbb += bz * bz;
  };
  if (i < g.nx && j < g.ny) {
    long long c = g.idx(i, j, k0);
    if (i + 1 < g.nx) {
      for (int k = k0; k < k1; ++k, c += g.sz) {
        const Pair<T> bv = ld_pair(b, c), dv = ld_pair(D, c);
        const Pair<T> bnv = bn ? ld_pair(bn, c) : bv;
        Pair<T> ro, zo;
        cell(bv.a, bnv.a, dv.a, ro.a, zo.a);
        cell(bv.b, bnv.b, dv.b, ro.b, zo.b);
        *reinterpret_cast<Pair<T>*>(r + c) = ro;
        *reinterpret_cast<Pair<T>*>(z + c) = zo;
      }
    } else {
      for (int k = k0; k < k1; ++k, c += g.sz) cell(b[c], bn ? bn[c] : b[c], D[c], r[c], z[c]);
    }
  }
  const double a = block_sum<NT>(srz, lds);
  const double bs = block_sum<NT>(szz, lds);
  const double cs = block_sum<NT>(sbb, lds);
  if (threadIdx.x == 0) {
    const unsigned int nbl = gridDim.x * gridDim.y * gridDim.z, bid = block_id();
    part[bid] = a;
    part[nbl + bid] = bs;
    part[2 * nbl + bid] = cs;
  }
}

// Operands of one layer of the paired z-march: pairs of z / p_prev at the y-neighbours and the
// cell below, the pair's coefficients, and (edge lanes only) z, p_prev, ax at the x-edge cell.
template <typename T>
struct SpmvPairLayer {
  Pair<T> zym, pym, zyp, pyp, zzp, pzp, d, ax, ay, aym, az;
  T ze, pe, axe;
};

template <typename T>
__device__ __forceinline__ void spmv_pair_load(const Geo& g, const StencilView<T>& A, const T* __restrict__ z,
                                               const T* __restrict__ pold, long long c, long long ce, bool edge,
                                               SpmvPairLayer<T>& v) {
  v.zym = ld_pair(z, c - g.sy);
  v.pym = ld_pair(pold, c - g.sy);
  v.zyp = ld_pair(z, c + g.sy);
  v.pyp = ld_pair(pold, c + g.sy);
  v.d = ld_pair(A.D, c);
  v.ax = ld_pair(A.ax, c);
  v.ay = ld_pair(A.ay, c);
  v.aym = ld_pair(A.ay, c - g.sy);
  v.zzp = ld_pair(z, c + g.sz);
  v.pzp = ld_pair(pold, c + g.sz);
  v.az = ld_pair(A.az, c);
  if (edge) {
    v.ze = z[ce];
    v.pe = pold[ce];
    v.axe = A.ax[ce];
  }
}

// CG SpMV, 3D, two cells per lane on 128 x 4 tiles (one row of 128 cells per wave). Rows start
// on 128-byte boundaries, so every full-wave load is one aligned 16-byte (FP64) or 8-byte (FP32)
// access per lane; the x-neighbours of a lane's pair come from the adjacent lanes by wave shifts
// (lane 0 loads cell i0 - 1, lane 63 cell i0 + 128), the z-neighbours from the registers of the
// z-march, the y-neighbours from L1/L2. Per cell the arithmetic is that of k_spmv. Lanes past the
// row end load the last in-row pair and store nothing.
template <typename T>
__global__ __launch_bounds__(NT) void k_spmv_pair(Geo g, int kchunk, StencilView<T> A, const T* __restrict__ z,
                                                   const T* __restrict__ pold, T* __restrict__ pnew,
                                                   T* __restrict__ q, double* part, KrylovCtl* ctl,
                                                   const double* part_u, unsigned int nb_u, int it) {
  __shared__ double lds[NT / kWave];
  __shared__ int sflag;
  const TileId b = xcd_tile();
  double beta;
  if (!cg_spmv_begin<NT>(ctl, part_u, nb_u, it, b.linear == 0, lds, &sflag, &beta)) return;
  const int lane = threadIdx.x & 63;
  const int i0 = b.x * 128, i = i0 + 2 * lane, j = b.y * TY + (threadIdx.x >> 6);
  const int k0 = b.z * kchunk, k1 = min(g.nz, k0 + kchunk);
  double acc = 0.0;
  if (j < g.ny) {  // whole wave
    const bool va = i < g.nx, vb = i + 1 < g.nx;
    long long c = g.idx(min(i, g.nx & ~1), j, k0);
    const bool edge = lane == 0 || lane == 63;
    const long long dce = (lane == 0 ? g.idx(i0 - 1, j, k0) : g.idx(min(i0 + 128, g.nx), j, k0)) - c;
    auto P = [&](T zz, T pp) -> double { return (double)(T)((double)zz + beta * (double)pp); };
    const Pair<T> zm = ld_pair(z, c - g.sz), pom = ld_pair(pold, c - g.sz), azm2 = ld_pair(A.az, c - g.sz);
    const Pair<T> z0 = ld_pair(z, c), p0 = ld_pair(pold, c);
    double pma = P(zm.a, pom.a), pmb = P(zm.b, pom.b);
    double azma = (double)azm2.a, azmb = (double)azm2.b;
    double pa = P(z0.a, p0.a), pb = P(z0.b, p0.b);
    SpmvPairLayer<T> cur, nxt;
    spmv_pair_load<T>(g, A, z, pold, c, c + dce, edge, cur);
    for (int k = k0; k < k1; ++k) {
      const bool more = k + 1 < k1;
      if (more) spmv_pair_load<T>(g, A, z, pold, c + g.sz, c + g.sz + dce, edge, nxt);
      const double ppa = P(cur.zzp.a, cur.pzp.a), ppb = P(cur.zzp.b, cur.pzp.b);
      const double pe = edge ? P(cur.ze, cur.pe) : 0.0;
      const double from_l = lane_from_left(pb), from_r = lane_from_right(pa);
      const double ax_l = lane_from_left((double)cur.ax.b);
      const double pxm = lane == 0 ? pe : from_l;              // p(i - 1)
      const double pxp = lane == 63 ? pe : from_r;             // p(i + 2)
      const double axm = lane == 0 ? (double)cur.axe : ax_l;   // ax(i - 1)
      double ya = (double)cur.d.a * pa + (double)cur.ax.a * pb + axm * pxm +
                  (double)cur.ay.a * P(cur.zyp.a, cur.pyp.a) + (double)cur.aym.a * P(cur.zym.a, cur.pym.a);
      ya += (double)cur.az.a * ppa + azma * pma;
      double yb = (double)cur.d.b * pb + (double)cur.ax.b * pxp + (double)cur.ax.a * pa +
                  (double)cur.ay.b * P(cur.zyp.b, cur.pyp.b) + (double)cur.aym.b * P(cur.zym.b, cur.pym.b);
      yb += (double)cur.az.b * ppb + azmb * pmb;
      const T ya_t = (T)ya, yb_t = (T)yb;
      if (vb) {
        *reinterpret_cast<Pair<T>*>(pnew + c) = Pair<T>{(T)pa, (T)pb};
        *reinterpret_cast<Pair<T>*>(q + c) = Pair<T>{ya_t, yb_t};
        acc += pa * (double)ya_t;
        acc += pb * (double)yb_t;
      } else if (va) {
        pnew[c] = (T)pa;
        q[c] = ya_t;
        acc += pa * (double)ya_t;
      }
      pma = pa;
      pmb = pb;
      pa = ppa;
      pb = ppb;
      azma = (double)cur.az.a;
      azmb = (double)cur.az.b;
      c += g.sz;
      if (more) cur = nxt;
    }
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part[block_id()] = s;
}

// CG update of iteration `it` on 128 x 4 tiles of cell pairs (paired aligned loads), kchunk
// layers per block (2D: one): x += alpha p, r -= alpha q, z = r / D, partials r.z and z.z to
// part_u[0..nb) and part_u[nb..2nb). A pair straddling the row end updates its first cell only.
// xfresh: x is taken as zero and not read (first update of a solve from x0 = 0).
template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_cg_update(Geo g, int kchunk, const T* __restrict__ D,
                                                   const T* __restrict__ p, const T* __restrict__ q,
                                                   T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                   double* __restrict__ part_u, KrylovCtl* ctl,
                                                   const double* __restrict__ part_s, unsigned int nb_s, int it,
                                                   int xfresh) {
  __shared__ double lds[2 * (NT / kWave)];
  __shared__ int sflag;
  const TileId b = xcd_tile();
  double alpha;
  if (!cg_update_begin<NT>(ctl, part_s, nb_s, it, b.linear == 0, lds, &sflag, &alpha)) return;
  const int i = b.x * 128 + 2 * (threadIdx.x & 63), j = b.y * TY + (threadIdx.x >> 6);
  const int k0 = THREE_D ? b.z * kchunk : 0, k1 = THREE_D ? min(g.nz, k0 + kchunk) : 1;
  double srz = 0.0, szz = 0.0;
  auto cell = [&](T xc, T pc, T rc, T qc, T dc, T& xo, T& ro, T& zo) {
    const double xv = (double)xc + alpha * (double)pc;
    const double rv = (double)rc - alpha * (double)qc;
    const T rT = (T)rv;
    const T zT = (T)((double)rT / (double)dc);
    xo = (T)xv;
    ro = rT;
    zo = zT;
    srz += (double)rT * (double)zT;
    szz += (double)zT * (double)zT;
  };
  if (i < g.nx && j < g.ny) {
    long long c = THREE_D ? g.idx(i, j, k0) : g.sidx(i, j);
    if (i + 1 < g.nx) {
#pragma unroll 2
      for (int k = k0; k < k1; ++k, c += g.sz) {
        const Pair<T> xv = xfresh ? Pair<T>{(T)0, (T)0} : ld_pair(x, c);
        const Pair<T> pv = ld_pair(p, c), rv = ld_pair(r, c), qv = ld_pair(q, c), dv = ld_pair(D, c);
        Pair<T> xo, ro, zo;
        cell(xv.a, pv.a, rv.a, qv.a, dv.a, xo.a, ro.a, zo.a);
        cell(xv.b, pv.b, rv.b, qv.b, dv.b, xo.b, ro.b, zo.b);
        *reinterpret_cast<Pair<T>*>(x + c) = xo;
        *reinterpret_cast<Pair<T>*>(r + c) = ro;
        *reinterpret_cast<Pair<T>*>(z + c) = zo;
      }
    } else {
      for (int k = k0; k < k1; ++k, c += g.sz) cell(xfresh ? (T)0 : x[c], p[c], r[c], q[c], D[c], x[c], r[c], z[c]);
    }
  }
  double v[2] = {srz, szz};
  block_sums<NT, 2>(v, lds);
  if (threadIdx.x == 0) {
    const unsigned int nb = gridDim.x * gridDim.y * gridDim.z;
    part_u[b.linear] = v[0];
    part_u[nb + b.linear] = v[1];
  }
}

// After the last update of a batch: the test the next SpMV would make (same reduction, same
// block size, so the same bits), so that the host sees the outcome of the batch.
__global__ __launch_bounds__(NT) void k_cg_check(KrylovCtl* ctl, const double* part_u, unsigned int nb_u, int it) {
  __shared__ double lds[NT / kWave];
  __shared__ int sflag;
  double beta;
  (void)cg_spmv_begin<NT>(ctl, part_u, nb_u, it, true, lds, &sflag, &beta);
}

// Iterative-refinement residual: xout = xin + d (d may be null; xout may be xin only then),
// r32 = b - A xout in FP64, partial ||D^-1 r||^2 (and ||D^-1 bn||^2 on the first sweep, which
// also sets the test).
template <bool THREE_D, bool SYM>
__global__ __launch_bounds__(NT) void k_refine(Geo g, int kchunk, StencilView<double> A,
                                                const double* b, const double* bn, const double* xin,
                                                const float* d, double* xout, float* r32,
                                                const unsigned char* fixed, double* part,
                                                unsigned int nblocks, int first) {
  __shared__ double lds[NT / kWave];
  const Tile t = tile_of(g, kchunk);
  double srr = 0.0, sbb = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    auto X = [&](long long cc) -> double { return xin[cc] + (d ? (double)d[cc] : 0.0); };
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const bool fx = fixed != nullptr && fixed[c] != 0;
      const double xc = X(c);
      double y;
      if (SYM) {
        y = A.D[c] * xc + A.ax[c] * X(c + 1) + A.ax[c - 1] * X(c - 1) + A.ay[c] * X(c + g.sy) +
            A.ay[c - g.sy] * X(c - g.sy);
        if (THREE_D) y += A.az[c] * X(c + g.sz) + A.az[c - g.sz] * X(c - g.sz);
      } else {
        y = apply_in_domain<double, THREE_D>(g, A, c, t.i, t.j, k, xc, X);
      }
      const double rv = fx ? 0.0 : b[c] - y;
      if (xout != xin) xout[c] = xc;
      r32[c] = (float)rv;
      const double rz = rv / A.D[c];
      srr += rz * rz;
      if (first) {
        const double bz = (bn ? bn[c] : b[c]) / A.D[c];
        sbb += bz * bz;
      }
    }
  }
  const double a = block_sum<NT>(srr, lds);
  const double bsum = block_sum<NT>(sbb, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = bsum;
  }
}

// Operands of one layer of the paired refinement march (see k_refine_pair).
struct RefinePairLayer {
  Pair<double> xzp, xym, xyp, d, ax, ay, aym, az, b;
  Pair<float> dzp, dym, dyp;
  double xe, axe;
  float de;
};

template <bool HAS_D>
__device__ __forceinline__ void refine_pair_load(const Geo& g, const StencilView<double>& A,
                                                 const double* __restrict__ b, const double* __restrict__ xin,
                                                 const float* __restrict__ d, long long c, long long ce, bool edge,
                                                 RefinePairLayer& v) {
  v.xzp = ld_pair(xin, c + g.sz);
  v.xym = ld_pair(xin, c - g.sy);
  v.xyp = ld_pair(xin, c + g.sy);
  v.d = ld_pair(A.D, c);
  v.ax = ld_pair(A.ax, c);
  v.ay = ld_pair(A.ay, c);
  v.aym = ld_pair(A.ay, c - g.sy);
  v.az = ld_pair(A.az, c);
  v.b = ld_pair(b, c);
  v.dzp = v.dym = v.dyp = Pair<float>{0.f, 0.f};
  if (HAS_D) {
    v.dzp = ld_pair(d, c + g.sz);
    v.dym = ld_pair(d, c - g.sy);
    v.dyp = ld_pair(d, c + g.sy);
  }
  v.xe = v.axe = 0.0;
  v.de = 0.f;
  if (edge) {
    v.xe = xin[ce];
    v.axe = A.ax[ce];
    if (HAS_D) v.de = d[ce];
  }
}

// k_refine for the 3D symmetric operator without fixed rows or a separate bn, on the 128 x 4 cell
// pair tiles and kchunk-layer z-march of k_spmv_pair (aligned 16-byte loads, x-neighbours by wave
// shifts, z-neighbours from registers). Per cell the arithmetic is that of k_refine. On the first
// sweep (no d) xout receives a copy of xin, so that the sweep after the first inner solve can write
// its iterate back into xin's buffer.
//
// With z32 non-null the sweep also starts the inner FP32 CG solve (k_init's work for b = r32,
// x0 = 0): z32 = r32 / D32 and the partials r.z, z.z and ||D32^-1 r32||^2 to part[2nb..5nb).
template <bool HAS_D>
__global__ __launch_bounds__(NT) void k_refine_pair(Geo g, int kchunk, StencilView<double> A,
                                                     const double* __restrict__ b, const double* __restrict__ xin,
                                                     const float* __restrict__ d, double* __restrict__ xout,
                                                     float* __restrict__ r32, const float* __restrict__ D32,
                                                     float* __restrict__ z32, double* part, int first) {
  __shared__ double lds[5 * (NT / kWave)];
  const TileId tb = xcd_tile();
  const int lane = threadIdx.x & 63;
  const int i0 = tb.x * 128, i = i0 + 2 * lane, j = tb.y * TY + (threadIdx.x >> 6);
  const int k0 = tb.z * kchunk, k1 = min(g.nz, k0 + kchunk);
  double srr = 0.0, sbb = 0.0, irz = 0.0, izz = 0.0, ibb = 0.0;
  const bool inner = z32 != nullptr;
  if (j < g.ny) {  // whole wave
    const bool va = i < g.nx, vb = i + 1 < g.nx;
    long long c = g.idx(min(i, g.nx & ~1), j, k0);
    const bool edge = lane == 0 || lane == 63;
    const long long dce = (lane == 0 ? g.idx(i0 - 1, j, k0) : g.idx(min(i0 + 128, g.nx), j, k0)) - c;
    // k_init's arithmetic for one cell of the inner solve (b = r32, x0 = 0)
    auto start = [&](float rT, float dd, float& zo) {
      const float zT = (float)((double)rT / (double)dd);
      zo = zT;
      irz += (double)rT * (double)zT;
      izz += (double)zT * (double)zT;
      const double bz = (double)rT / (double)dd;
      ibb += bz * bz;
    };
    auto add = [](Pair<double> x, Pair<float> e) -> Pair<double> {
      return HAS_D ? Pair<double>{x.a + (double)e.a, x.b + (double)e.b} : x;
    };
    Pair<float> dz0{0.f, 0.f}, dzm{0.f, 0.f};
    if (HAS_D) {
      dz0 = ld_pair(d, c);
      dzm = ld_pair(d, c - g.sz);
    }
    const Pair<double> xm = add(ld_pair(xin, c - g.sz), dzm);
    const Pair<double> azm2 = ld_pair(A.az, c - g.sz);
    const Pair<double> x0 = add(ld_pair(xin, c), dz0);
    double xma = xm.a, xmb = xm.b, azma = azm2.a, azmb = azm2.b, xa = x0.a, xb = x0.b;
    RefinePairLayer cur, nxt;
    refine_pair_load<HAS_D>(g, A, b, xin, d, c, c + dce, edge, cur);
    for (int k = k0; k < k1; ++k) {
      const bool more = k + 1 < k1;
      if (more) refine_pair_load<HAS_D>(g, A, b, xin, d, c + g.sz, c + g.sz + dce, edge, nxt);
      const Pair<double> xp = add(cur.xzp, cur.dzp), xym = add(cur.xym, cur.dym), xyp = add(cur.xyp, cur.dyp);
      const double xe = edge ? cur.xe + (HAS_D ? (double)cur.de : 0.0) : 0.0;
      const double from_l = lane_from_left(xb), from_r = lane_from_right(xa);
      const double ax_l = lane_from_left(cur.ax.b);
      const double xxm = lane == 0 ? xe : from_l;         // X(i - 1)
      const double xxp = lane == 63 ? xe : from_r;        // X(i + 2)
      const double axm = lane == 0 ? cur.axe : ax_l;      // ax(i - 1)
      double ya = cur.d.a * xa + cur.ax.a * xb + axm * xxm + cur.ay.a * xyp.a + cur.aym.a * xym.a;
      ya += cur.az.a * xp.a + azma * xma;
      double yb = cur.d.b * xb + cur.ax.b * xxp + cur.ax.a * xa + cur.ay.b * xyp.b + cur.aym.b * xym.b;
      yb += cur.az.b * xp.b + azmb * xmb;
      const double ra = cur.b.a - ya, rb = cur.b.b - yb;
      const double za = ra / cur.d.a, zb = rb / cur.d.b;
      const double ba = cur.b.a / cur.d.a, bbz = cur.b.b / cur.d.b;
      if (vb) {
        const Pair<float> rT{(float)ra, (float)rb};
        *reinterpret_cast<Pair<double>*>(xout + c) = Pair<double>{xa, xb};
        *reinterpret_cast<Pair<float>*>(r32 + c) = rT;
        srr += za * za;
        srr += zb * zb;
        if (first) {
          sbb += ba * ba;
          sbb += bbz * bbz;
        }
        if (inner) {
          const Pair<float> dd = ld_pair(D32, c);
          Pair<float> zo;
          start(rT.a, dd.a, zo.a);
          start(rT.b, dd.b, zo.b);
          *reinterpret_cast<Pair<float>*>(z32 + c) = zo;
        }
      } else if (va) {
        xout[c] = xa;
        r32[c] = (float)ra;
        srr += za * za;
        if (first) sbb += ba * ba;
        if (inner) start((float)ra, D32[c], z32[c]);
      }
      xma = xa;
      xmb = xb;
      xa = xp.a;
      xb = xp.b;
      azma = cur.az.a;
      azmb = cur.az.b;
      c += g.sz;
      if (more) cur = nxt;
    }
  }
  const unsigned int bid = block_id(), nbl = gridDim.x * gridDim.y * gridDim.z;
  if (inner) {
    double v[5] = {srr, sbb, irz, izz, ibb};
    block_sums<NT, 5>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 5; ++q) part[q * nbl + bid] = v[q];
    }
  } else {
    double v[2] = {srr, sbb};
    block_sums<NT, 2>(v, lds);
    if (threadIdx.x == 0) {
      part[bid] = v[0];
      part[nbl + bid] = v[1];
    }
  }
}

// Outer test of a refinement sweep; with `inner` non-null also k_fin_init's start of the inner CG
// solve from the sweep's partials, its tolerance relative to ||D32^-1 r32|| being the one the
// refinement loop asks for: min(0.5, max(inner_rtol, gap / 2)), gap = outer target / residual.
__global__ __launch_bounds__(kRedThreads) void k_fin_refine(const double* part, unsigned int nblocks, KrylovCtl* ctl,
                                                            int first, double rtol, double atol, int guard,
                                                            KrylovCtl* inner, double inner_rtol, int inner_max_it) {
  __shared__ double lds[kRedThreads / kWave];
  const double rr = block_sum_partials<kRedThreads>(part, nblocks, lds);
  const double bb = first ? block_sum_partials<kRedThreads>(part + nblocks, nblocks, lds) : 0.0;
  double irz = 0.0, izz = 0.0, ibb = 0.0;
  if (inner) {
    irz = block_sum_partials<kRedThreads>(part + 2 * nblocks, nblocks, lds);
    izz = block_sum_partials<kRedThreads>(part + 3 * nblocks, nblocks, lds);
    ibb = block_sum_partials<kRedThreads>(part + 4 * nblocks, nblocks, lds);
  }
  if (threadIdx.x == 0) {
    if (first) {
      const double ttol = fmax(rtol * sqrt(bb), atol);
      ctl->bb = bb;
      ctl->ttol2 = ttol * ttol;
      ctl->rr_guess = rr;
      ctl->guess_rejected = (guard && !(rr < bb)) ? 1 : 0;
    }
    ctl->zz = rr;
    ctl->converged = (rr <= ctl->ttol2) ? 1 : 0;
    ctl->done = ctl->converged;
    if (inner) {
      const double gap = sqrt(ctl->ttol2) / fmax(sqrt(rr), 1e-300);
      const double irt = fmin(0.5, fmax(inner_rtol, 0.5 * gap));
      const double ttol = fmax(irt * sqrt(ibb), 0.0);
      inner->rz = irz;
      inner->rzb[0] = irz;
      inner->rzb[1] = irz;
      inner->zz = izz;
      inner->bb = ibb;
      inner->ttol2 = ttol * ttol;
      inner->iters = 0;
      inner->beta = 0.0;
      inner->alpha = 0.0;
      inner->breakdown = 0;
      inner->max_it = inner_max_it;
      inner->rr_guess = izz;
      inner->guess_rejected = 0;
      inner->converged = (izz <= inner->ttol2) ? 1 : 0;
      inner->done = inner->converged;
    }
  }
}

// ---- BiCGSTAB ---------------------------------------------------------------------------

// r = b - A x (or b when x0 = 0), rh = r, p_prev = v_prev = 0; partials rh.r, ||D^-1 r||^2,
// ||D^-1 bn||^2.
template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_bcg_init(Geo g, int kchunk, StencilView<T> A, const T* b, const T* bn,
                                                  T* x, T* r, T* rh, T* p, T* v, const unsigned char* fixed,
                                                  int x0_zero, double* part, unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  const Tile t = tile_of(g, kchunk);
  double srr = 0.0, szz = 0.0, sbb = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const bool fx = fixed != nullptr && fixed[c] != 0;
      double ax = 0.0;
      if (!x0_zero)
        ax = apply_in_domain<T, THREE_D>(g, A, c, t.i, t.j, k, (double)x[c],
                                         [&](long long cc) { return (double)x[cc]; });
      const T rT = (T)(fx ? 0.0 : (double)b[c] - ax);
      r[c] = rT;
      rh[c] = rT;
      p[c] = (T)0;
      v[c] = (T)0;
      const double zr = (double)rT / (double)A.D[c];
      srr += (double)rT * (double)rT;
      szz += zr * zr;
      const double bz = (double)(bn ? bn[c] : b[c]) / (double)A.D[c];
      sbb += bz * bz;
    }
  }
  if (x0_zero && t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz)
      if (!(fixed != nullptr && fixed[c] != 0)) x[c] = (T)0;
  }
  const double a = block_sum<NT>(srr, lds);
  const double bsum = block_sum<NT>(szz, lds);
  const double cs = block_sum<NT>(sbb, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = bsum;
    part[2 * nblocks + bid] = cs;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_fin_bcg_init(const double* part, unsigned int nblocks,
                                                              KrylovCtl* ctl, double rtol, double atol, int max_it,
                                                              int guard, int x0_zero) {
  __shared__ double lds[kRedThreads / kWave];
  const double rho = block_sum_partials<kRedThreads>(part, nblocks, lds);
  const double zz = block_sum_partials<kRedThreads>(part + nblocks, nblocks, lds);
  const double bb = block_sum_partials<kRedThreads>(part + 2 * nblocks, nblocks, lds);
  if (threadIdx.x == 0) {
    const double ttol = fmax(rtol * sqrt(bb), atol);
    ctl->rz = rho;
    ctl->zz = zz;
    ctl->bb = bb;
    ctl->ttol2 = ttol * ttol;
    ctl->iters = 0;
    ctl->beta = 0.0;  // p = r on the first iteration (p_prev = v_prev = 0)
    ctl->alpha = 1.0;
    ctl->omega = 1.0;
    ctl->pq = 0.0;
    ctl->half = 0;
    ctl->breakdown = 0;
    ctl->max_it = max_it;
    ctl->rr_guess = zz;
    ctl->guess_rejected = (guard && !x0_zero && !(zz < bb)) ? 1 : 0;
    ctl->converged = (zz <= ctl->ttol2) ? 1 : 0;
    ctl->done = (ctl->converged || ctl->guess_rejected) ? 1 : 0;
    if (!ctl->done && (rho == 0.0 || !isfinite(rho))) {
      ctl->breakdown = 3;
      ctl->done = 1;
    }
  }
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_bcg_spmv1(Geo g, int kchunk, StencilView<T> A, const T* r, const T* rh,
                                                   const T* pold, const T* vold, T* pnew, T* vnew, double* part,
                                                   const KrylovCtl* ctl) {
  __shared__ double lds[NT / kWave];
  if (ctl->done) return;
  const double beta = ctl->beta, omega = ctl->omega;
  const Tile t = tile_of(g, kchunk);
  double acc = 0.0;
  if (t.valid) {
    auto P = [&](long long cc) -> T {
      return (T)((double)r[cc] + beta * ((double)pold[cc] - omega * (double)vold[cc]));
    };
    auto Y = [&](long long cc) -> double { return (double)(T)((double)P(cc) / (double)A.D[cc]); };
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const T pc = P(c);
      const double yc = (double)(T)((double)pc / (double)A.D[c]);
      const T vT = (T)apply_in_domain<T, THREE_D>(g, A, c, t.i, t.j, k, yc, Y);
      pnew[c] = pc;
      vnew[c] = vT;
      acc += (double)rh[c] * (double)vT;
    }
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part[block_id()] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_fin_bcg1(const double* part, unsigned int nblocks, KrylovCtl* ctl) {
  __shared__ double lds[kRedThreads / kWave];
  if (ctl->done) return;
  const double rv = block_sum_partials<kRedThreads>(part, nblocks, lds);
  if (threadIdx.x == 0) {
    ctl->pq = rv;
    if (rv == 0.0 || !isfinite(rv)) {
      ctl->breakdown = 1;
      ctl->done = 1;
    } else {
      ctl->alpha = ctl->rz / rv;
    }
  }
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_bcg_spmv2(Geo g, int kchunk, StencilView<T> A, const T* r, const T* v,
                                                   T* s, T* tv, double* part, const KrylovCtl* ctl,
                                                   unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  if (ctl->done) return;
  const double alpha = ctl->alpha;
  const Tile t = tile_of(g, kchunk);
  double sts = 0.0, stt = 0.0, sss = 0.0;
  if (t.valid) {
    auto S = [&](long long cc) -> T { return (T)((double)r[cc] - alpha * (double)v[cc]); };
    auto Z = [&](long long cc) -> double { return (double)(T)((double)S(cc) / (double)A.D[cc]); };
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const T sc = S(c);
      const double zc = (double)(T)((double)sc / (double)A.D[c]);
      const T tT = (T)apply_in_domain<T, THREE_D>(g, A, c, t.i, t.j, k, zc, Z);
      s[c] = sc;
      tv[c] = tT;
      sts += (double)tT * (double)sc;
      stt += (double)tT * (double)tT;
      const double zn = (double)sc / (double)A.D[c];
      sss += zn * zn;
    }
  }
  const double a = block_sum<NT>(sts, lds);
  const double b = block_sum<NT>(stt, lds);
  const double cc = block_sum<NT>(sss, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = b;
    part[2 * nblocks + bid] = cc;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_fin_bcg2(const double* part, unsigned int nblocks, KrylovCtl* ctl) {
  __shared__ double lds[kRedThreads / kWave];
  if (ctl->done) return;
  const double ts = block_sum_partials<kRedThreads>(part, nblocks, lds);
  const double tt = block_sum_partials<kRedThreads>(part + nblocks, nblocks, lds);
  const double ss = block_sum_partials<kRedThreads>(part + 2 * nblocks, nblocks, lds);
  if (threadIdx.x == 0) {
    ctl->iters = ctl->iters + 1;
    if (ss <= ctl->ttol2) {
      ctl->half = 1;  // x += alpha y, r = s
      ctl->omega = 0.0;
    } else {
      ctl->omega = tt > 0.0 ? ts / tt : 0.0;
    }
  }
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_bcg_update(Geo g, int kchunk, const T* D, const T* p, const T* s,
                                                    const T* tv, const T* rh, T* x, T* r, double* part,
                                                    const KrylovCtl* ctl, unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  if (ctl->done) return;
  const double alpha = ctl->alpha, omega = ctl->omega;
  const Tile t = tile_of(g, kchunk);
  double srr = 0.0, szz = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const double d = (double)D[c];
      const double yc = (double)(T)((double)p[c] / d);
      const double zc = (double)(T)((double)s[c] / d);
      const T xT = (T)((double)x[c] + (alpha * yc + omega * zc));
      const T rT = (T)((double)s[c] - omega * (double)tv[c]);
      x[c] = xT;
      r[c] = rT;
      srr += (double)rh[c] * (double)rT;
      const double zr = (double)rT / d;
      szz += zr * zr;
    }
  }
  const double a = block_sum<NT>(srr, lds);
  const double b = block_sum<NT>(szz, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = b;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_fin_bcg3(const double* part, unsigned int nblocks, KrylovCtl* ctl) {
  __shared__ double lds[kRedThreads / kWave];
  if (ctl->done) return;
  const double rho = block_sum_partials<kRedThreads>(part, nblocks, lds);
  const double rr = block_sum_partials<kRedThreads>(part + nblocks, nblocks, lds);
  if (threadIdx.x == 0) {
    ctl->zz = rr;
    if (ctl->half || rr <= ctl->ttol2) {
      ctl->converged = 1;
      ctl->done = 1;
    } else if (ctl->omega == 0.0) {
      ctl->breakdown = 2;
      ctl->done = 1;
    } else if (ctl->iters >= ctl->max_it) {
      ctl->done = 1;
    } else if (rho == 0.0 || !isfinite(rho)) {
      ctl->breakdown = 3;
      ctl->done = 1;
    } else {
      ctl->beta = (rho / ctl->rz) * (ctl->alpha / ctl->omega);
      ctl->rz = rho;
    }
  }
}

// x = 0 on non-fixed interior rows (rejected initial guess).
template <bool THREE_D, typename T>
__global__ __launch_bounds__(NT) void k_zero_interior(Geo g, int kchunk, T* x,
                                                       const unsigned char* fixed) {
  const Tile t = tile_of(g, kchunk);
  if (!t.valid) return;
  long long c = base_index<THREE_D>(g, t);
  for (int k = t.k0; k < t.k1; ++k, c += g.sz)
    if (!(fixed != nullptr && fixed[c] != 0)) x[c] = (T)0;
}

}  // namespace

Krylov::Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof)
    : g_(g), three_d_(three_d), s_(stream), prof_(prof) {
  try {
    build();
  } catch (...) {
    release();  // a failed set-up frees what it allocated
    throw;
  }
}

Krylov::Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof, const KrylovWorkspace& ws)
    : g_(g), three_d_(three_d), s_(stream), prof_(prof), borrowed_(true) {
  setup_shapes();
  r_ = ws.r;
  z_ = ws.z;
  p0_ = ws.p0;
  p1_ = ws.p1;
  q_ = ws.q;
  part_ = ws.part;
  part_u_ = ws.part_u;
  ctl_ = ws.ctl;
  host_ctl_ = &host_ctl_own_;
  // no autotuning (it would time launches with events): two columns per lane, whole columns per
  // block in 3D, unless FREHG_SPMV_SHAPE pins a shape
  if (!tuned_[0]) {
    for (int prec = 0; prec < 2; ++prec) {
      set_spmv_shape(prec, three_d_ ? 2 : 1, three_d_ ? g_.nz : 1);
      tuned_[prec] = true;
    }
  }
}

namespace {
inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

size_t Krylov::workspace_bytes(const Geo& g0, bool three_d) {
  Geo g = g0;
  if (!three_d) g.nz = 1;
  const long long n = three_d ? g.ncell() : g.nsurf();
  const size_t cap = ceil_div(g.nx, TX) * ceil_div(g.ny, TY) * (size_t)(three_d ? g.nz : 1);
  return 5 * align256((size_t)n * sizeof(double)) + align256(5 * cap * sizeof(double)) +
         align256(2 * cap * sizeof(double)) + align256(3 * sizeof(KrylovCtl));
}

KrylovWorkspace Krylov::carve_workspace(void* base, const Geo& g0, bool three_d) {
  Geo g = g0;
  if (!three_d) g.nz = 1;
  const long long n = three_d ? g.ncell() : g.nsurf();
  const size_t cap = ceil_div(g.nx, TX) * ceil_div(g.ny, TY) * (size_t)(three_d ? g.nz : 1);
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += align256(bytes);
    return q;
  };
  KrylovWorkspace w;
  w.r = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  w.z = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  w.p0 = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  w.p1 = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  w.q = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  w.part = reinterpret_cast<double*>(take(5 * cap * sizeof(double)));
  w.part_u = reinterpret_cast<double*>(take(2 * cap * sizeof(double)));
  w.ctl = reinterpret_cast<KrylovCtl*>(take(3 * sizeof(KrylovCtl)));
  w.n = n;
  return w;
}

void Krylov::setup_shapes() {
  if (!three_d_) g_.nz = 1;
  n_ = three_d_ ? g_.ncell() : g_.nsurf();
  kc_f_ = 1;
  grid_f_ = dim3(ceil_div(g_.nx, TX), ceil_div(g_.ny, TY), ceil_div(three_d_ ? g_.nz : 1, kc_f_));
  nb_f_ = grid_f_.x * grid_f_.y * grid_f_.z;
  // partials sized for the largest grid any kernel (or candidate) uses: 64 x 4 tiles, one layer each
  part_cap_ = ceil_div(g_.nx, TX) * ceil_div(g_.ny, TY) * (unsigned int)(three_d_ ? g_.nz : 1);
  // CG update: whole columns per block in 3D (its partials are reduced by every SpMV block, so
  // their number is kept small)
  kc_u_ = three_d_ ? g_.nz : 1;
  if (const char* env = std::getenv("FREHG_CG_UPDATE_KC")) kc_u_ = std::max(1, std::min(three_d_ ? g_.nz : 1, std::atoi(env)));
  grid_u_ = dim3(ceil_div(g_.nx, 2 * TX), ceil_div(g_.ny, TY), three_d_ ? ceil_div(g_.nz, kc_u_) : 1);
  nb_u_ = grid_u_.x * grid_u_.y * grid_u_.z;
  for (int prec = 0; prec < 2; ++prec) set_spmv_shape(prec, 1, three_d_ ? std::min(g_.nz, 4) : 1);
  // FREHG_SPMV_SHAPE="cpl:kc" pins the SpMV shape (cells per lane 1 or 2, kc layers per block)
  // and skips the autotuner.
  if (const char* env = std::getenv("FREHG_SPMV_SHAPE")) {
    int cpl = -1, kc = -1;
    if (std::sscanf(env, "%d:%d", &cpl, &kc) != 2 || !(cpl == 1 || cpl == 2) || kc < 1)
      throw ArgError(std::string("FREHG_SPMV_SHAPE must be \"cpl:kc\" with cpl in {1,2} and kc >= 1, got \"") +
                     env + "\"");
    for (int prec = 0; prec < 2; ++prec) {
      set_spmv_shape(prec, cpl, kc);
      tuned_[prec] = true;
    }
  }
}

void Krylov::build() {
  setup_shapes();
  FREHG_HIP(hipMalloc(&part_, 5 * (size_t)part_cap_ * sizeof(double)));  // up to 5 sums per block
  FREHG_HIP(hipMalloc(&part_u_, 2 * (size_t)part_cap_ * sizeof(double)));
  const size_t bytes = (size_t)n_ * sizeof(double);
  for (double** p : {&r_, &z_, &p0_, &p1_, &q_, &xw_}) {
    FREHG_HIP(hipMalloc(p, bytes));
    FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
  }
  FREHG_HIP(hipMalloc(&ctl_, 3 * sizeof(KrylovCtl)));
  FREHG_HIP(hipMemsetAsync(ctl_, 0, 3 * sizeof(KrylovCtl), s_));
  FREHG_HIP(hipHostMalloc(&host_ctl_, sizeof(KrylovCtl), hipHostMallocDefault));
  FREHG_HIP(hipStreamSynchronize(s_));
}

Krylov::~Krylov() { release(); }

void Krylov::release() {
  if (borrowed_) return;  // the workspace belongs to the caller
  for (void* p : {(void*)r_, (void*)z_, (void*)p0_, (void*)p1_, (void*)q_, (void*)xw_, (void*)b32_,
                  (void*)r32_, (void*)z32_, (void*)p032_, (void*)p132_, (void*)q32_, (void*)d32_,
                  (void*)rh_, (void*)v0_, (void*)v1_, (void*)sv_, (void*)tv_, (void*)rh32_, (void*)v032_,
                  (void*)v132_, (void*)sv32_, (void*)tv32_, (void*)part_, (void*)part_u_, (void*)ctl_,
                  (void*)stage_})
    if (p) (void)hipFree(p);
  if (host_ctl_) (void)hipHostFree(host_ctl_);
}

void Krylov::set_spmv_shape(int prec, int cpl, int kc) {
  const int nz = three_d_ ? g_.nz : 1;
  if (!three_d_) cpl = 1;
  if (cpl != 1 && cpl != 2) throw ArgError("Krylov: SpMV cells per lane must be 1 or 2");
  SpmvShape& sh = shape_[prec];
  sh.cpl = cpl;
  sh.kc = std::max(1, std::min(kc, nz));
  sh.grid = dim3(ceil_div(g_.nx, TX * cpl), ceil_div(g_.ny, TY), ceil_div(nz, sh.kc));
  sh.nb = sh.grid.x * sh.grid.y * sh.grid.z;
  if (sh.nb > part_cap_) throw HipError("Krylov: SpMV shape needs more partials than allocated");
}

template <typename T>
void Krylov::launch_spmv(const StencilView<T>& A, const T* z, const T* pold, T* pnew, T* q, KrylovCtl* ctl, int it,
                         hipEvent_t ev0, hipEvent_t ev1) {
  const SpmvShape& sh = shape_[sizeof(T) == 4 ? 1 : 0];
  const double* pu = part_u_;
  if (!three_d_) {
    hipExtLaunchKernelGGL((k_spmv<T, false>), sh.grid, dim3(NT), 0, s_, ev0, ev1, 0, g_, sh.kc, A, z, pold, pnew, q,
                          part_, ctl, pu, nb_u_, it);
  } else if (sh.cpl == 1) {
    hipExtLaunchKernelGGL((k_spmv<T, true>), sh.grid, dim3(NT), 0, s_, ev0, ev1, 0, g_, sh.kc, A, z, pold, pnew, q,
                          part_, ctl, pu, nb_u_, it);
  } else {
    hipExtLaunchKernelGGL((k_spmv_pair<T>), sh.grid, dim3(NT), 0, s_, ev0, ev1, 0, g_, sh.kc, A, z, pold, pnew, q,
                          part_, ctl, pu, nb_u_, it);
  }
  FREHG_LAUNCH_CHECK();
}

template <typename T>
void Krylov::autotune(const StencilView<T>& A) {
  const int prec = sizeof(T) == 4 ? 1 : 0;
  tuned_[prec] = true;
  if (borrowed_) return;
  if (!three_d_) return;
  if (prec == 1) ensure_mixed();
  KrylovCtl* tctl = ctl_ + 2;
  FREHG_HIP(hipMemsetAsync(tctl, 0, sizeof(KrylovCtl), s_));
  T* p0 = prec ? (T*)p032_ : (T*)p0_;
  T* p1 = prec ? (T*)p132_ : (T*)p1_;
  T* qq = prec ? (T*)q32_ : (T*)q_;
  hipEvent_t e0, e1;
  FREHG_HIP(hipEventCreate(&e0));
  FREHG_HIP(hipEventCreate(&e1));
  const int nz = g_.nz;
  std::vector<int> kcs;
  for (int kc : {1, 2, 4, 8, 15, nz})
    if (kc >= 1 && kc <= nz && std::find(kcs.begin(), kcs.end(), kc) == kcs.end()) kcs.push_back(kc);
  int best_cpl = shape_[prec].cpl, best_kc = shape_[prec].kc;
  float best_ms = 1e30f;
  for (int cpl : {1, 2}) {
    for (int kc : kcs) {
      set_spmv_shape(prec, cpl, kc);
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {  // the first round warms caches and clocks
        FREHG_HIP(hipEventRecord(e0, s_));
        for (int it = 0; it < 3; ++it) launch_spmv<T>(A, A.D, p0, p1, qq, tctl, 0, nullptr, nullptr);
        FREHG_HIP(hipEventRecord(e1, s_));
        FREHG_HIP(hipEventSynchronize(e1));
        FREHG_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_cpl = cpl;
        best_kc = kc;
      }
    }
  }
  set_spmv_shape(prec, best_cpl, best_kc);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

Krylov::SweepTarget Krylov::first_sweep_target() {
  ensure_mixed();
  return SweepTarget{r32_, z32_, xw_, part_, part_cap_};
}

void Krylov::ensure_mixed() {
  if (b32_) return;
  if (borrowed_) throw ArgError("Krylov: a solve on a borrowed workspace is FP64 CG only");
  const size_t bytes = (size_t)n_ * sizeof(float);
  for (float** p : {&b32_, &r32_, &z32_, &p032_, &p132_, &q32_, &d32_}) {
    FREHG_HIP(hipMalloc(p, bytes));
    FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
  }
  FREHG_HIP(hipMalloc(&stage_, 5 * kRedStage * sizeof(double)));
}

void Krylov::ensure_bicgstab(bool fp32) {
  if (borrowed_) throw ArgError("Krylov: a solve on a borrowed workspace is FP64 CG only");
  if (fp32) {
    if (rh32_) return;
    const size_t bytes = (size_t)n_ * sizeof(float);
    for (float** p : {&rh32_, &v032_, &v132_, &sv32_, &tv32_}) {
      FREHG_HIP(hipMalloc(p, bytes));
      FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
    }
  } else {
    if (rh_) return;
    const size_t bytes = (size_t)n_ * sizeof(double);
    for (double** p : {&rh_, &v0_, &v1_, &sv_, &tv_}) {
      FREHG_HIP(hipMalloc(p, bytes));
      FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
    }
  }
}

void Krylov::read_ctl(KrylovCtl* dev, KrylovCtl* host) {
  FREHG_HIP(hipMemcpyAsync(host, dev, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Krylov::cg_start(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed, bool zero,
                      bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl, T* r, T* z, T* p0) {
  const dim3 blk(NT);
  unsigned int nb = nb_f_;
  if (three_d_ && zero && fixed == nullptr) {
    // x0 = 0 without fixed rows: elementwise, on the SpMV's tiles and layers per block
    const int kc = shape_[sizeof(T) == 4 ? 1 : 0].kc;
    const dim3 grid(ceil_div(g_.nx, 2 * TX), ceil_div(g_.ny, TY), ceil_div(g_.nz, kc));
    nb = grid.x * grid.y * grid.z;
    hipLaunchKernelGGL((k_init_zero_pair<T>), grid, blk, 0, s_, g_, kc, A.D, b, bn, r, z, part_);
  } else if (three_d_)
    hipLaunchKernelGGL((k_init<T, true>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, x, r, z, fixed, zero ? 1 : 0,
                       part_, nb);
  else
    hipLaunchKernelGGL((k_init<T, false>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, x, r, z, fixed, zero ? 1 : 0,
                       part_, nb);
  FREHG_LAUNCH_CHECK();
  (void)p0;
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(kRedThreads), 0, s_, part_, nb, ctl, rtol, atol, max_it,
                     guard ? 1 : 0, zero ? 1 : 0);
  FREHG_LAUNCH_CHECK();
}

template <typename T>
void Krylov::zero_interior(T* x, const unsigned char* fixed) {
  if (three_d_) hipLaunchKernelGGL((k_zero_interior<true, T>), grid_f_, dim3(NT), 0, s_, g_, kc_f_, x, fixed);
  else hipLaunchKernelGGL((k_zero_interior<false, T>), grid_f_, dim3(NT), 0, s_, g_, kc_f_, x, fixed);
  FREHG_LAUNCH_CHECK();
}

// CG iterations from .. from + count - 1 (p_prev of iteration it is p[it & 1]), then the test of
// the last update (k_cg_check), all without host synchronisation.
template <typename T>
void Krylov::cg_iterations(const StencilView<T>& A, T* x, T* r, T* z, T* p0, T* p1, T* q, KrylovCtl* ctl, int from,
                           int count, int kt_spmv, bool x_fresh) {
  const dim3 blk(NT);
  T* pbuf[2] = {p0, p1};
  for (int it = from; it < from + count; ++it) {
    // iteration 0 has beta = 0: p = z + 0 p_prev, with z itself as p_prev (no cleared buffer read)
    T* pold = it == 0 ? z : pbuf[it & 1];
    T* pnew = pbuf[(it & 1) ^ 1];
    // Only the 3D (groundwater) launches are timed: the 2D surface solve is launch-bound, and
    // launches with timing events take the host several microseconds longer to issue.
    {
      hipEvent_t ev0 = three_d_ ? prof_->start(kt_spmv) : nullptr;
      hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
      launch_spmv<T>(A, z, pold, pnew, q, ctl, it, ev0, ev1);
    }
    {
      hipEvent_t ev0 = three_d_ ? prof_->start(FREHG_KT_GW_UPDATE) : nullptr;
      hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
      const unsigned int nbs = shape_[sizeof(T) == 4 ? 1 : 0].nb;
      const int xf = (it == 0 && x_fresh) ? 1 : 0;
      if (three_d_)
        hipExtLaunchKernelGGL((k_cg_update<T, true>), grid_u_, blk, 0, s_, ev0, ev1, 0, g_, kc_u_, A.D, (const T*)pnew,
                              (const T*)q, x, r, z, part_u_, ctl, (const double*)part_, nbs, it, xf);
      else
        hipExtLaunchKernelGGL((k_cg_update<T, false>), grid_u_, blk, 0, s_, ev0, ev1, 0, g_, kc_u_, A.D,
                              (const T*)pnew, (const T*)q, x, r, z, part_u_, ctl, (const double*)part_, nbs, it, xf);
      FREHG_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(NT), 0, s_, ctl, (const double*)part_u_, nb_u_, from + count);
  FREHG_LAUNCH_CHECK();
}

// Host loop after `launched` queued iterations: wait, and queue further batches until the solve
// has ended; then the outcome (throws on breakdown).
template <typename T>
int Krylov::cg_finish(const StencilView<T>& A, T* x, T* r, T* z, T* p0, T* p1, T* q, KrylovCtl* ctl, int max_it,
                      int launched, int batch, double* res_norm, double* rhs_norm, bool* converged, int kt_spmv,
                      bool x_fresh) {
  for (;;) {
    read_ctl(ctl, host_ctl_);
    if (host_ctl_->done) break;
    if (launched >= max_it + 2) break;
    batch = std::max(2, std::min(16, batch / 2));
    cg_iterations<T>(A, x, r, z, p0, p1, q, ctl, launched, batch, kt_spmv, x_fresh);
    launched += batch;
  }
  // a solve from an uncleared x0 = 0 that ended before its first update
  if (x_fresh && host_ctl_->iters == 0) zero_interior<T>(x, nullptr);
  if (host_ctl_->breakdown)
    throw ConvergenceError(std::string("Krylov ") + (three_d_ ? "3D" : "2D") + (sizeof(T) == 4 ? " fp32" : " fp64") +
                           ": breakdown (p.Ap = " + std::to_string(host_ctl_->pq) + " at iteration " +
                           std::to_string(host_ctl_->iters + 1) + ", ||D^-1 r||^2 = " + std::to_string(host_ctl_->zz) +
                           ")");
  const int iters = host_ctl_->iters;
  // each launched 3D iteration recorded one spmv and one update launch
  if (three_d_ && launched > iters) {
    prof_->discard_last(kt_spmv, (size_t)(launched - iters));
    prof_->discard_last(FREHG_KT_GW_UPDATE, (size_t)(launched - iters));
  }
  if (sizeof(T) == 8) last_iters_ = iters;
  else last_inner_iters_ = iters;
  *res_norm = std::sqrt(host_ctl_->zz);
  *rhs_norm = std::sqrt(host_ctl_->bb);
  *converged = host_ctl_->converged != 0;
  return iters;
}

template <typename T>
int Krylov::run_pcg(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
                    bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl,
                    T* r, T* z, T* p0, T* p1, T* q, bool* rejected, double* res_norm,
                    double* rhs_norm, bool* converged, int kt_spmv) {
  cg_start<T>(A, b, bn, x, fixed, x0_zero, guard, rtol, atol, max_it, ctl, r, z, p0);
  *rejected = false;
  bool zero = x0_zero;
  if (guard && !x0_zero) {
    read_ctl(ctl, host_ctl_);
    if (host_ctl_->guess_rejected) {
      *rejected = true;
      zero = true;
      cg_start<T>(A, b, bn, x, fixed, true, guard, rtol, atol, max_it, ctl, r, z, p0);
    }
  }
  const bool fresh = zero && fixed == nullptr;
  const int est = (sizeof(T) == 8) ? last_iters_ : last_inner_iters_;
  const int batch = std::max(2, est + 1);
  cg_iterations<T>(A, x, r, z, p0, p1, q, ctl, 0, batch, kt_spmv, fresh);
  return cg_finish<T>(A, x, r, z, p0, p1, q, ctl, max_it, batch, batch, res_norm, rhs_norm, converged, kt_spmv,
                      fresh);
}

void Krylov::enqueue_cg(const StencilView<double>& A, const double* b, const double* bn, double* x,
                        const unsigned char* fixed, bool x0_zero, const KrylovConfig& cfg, int budget) {
  if (cfg.ksp != FREHG_KSP_CG || cfg.precision != 0) throw ArgError("Krylov::enqueue_cg: FP64 CG only");
  if (three_d_ && !tuned_[0]) throw ArgError("Krylov::enqueue_cg: 3D SpMV shape not tuned yet");
  cg_start<double>(A, b, bn, x, fixed, x0_zero, false, cfg.rtol, cfg.atol, cfg.max_it, ctl_, r_, z_, p0_);
  enq_fresh_ = x0_zero && fixed == nullptr;
  cg_iterations<double>(A, x, r_, z_, p0_, p1_, q_, ctl_, 0, budget, three_d_ ? FREHG_KT_GW_SPMV : FREHG_KT_SW_SPMV,
                        enq_fresh_);
}

SolveResult Krylov::finish_cg(const StencilView<double>& A, double* x, const KrylovConfig& cfg, int budget) {
  SolveResult res;
  bool conv = false;
  res.iterations = cg_finish<double>(A, x, r_, z_, p0_, p1_, q_, ctl_, cfg.max_it, budget, budget,
                                     &res.residual_norm, &res.rhs_norm, &conv,
                                     three_d_ ? FREHG_KT_GW_SPMV : FREHG_KT_SW_SPMV, enq_fresh_);
  res.converged = conv;
  if (!conv)
    throw ConvergenceError(std::string("Krylov ") + (three_d_ ? "3D" : "2D") + ": not converged within max_it (" +
                           std::to_string(res.iterations) + " iterations, ||D^-1 r|| = " +
                           std::to_string(res.residual_norm) + ", ||D^-1 b|| = " + std::to_string(res.rhs_norm) + ")");
  return res;
}

template <typename T>
int Krylov::run_bicgstab(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
                         bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl,
                         bool* rejected, double* res_norm, double* rhs_norm, bool* converged, int kt_spmv) {
  const bool fp32 = sizeof(T) == 4;
  ensure_bicgstab(fp32);
  T* r = fp32 ? (T*)r32_ : (T*)r_;
  T* rh = fp32 ? (T*)rh32_ : (T*)rh_;
  T* pbuf[2] = {fp32 ? (T*)p032_ : (T*)p0_, fp32 ? (T*)p132_ : (T*)p1_};
  T* vbuf[2] = {fp32 ? (T*)v032_ : (T*)v0_, fp32 ? (T*)v132_ : (T*)v1_};
  T* sv = fp32 ? (T*)sv32_ : (T*)sv_;
  T* tv = fp32 ? (T*)tv32_ : (T*)tv_;
  const dim3 blk(NT);
  auto init = [&](bool zero) {
    if (three_d_)
      hipLaunchKernelGGL((k_bcg_init<T, true>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, x, r, rh, pbuf[0],
                         vbuf[0], fixed, zero ? 1 : 0, part_, nb_f_);
    else
      hipLaunchKernelGGL((k_bcg_init<T, false>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, x, r, rh, pbuf[0],
                         vbuf[0], fixed, zero ? 1 : 0, part_, nb_f_);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fin_bcg_init, dim3(1), dim3(kRedThreads), 0, s_, part_, nb_f_, ctl, rtol, atol, max_it,
                       guard ? 1 : 0, zero ? 1 : 0);
    FREHG_LAUNCH_CHECK();
  };
  init(x0_zero);
  *rejected = false;
  if (guard && !x0_zero) {
    read_ctl(ctl, host_ctl_);
    if (host_ctl_->guess_rejected) {
      *rejected = true;
      init(true);
    }
  }
  int parity = 0;
  int& last = fp32 ? last_bcg_inner_iters_ : last_bcg_iters_;
  int batch = std::max(2, last + 1);
  int launched = 0;
  for (;;) {
    for (int it = 0; it < batch; ++it) {
      const T* pold = pbuf[parity];
      T* pnew = pbuf[parity ^ 1];
      const T* vold = vbuf[parity];
      T* vnew = vbuf[parity ^ 1];
      {
        hipEvent_t ev0 = three_d_ ? prof_->start(kt_spmv) : nullptr;
        hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
        if (three_d_)
          hipExtLaunchKernelGGL((k_bcg_spmv1<T, true>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, (const T*)r, (const T*)rh, pold,
                             vold, pnew, vnew, part_, ctl);
        else
          hipExtLaunchKernelGGL((k_bcg_spmv1<T, false>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, (const T*)r, (const T*)rh,
                             pold, vold, pnew, vnew, part_, ctl);
        FREHG_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_fin_bcg1, dim3(1), dim3(kRedThreads), 0, s_, part_, nb_f_, ctl);
      FREHG_LAUNCH_CHECK();
      {
        hipEvent_t ev0 = three_d_ ? prof_->start(kt_spmv) : nullptr;
        hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
        if (three_d_)
          hipExtLaunchKernelGGL((k_bcg_spmv2<T, true>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, (const T*)r, (const T*)vnew, sv,
                             tv, part_, ctl, nb_f_);
        else
          hipExtLaunchKernelGGL((k_bcg_spmv2<T, false>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, (const T*)r, (const T*)vnew,
                             sv, tv, part_, ctl, nb_f_);
        FREHG_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_fin_bcg2, dim3(1), dim3(kRedThreads), 0, s_, part_, nb_f_, ctl);
      FREHG_LAUNCH_CHECK();
      {
        hipEvent_t ev0 = three_d_ ? prof_->start(FREHG_KT_GW_UPDATE) : nullptr;
        hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
        if (three_d_)
          hipExtLaunchKernelGGL((k_bcg_update<T, true>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A.D, (const T*)pnew, (const T*)sv,
                             (const T*)tv, (const T*)rh, x, r, part_, ctl, nb_f_);
        else
          hipExtLaunchKernelGGL((k_bcg_update<T, false>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A.D, (const T*)pnew,
                             (const T*)sv, (const T*)tv, (const T*)rh, x, r, part_, ctl, nb_f_);
        FREHG_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_fin_bcg3, dim3(1), dim3(kRedThreads), 0, s_, part_, nb_f_, ctl);
      FREHG_LAUNCH_CHECK();
      parity ^= 1;
    }
    launched += batch;
    read_ctl(ctl, host_ctl_);
    if (host_ctl_->done) break;
    if (launched >= max_it + 2) break;
    batch = std::max(2, std::min(16, batch / 2));
  }
  if (host_ctl_->breakdown) {
    static const char* what[] = {"", "rh.v = 0", "omega = 0", "rho = 0"};
    throw ConvergenceError(std::string("BiCGSTAB ") + (three_d_ ? "3D" : "2D") + (fp32 ? " fp32" : " fp64") +
                           ": breakdown (" + what[host_ctl_->breakdown & 3] + ") at iteration " +
                           std::to_string(host_ctl_->iters + 1));
  }
  const int iters = host_ctl_->iters;
  // each launched 3D iteration recorded two spmv launches and one update launch
  if (three_d_ && launched > iters) {
    prof_->discard_last(kt_spmv, 2 * (size_t)(launched - iters));
    prof_->discard_last(FREHG_KT_GW_UPDATE, (size_t)(launched - iters));
  }
  last = iters;
  *res_norm = std::sqrt(host_ctl_->zz);
  *rhs_norm = std::sqrt(host_ctl_->bb);
  *converged = host_ctl_->converged != 0;
  return iters;
}

template <typename T>
int Krylov::run(int ksp, const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
                bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl, bool* rejected,
                double* res_norm, double* rhs_norm, bool* converged, int kt_spmv) {
  if (ksp == FREHG_KSP_BICGSTAB)
    return run_bicgstab<T>(A, b, bn, x, fixed, x0_zero, guard, rtol, atol, max_it, ctl, rejected, res_norm,
                           rhs_norm, converged, kt_spmv);
  if (!A.symmetric()) throw ArgError("Krylov: CG needs a symmetric operator (use BiCGSTAB)");
  if (sizeof(T) == 4)
    return run_pcg<T>(A, b, bn, x, fixed, x0_zero, guard, rtol, atol, max_it, ctl, (T*)r32_, (T*)z32_, (T*)p032_,
                      (T*)p132_, (T*)q32_, rejected, res_norm, rhs_norm, converged, kt_spmv);
  return run_pcg<T>(A, b, bn, x, fixed, x0_zero, guard, rtol, atol, max_it, ctl, (T*)r_, (T*)z_, (T*)p0_, (T*)p1_,
                    (T*)q_, rejected, res_norm, rhs_norm, converged, kt_spmv);
}

SolveResult Krylov::solve(const StencilView<double>& A, const StencilView<float>* A32, const double* b,
                          const double* bn, double* x, const unsigned char* fixed, bool x0_zero,
                          bool guard_guess, const KrylovConfig& cfg, unsigned int swept_blocks) {
  SolveResult res;
  // (the autotuner's SpMV launches write part_, which holds a caller's sweep partials)
  if (!tuned_[0] && !swept_blocks) autotune<double>(A);
  if (swept_blocks && (cfg.precision == 0 || A32 == nullptr))
    throw ArgError("Krylov: a pre-swept solve is a mixed-precision solve");
  if (cfg.precision == 0 || A32 == nullptr) {
    bool rej = false, conv = false;
    res.iterations = run<double>(cfg.ksp, A, b, bn, x, fixed, x0_zero, guard_guess, cfg.rtol, cfg.atol,
                                 cfg.max_it, ctl_, &rej, &res.residual_norm, &res.rhs_norm, &conv,
                                 three_d_ ? FREHG_KT_GW_SPMV : FREHG_KT_SW_SPMV);
    res.guess_rejected = rej;
    res.converged = conv;
    if (!conv)
      throw ConvergenceError(std::string("Krylov ") + (three_d_ ? "3D" : "2D") + ": not converged within max_it (" +
                             std::to_string(res.iterations) + " iterations, ||D^-1 r|| = " +
                             std::to_string(res.residual_norm) + ", ||D^-1 b|| = " + std::to_string(res.rhs_norm) + ")");
    return res;
  }
  // FP32 inner / FP64 outer iterative refinement.
  ensure_mixed();
  const dim3 blk(NT);
  const unsigned int nb = nb_f_;
  KrylovCtl* outer = ctl_;
  KrylovCtl* inner = ctl_ + 1;
  if (x0_zero) {
    zero_interior<double>(x, fixed);
  }
  // The iterate alternates between x and xw_: the first sweep copies x to xw_, and each later
  // sweep reads one and writes the other, so that after one refinement it is back in x. xw_'s
  // ghost entries are never written: they stay finite (zero or earlier heads) and meet zero
  // coefficients only.
  // Paired sweep (3D, symmetric, no fixed rows, bn = b); with CG as the inner solver it also
  // starts the inner solve (residual straight into r32_, z32_, inner control block), so the
  // inner CG begins with its first iteration.
  const bool paired = three_d_ && A.symmetric() && fixed == nullptr && bn == nullptr;
  const bool fused = paired && cfg.ksp == FREHG_KSP_CG;
  auto refine = [&](const double* xin, const float* d, double* xout, bool first) {
    const bool sym = A.symmetric();
    hipEvent_t ev0 = three_d_ ? prof_->start(FREHG_KT_GW_REFINE) : nullptr;
    hipEvent_t ev1 = three_d_ ? prof_->stop() : nullptr;
    unsigned int nbr = nb;
    if (paired) {
      const int kc = shape_[0].kc;
      const dim3 grid(ceil_div(g_.nx, 2 * TX), ceil_div(g_.ny, TY), ceil_div(g_.nz, kc));
      nbr = grid.x * grid.y * grid.z;
      float* r32 = fused ? r32_ : b32_;
      const float* D32 = fused ? A32->D : nullptr;
      float* z32 = fused ? z32_ : nullptr;
      if (d)
        hipExtLaunchKernelGGL((k_refine_pair<true>), grid, blk, 0, s_, ev0, ev1, 0, g_, kc, A, b, xin, d, xout, r32,
                              D32, z32, part_, first ? 1 : 0);
      else
        hipExtLaunchKernelGGL((k_refine_pair<false>), grid, blk, 0, s_, ev0, ev1, 0, g_, kc, A, b, xin, d, xout,
                              r32, D32, z32, part_, first ? 1 : 0);
    } else if (three_d_ && sym)
      hipExtLaunchKernelGGL((k_refine<true, true>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, b, bn, xin, d,
                            xout, b32_, fixed, part_, nb, first ? 1 : 0);
    else if (three_d_)
      hipExtLaunchKernelGGL((k_refine<true, false>), grid_f_, blk, 0, s_, ev0, ev1, 0, g_, kc_f_, A, b, bn, xin, d,
                            xout, b32_, fixed, part_, nb, first ? 1 : 0);
    else if (sym)
      hipLaunchKernelGGL((k_refine<false, true>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, xin, d, xout, b32_,
                         fixed, part_, nb, first ? 1 : 0);
    else
      hipLaunchKernelGGL((k_refine<false, false>), grid_f_, blk, 0, s_, g_, kc_f_, A, b, bn, xin, d, xout, b32_,
                         fixed, part_, nb, first ? 1 : 0);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fin_refine, dim3(1), dim3(kRedThreads), 0, s_, part_, nbr, outer, first ? 1 : 0, cfg.rtol,
                       cfg.atol, (guard_guess && !x0_zero) ? 1 : 0, fused ? inner : nullptr, cfg.inner_rtol,
                       cfg.max_it);
    FREHG_LAUNCH_CHECK();
  };
  double* xcur = xw_;
  double* xalt = x;
  if (swept_blocks) {
    if (!fused || x0_zero || swept_blocks > part_cap_)
      throw ArgError("Krylov: a pre-swept solve needs the paired 3D CG refinement from x0 = x");
    // one block of partials per tile and layer: summed in two stages
    hipLaunchKernelGGL(k_reduce_sum_stage, dim3(kRedStage, 5), dim3(256), 0, s_, (const double*)part_, swept_blocks,
                       stage_);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fin_refine, dim3(1), dim3(kRedThreads), 0, s_, (const double*)stage_, kRedStage, outer, 1,
                       cfg.rtol, cfg.atol, guard_guess ? 1 : 0, inner, cfg.inner_rtol, cfg.max_it);
    FREHG_LAUNCH_CHECK();
    if (!tuned_[0]) autotune<double>(A);
  } else {
    refine(x, nullptr, xw_, true);
  }
  read_ctl(outer, host_ctl_);
  if (host_ctl_->guess_rejected) {
    res.guess_rejected = true;
    zero_interior<double>(x, fixed);
    refine(x, nullptr, xw_, true);
    read_ctl(outer, host_ctl_);
  }
  res.rhs_norm = std::sqrt(host_ctl_->bb);
  res.residual_norm = std::sqrt(host_ctl_->zz);
  bool conv = host_ctl_->converged != 0;
  const StencilView<float> A32v = *A32;
  if (!tuned_[1]) autotune<float>(A32v);
  while (!conv && res.refinements < cfg.max_refine) {
    bool rej = false, iconv = false;
    double rn = 0, bnrm = 0;
    // The inner solve only has to close the gap to the outer tolerance (with a factor-2 margin):
    // a good initial guess leaves a small gap and needs fewer inner iterations.
    const double gap = std::sqrt(host_ctl_->ttol2) / std::max(res.residual_norm, 1e-300);
    const double inner_rt = std::min(0.5, std::max(cfg.inner_rtol, 0.5 * gap));
    if (fused) {
      // started by the sweep (same tolerance, computed on the device)
      const int batch = std::max(2, last_inner_iters_ + 1);
      cg_iterations<float>(A32v, d32_, r32_, z32_, p032_, p132_, q32_, inner, 0, batch, FREHG_KT_GW_SPMV32, true);
      res.iterations += cg_finish<float>(A32v, d32_, r32_, z32_, p032_, p132_, q32_, inner, cfg.max_it, batch, batch,
                                         &rn, &bnrm, &iconv, FREHG_KT_GW_SPMV32, true);
    } else {
      res.iterations += run<float>(cfg.ksp, A32v, b32_, nullptr, d32_, fixed, true, false, inner_rt, 0.0,
                                   cfg.max_it, inner, &rej, &rn, &bnrm, &iconv,
                                   three_d_ ? FREHG_KT_GW_SPMV32 : FREHG_KT_SW_SPMV);
    }
    refine(xcur, d32_, xalt, false);
    std::swap(xcur, xalt);
    res.refinements += 1;
    read_ctl(outer, host_ctl_);
    res.residual_norm = std::sqrt(host_ctl_->zz);
    conv = host_ctl_->converged != 0;
  }
  // without a refinement x itself holds the iterate (xw_ holds a copy)
  if (res.refinements > 0 && xcur != x)
    FREHG_HIP(hipMemcpyAsync(x, xcur, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  if (!conv) {
    // Refinement stalled: finish in FP64 from the current iterate.
    bool rej = false, c64 = false;
    res.fallback_fp64 = true;
    res.iterations += run<double>(cfg.ksp, A, b, bn, x, fixed, false, false, cfg.rtol, cfg.atol, cfg.max_it, ctl_,
                                  &rej, &res.residual_norm, &res.rhs_norm, &c64,
                                  three_d_ ? FREHG_KT_GW_SPMV : FREHG_KT_SW_SPMV);
    conv = c64;
    // the FP64 test above is relative to ||D^-1 b||, the same normalisation as the outer test
  }
  res.converged = conv;
  if (!conv) throw ConvergenceError("Krylov: mixed-precision solve not converged");
  return res;
}

template int Krylov::run_pcg<double>(const StencilView<double>&, const double*, const double*, double*,
                                     const unsigned char*, bool, bool, double, double, int, KrylovCtl*,
                                     double*, double*, double*, double*, double*, bool*, double*, double*,
                                     bool*, int);
template int Krylov::run_pcg<float>(const StencilView<float>&, const float*, const float*, float*,
                                    const unsigned char*, bool, bool, double, double, int, KrylovCtl*, float*,
                                    float*, float*, float*, float*, bool*, double*, double*, bool*, int);

}  // namespace frehg
