// Batched ensemble of independent Richards boxes (see ensemble.hpp).
//
// Tile spaces: "cell" tiles are 64 x 4 columns of one layer (ntx * nty * nz per box); "SpMV"
// tiles are 64 x 4 columns of kc layers. A launch covers the tiles of every box; block -> tile goes
// through xcd_remap (neighbouring tiles share an XCD) and tile -> box through a prefix table.
// Per-block partial sums are indexed by tile, so the partials of one box are contiguous and a
// one-block-per-box kernel reduces them in a fixed order.
//
// A step is three kernel sequences, each captured once into a HIP graph:
//   A  save h^n / theta^n, K and C, assembly, initial guess, Krylov solve with an iteration budget
//   B1 head BCs, fluxes, corrector (+ ledger), a budget of saturation fixed-point sweeps
//   B2 reallocation branches, clamp (+ ledger), per-box sums
// with a convergence check on the host after A and B1; a box that needs more than the budget is
// finished by eager launches (the budget grows for the next capture).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ensemble.hpp"

namespace frehg {

namespace {

using namespace redev;
constexpr int NT = kReNT;
constexpr int NSUM = 6;  // ledger partials per tile: before, storage, boundary, after, adjust, max dh

struct Pools {
  double *h, *hn, *wc, *wcn, *wch, *hwc, *ch, *kr, *Kx, *Ky, *Kz, *qx, *qy, *qz, *D, *ax, *ay, *az, *rhs;
  // extrapolated start only: the post-solve head of the last step, then the start of the next
  double* hp = nullptr;
};

__device__ __forceinline__ int find_box(const unsigned int* t0, int nbox, unsigned int tile) {
  int lo = 0, hi = nbox;  // t0[lo] <= tile < t0[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (t0[mid] <= tile) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Box and cell of this thread in a cell-tile launch.
struct BoxCell {
  int b;
  unsigned int tile;
  long long off;
  ReK P;
  Cell t;
};

__device__ __forceinline__ BoxCell box_cell(const EnsBox* boxes, const unsigned int* t0, int nbox) {
  BoxCell r;
  r.tile = xcd_remap(blockIdx.x, gridDim.x);
  r.b = find_box(t0, nbox, r.tile);
  const EnsBox& B = boxes[r.b];
  r.P = B.P;
  r.off = B.off;
  const unsigned int local = r.tile - t0[r.b];
  const unsigned int bx = local % B.ntx, rest = local / B.ntx;
  r.t = cell_at(r.P.g, bx, rest % B.nty, rest / B.nty);
  return r;
}

// box_cell with the box parameters in soil variant K (see with_soil_variant_code) and the tile
// order of the single-box kernels: ZF = 1 takes the layer fastest within the box's tiles (the
// layers above and below a tile were just read through the same L2), ZF = 0 x fastest.
template <class K>
struct BoxCellT {
  int b;
  unsigned int tile;
  long long off;
  K P;
  Cell t;
};

template <class K, int ZF>
__device__ __forceinline__ BoxCellT<K> box_cell_as(const EnsBox* boxes, const unsigned int* t0, int nbox) {
  BoxCellT<K> r;
  r.tile = xcd_remap(blockIdx.x, gridDim.x);
  r.b = find_box(t0, nbox, r.tile);
  const EnsBox& B = boxes[r.b];
  r.P = rek_to<K>(B.P);
  r.off = B.off;
  const unsigned int local = r.tile - t0[r.b];
  unsigned int bx, by, bz;
  if (ZF) {
    const unsigned int nz = (unsigned int)r.P.g.nz, rest = local / nz;
    bz = local - rest * nz;
    bx = rest % B.ntx;
    by = rest / B.ntx;
  } else {
    const unsigned int rest = local / B.ntx;
    bx = local - rest * B.ntx;
    by = rest % B.nty;
    bz = rest / B.nty;
  }
  r.t = cell_at(r.P.g, bx, by, bz);
  return r;
}

// Column run of this thread in an SpMV-tile launch: cell (i, j) of the box, layers [k0, k1).
struct BoxCol {
  int b;
  unsigned int tile;
  long long off;
  Geo g;
  int i, j, k0, k1;
  bool valid;
};

__device__ __forceinline__ BoxCol box_col(const EnsBox* boxes, const unsigned int* s0, int nbox) {
  BoxCol r;
  r.tile = xcd_remap(blockIdx.x, gridDim.x);
  r.b = find_box(s0, nbox, r.tile);
  const EnsBox& B = boxes[r.b];
  r.off = B.off;
  r.g = B.P.g;
  const unsigned int local = r.tile - s0[r.b];
  const unsigned int bx = local % B.ntx, rest = local / B.ntx, by = rest % B.nty, bz = rest / B.nty;
  r.i = bx * 64 + (threadIdx.x & 63);
  r.j = by * 4 + (threadIdx.x >> 6);
  r.k0 = bz * B.kc;
  r.k1 = min(r.g.nz, r.k0 + B.kc);
  r.valid = r.i < r.g.nx && r.j < r.g.ny;
  return r;
}

// ---- Richards kernels ------------------------------------------------------------------------

__global__ __launch_bounds__(NT) void kb_save(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (!q.t.valid) return;
  const long long c = q.off + q.t.c;
  F.hn[c] = F.h[c];
  F.wcn[c] = F.wc[c];
}

// Ghost cells of h^n / theta^n: the face ghosts of every boundary cell (corner ghosts are unused).
__global__ __launch_bounds__(NT) void kb_save_ghosts(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (!q.t.valid) return;
  const Geo& g = q.P.g;
  const Cell& t = q.t;
  const long long gh[6] = {t.i == 0 ? t.c - 1 : -1,    t.i == g.nx - 1 ? t.c + 1 : -1,
                           t.j == 0 ? t.c - g.sy : -1, t.j == g.ny - 1 ? t.c + g.sy : -1,
                           t.k == 0 ? t.c - g.sz : -1, t.k == g.nz - 1 ? t.c + g.sz : -1};
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (gh[k] >= 0) {
      F.hn[q.off + gh[k]] = F.h[q.off + gh[k]];
      F.wcn[q.off + gh[k]] = F.wc[q.off + gh[k]];
    }
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void kb_pre(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F) {
  const BoxCellT<K> q = box_cell_as<K, ZF>(boxes, t0, nbox);
  if (q.t.valid) re_pre(q.P, q.t, F.hn + q.off, F.kr + q.off, F.ch + q.off);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void kb_assemble(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                   float* D32, float* ax32, float* ay32, float* az32) {
  const BoxCellT<K> q = box_cell_as<K, ZF>(boxes, t0, nbox);
  if (!q.t.valid) return;
  const long long o = q.off;
  re_assemble(q.P, q.t, F.kr + o, F.ch + o, F.wcn + o, F.hn + o, F.D + o, F.ax + o, F.ay + o, F.az + o, F.rhs + o,
              D32 ? D32 + o : nullptr, ax32 + o, ay32 + o, az32 + o);
}

// kb_save for the extrapolated start with a previous increment: before h^n is overwritten (it
// still holds the last step's starting head), the start h^n + (post-solve head - that head) goes
// to hp, in the single-box arithmetic (k_extrap_interior).
__global__ __launch_bounds__(NT) void kb_save_extrap(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (!q.t.valid) return;
  const long long c = q.off + q.t.c;
  const double h = F.h[c];
  F.hp[c] = h + (F.hp[c] - F.hn[c]);
  F.hn[c] = h;
  F.wcn[c] = F.wc[c];
}

// x0 on interior cells: 0 (mode 0), h^n (1), or the extrapolated start kept in hp (2).
__global__ __launch_bounds__(NT) void kb_guess(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                int mode) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (!q.t.valid) return;
  const long long c = q.off + q.t.c;
  F.h[c] = mode == 2 ? F.hp[c] : (mode == 1 ? F.hn[c] : 0.0);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void kb_post(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F) {
  const BoxCellT<K> q = box_cell_as<K, ZF>(boxes, t0, nbox);
  if (!q.t.valid) return;
  if (F.hp) F.hp[q.off + q.t.c] = F.h[q.off + q.t.c];  // the post-solve head (extrapolated start)
  re_post(q.P, q.t, F.h + q.off, F.kr + q.off, F.ch + q.off, F.wch + q.off);
}

// Fluxes at h^{n+1} and the corrector in one pass (re_flux_correct).
template <class K, int ZF>
__global__ __launch_bounds__(NT) void kb_flux_correct(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                       unsigned char* sat0, unsigned char* cand, double* part,
                                                       unsigned int ntile) {
  __shared__ double lds[NT / kWave];
  const BoxCellT<K> q = box_cell_as<K, ZF>(boxes, t0, nbox);
  double a = 0.0, b = 0.0, d = 0.0;
  if (q.t.valid) {
    const long long o = q.off;
    re_flux_correct(q.P, q.t, F.h + o, F.kr + o, F.hn + o, F.wcn + o, F.wch + o, F.Kx + o, F.Ky + o, F.Kz + o,
                    F.qx + o, F.qy + o, F.qz + o, F.wc + o, F.hwc + o, sat0 + o, cand + o, nullptr, a, b, d);
  }
  a = block_sum<NT>(a, lds);
  b = block_sum<NT>(b, lds);
  d = block_sum<NT>(d, lds);
  if (threadIdx.x == 0) {
    part[q.tile] = a;
    part[ntile + q.tile] = b;
    part[2 * ntile + q.tile] = d;
  }
}

// Zero the sweep-change flags (a kernel, so that graphs contain no memset nodes).
__global__ void kb_clear_flags(int* flags, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) flags[i] = 0;
}

// Sweep s of the saturation fixed point; once a sweep changed nothing, later sweeps only copy.
__global__ __launch_bounds__(NT) void kb_realloc(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                  const unsigned char* sat0, const unsigned char* cand,
                                                  const unsigned char* prev, unsigned char* next, int* flags, int s) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (!q.t.valid) return;
  const long long o = q.off;
  if (s > 0 && flags[s - 1] == 0) {
    next[o + q.t.c] = prev[o + q.t.c];
    return;
  }
  re_realloc(q.P, q.t, FacesK{F.Kx + o, F.Ky + o}, sat0 + o, cand + o, prev + o, next + o, flags + s);
}

template <class K, int ZF>
__global__ __launch_bounds__(NT) void kb_finalize(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                   const unsigned char* sat0, const unsigned char* fin, double* part,
                                                   unsigned int ntile) {
  __shared__ double lds[NT / kWave];
  __shared__ double lmax[NT / kWave];
  const BoxCellT<K> q = box_cell_as<K, ZF>(boxes, t0, nbox);
  double a = 0.0, b = 0.0, dh = 0.0;
  if (q.t.valid) {
    const long long o = q.off;
    re_finalize(q.P, q.t, FacesK{F.Kx + o, F.Ky + o}, sat0 + o, fin + o, F.wch + o, F.hwc + o, F.hn + o,
                F.wc + o, F.h + o, nullptr, nullptr, a, b, dh);
  }
  a = block_sum<NT>(a, lds);
  b = block_sum<NT>(b, lds);
  dh = wave_max(dh);
  if ((threadIdx.x % kWave) == 0) lmax[threadIdx.x / kWave] = dh;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lmax[0];
    for (int k = 1; k < NT / kWave; ++k) m = fmax(m, lmax[k]);
    part[3 * ntile + q.tile] = a;
    part[4 * ntile + q.tile] = b;
    part[5 * ntile + q.tile] = m;
  }
}

__global__ __launch_bounds__(NT) void kb_initial(const EnsBox* boxes, const unsigned int* t0, int nbox, Pools F,
                                                  int box, int from_head) {
  const BoxCell q = box_cell(boxes, t0, nbox);
  if (q.b != box || !q.t.valid) return;
  const long long o = q.off;
  re_initial(q.P, q.t, from_head, F.h + o, F.wc + o, F.hn + o, F.wcn + o, F.wch + o, F.hwc + o, F.ch + o);
}

// Per-box sums (q < 5) and max (q = 5) of the ledger partials, one block per box.
__global__ __launch_bounds__(NT) void kb_box_sums(const double* part, unsigned int ntile, const unsigned int* t0,
                                                   double* out) {
  __shared__ double lds[NT / kWave];
  const int b = blockIdx.x;
  const unsigned int lo = t0[b], hi = t0[b + 1];
  for (int q = 0; q < NSUM; ++q) {
    double v = q == 5 ? 0.0 : 0.0;
    for (unsigned int i = lo + threadIdx.x; i < hi; i += NT) {
      const double x = part[(size_t)q * ntile + i];
      v = q == 5 ? fmax(v, x) : v + x;
    }
    double r;
    if (q == 5) {
      v = wave_max(v);
      __syncthreads();
      if ((threadIdx.x % kWave) == 0) lds[threadIdx.x / kWave] = v;
      __syncthreads();
      r = lds[0];
      for (int k = 1; k < NT / kWave; ++k) r = fmax(r, lds[k]);
    } else {
      r = block_sum<NT>(v, lds);
    }
    if (threadIdx.x == 0) out[(size_t)b * 8 + q] = r;
    __syncthreads();
  }
}

// x = xa for boxes whose refinement left the iterate in the alternate buffer.
__global__ __launch_bounds__(NT) void kb_take_x(const EnsBox* boxes, const unsigned int* s0, int nbox, double* h,
                                                 const double* xa, const EnsCtl* ctl) {
  const BoxCol q = box_col(boxes, s0, nbox);
  if (!q.valid || ctl[q.b].parity == 0) return;
  for (int k = q.k0; k < q.k1; ++k) {
    const long long c = q.off + q.g.idx(q.i, q.j, k);
    h[c] = xa[c];
  }
}

// ---- Krylov kernels (per-box control blocks; a block of a finished box returns at once) -------
// They run over the SpMV tiles (64 x 4 columns x kc layers), so a launch has few blocks and the
// blocks of finished boxes cost little.




template <typename T>
struct Sten {
  const T *D, *ax, *ay, *az;
};


// r = b - A x (or b), z = r / D, p_prev = 0; partials r.z, z.z, (b/D)^2.
template <typename T>
__global__ __launch_bounds__(NT) void kb_cg_init(const EnsBox* boxes, const unsigned int* s0, int nbox, Sten<T> A,
                                                  const T* bvec, T* x, T* r, T* z, T* pprev, EnsCtl* ctl, int inner,
                                                  int x0_zero, double* part, unsigned int ntile) {
  __shared__ double lds[NT / kWave];
  const BoxCol q = box_col(boxes, s0, nbox);
  if (inner && ctl[q.b].o.done) return;
  double srz = 0.0, szz = 0.0, sbb = 0.0;
  if (q.valid) {
    const Geo& g = q.g;
    for (int k = q.k0; k < q.k1; ++k) {
      const long long c = q.off + g.idx(q.i, q.j, k);
      double ax = 0.0;
      if (!x0_zero) {
        ax = (double)A.D[c] * (double)x[c] + (double)A.ax[c] * (double)x[c + 1] +
             (double)A.ax[c - 1] * (double)x[c - 1] + (double)A.ay[c] * (double)x[c + g.sy] +
             (double)A.ay[c - g.sy] * (double)x[c - g.sy] + (double)A.az[c] * (double)x[c + g.sz] +
             (double)A.az[c - g.sz] * (double)x[c - g.sz];
      }
      const T rT = (T)((double)bvec[c] - ax);
      const T zT = (T)((double)rT / (double)A.D[c]);
      r[c] = rT;
      z[c] = zT;
      pprev[c] = (T)0;
      srz += (double)rT * (double)zT;
      szz += (double)zT * (double)zT;
      const double bz = (double)bvec[c] / (double)A.D[c];
      sbb += bz * bz;
    }
    if (x0_zero)
      for (int k = q.k0; k < q.k1; ++k) x[q.off + g.idx(q.i, q.j, k)] = (T)0;
  }
  const double a = block_sum<NT>(srz, lds);
  const double b2 = block_sum<NT>(szz, lds);
  const double c2 = block_sum<NT>(sbb, lds);
  if (threadIdx.x == 0) {
    part[q.tile] = a;
    part[ntile + q.tile] = b2;
    part[2 * ntile + q.tile] = c2;
  }
}

__device__ inline double range_sum(const double* p, unsigned int lo, unsigned int hi, double* lds) {
  double s = 0.0;
  for (unsigned int i = lo + threadIdx.x; i < hi; i += NT) s += p[i];
  return block_sum<NT>(s, lds);
}

__global__ __launch_bounds__(NT) void kb_fin_init(const double* part, unsigned int ntile, const unsigned int* t0,
                                                   EnsCtl* ctl, int inner, double rtol, double atol, int max_it,
                                                   double inner_rtol) {
  __shared__ double lds[NT / kWave];
  const int b = blockIdx.x;
  EnsCtl& E = ctl[b];
  if (inner && E.o.done) {
    if (threadIdx.x == 0) E.i.done = 1;
    return;
  }
  const double rz = range_sum(part, t0[b], t0[b + 1], lds);
  const double zz = range_sum(part + ntile, t0[b], t0[b + 1], lds);
  const double bb = range_sum(part + 2 * (size_t)ntile, t0[b], t0[b + 1], lds);
  if (threadIdx.x == 0) {
    KrylovCtl& c = inner ? E.i : E.o;
    double ttol;
    if (inner) {
      const double gap = sqrt(E.o.ttol2) / fmax(sqrt(E.o.zz), 1e-300);
      ttol = fmin(0.5, fmax(inner_rtol, 0.5 * gap)) * sqrt(bb);
    } else {
      ttol = fmax(rtol * sqrt(bb), atol);
    }
    c.rz = rz;
    c.zz = zz;
    c.bb = bb;
    c.ttol2 = ttol * ttol;
    c.iters = 0;
    c.beta = 0.0;
    c.alpha = 0.0;
    c.breakdown = 0;
    c.max_it = max_it;
    c.guess_rejected = 0;
    c.converged = (zz <= c.ttol2) ? 1 : 0;
    c.done = c.converged;
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_spmv(const EnsBox* boxes, const unsigned int* s0, int nbox, Sten<T> A,
                                               const T* __restrict__ z, const T* __restrict__ pold,
                                               T* __restrict__ pnew, T* __restrict__ q, EnsCtl* ctl, int inner,
                                               double* part) {
  __shared__ double lds[NT / kWave];
  const unsigned int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int b = find_box(s0, nbox, tile);
  const KrylovCtl& kc = inner ? ctl[b].i : ctl[b].o;
  if (kc.done) return;
  const double beta = kc.beta;
  const EnsBox& B = boxes[b];
  const Geo g = B.P.g;
  const long long o = B.off;
  const unsigned int local = tile - s0[b];
  const unsigned int bx = local % B.ntx, rest = local / B.ntx, by = rest % B.nty, bz = rest / B.nty;
  const int i = bx * 64 + (threadIdx.x & 63), j = by * 4 + (threadIdx.x >> 6);
  const int k0 = bz * B.kc, k1 = min(g.nz, k0 + B.kc);
  double acc = 0.0;
  if (i < g.nx && j < g.ny) {
    const T* D = A.D + o;
    const T* ax = A.ax + o;
    const T* ay = A.ay + o;
    const T* az = A.az + o;
    const T* zz = z + o;
    const T* pp0 = pold + o;
    auto P = [&](long long cc) -> double { return (double)(T)((double)zz[cc] + beta * (double)pp0[cc]); };
    long long c = g.idx(i, j, k0);
    double pm = P(c - g.sz), azm = (double)az[c - g.sz];
    double pc = P(c);
    for (int k = k0; k < k1; ++k) {
      const double pp = P(c + g.sz);
      const double azc = (double)az[c];
      const double y = (double)D[c] * pc + (double)ax[c] * P(c + 1) + (double)ax[c - 1] * P(c - 1) +
                       (double)ay[c] * P(c + g.sy) + (double)ay[c - g.sy] * P(c - g.sy) + (azc * pp + azm * pm);
      const T yT = (T)y;
      pnew[o + c] = (T)pc;
      q[o + c] = yT;
      acc += pc * (double)yT;
      pm = pc;
      pc = pp;
      azm = azc;
      c += g.sz;
    }
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part[tile] = s;
}

__global__ __launch_bounds__(NT) void kb_fin_spmv(const double* part, const unsigned int* s0, EnsCtl* ctl, int inner) {
  __shared__ double lds[NT / kWave];
  const int b = blockIdx.x;
  KrylovCtl& c = inner ? ctl[b].i : ctl[b].o;
  if (c.done) return;
  const double pq = range_sum(part, s0[b], s0[b + 1], lds);
  if (threadIdx.x == 0) {
    c.pq = pq;
    if (!(pq > 0.0) || !isfinite(pq)) {
      c.breakdown = 1;
      c.done = 1;
    } else {
      c.alpha = c.rz / pq;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_update(const EnsBox* boxes, const unsigned int* s0, int nbox, const T* D,
                                                 const T* p, const T* qv, T* x, T* r, T* z, EnsCtl* ctl, int inner,
                                                 double* part, unsigned int ntile) {
  __shared__ double lds[NT / kWave];
  const BoxCol q = box_col(boxes, s0, nbox);
  const KrylovCtl& kc = inner ? ctl[q.b].i : ctl[q.b].o;
  if (kc.done) return;
  const double alpha = kc.alpha;
  double srz = 0.0, szz = 0.0;
  if (q.valid) {
    for (int k = q.k0; k < q.k1; ++k) {
      const long long c = q.off + q.g.idx(q.i, q.j, k);
      const double xv = (double)x[c] + alpha * (double)p[c];
      const double rv = (double)r[c] - alpha * (double)qv[c];
      const T rT = (T)rv;
      const T zT = (T)((double)rT / (double)D[c]);
      x[c] = (T)xv;
      r[c] = rT;
      z[c] = zT;
      srz += (double)rT * (double)zT;
      szz += (double)zT * (double)zT;
    }
  }
  const double a = block_sum<NT>(srz, lds);
  const double b = block_sum<NT>(szz, lds);
  if (threadIdx.x == 0) {
    part[q.tile] = a;
    part[ntile + q.tile] = b;
  }
}

__global__ __launch_bounds__(NT) void kb_fin_update(const double* part, unsigned int ntile, const unsigned int* t0,
                                                     EnsCtl* ctl, int inner) {
  __shared__ double lds[NT / kWave];
  const int b = blockIdx.x;
  KrylovCtl& c = inner ? ctl[b].i : ctl[b].o;
  if (c.done) return;
  const double rz = range_sum(part, t0[b], t0[b + 1], lds);
  const double zz = range_sum(part + ntile, t0[b], t0[b + 1], lds);
  if (threadIdx.x == 0) {
    const int it = c.iters + 1;
    c.iters = it;
    if (inner) {
      ctl[b].inner_total += 1;
      int& m = ctl[b].inner_max[ctl[b].refinements == 0 ? 0 : 1];
      m = max(m, it);
    }
    c.zz = zz;
    if (zz <= c.ttol2) {
      c.converged = 1;
      c.done = 1;
    } else if (it >= c.max_it) {
      c.done = 1;
    } else {
      c.beta = rz / c.rz;
    }
    c.rz = rz;
  }
}

// Refinement sweep: x_new = x_cur + d (first sweep: x_cur itself), r32 = b - A x_new in FP64,
// partials ||D^-1 r||^2 (and ||D^-1 b||^2 on the first sweep). x_cur is h or xa by the box's parity.
__global__ __launch_bounds__(NT) void kb_refine(const EnsBox* boxes, const unsigned int* s0, int nbox, Sten<double> A,
                                                 const double* bvec, double* h, double* xa, const float* d,
                                                 float* r32, EnsCtl* ctl, int first, double* part, unsigned int ntile) {
  __shared__ double lds[NT / kWave];
  const BoxCol q = box_col(boxes, s0, nbox);
  if (!first && ctl[q.b].o.done) return;
  const int par = first ? 0 : ctl[q.b].parity;
  const double* xin = par ? xa : h;
  double* xout = par ? h : xa;
  double srr = 0.0, sbb = 0.0;
  if (q.valid) {
    const Geo& g = q.g;
    auto X = [&](long long cc) -> double { return xin[cc] + (first ? 0.0 : (double)d[cc]); };
    for (int k = q.k0; k < q.k1; ++k) {
      const long long c = q.off + g.idx(q.i, q.j, k);
      const double xc = X(c);
      const double y = A.D[c] * xc + A.ax[c] * X(c + 1) + A.ax[c - 1] * X(c - 1) + A.ay[c] * X(c + g.sy) +
                       A.ay[c - g.sy] * X(c - g.sy) + A.az[c] * X(c + g.sz) + A.az[c - g.sz] * X(c - g.sz);
      const double rv = bvec[c] - y;
      if (!first) xout[c] = xc;
      r32[c] = (float)rv;
      const double rz = rv / A.D[c];
      srr += rz * rz;
      if (first) {
        const double bz = bvec[c] / A.D[c];
        sbb += bz * bz;
      }
    }
  }
  const double a = block_sum<NT>(srr, lds);
  const double b = block_sum<NT>(sbb, lds);
  if (threadIdx.x == 0) {
    part[q.tile] = a;
    part[ntile + q.tile] = b;
  }
}

__global__ __launch_bounds__(NT) void kb_fin_refine(const double* part, unsigned int ntile, const unsigned int* t0,
                                                     EnsCtl* ctl, int first, double rtol, double atol) {
  __shared__ double lds[NT / kWave];
  const int b = blockIdx.x;
  EnsCtl& E = ctl[b];
  if (!first && E.o.done) return;
  const double rr = range_sum(part, t0[b], t0[b + 1], lds);
  const double bb = first ? range_sum(part + ntile, t0[b], t0[b + 1], lds) : 0.0;
  if (threadIdx.x == 0) {
    if (first) {
      const double ttol = fmax(rtol * sqrt(bb), atol);
      E.o.bb = bb;
      E.o.ttol2 = ttol * ttol;
      E.refinements = 0;
      E.inner_total = 0;
      E.inner_max[0] = 0;
      E.inner_max[1] = 0;
      E.parity = 0;
    } else {
      E.parity ^= 1;
      E.refinements += 1;
    }
    E.o.zz = rr;
    E.o.converged = (rr <= E.o.ttol2) ? 1 : 0;
    E.o.done = E.o.converged;
  }
}

// flags[0] = boxes not converged, flags[1] = boxes whose inner solve is running, flags[2] = breakdowns.
__global__ void kb_status(const EnsCtl* ctl, int nbox, int* flags) {
  __shared__ int a[3];
  if (threadIdx.x < 3) a[threadIdx.x] = 0;
  __syncthreads();
  for (int b = threadIdx.x; b < nbox; b += blockDim.x) {
    if (!ctl[b].o.converged) atomicAdd(&a[0], 1);
    if (!ctl[b].i.done) atomicAdd(&a[1], 1);
    if (ctl[b].o.breakdown || ctl[b].i.breakdown) atomicAdd(&a[2], 1);
  }
  __syncthreads();
  if (threadIdx.x < 3) flags[threadIdx.x] = a[threadIdx.x];
}

}  // namespace

// ---------------------------------------------------------------------------------------------

Ensemble::Ensemble(int nbox, const int32_t* dims, const frehg_re_params* params, hipStream_t s, Profiler* prof)
    : nbox_(nbox), s_(s), prof_(prof) {
  try {
    build(nbox, dims, params);
  } catch (...) {
    release();  // a failed set-up frees what it allocated
    throw;
  }
}

void Ensemble::build(int nbox, const int32_t* dims, const frehg_re_params* params) {
  if (nbox < 1) throw ArgError("frehg_ens_init: need at least one box");
  if (!dims || !params) throw ArgError("frehg_ens_init: null argument");
  dims_.assign(dims, dims + 3 * nbox);
  params_.assign(params, params + nbox);
  const frehg_re_params& p0 = params[0];
  kc_ = krylov_config(p0.solver, "frehg_ens_init");
  if (kc_.ksp != FREHG_KSP_CG) throw ArgError("frehg_ens_init: the ensemble solves with CG");
  mixed_ = kc_.precision == 1;
  if (p0.initial_guess != FREHG_GUESS_ZERO && p0.initial_guess != FREHG_GUESS_PREVIOUS &&
      p0.initial_guess != FREHG_GUESS_EXTRAPOLATED)
    throw ArgError("frehg_ens_init: initial guess must be zero or previous-step");
  guess_ = p0.initial_guess;
  hbox_.resize(nbox);
  tile0_.assign(nbox + 1, 0);
  spmv0_.assign(nbox + 1, 0);
  classes_.assign(nbox, nullptr);
  cls_.assign(nbox, nullptr);
  ks_.assign(nbox, nullptr);
  for (int b = 0; b < nbox; ++b) {
    const int nx = dims[3 * b], ny = dims[3 * b + 1], nz = dims[3 * b + 2];
    if (nx < 1 || ny < 1 || nz < 1) throw ArgError("frehg_ens_init: box sizes must be >= 1");
    const frehg_re_params& p = params[b];
    if (p.adaptive_dt) throw ArgError("frehg_ens_init: the ensemble steps with a fixed dt (adaptive_dt = 0)");
    for (int f = 0; f < 4; ++f)
      if (p.bc_type[f] != 0) throw ArgError("frehg_ens_init: lateral boundaries must be no-flux (0)");
    if (p.bc_type[4] < 0 || p.bc_type[4] > 3 || p.bc_type[5] < 0 || p.bc_type[5] > 2)
      throw ArgError("frehg_ens_init: bottom bc must be 0..3, top bc 0..2");
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.dt > 0)) throw ArgError("frehg_ens_init: dx, dy, dz, dt must be > 0");
    if (p.picard_max != 1) throw ArgError("frehg_ens_init: picard_max must be 1");
    if (std::memcmp(&p.solver, &p0.solver, sizeof(p.solver)) != 0 || p.initial_guess != p0.initial_guess)
      throw ArgError("frehg_ens_init: every box must use the same solver configuration and initial guess");
    const Geo g = make_geo(nx, ny, nz);
    EnsBox& B = hbox_[b];
    B.off = storage_;
    B.ntx = ceil_div(nx, 64);
    B.nty = ceil_div(ny, 4);
    B.kc = nz;
    storage_ += g.ncell();
    cells_ += (long long)nx * ny * nz;
    tile0_[b + 1] = tile0_[b] + (unsigned int)(B.ntx * B.nty * nz);
    // default soil: the params' class, uniform
    SoilClass sc = make_soil_class(p.soil.alpha, p.soil.n, p.soil.theta_s, p.soil.theta_r, p.soil.Ks_x, p.soil.Ks_y,
                                   p.soil.Ks_z, p.soil.Ss, p.soil.use_vg, p.soil.use_mvg, p.soil.air_entry);
    FREHG_HIP(hipMalloc(&classes_[b], sizeof(SoilClass)));
    FREHG_HIP(hipMemcpy(classes_[b], &sc, sizeof(SoilClass), hipMemcpyHostToDevice));
    B.P = redev::make_rek(g, SoilView{classes_[b], nullptr, nullptr, sc}, p, nullptr);
  }
  ntile_ = tile0_[nbox];
  // pools
  auto dalloc = [&](double** p) {
    FREHG_HIP(hipMalloc(p, (size_t)storage_ * sizeof(double)));
    FREHG_HIP(hipMemsetAsync(*p, 0, (size_t)storage_ * sizeof(double), s_));
  };
  auto falloc = [&](float** p) {
    FREHG_HIP(hipMalloc(p, (size_t)storage_ * sizeof(float)));
    FREHG_HIP(hipMemsetAsync(*p, 0, (size_t)storage_ * sizeof(float), s_));
  };
  for (double** p : {&F_.h, &F_.hn, &F_.wc, &F_.wcn, &F_.wch, &F_.hwc, &F_.ch, &F_.kr, &F_.Kx, &F_.Ky, &F_.Kz, &F_.qx,
                     &F_.qy, &F_.qz, &F_.D, &F_.ax, &F_.ay, &F_.az, &F_.rhs, &F_.r, &F_.z, &F_.p0, &F_.p1, &F_.q})
    dalloc(p);
  if (guess_ == FREHG_GUESS_EXTRAPOLATED) dalloc(&F_.hp);
  if (mixed_) {
    dalloc(&F_.xa);
    for (float** p : {&F_.D32, &F_.ax32, &F_.ay32, &F_.az32, &F_.b32, &F_.r32, &F_.z32, &F_.p032, &F_.p132, &F_.q32,
                      &F_.d32})
      falloc(p);
  }
  for (unsigned char** p : {&F_.sat0, &F_.cand, &F_.satA, &F_.satB}) {
    FREHG_HIP(hipMalloc(p, (size_t)storage_));
    FREHG_HIP(hipMemsetAsync(*p, 0, (size_t)storage_, s_));
  }
  FREHG_HIP(hipMalloc(&dbox_, nbox * sizeof(EnsBox)));
  FREHG_HIP(hipMalloc(&dtile0_, (nbox + 1) * sizeof(unsigned int)));
  FREHG_HIP(hipMalloc(&dspmv0_, (nbox + 1) * sizeof(unsigned int)));
  FREHG_HIP(hipMalloc(&ctl_, nbox * sizeof(EnsCtl)));
  FREHG_HIP(hipMemsetAsync(ctl_, 0, nbox * sizeof(EnsCtl), s_));
  FREHG_HIP(hipMalloc(&part_, (size_t)NSUM * ntile_ * sizeof(double)));
  FREHG_HIP(hipMalloc(&boxsum_, (size_t)nbox * 8 * sizeof(double)));
  FREHG_HIP(hipMalloc(&flags_, (3 + kMaxSweepFlags) * sizeof(int)));
  FREHG_HIP(hipMemsetAsync(flags_, 0, (3 + kMaxSweepFlags) * sizeof(int), s_));
  FREHG_HIP(hipHostMalloc(&hflags_, (3 + kMaxSweepFlags) * sizeof(int), hipHostMallocDefault));
  FREHG_HIP(hipHostMalloc(&hboxsum_, (size_t)nbox * 8 * sizeof(double), hipHostMallocDefault));
  FREHG_HIP(hipHostMalloc(&hctl_, nbox * sizeof(EnsCtl), hipHostMallocDefault));
  upload_boxes();
  FREHG_HIP(hipStreamSynchronize(s_));
}

Ensemble::~Ensemble() { release(); }

void Ensemble::release() {
  for (auto& g : graph_)
    if (g) (void)hipGraphExecDestroy(g);
  std::vector<void*> ptrs = {F_.h,   F_.hn,   F_.wc,  F_.wcn,  F_.wch,  F_.hwc,  F_.ch,   F_.kr,   F_.Kx,   F_.Ky,
                             F_.Kz,  F_.qx,   F_.qy,  F_.qz,   F_.D,    F_.ax,   F_.ay,   F_.az,   F_.rhs,  F_.r,
                             F_.z,   F_.p0,   F_.p1,  F_.q,    F_.xa,   F_.D32,  F_.ax32, F_.ay32, F_.az32, F_.b32,
                             F_.r32, F_.z32,  F_.p032, F_.p132, F_.q32, F_.d32, F_.sat0, F_.cand, F_.satA, F_.satB,
                             dbox_,  dtile0_, dspmv0_, ctl_,   part_,  boxsum_, flags_, F_.hp};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto* p : classes_)
    if (p) (void)hipFree(p);
  for (auto* p : cls_)
    if (p) (void)hipFree(p);
  for (auto* p : ks_)
    if (p) (void)hipFree(p);
  if (hflags_) (void)hipHostFree(hflags_);
  if (hboxsum_) (void)hipHostFree(hboxsum_);
  if (hctl_) (void)hipHostFree(hctl_);
  for (auto& g : graph_) g = nullptr;
  F_ = Pool{};
  dbox_ = nullptr;
  dtile0_ = dspmv0_ = nullptr;
  ctl_ = nullptr;
  part_ = boxsum_ = nullptr;
  flags_ = hflags_ = nullptr;
  hboxsum_ = nullptr;
  hctl_ = nullptr;
  classes_.clear();
  cls_.clear();
  ks_.clear();
}

void Ensemble::upload_boxes() {
  // one soil variant for the per-cell kernels when every box has the same (else run-time checks)
  variant_ = nbox_ > 0 ? soil_variant_code(hbox_[0].P) : 0;
  for (int b = 1; b < nbox_; ++b)
    if (soil_variant_code(hbox_[b].P) != variant_) variant_ = 0;
  for (int b = 0; b < nbox_; ++b) {
    const EnsBox& B = hbox_[b];
    spmv0_[b + 1] = spmv0_[b] + (unsigned int)(B.ntx * B.nty * ceil_div(B.P.g.nz, B.kc));
  }
  nspmv_ = spmv0_[nbox_];
  FREHG_HIP(hipMemcpyAsync(dbox_, hbox_.data(), nbox_ * sizeof(EnsBox), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(dtile0_, tile0_.data(), (nbox_ + 1) * sizeof(unsigned int), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipMemcpyAsync(dspmv0_, spmv0_.data(), (nbox_ + 1) * sizeof(unsigned int), hipMemcpyHostToDevice, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  for (auto& g : graph_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void Ensemble::set_soil(int box, const frehg_soil_params* classes, int ncls, const int32_t* class_idx, int idx3d,
                        const double* ks_cell) {
  if (box < 0 || box >= nbox_) throw ArgError("frehg_ens_set_soil: box out of range");
  if (ncls < 1 || !classes) throw ArgError("frehg_ens_set_soil: need at least one soil class");
  EnsBox& B = hbox_[box];
  const Geo g = B.P.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  std::vector<SoilClass> hc;
  for (int c = 0; c < ncls; ++c) {
    const frehg_soil_params& s = classes[c];
    if (!(s.n > 1.0) || !(s.theta_s > s.theta_r) || !(s.alpha > 0))
      throw ArgError("frehg_ens_set_soil: invalid class (need n > 1, theta_s > theta_r, alpha > 0)");
    hc.push_back(make_soil_class(s.alpha, s.n, s.theta_s, s.theta_r, s.Ks_x, s.Ks_y, s.Ks_z, s.Ss, s.use_vg,
                                 s.use_mvg, s.air_entry));
  }
  auto clamp = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  std::vector<int32_t> pc;
  if (class_idx) {
    pc.assign((size_t)g.ncell(), 0);
    for (int k = -1; k <= nz; ++k)
      for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i) {
          const int ci = clamp(i, nx), cj = clamp(j, ny), ck = clamp(k, nz);
          const size_t o = idx3d ? (size_t)ci + (size_t)cj * nx + (size_t)ck * nx * ny : (size_t)ci + (size_t)cj * nx;
          const int32_t v = class_idx[o];
          if (v < 0 || v >= ncls) throw ArgError("frehg_ens_set_soil: class index out of range");
          pc[(size_t)g.idx(i, j, k)] = v;
        }
  }
  std::vector<double> pk;
  if (ks_cell) {
    pk.assign((size_t)g.ncell(), 0.0);
    for (int k = -1; k <= nz; ++k)
      for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i)
          pk[(size_t)g.idx(i, j, k)] =
              ks_cell[(size_t)clamp(i, nx) + (size_t)clamp(j, ny) * nx + (size_t)clamp(k, nz) * nx * ny];
  }
  for (void* p : {(void*)classes_[box], (void*)cls_[box], (void*)ks_[box]})
    if (p) FREHG_HIP(hipFree(p));
  classes_[box] = nullptr;
  cls_[box] = nullptr;
  ks_[box] = nullptr;
  FREHG_HIP(hipMalloc(&classes_[box], hc.size() * sizeof(SoilClass)));
  FREHG_HIP(hipMemcpy(classes_[box], hc.data(), hc.size() * sizeof(SoilClass), hipMemcpyHostToDevice));
  if (!pc.empty()) {
    FREHG_HIP(hipMalloc(&cls_[box], pc.size() * sizeof(int32_t)));
    FREHG_HIP(hipMemcpy(cls_[box], pc.data(), pc.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  if (!pk.empty()) {
    FREHG_HIP(hipMalloc(&ks_[box], pk.size() * sizeof(double)));
    FREHG_HIP(hipMemcpy(ks_[box], pk.data(), pk.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  B.P.soil = SoilView{classes_[box], cls_[box], ks_[box], hc[0]};
  upload_boxes();
}

void Ensemble::set_initial(int box, const double* h_int, const double* wc_int) {
  if (box < 0 || box >= nbox_) throw ArgError("frehg_ens_set_initial: box out of range");
  if ((h_int == nullptr) == (wc_int == nullptr))
    throw ArgError("frehg_ens_set_initial: pass exactly one of h / wc");
  have_inc_ = false;  // the next step starts from h^n (as frehg_re_set_initial)
  const EnsBox& B = hbox_[box];
  const Geo g = B.P.g;
  std::vector<double> pad((size_t)g.ncell(), 0.0);
  double* dst = (h_int ? F_.h : F_.wc) + B.off;
  FREHG_HIP(hipMemcpy(pad.data(), dst, pad.size() * sizeof(double), hipMemcpyDeviceToHost));
  const double* src = h_int ? h_int : wc_int;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        pad[(size_t)g.idx(i, j, k)] = src[(size_t)i + (size_t)j * g.nx + (size_t)k * g.nx * g.ny];
  FREHG_HIP(hipMemcpy(dst, pad.data(), pad.size() * sizeof(double), hipMemcpyHostToDevice));
  Pools P{F_.h, F_.hn, F_.wc, F_.wcn, F_.wch, F_.hwc, F_.ch, F_.kr, F_.Kx, F_.Ky, F_.Kz, F_.qx, F_.qy, F_.qz,
          F_.D, F_.ax, F_.ay, F_.az, F_.rhs};
  hipLaunchKernelGGL(kb_initial, dim3(ntile_), dim3(NT), 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dtile0_,
                     nbox_, P, box, h_int ? 1 : 0);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipStreamSynchronize(s_));
}

double* Ensemble::field(const std::string& n) {
  if (n == "h") return F_.h;
  if (n == "hn") return F_.hn;
  if (n == "wc") return F_.wc;
  if (n == "wcn") return F_.wcn;
  if (n == "wch") return F_.wch;
  if (n == "hwc") return F_.hwc;
  if (n == "ch") return F_.ch;
  if (n == "Kx") return F_.Kx;
  if (n == "Ky") return F_.Ky;
  if (n == "Kz") return F_.Kz;
  if (n == "qx") return F_.qx;
  if (n == "qy") return F_.qy;
  if (n == "qz") return F_.qz;
  if (n == "Gct") return F_.D;
  if (n == "Grhs") return F_.rhs;
  throw ArgError("unknown GW field '" + n + "'");
}

void Ensemble::get_field(int box, const std::string& name, double* dst) {
  if (box < 0 || box >= nbox_) throw ArgError("frehg_ens_get_field: box out of range");
  const EnsBox& B = hbox_[box];
  copy_field_out(B.P.g, B.P.g.nz + 2, dst, field(name) + B.off, s_);
  FREHG_HIP(hipStreamSynchronize(s_));
}

// ---- kernel sequences ----------------------------------------------------------------------

void Ensemble::enq_prepare() {
  const Pools P{F_.h, F_.hn, F_.wc, F_.wcn, F_.wch, F_.hwc, F_.ch, F_.kr, F_.Kx, F_.Ky, F_.Kz, F_.qx, F_.qy, F_.qz,
                F_.D, F_.ax, F_.ay, F_.az, F_.rhs, F_.hp};
  const dim3 grid(ntile_), blk(NT);
  const EnsBox* bx = dbox_;
  const unsigned int* t0 = dtile0_;
  const int mode = prep_mode();
  if (mode == 2)
    hipLaunchKernelGGL(kb_save_extrap, grid, blk, 0, s_, bx, t0, nbox_, P);
  else
    hipLaunchKernelGGL(kb_save, grid, blk, 0, s_, bx, t0, nbox_, P);
  hipLaunchKernelGGL(kb_save_ghosts, grid, blk, 0, s_, bx, t0, nbox_, P);
  with_soil_variant_code(variant_, [&](auto tag) {
    using K = decltype(tag);
    hipLaunchKernelGGL((kb_pre<K, 1>), grid, blk, 0, s_, bx, t0, nbox_, P);
    hipLaunchKernelGGL((kb_assemble<K, 1>), grid, blk, 0, s_, bx, t0, nbox_, P, mixed_ ? F_.D32 : nullptr, F_.ax32,
                       F_.ay32, F_.az32);
  });
  hipLaunchKernelGGL(kb_guess, grid, blk, 0, s_, bx, t0, nbox_, P, mode);
  FREHG_LAUNCH_CHECK();
  // Krylov start: FP64 CG init, or the first refinement residual
  if (!mixed_) {
    Sten<double> A{F_.D, F_.ax, F_.ay, F_.az};
    hipLaunchKernelGGL((kb_cg_init<double>), dim3(nspmv_), blk, 0, s_, bx, (const unsigned int*)dspmv0_, nbox_, A, (const double*)F_.rhs, F_.h, F_.r,
                       F_.z, F_.p0, ctl_, 0, mode == 0 ? 1 : 0, part_, nspmv_);
    hipLaunchKernelGGL(kb_fin_init, dim3(nbox_), blk, 0, s_, (const double*)part_, nspmv_, (const unsigned int*)dspmv0_, ctl_, 0, kc_.rtol,
                       kc_.atol, kc_.max_it, kc_.inner_rtol);
  } else {
    Sten<double> A{F_.D, F_.ax, F_.ay, F_.az};
    hipLaunchKernelGGL(kb_refine, dim3(nspmv_), blk, 0, s_, bx, (const unsigned int*)dspmv0_, nbox_, A, (const double*)F_.rhs,
                       F_.h, F_.xa, (const float*)F_.d32, F_.b32, ctl_, 1, part_, nspmv_);
    hipLaunchKernelGGL(kb_fin_refine, dim3(nbox_), blk, 0, s_, (const double*)part_, nspmv_, (const unsigned int*)dspmv0_, ctl_, 1, kc_.rtol,
                       kc_.atol);
  }
  FREHG_LAUNCH_CHECK();
}

// inner_iters: CG iterations per inner solve (mixed) or in total (FP64); refinements: mixed only.
void Ensemble::enq_solve(int inner_iters, int refinements, int inner_first) {
  const dim3 grid(ntile_), gsp(nspmv_), blk(NT), gb(nbox_);
  const EnsBox* bx = dbox_;
  const unsigned int* s0 = dspmv0_;
  auto iterate = [&](auto tag, Sten<decltype(tag)> A, decltype(tag)* x, decltype(tag)* r, decltype(tag)* z,
                     decltype(tag)* p0, decltype(tag)* p1, decltype(tag)* q, int inner, int n) {
    using T = decltype(tag);
    T* pb[2] = {p0, p1};
    for (int it = 0; it < n; ++it) {
      T* pold = pb[it & 1];
      T* pnew = pb[(it & 1) ^ 1];
      hipLaunchKernelGGL((kb_spmv<T>), gsp, blk, 0, s_, bx, s0, nbox_, A, (const T*)z, (const T*)pold, pnew, q, ctl_,
                         inner, part_);
      hipLaunchKernelGGL(kb_fin_spmv, gb, blk, 0, s_, (const double*)part_, s0, ctl_, inner);
      hipLaunchKernelGGL((kb_update<T>), gsp, blk, 0, s_, bx, s0, nbox_, A.D, (const T*)pnew, (const T*)q, x, r, z,
                         ctl_, inner, part_, nspmv_);
      hipLaunchKernelGGL(kb_fin_update, gb, blk, 0, s_, (const double*)part_, nspmv_, s0, ctl_, inner);
    }
    FREHG_LAUNCH_CHECK();
  };
  if (!mixed_) {
    // odd counts would leave p in p1: iterate in pairs
    const int n = inner_iters + (inner_iters & 1);
    iterate(0.0, Sten<double>{F_.D, F_.ax, F_.ay, F_.az}, F_.h, F_.r, F_.z, F_.p0, F_.p1, F_.q, 0, n);
    return;
  }
  Sten<float> A32{F_.D32, F_.ax32, F_.ay32, F_.az32};
  Sten<double> A{F_.D, F_.ax, F_.ay, F_.az};
  for (int rf = 0; rf < refinements; ++rf) {
    hipLaunchKernelGGL((kb_cg_init<float>), gsp, blk, 0, s_, bx, s0, nbox_, A32, (const float*)F_.b32, F_.d32,
                       F_.r32, F_.z32, F_.p032, ctl_, 1, 1, part_, nspmv_);
    hipLaunchKernelGGL(kb_fin_init, gb, blk, 0, s_, (const double*)part_, nspmv_, s0, ctl_, 1, kc_.rtol, kc_.atol,
                       kc_.max_it, kc_.inner_rtol);
    FREHG_LAUNCH_CHECK();
    iterate(0.0f, A32, F_.d32, F_.r32, F_.z32, F_.p032, F_.p132, F_.q32, 1,
            rf == 0 && inner_first >= 0 ? inner_first : inner_iters);
    hipLaunchKernelGGL(kb_refine, gsp, blk, 0, s_, bx, s0, nbox_, A, (const double*)F_.rhs, F_.h, F_.xa,
                       (const float*)F_.d32, F_.b32, ctl_, 0, part_, nspmv_);
    hipLaunchKernelGGL(kb_fin_refine, gb, blk, 0, s_, (const double*)part_, nspmv_, s0, ctl_, 0, kc_.rtol, kc_.atol);
    FREHG_LAUNCH_CHECK();
  }
}

void Ensemble::enq_update(int sweeps) {
  const Pools P{F_.h, F_.hn, F_.wc, F_.wcn, F_.wch, F_.hwc, F_.ch, F_.kr, F_.Kx, F_.Ky, F_.Kz, F_.qx, F_.qy, F_.qz,
                F_.D, F_.ax, F_.ay, F_.az, F_.rhs, F_.hp};
  const dim3 grid(ntile_), blk(NT);
  const EnsBox* bx = dbox_;
  const unsigned int* t0 = dtile0_;
  if (mixed_)
    hipLaunchKernelGGL(kb_take_x, dim3(nspmv_), blk, 0, s_, bx, (const unsigned int*)dspmv0_, nbox_, F_.h,
                       (const double*)F_.xa, ctl_);
  with_soil_variant_code(variant_, [&](auto tag) {
    using K = decltype(tag);
    hipLaunchKernelGGL((kb_post<K, 1>), grid, blk, 0, s_, bx, t0, nbox_, P);
    hipLaunchKernelGGL((kb_flux_correct<K, 1>), grid, blk, 0, s_, bx, t0, nbox_, P, F_.sat0, F_.cand, part_, ntile_);
  });
  hipLaunchKernelGGL(kb_clear_flags, dim3(1), dim3(64), 0, s_, flags_ + 3, kMaxSweepFlags);
  for (int sw = 0; sw < sweeps; ++sw) {
    // sweep 0 reads sat0 directly; later sweeps ping-pong between satA and satB
    const unsigned char* prev = sw == 0 ? F_.sat0 : (sw & 1 ? F_.satA : F_.satB);
    unsigned char* next = sw & 1 ? F_.satB : F_.satA;
    hipLaunchKernelGGL(kb_realloc, grid, blk, 0, s_, bx, t0, nbox_, P, (const unsigned char*)F_.sat0,
                       (const unsigned char*)F_.cand, prev, next, flags_ + 3, sw);
  }
  FREHG_LAUNCH_CHECK();
}

void Ensemble::enq_finish(const unsigned char* fin) {
  const Pools P{F_.h, F_.hn, F_.wc, F_.wcn, F_.wch, F_.hwc, F_.ch, F_.kr, F_.Kx, F_.Ky, F_.Kz, F_.qx, F_.qy, F_.qz,
                F_.D, F_.ax, F_.ay, F_.az, F_.rhs};
  const dim3 grid(ntile_), blk(NT);
  // after an odd number of sweeps the state is in satA, after an even number in satB
  with_soil_variant_code(variant_, [&](auto tag) {
    using K = decltype(tag);
    hipLaunchKernelGGL((kb_finalize<K, 0>), grid, blk, 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dtile0_,
                       nbox_, P, (const unsigned char*)F_.sat0, fin, part_, ntile_);
  });
  hipLaunchKernelGGL(kb_box_sums, dim3(nbox_), blk, 0, s_, (const double*)part_, ntile_, (const unsigned int*)dtile0_,
                     boxsum_);
  FREHG_LAUNCH_CHECK();
}

void Ensemble::enq_status() {
  hipLaunchKernelGGL(kb_status, dim3(1), dim3(256), 0, s_, (const EnsCtl*)ctl_, nbox_, flags_);
  FREHG_LAUNCH_CHECK();
}

void Ensemble::read_status() {
  FREHG_HIP(hipMemcpyAsync(hflags_, flags_, (3 + kMaxSweepFlags) * sizeof(int), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

void Ensemble::autotune() {
  tuned_ = true;
  // SpMV layers per block: the whole column, or chunks of 8 / 16 layers; pick the fastest.
  hipEvent_t e0, e1;
  FREHG_HIP(hipEventCreate(&e0));
  FREHG_HIP(hipEventCreate(&e1));
  EnsCtl* tctl = ctl_;
  FREHG_HIP(hipMemsetAsync(tctl, 0, nbox_ * sizeof(EnsCtl), s_));
  int best = 0;
  float best_ms = 1e30f;
  const int cands[3] = {0, 16, 8};
  for (int ci = 0; ci < 3; ++ci) {
    for (auto& B : hbox_) B.kc = cands[ci] == 0 ? B.P.g.nz : std::min(B.P.g.nz, cands[ci]);
    upload_boxes();
    Sten<double> A{F_.D, F_.ax, F_.ay, F_.az};
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      FREHG_HIP(hipEventRecord(e0, s_));
      for (int it = 0; it < 3; ++it)
        hipLaunchKernelGGL((kb_spmv<double>), dim3(nspmv_), dim3(NT), 0, s_, (const EnsBox*)dbox_,
                           (const unsigned int*)dspmv0_, nbox_, A, (const double*)F_.D, (const double*)F_.p0, F_.p1,
                           F_.q, tctl, 0, part_);
      FREHG_LAUNCH_CHECK();
      FREHG_HIP(hipEventRecord(e1, s_));
      FREHG_HIP(hipEventSynchronize(e1));
      FREHG_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = cands[ci];
    }
  }
  for (auto& B : hbox_) B.kc = best == 0 ? B.P.g.nz : std::min(B.P.g.nz, best);
  upload_boxes();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void Ensemble::run_graph(int which, int budget_a, int budget_b, int budget_c) {
  const int mode = which == 0 ? prep_mode() : 0;
  if (!graph_[which] || graph_budget_[which] != budget_a || graph_budget2_[which] != budget_b ||
      graph_budget3_[which] != budget_c || graph_mode_[which] != mode) {
    if (graph_[which]) FREHG_HIP(hipGraphExecDestroy(graph_[which]));
    graph_[which] = nullptr;
    hipGraph_t g;
    FREHG_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      if (which == 0) {
        enq_prepare();
        enq_solve(budget_a, budget_b, budget_c);
      } else if (which == 1) {
        enq_update(budget_a);
      } else {
        enq_finish(budget_a & 1 ? F_.satA : F_.satB);
      }
    } catch (...) {
      (void)hipStreamEndCapture(s_, &g);
      throw;
    }
    FREHG_HIP(hipStreamEndCapture(s_, &g));
    FREHG_HIP(hipGraphInstantiate(&graph_[which], g, nullptr, nullptr, 0));
    FREHG_HIP(hipGraphDestroy(g));
    graph_budget_[which] = budget_a;
    graph_budget2_[which] = budget_b;
    graph_budget3_[which] = budget_c;
    graph_mode_[which] = mode;
  }
  FREHG_HIP(hipGraphLaunch(graph_[which], s_));
}

void Ensemble::step(int use_graph, frehg_ens_stats* st, frehg_re_stats* per_box) {
  if (!tuned_) autotune();
  const bool graphs = use_graph && budget_inner_ > 0;
  int extra = 0;
  // phase A: prepare + solve
  const int chunk = 8;
  if (graphs) {
    run_graph(0, mixed_ ? budget_inner_ : budget_fp64_, mixed_ ? budget_refine_ : 0, mixed_ ? budget_inner0_ : 0);
  } else {
    enq_prepare();
  }
  enq_status();
  read_status();
  int fp64_done = graphs ? budget_fp64_ : 0;
  if (!mixed_) {
    while (hflags_[0] > 0 && hflags_[2] == 0) {
      if (fp64_done >= kc_.max_it + 2) break;
      enq_solve(chunk, 0);
      fp64_done += chunk;
      if (graphs) ++extra;
      enq_status();
      read_status();
    }
  } else {
    int rounds = 0;
    while (hflags_[0] > 0 && hflags_[2] == 0 && rounds < kc_.max_refine + 8) {
      // one refinement with an inner solve run to its own tolerance
      const dim3 grid(ntile_), blk(NT), gb(nbox_);
      Sten<float> A32{F_.D32, F_.ax32, F_.ay32, F_.az32};
      hipLaunchKernelGGL((kb_cg_init<float>), dim3(nspmv_), blk, 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dspmv0_,
                         nbox_, A32, (const float*)F_.b32, F_.d32, F_.r32, F_.z32, F_.p032, ctl_, 1, 1, part_, nspmv_);
      hipLaunchKernelGGL(kb_fin_init, gb, blk, 0, s_, (const double*)part_, nspmv_, (const unsigned int*)dspmv0_, ctl_,
                         1, kc_.rtol, kc_.atol, kc_.max_it, kc_.inner_rtol);
      FREHG_LAUNCH_CHECK();
      int inner_done = 0;
      for (;;) {
        // chunk iterations (even, so p ends in p0 for the next chunk)
        const dim3 gsp(nspmv_);
        float* pb[2] = {F_.p032, F_.p132};
        for (int it = 0; it < chunk; ++it) {
          float* pold = pb[it & 1];
          float* pnew = pb[(it & 1) ^ 1];
          hipLaunchKernelGGL((kb_spmv<float>), gsp, blk, 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dspmv0_,
                             nbox_, A32, (const float*)F_.z32, (const float*)pold, pnew, F_.q32, ctl_, 1, part_);
          hipLaunchKernelGGL(kb_fin_spmv, gb, blk, 0, s_, (const double*)part_, (const unsigned int*)dspmv0_, ctl_, 1);
          hipLaunchKernelGGL((kb_update<float>), gsp, blk, 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dspmv0_,
                             nbox_, (const float*)F_.D32, (const float*)pnew, (const float*)F_.q32, F_.d32, F_.r32,
                             F_.z32, ctl_, 1, part_, nspmv_);
          hipLaunchKernelGGL(kb_fin_update, gb, blk, 0, s_, (const double*)part_, nspmv_, (const unsigned int*)dspmv0_,
                             ctl_, 1);
        }
        FREHG_LAUNCH_CHECK();
        inner_done += chunk;
        enq_status();
        read_status();
        if (hflags_[1] == 0 || hflags_[2] != 0 || inner_done >= kc_.max_it + 2) break;
      }
      Sten<double> A{F_.D, F_.ax, F_.ay, F_.az};
      hipLaunchKernelGGL(kb_refine, dim3(nspmv_), blk, 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dspmv0_, nbox_,
                         A, (const double*)F_.rhs, F_.h, F_.xa, (const float*)F_.d32, F_.b32, ctl_, 0, part_, nspmv_);
      hipLaunchKernelGGL(kb_fin_refine, gb, blk, 0, s_, (const double*)part_, nspmv_, (const unsigned int*)dspmv0_,
                         ctl_, 0, kc_.rtol, kc_.atol);
      FREHG_LAUNCH_CHECK();
      if (graphs) ++extra;
      ++rounds;
      enq_status();
      read_status();
    }
  }
  if (hflags_[2] != 0) throw ConvergenceError("ensemble Krylov: breakdown (p.Ap <= 0) in at least one box");
  if (hflags_[0] != 0) throw ConvergenceError("ensemble Krylov: not converged within the iteration limits");
  // record the budgets (from the per-box counters) for the graphs of later steps
  FREHG_HIP(hipMemcpyAsync(hctl_, ctl_, nbox_ * sizeof(EnsCtl), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  int max_it = 0, max_ref = 0, max_inner = 0, max_inner0 = 0;
  for (int b = 0; b < nbox_; ++b) {
    max_it = std::max(max_it, mixed_ ? hctl_[b].inner_total : hctl_[b].o.iters);
    max_ref = std::max(max_ref, hctl_[b].refinements);
    if (mixed_) {
      max_inner0 = std::max(max_inner0, hctl_[b].inner_max[0]);
      max_inner = std::max(max_inner, hctl_[b].inner_max[1]);
    }
  }
  // phase B1: BCs, fluxes, corrector, saturation sweeps
  if (graphs) {
    run_graph(1, budget_sweeps_, 0);
  } else {
    enq_update(budget_sweeps_);
  }
  read_status();
  int sweeps = budget_sweeps_;
  const unsigned char* fin = budget_sweeps_ & 1 ? F_.satA : F_.satB;  // where the budget's last sweep wrote
  bool eager_finish = !graphs;
  if (hflags_[3 + budget_sweeps_ - 1] != 0) {
    // still changing after the budget: continue eagerly from the last sweep's output
    const Pools P{F_.h, F_.hn, F_.wc, F_.wcn, F_.wch, F_.hwc, F_.ch, F_.kr, F_.Kx, F_.Ky, F_.Kz, F_.qx, F_.qy, F_.qz,
                  F_.D, F_.ax, F_.ay, F_.az, F_.rhs};
    int sw = budget_sweeps_;
    for (; sw < kMaxSweepFlags; ++sw) {
      const unsigned char* prev = sw & 1 ? F_.satA : F_.satB;
      unsigned char* next = sw & 1 ? F_.satB : F_.satA;
      hipLaunchKernelGGL(kb_realloc, dim3(ntile_), dim3(NT), 0, s_, (const EnsBox*)dbox_, (const unsigned int*)dtile0_,
                         nbox_, P, (const unsigned char*)F_.sat0, (const unsigned char*)F_.cand, prev, next,
                         flags_ + 3, sw);
      FREHG_LAUNCH_CHECK();
      ++extra;
      read_status();
      if (hflags_[3 + sw] == 0) break;
    }
    if (sw >= kMaxSweepFlags) throw ConvergenceError("ensemble: saturation fixed point did not settle");
    sweeps = sw + 1;
    fin = sweeps & 1 ? F_.satA : F_.satB;
    eager_finish = true;
    budget_sweeps_ = std::min(kMaxSweepFlags, sweeps + 1);
  } else {
    sweeps = 1;
    while (sweeps < budget_sweeps_ && hflags_[3 + sweeps - 1] != 0) ++sweeps;
  }
  // phase B2: reallocation branches, clamp, ledger
  if (eager_finish) {
    enq_finish(fin);
  } else {
    run_graph(2, budget_sweeps_, 0);
  }
  FREHG_HIP(hipMemcpyAsync(hboxsum_, boxsum_, (size_t)nbox_ * 8 * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  have_inc_ = guess_ == FREHG_GUESS_EXTRAPOLATED;  // hp holds this step's post-solve head
  // budgets for the next captured step
  if (mixed_) {
    budget_refine_ = std::max(budget_refine_, max_ref);
    budget_inner0_ = std::max(budget_inner0_, max_inner0 + 1);
    budget_inner_ = std::max(budget_inner_, max_inner + 1);
  } else {
    budget_fp64_ = std::max(budget_fp64_, max_it);
    budget_inner_ = std::max(budget_inner_, 1);
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->max_iterations = max_it;
    st->max_refinements = max_ref;
    st->converged = 1;
    st->graph = graphs ? 1 : 0;
    st->extra_launches = extra;
    st->realloc_sweeps = sweeps;
  }
  double worst = 0.0, mb = 0.0, ma = 0.0;
  for (int b = 0; b < nbox_; ++b) {
    const double* s = hboxsum_ + (size_t)b * 8;
    const double res = s[3] - s[0] - s[2] - s[1] - s[4];
    worst = std::max(worst, std::fabs(res) / s[0]);
    mb += s[0];
    ma += s[3];
    if (per_box) {
      frehg_re_stats& r = per_box[b];
      std::memset(&r, 0, sizeof(r));
      r.solve.iterations = mixed_ ? hctl_[b].inner_total : hctl_[b].o.iters;
      r.solve.refinements = hctl_[b].refinements;
      r.solve.residual_norm = std::sqrt(hctl_[b].o.zz);
      r.solve.rhs_norm = std::sqrt(hctl_[b].o.bb);
      r.solve.converged = hctl_[b].o.converged;
      r.dt_used = params_[b].dt;
      r.dt_next = params_[b].dt;
      r.picard_iterations = 1;
      r.picard_converged = 1;
      r.max_dh = s[5];
      r.mass_before = s[0];
      r.mass_storage_ss = s[1];
      r.mass_boundary_in = s[2];
      r.mass_after = s[3];
      r.mass_adjust = s[4];
      r.mass_residual = res;
    }
  }
  if (st) {
    st->max_rel_mass_residual = worst;
    st->mass_before = mb;
    st->mass_after = ma;
  }
}

}  // namespace frehg
