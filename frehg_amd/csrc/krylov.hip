// Jacobi-PCG kernels for the 7-/5-point stencil systems of the Richards and shallow-water
// steps. One iteration is two kernels:
//   spmv:   p = z + beta p_prev (recomputed at each stencil point), q = A p, partial p.q;
//           last block: alpha = rz / p.q
//   update: x += alpha p, r -= alpha q, z = r / D, partials r.z and z.z;
//           last block: convergence test, beta, rz
// Each thread owns one (i, j) column of a 64 x 4 tile and marches down k over a chunk of
// layers, keeping the k-1 / k / k+1 values of p and the az coupling of the layer above in
// registers, so every array is streamed from HBM once per iteration.
#include <algorithm>
#include <cstring>

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
  Tile t;
  t.i = blockIdx.x * TX + (threadIdx.x & (TX - 1));
  t.j = blockIdx.y * TY + (threadIdx.x / TX);
  t.k0 = blockIdx.z * kchunk;
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

// r = b - A x (or b when x0 = 0), z = r / D, p_prev = 0; partials r.z, z.z, (bn/D)^2.
template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_init(Geo g, int kchunk, StencilView<T> A, const T* b,
                                              const T* bn, T* x, T* r, T* z, T* pprev,
                                              const unsigned char* fixed, int x0_zero,
                                              double* part, KrylovCtl* ctl, unsigned int nblocks,
                                              double rtol, double atol, int max_it, int guard) {
  __shared__ double lds[NT / kWave];
  __shared__ int flag;
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
      pprev[c] = (T)0;
      srz += (double)rT * (double)zT;
      szz += (double)zT * (double)zT;
      const double bz = (double)(bn ? bn[c] : b[c]) / (double)A.D[c];
      sbb += bz * bz;
    }
  }
  // x0 = 0 is applied after every thread has read its neighbours' x (only when x0_zero, in
  // which case x is not read at all).
  if (x0_zero && t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz)
      if (!(fixed != nullptr && fixed[c] != 0)) x[c] = (T)0;
  }
  const double a = block_sum<NT>(srz, lds);
  const double bsum = block_sum<NT>(szz, lds);
  const double cs = block_sum<NT>(sbb, lds);
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = bsum;
    part[2 * nblocks + bid] = cs;
  }
  if (last_block(&ctl->counter[0], nblocks, &flag)) {
    const double rz = block_sum_partials<NT>(part, nblocks, lds);
    const double zz = block_sum_partials<NT>(part + nblocks, nblocks, lds);
    const double bb = block_sum_partials<NT>(part + 2 * nblocks, nblocks, lds);
    if (threadIdx.x == 0) {
      const double ttol = fmax(rtol * sqrt(bb), atol);
      ctl->rz = rz;
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
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_spmv(Geo g, int kchunk, StencilView<T> A, const T* z,
                                              const T* pold, T* pnew, T* q, double* part,
                                              KrylovCtl* ctl, unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  __shared__ int flag;
  if (ctl->done) return;
  const double beta = ctl->beta;
  const Tile t = tile_of(g, kchunk);
  double acc = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    auto P = [&](long long cc) -> double { return (double)(T)((double)z[cc] + beta * (double)pold[cc]); };
    double pm = 0.0, azm = 0.0;
    if (THREE_D) {
      pm = P(c - g.sz);
      azm = (double)A.az[c - g.sz];
    }
    double pc = P(c);
    for (int k = t.k0; k < t.k1; ++k) {
      double pp = 0.0, azc = 0.0;
      if (THREE_D) {
        pp = P(c + g.sz);
        azc = (double)A.az[c];
      }
      double y = (double)A.D[c] * pc + (double)A.ax[c] * P(c + 1) + (double)A.ax[c - 1] * P(c - 1) +
                 (double)A.ay[c] * P(c + g.sy) + (double)A.ay[c - g.sy] * P(c - g.sy);
      if (THREE_D) y += azc * pp + azm * pm;
      const T yT = (T)y;
      pnew[c] = (T)pc;
      q[c] = yT;
      acc += pc * (double)yT;
      pm = pc;
      pc = pp;
      azm = azc;
      c += g.sz;
    }
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part[block_id()] = s;
  if (last_block(&ctl->counter[1], nblocks, &flag)) {
    const double pq = block_sum_partials<NT>(part, nblocks, lds);
    if (threadIdx.x == 0) {
      ctl->pq = pq;
      if (!(pq > 0.0) || !isfinite(pq)) {
        ctl->breakdown = 1;
        ctl->done = 1;
      } else {
        ctl->alpha = ctl->rz / pq;
      }
    }
  }
}

template <typename T, bool THREE_D>
__global__ __launch_bounds__(NT) void k_update(Geo g, int kchunk, const T* D, const T* p, const T* q,
                                                T* x, T* r, T* z, double* part, KrylovCtl* ctl,
                                                unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  __shared__ int flag;
  if (ctl->done) return;
  const double alpha = ctl->alpha;
  const Tile t = tile_of(g, kchunk);
  double srz = 0.0, szz = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const double xv = (double)x[c] + alpha * (double)p[c];
      const double rv = (double)r[c] - alpha * (double)q[c];
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
  const unsigned int bid = block_id();
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = b;
  }
  if (last_block(&ctl->counter[2], nblocks, &flag)) {
    const double rz = block_sum_partials<NT>(part, nblocks, lds);
    const double zz = block_sum_partials<NT>(part + nblocks, nblocks, lds);
    if (threadIdx.x == 0) {
      const int it = ctl->iters + 1;
      ctl->iters = it;
      ctl->zz = zz;
      if (zz <= ctl->ttol2) {
        ctl->converged = 1;
        ctl->done = 1;
      } else if (it >= ctl->max_it) {
        ctl->done = 1;
      } else {
        ctl->beta = rz / ctl->rz;
      }
      ctl->rz = rz;
    }
  }
}

// Iterative-refinement residual: xout = xin + d (d may be null), r32 = b - A xout in FP64,
// partial ||D^-1 r||^2 (and ||D^-1 bn||^2 on the first sweep, which also sets the test).
template <bool THREE_D>
__global__ __launch_bounds__(NT) void k_refine(Geo g, int kchunk, StencilView<double> A,
                                                const double* b, const double* bn, const double* xin,
                                                const float* d, double* xout, float* r32,
                                                const unsigned char* fixed, double* part,
                                                KrylovCtl* ctl, unsigned int nblocks, int first,
                                                double rtol, double atol, int guard) {
  __shared__ double lds[NT / kWave];
  __shared__ int flag;
  const Tile t = tile_of(g, kchunk);
  double srr = 0.0, sbb = 0.0;
  if (t.valid) {
    long long c = base_index<THREE_D>(g, t);
    auto X = [&](long long cc) -> double { return xin[cc] + (d ? (double)d[cc] : 0.0); };
    for (int k = t.k0; k < t.k1; ++k, c += g.sz) {
      const bool fx = fixed != nullptr && fixed[c] != 0;
      const double xc = X(c);
      double y = A.D[c] * xc + A.ax[c] * X(c + 1) + A.ax[c - 1] * X(c - 1) + A.ay[c] * X(c + g.sy) +
                 A.ay[c - g.sy] * X(c - g.sy);
      if (THREE_D) y += A.az[c] * X(c + g.sz) + A.az[c - g.sz] * X(c - g.sz);
      const double rv = fx ? 0.0 : b[c] - y;
      if (d) xout[c] = xc;
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
  if (last_block(&ctl->counter[3], nblocks, &flag)) {
    const double rr = block_sum_partials<NT>(part, nblocks, lds);
    const double bb = first ? block_sum_partials<NT>(part + nblocks, nblocks, lds) : 0.0;
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
    }
  }
}

// x = 0 on non-fixed interior rows (rejected initial guess).
template <bool THREE_D>
__global__ __launch_bounds__(NT) void k_zero_interior(Geo g, int kchunk, double* x,
                                                       const unsigned char* fixed) {
  const Tile t = tile_of(g, kchunk);
  if (!t.valid) return;
  long long c = base_index<THREE_D>(g, t);
  for (int k = t.k0; k < t.k1; ++k, c += g.sz)
    if (!(fixed != nullptr && fixed[c] != 0)) x[c] = 0.0;
}

}  // namespace

Krylov::Krylov(const Geo& g, bool three_d, hipStream_t stream, Profiler* prof)
    : g_(g), three_d_(three_d), s_(stream), prof_(prof) {
  if (!three_d_) g_.nz = 1;
  n_ = three_d_ ? g_.ncell() : g_.nsurf();
  set_kchunk(three_d_ ? g_.nz : 1);
  const size_t bytes = (size_t)n_ * sizeof(double);
  for (double** p : {&r_, &z_, &p0_, &p1_, &q_, &xw_}) {
    FREHG_HIP(hipMalloc(p, bytes));
    FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
  }
  FREHG_HIP(hipMalloc(&ctl_, 2 * sizeof(KrylovCtl)));
  FREHG_HIP(hipMemsetAsync(ctl_, 0, 2 * sizeof(KrylovCtl), s_));
  FREHG_HIP(hipHostMalloc(&host_ctl_, sizeof(KrylovCtl), hipHostMallocDefault));
  FREHG_HIP(hipStreamSynchronize(s_));
}

Krylov::~Krylov() {
  for (void* p : {(void*)r_, (void*)z_, (void*)p0_, (void*)p1_, (void*)q_, (void*)xw_, (void*)b32_,
                  (void*)r32_, (void*)z32_, (void*)p032_, (void*)p132_, (void*)q32_, (void*)d32_,
                  (void*)part_, (void*)ctl_})
    if (p) (void)hipFree(p);
  if (host_ctl_) (void)hipHostFree(host_ctl_);
}

void Krylov::set_kchunk(int kc) {
  kc = std::max(1, std::min(kc, three_d_ ? g_.nz : 1));
  if (kc == kchunk_ && part_) return;
  kchunk_ = kc;
  grid_ = dim3(ceil_div(g_.nx, TX), ceil_div(g_.ny, TY), ceil_div(three_d_ ? g_.nz : 1, kchunk_));
  const unsigned int nb = grid_.x * grid_.y * grid_.z;
  if (nb > nblocks_) {
    if (part_) FREHG_HIP(hipFree(part_));
    FREHG_HIP(hipMalloc(&part_, 3 * (size_t)nb * sizeof(double)));
  }
  nblocks_ = nb;
}

void Krylov::ensure_mixed() {
  if (b32_) return;
  const size_t bytes = (size_t)n_ * sizeof(float);
  for (float** p : {&b32_, &r32_, &z32_, &p032_, &p132_, &q32_, &d32_}) {
    FREHG_HIP(hipMalloc(p, bytes));
    FREHG_HIP(hipMemsetAsync(*p, 0, bytes, s_));
  }
}

void Krylov::read_ctl(KrylovCtl* dev, KrylovCtl* host) {
  FREHG_HIP(hipMemcpyAsync(host, dev, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
}

template <typename T>
int Krylov::run_pcg(const StencilView<T>& A, const T* b, const T* bn, T* x, const unsigned char* fixed,
                    bool x0_zero, bool guard, double rtol, double atol, int max_it, KrylovCtl* ctl,
                    T* r, T* z, T* p0, T* p1, T* q, bool* rejected, double* res_norm,
                    double* rhs_norm, bool* converged, int kt_spmv) {
  const dim3 blk(NT);
  const unsigned int nb = nblocks_;
  auto init = [&](bool zero) {
    if (three_d_)
      hipLaunchKernelGGL((k_init<T, true>), grid_, blk, 0, s_, g_, kchunk_, A, b, bn, x, r, z, p0, fixed,
                         zero ? 1 : 0, part_, ctl, nb, rtol, atol, max_it, guard ? 1 : 0);
    else
      hipLaunchKernelGGL((k_init<T, false>), grid_, blk, 0, s_, g_, kchunk_, A, b, bn, x, r, z, p0, fixed,
                         zero ? 1 : 0, part_, ctl, nb, rtol, atol, max_it, guard ? 1 : 0);
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
  T* pbuf[2] = {p0, p1};
  int parity = 0;
  int est = (sizeof(T) == 8) ? last_iters_ : last_inner_iters_;
  int batch = std::max(2, est + 1);
  int launched = 0;
  for (;;) {
    for (int it = 0; it < batch; ++it) {
      T* pold = pbuf[parity];
      T* pnew = pbuf[parity ^ 1];
      prof_->begin(kt_spmv);
      if (three_d_)
        hipLaunchKernelGGL((k_spmv<T, true>), grid_, blk, 0, s_, g_, kchunk_, A, z, pold, pnew, q, part_, ctl, nb);
      else
        hipLaunchKernelGGL((k_spmv<T, false>), grid_, blk, 0, s_, g_, kchunk_, A, z, pold, pnew, q, part_, ctl, nb);
      FREHG_LAUNCH_CHECK();
      prof_->end();
      prof_->begin(2);
      if (three_d_)
        hipLaunchKernelGGL((k_update<T, true>), grid_, blk, 0, s_, g_, kchunk_, A.D, pnew, q, x, r, z, part_, ctl, nb);
      else
        hipLaunchKernelGGL((k_update<T, false>), grid_, blk, 0, s_, g_, kchunk_, A.D, pnew, q, x, r, z, part_, ctl, nb);
      FREHG_LAUNCH_CHECK();
      prof_->end();
      parity ^= 1;
    }
    launched += batch;
    read_ctl(ctl, host_ctl_);
    if (host_ctl_->done) break;
    if (launched >= max_it + 2) break;
    batch = std::max(2, std::min(16, batch / 2));
  }
  if (host_ctl_->breakdown) throw ConvergenceError("Krylov: breakdown (p.Ap <= 0 or non-finite)");
  const int iters = host_ctl_->iters;
  // each launched iteration recorded one spmv and one update launch
  if (launched > iters) prof_->discard_tail(2 * (size_t)(launched - iters));
  if (sizeof(T) == 8) last_iters_ = iters;
  else last_inner_iters_ = iters;
  *res_norm = std::sqrt(host_ctl_->zz);
  *rhs_norm = std::sqrt(host_ctl_->bb);
  *converged = host_ctl_->converged != 0;
  return iters;
}

SolveResult Krylov::solve(const StencilView<double>& A, const StencilView<float>* A32, const double* b,
                          const double* bn, double* x, const unsigned char* fixed, bool x0_zero,
                          bool guard_guess, const KrylovConfig& cfg) {
  SolveResult res;
  if (cfg.precision == 0 || A32 == nullptr) {
    bool rej = false, conv = false;
    res.iterations = run_pcg<double>(A, b, bn, x, fixed, x0_zero, guard_guess, cfg.rtol, cfg.atol, cfg.max_it,
                                     ctl_, r_, z_, p0_, p1_, q_, &rej, &res.residual_norm, &res.rhs_norm,
                                     &conv, 0);
    res.guess_rejected = rej;
    res.converged = conv;
    if (!conv) throw ConvergenceError("Krylov: not converged within max_it");
    return res;
  }
  // FP32 inner / FP64 outer iterative refinement.
  ensure_mixed();
  const dim3 blk(NT);
  const unsigned int nb = nblocks_;
  KrylovCtl* outer = ctl_;
  KrylovCtl* inner = ctl_ + 1;
  if (x0_zero) {
    if (three_d_) hipLaunchKernelGGL((k_zero_interior<true>), grid_, blk, 0, s_, g_, kchunk_, x, fixed);
    else hipLaunchKernelGGL((k_zero_interior<false>), grid_, blk, 0, s_, g_, kchunk_, x, fixed);
    FREHG_LAUNCH_CHECK();
  }
  FREHG_HIP(hipMemcpyAsync(xw_, x, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  double* xcur = x;
  double* xalt = xw_;
  auto refine = [&](const double* xin, const float* d, double* xout, bool first) {
    if (three_d_)
      hipLaunchKernelGGL((k_refine<true>), grid_, blk, 0, s_, g_, kchunk_, A, b, bn, xin, d, xout, b32_, fixed,
                         part_, outer, nb, first ? 1 : 0, cfg.rtol, cfg.atol, (guard_guess && !x0_zero) ? 1 : 0);
    else
      hipLaunchKernelGGL((k_refine<false>), grid_, blk, 0, s_, g_, kchunk_, A, b, bn, xin, d, xout, b32_, fixed,
                         part_, outer, nb, first ? 1 : 0, cfg.rtol, cfg.atol, (guard_guess && !x0_zero) ? 1 : 0);
    FREHG_LAUNCH_CHECK();
  };
  refine(xcur, nullptr, xcur, true);
  read_ctl(outer, host_ctl_);
  if (host_ctl_->guess_rejected) {
    res.guess_rejected = true;
    if (three_d_) hipLaunchKernelGGL((k_zero_interior<true>), grid_, blk, 0, s_, g_, kchunk_, xcur, fixed);
    else hipLaunchKernelGGL((k_zero_interior<false>), grid_, blk, 0, s_, g_, kchunk_, xcur, fixed);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(xw_, x, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    refine(xcur, nullptr, xcur, true);
    read_ctl(outer, host_ctl_);
  }
  res.rhs_norm = std::sqrt(host_ctl_->bb);
  res.residual_norm = std::sqrt(host_ctl_->zz);
  bool conv = host_ctl_->converged != 0;
  const StencilView<float> A32v = *A32;
  while (!conv && res.refinements < cfg.max_refine) {
    bool rej = false, iconv = false;
    double rn = 0, bnrm = 0;
    res.iterations += run_pcg<float>(A32v, b32_, nullptr, d32_, fixed, true, false, cfg.inner_rtol, 0.0,
                                     cfg.max_it, inner, r32_, z32_, p032_, p132_, q32_, &rej, &rn, &bnrm,
                                     &iconv, 1);
    refine(xcur, d32_, xalt, false);
    std::swap(xcur, xalt);
    res.refinements += 1;
    read_ctl(outer, host_ctl_);
    res.residual_norm = std::sqrt(host_ctl_->zz);
    conv = host_ctl_->converged != 0;
  }
  if (xcur != x) FREHG_HIP(hipMemcpyAsync(x, xcur, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  if (!conv) {
    // Refinement stalled: finish in FP64 from the current iterate.
    bool rej = false, c64 = false;
    res.fallback_fp64 = true;
    res.iterations += run_pcg<double>(A, b, bn, x, fixed, false, false, cfg.rtol, cfg.atol, cfg.max_it, ctl_,
                                      r_, z_, p0_, p1_, q_, &rej, &res.residual_norm, &res.rhs_norm, &c64, 0);
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
