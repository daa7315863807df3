// Passive solute on the GPU (see solute.hpp). Tiles of 64 (x) x 4 (y) cells per 256-thread
// block, one soil layer per block in z, as the flow kernels; sums are block partials reduced by
// one block in a fixed order, maxima are order-free.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "re.hpp"
#include "solute.hpp"
#include "swe.hpp"

namespace frehg {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ int tile_i() { return blockIdx.x * 64 + (threadIdx.x & 63); }
__device__ __forceinline__ int tile_j() { return blockIdx.y * 4 + (threadIdx.x >> 6); }
__device__ __forceinline__ unsigned int tile_block() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

__device__ __forceinline__ double block_max(double v, double* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double m = lds[0];
#pragma unroll
  for (int q = 1; q < NT / kWave; ++q) m = fmax(m, lds[q]);
  return m;
}

// Rain mixing (Orchestrator.cpp:1001-1023): cells of every row but the last, wet at the start of
// the step (depth now minus the rain height the surface step added), mix in rain of c_rain.
__global__ __launch_bounds__(NT) void k_sol_rain(Geo g, const double* __restrict__ eta,
                                                 const double* __restrict__ bot, double dw, double c_rain,
                                                 double min_depth, double area, double* __restrict__ cs,
                                                 double* __restrict__ part) {
  __shared__ double lds[NT / kWave];
  const int i = tile_i(), j = tile_j();
  double added = 0.0;
  if (i < g.nx && j < g.ny - 1) {
    const long long c = g.sidx(i, j);
    const double post = fmax(0.0, eta[c] - bot[c]);
    const double d0 = fmax(0.0, post - dw);
    if (d0 > min_depth) {
      cs[c] = (d0 * cs[c] + dw * c_rain) / (d0 + dw);
      added = dw * c_rain * area;
    }
  }
  const double s = block_sum<NT>(added, lds);
  if (threadIdx.x == 0) part[tile_block()] = s;
}

// C *= fac on interior cells (applyDecaySurface / applyDecaySubsurface, SourceSink.cpp:60-86).
__global__ __launch_bounds__(NT) void k_sol_scale(Geo g, int layers, double fac, double* __restrict__ c) {
  const int i = tile_i(), j = tile_j(), k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const long long q = layers > 1 ? g.idx(i, j, k) : g.sidx(i, j);
  c[q] *= fac;
}

// Largest sum of outgoing face Courant numbers (Advection.cpp:39-57, 116-137); interior faces
// only. part[nb + b] = 0 (the unused minimum of k_reduce_maxmin).
__global__ __launch_bounds__(NT) void k_sol_courant(Geo g, int three, const double* __restrict__ fx,
                                                    const double* __restrict__ fy, const double* __restrict__ fz,
                                                    double dt, double dx, double dy, double dz,
                                                    double* __restrict__ part, unsigned int nb) {
  __shared__ double lds[NT / kWave];
  const int i = tile_i(), j = tile_j(), k = blockIdx.z;
  double c = 0.0;
  if (i < g.nx && j < g.ny) {
    if (three) {
      const double qE = (i + 1 < g.nx) ? fx[g.idx(i, j, k)] : 0.0;
      const double qW = (i - 1 >= 0) ? fx[g.idx(i - 1, j, k)] : 0.0;
      const double qN = (j + 1 < g.ny) ? fy[g.idx(i, j, k)] : 0.0;
      const double qS = (j - 1 >= 0) ? fy[g.idx(i, j - 1, k)] : 0.0;
      const double qD = (k + 1 < g.nz) ? fz[g.idx(i, j, k)] : 0.0;
      const double qU = (k - 1 >= 0) ? fz[g.idx(i, j, k - 1)] : 0.0;
      const double ox = fmax(qE, 0.0) + fmax(-qW, 0.0);
      const double oy = fmax(qN, 0.0) + fmax(-qS, 0.0);
      const double oz = fmax(qD, 0.0) + fmax(-qU, 0.0);
      c = ox * dt / dx + oy * dt / dy + oz * dt / dz;
    } else {
      const double uE = (i + 1 < g.nx) ? fx[g.sidx(i, j)] : 0.0;
      const double uW = (i - 1 >= 0) ? fx[g.sidx(i - 1, j)] : 0.0;
      const double vN = (j + 1 < g.ny) ? fy[g.sidx(i, j)] : 0.0;
      const double vS = (j - 1 >= 0) ? fy[g.sidx(i, j - 1)] : 0.0;
      const double ox = fmax(uE, 0.0) + fmax(-uW, 0.0);
      const double oy = fmax(vN, 0.0) + fmax(-vS, 0.0);
      c = ox * dt / dx + oy * dt / dy;
    }
  }
  const double m = block_max(c, lds);
  if (threadIdx.x == 0) {
    part[tile_block()] = m;
    part[nb + tile_block()] = 0.0;
  }
}

__device__ __forceinline__ double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

// Upwinded face value, MUSCL-reconstructed with the minmod slope (Advection.cpp:16-27).
__device__ __forceinline__ double face_value(int muscl, double uf, double cLL, double cL, double cR, double cRR) {
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

// Conservative update from the face fluxes (Advection.cpp:67-100, 147-194); `co` holds the
// concentrations before the update.
__global__ __launch_bounds__(NT) void k_sol_advect(Geo g, int three, int muscl, const double* __restrict__ fx,
                                                   const double* __restrict__ fy, const double* __restrict__ fz,
                                                   double dt, double dx, double dy, double dz,
                                                   const double* __restrict__ co, double* __restrict__ c) {
  const int i = tile_i(), j = tile_j(), k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = three ? g.nz : 1;
  auto at = [&](int ii, int jj, int kk) -> long long { return three ? g.idx(ii, jj, kk) : g.sidx(ii, jj); };
  auto C = [&](int ii, int jj, int kk) { return co[at(ii, jj, kk)]; };
  double Fe = 0.0, Fw = 0.0, Fn = 0.0, Fs = 0.0, Fd = 0.0, Fu = 0.0;
  if (i + 1 < nx) {
    const double uf = fx[at(i, j, k)];
    const double cLL = (i - 1 >= 0) ? C(i - 1, j, k) : C(i, j, k);
    const double cRR = (i + 2 < nx) ? C(i + 2, j, k) : C(i + 1, j, k);
    Fe = uf * face_value(muscl, uf, cLL, C(i, j, k), C(i + 1, j, k), cRR);
  }
  if (i - 1 >= 0) {
    const double uf = fx[at(i - 1, j, k)];
    const double cLL = (i - 2 >= 0) ? C(i - 2, j, k) : C(i - 1, j, k);
    const double cRR = (i + 1 < nx) ? C(i + 1, j, k) : C(i, j, k);
    Fw = uf * face_value(muscl, uf, cLL, C(i - 1, j, k), C(i, j, k), cRR);
  }
  if (j + 1 < ny) {
    const double uf = fy[at(i, j, k)];
    const double cLL = (j - 1 >= 0) ? C(i, j - 1, k) : C(i, j, k);
    const double cRR = (j + 2 < ny) ? C(i, j + 2, k) : C(i, j + 1, k);
    Fn = uf * face_value(muscl, uf, cLL, C(i, j, k), C(i, j + 1, k), cRR);
  }
  if (j - 1 >= 0) {
    const double uf = fy[at(i, j - 1, k)];
    const double cLL = (j - 2 >= 0) ? C(i, j - 2, k) : C(i, j - 1, k);
    const double cRR = (j + 1 < ny) ? C(i, j + 1, k) : C(i, j, k);
    Fs = uf * face_value(muscl, uf, cLL, C(i, j - 1, k), C(i, j, k), cRR);
  }
  if (three) {
    if (k + 1 < nz) {
      const double uf = fz[at(i, j, k)];
      const double cLL = (k - 1 >= 0) ? C(i, j, k - 1) : C(i, j, k);
      const double cRR = (k + 2 < nz) ? C(i, j, k + 2) : C(i, j, k + 1);
      Fd = uf * face_value(muscl, uf, cLL, C(i, j, k), C(i, j, k + 1), cRR);
    }
    if (k - 1 >= 0) {
      const double uf = fz[at(i, j, k - 1)];
      const double cLL = (k - 2 >= 0) ? C(i, j, k - 2) : C(i, j, k - 1);
      const double cRR = (k + 1 < nz) ? C(i, j, k + 1) : C(i, j, k);
      Fu = uf * face_value(muscl, uf, cLL, C(i, j, k - 1), C(i, j, k), cRR);
    }
    c[at(i, j, k)] = C(i, j, k) - dt / dx * (Fe - Fw) - dt / dy * (Fn - Fs) - dt / dz * (Fd - Fu);
  } else {
    c[at(i, j, k)] = C(i, j, k) - dt / dx * (Fe - Fw) - dt / dy * (Fn - Fs);
  }
}

// The implicit diffusion operator I - dt D L with Neumann walls (Diffusion.cpp:38-78, 105-152):
// diag 1 + the coefficients of the faces that exist, -c on each interior face. Ghost entries
// stay zero.
__global__ __launch_bounds__(NT) void k_sol_diff_op(Geo g, int three, double cx, double cy, double cz,
                                                    double* __restrict__ D, double* __restrict__ ax,
                                                    double* __restrict__ ay, double* __restrict__ az) {
  const int i = tile_i(), j = tile_j(), k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const long long c = three ? g.idx(i, j, k) : g.sidx(i, j);
  double diag = 1.0;
  if (i - 1 >= 0) diag += cx;
  if (i + 1 < g.nx) diag += cx;
  if (j - 1 >= 0) diag += cy;
  if (j + 1 < g.ny) diag += cy;
  if (three) {
    if (k - 1 >= 0) diag += cz;
    if (k + 1 < g.nz) diag += cz;
    az[c] = (k + 1 < g.nz) ? -cz : 0.0;
  }
  D[c] = diag;
  ax[c] = (i + 1 < g.nx) ? -cx : 0.0;
  ay[c] = (j + 1 < g.ny) ? -cy : 0.0;
}

// applyInterfaceExchange (SourceSink.cpp:22-53) with the column's exchanged height
// vex = -q dt / area (+ = infiltration), the post-exchange depth and top-cell water height.
__global__ __launch_bounds__(NT) void k_sol_exchange(Geo gs, Geo gg, const double* __restrict__ q, double dt,
                                                     double area, double dz, const double* __restrict__ eta,
                                                     const double* __restrict__ bot, const double* __restrict__ wc,
                                                     double* __restrict__ cs, double* __restrict__ cg) {
  const int i = tile_i(), j = tile_j();
  if (i >= gs.nx || j >= gs.ny) return;
  const long long c = gs.sidx(i, j), t = gg.idx(i, j, 0);
  const double v = -q[(size_t)i + (size_t)j * gs.nx] * dt / area;
  if (v > 0.0) {
    const double h = wc[t] * dz;
    if (h > 0.0) {
      double h0 = h - v;
      if (h0 < 0.0) h0 = 0.0;
      cg[t] = (cg[t] * h0 + cs[c] * v) / h;
    }
  } else if (v < 0.0) {
    const double s = -v;
    const double d = fmax(0.0, eta[c] - bot[c]);
    if (d > 0.0) {
      double d0 = d - s;
      if (d0 < 0.0) d0 = 0.0;
      cs[c] = (cs[c] * d0 + cg[t] * s) / d;
    }
  }
}

// Block partials of the solute mass: surface C max(eta - bottom, 0) area, soil C theta vol.
__global__ __launch_bounds__(NT) void k_sol_mass(Geo g, int three, const double* __restrict__ c,
                                                 const double* __restrict__ eta, const double* __restrict__ bot,
                                                 const double* __restrict__ wc, double w, double* __restrict__ part) {
  __shared__ double lds[NT / kWave];
  const int i = tile_i(), j = tile_j(), k = blockIdx.z;
  double m = 0.0;
  if (i < g.nx && j < g.ny) {
    if (three) {
      const long long q = g.idx(i, j, k);
      m = c[q] * wc[q] * w;
    } else {
      const long long q = g.sidx(i, j);
      m = c[q] * fmax(0.0, eta[q] - bot[q]) * w;
    }
  }
  const double s = block_sum<NT>(m, lds);
  if (threadIdx.x == 0) part[tile_block()] = s;
}

}  // namespace

Solute::Solute(const frehg_solute_params& p, SweState* swe, ReState* re, hipStream_t s, Profiler* prof)
    : p_(p), swe_(swe), re_(re), s_(s), prof_(prof) {
  if (!swe && !re) throw ArgError("frehg_solute_init: initialise the surface and/or groundwater solver first");
  if (!(p.cfl_max > 0.0)) throw ArgError("frehg_solute_init: cfl_max must be > 0");
  if (p.D < 0.0 || p.k_decay < 0.0) throw ArgError("frehg_solute_init: D and k_decay must be >= 0");
  if (p.substeps < 1) throw ArgError("frehg_solute_init: substeps must be >= 1");
  (void)krylov_config(p.solver, "frehg_solute_init");
  if (p.solver.precision != FREHG_PREC_FP64) throw ArgError("frehg_solute_init: the diffusion solve runs in FP64");
  try {
    if (swe) {
      gs_ = swe->geo();
      dx_ = swe->dx();
      dy_ = swe->dy();
    }
    if (re) {
      gg_ = re->geo();
      dx_ = re->params().dx;
      dy_ = re->params().dy;
      dz_ = re->params().dz;
      if (!swe) gs_ = make_geo(gg_.nx, gg_.ny, 1);
    }
    if (swe && re && (gs_.nx != gg_.nx || gs_.ny != gg_.ny))
      throw ArgError("frehg_solute_init: surface and groundwater grids differ");
    ns_ = gs_.nsurf();
    ng_ = re ? gg_.ncell() : 0;
    const unsigned int nb2 = ceil_div(gs_.nx, 64) * ceil_div(gs_.ny, 4);
    nb_ = re ? nb2 * (unsigned)gg_.nz : nb2;
    FREHG_HIP(hipMalloc(&part_, 2 * (size_t)nb_ * sizeof(double)));
    FREHG_HIP(hipMalloc(&dres_, 4 * sizeof(double)));
    FREHG_HIP(hipHostMalloc(&hres_, 4 * sizeof(double), hipHostMallocDefault));
    auto zeros = [&](double** p, long long n) {
      FREHG_HIP(hipMalloc(p, (size_t)n * sizeof(double)));
      FREHG_HIP(hipMemsetAsync(*p, 0, (size_t)n * sizeof(double), s_));
    };
    if (swe) {
      for (double** f : {&cs_, &cs_old_, &Ds_, &axs_, &ays_}) zeros(f, ns_);
      if (p.implicit_diffusion && p.D > 0.0) kry_s_.reset(new Krylov(gs_, false, s_, prof_));
    }
    if (re) {
      for (double** f : {&cg_, &cg_old_, &Dg_, &axg_, &ayg_, &azg_}) zeros(f, ng_);
      if (p.implicit_diffusion && p.D > 0.0) kry_g_.reset(new Krylov(gg_, true, s_, prof_));
    }
    FREHG_HIP(hipStreamSynchronize(s_));
  } catch (...) {
    for (void* q : {(void*)cs_, (void*)cs_old_, (void*)Ds_, (void*)axs_, (void*)ays_, (void*)cg_, (void*)cg_old_,
                    (void*)Dg_, (void*)axg_, (void*)ayg_, (void*)azg_, (void*)part_, (void*)dres_})
      if (q) (void)hipFree(q);
    if (hres_) (void)hipHostFree(hres_);
    throw;
  }
}

Solute::~Solute() {
  for (void* q : {(void*)cs_, (void*)cs_old_, (void*)Ds_, (void*)axs_, (void*)ays_, (void*)cg_, (void*)cg_old_,
                  (void*)Dg_, (void*)axg_, (void*)ayg_, (void*)azg_, (void*)part_, (void*)dres_})
    if (q) (void)hipFree(q);
  if (hres_) (void)hipHostFree(hres_);
}

void Solute::set_field(const std::string& name, const double* src) {
  if (name == "surface" && cs_) copy_field_in(gs_, 1, cs_, src, s_);
  else if (name == "subsurface" && cg_) copy_field_in(gg_, gg_.nz + 2, cg_, src, s_);
  else throw ArgError("unknown or absent solute field '" + name + "'");
  FREHG_HIP(hipStreamSynchronize(s_));
}

void Solute::get_field(const std::string& name, double* dst) {
  if (name == "surface" && cs_) copy_field_out(gs_, 1, dst, cs_, s_);
  else if (name == "subsurface" && cg_) copy_field_out(gg_, gg_.nz + 2, dst, cg_, s_);
  else throw ArgError("unknown or absent solute field '" + name + "'");
  FREHG_HIP(hipStreamSynchronize(s_));
}

void Solute::set_initial(const double* surface, const double* subsurface) {
  if (surface && cs_) {
    std::vector<double> pad((size_t)(gs_.nx + 2) * (gs_.ny + 2), 0.0);
    for (int j = 0; j < gs_.ny; ++j)
      for (int i = 0; i < gs_.nx; ++i)
        pad[(size_t)(i + 1) + (size_t)(j + 1) * (gs_.nx + 2)] = surface[(size_t)i + (size_t)j * gs_.nx];
    set_field("surface", pad.data());
  }
  if (subsurface && cg_) {
    const size_t px = (size_t)gg_.nx + 2, py = (size_t)gg_.ny + 2;
    std::vector<double> pad(px * py * ((size_t)gg_.nz + 2), 0.0);
    for (int k = 0; k < gg_.nz; ++k)
      for (int j = 0; j < gg_.ny; ++j)
        for (int i = 0; i < gg_.nx; ++i)
          pad[(size_t)(i + 1) + (size_t)(j + 1) * px + (size_t)(k + 1) * px * py] =
              subsurface[(size_t)i + (size_t)j * gg_.nx + (size_t)k * gg_.nx * gg_.ny];
    set_field("subsurface", pad.data());
  }
}

double Solute::mass() {
  double total = 0.0;
  const dim3 blk(NT);
  if (swe_) {
    const dim3 grid(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4));
    hipLaunchKernelGGL(k_sol_mass, grid, blk, 0, s_, gs_, 0, (const double*)cs_, (const double*)swe_->d("eta"),
                       (const double*)swe_->d("bottom"), (const double*)nullptr, dx_ * dy_, part_);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_, grid.x * grid.y, 1,
                       dres_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    total += hres_[0];
  }
  if (re_) {
    const dim3 grid(ceil_div(gg_.nx, 64), ceil_div(gg_.ny, 4), gg_.nz);
    hipLaunchKernelGGL(k_sol_mass, grid, blk, 0, s_, gg_, 1, (const double*)cg_, (const double*)nullptr,
                       (const double*)nullptr, (const double*)re_->d_wc(), dx_ * dy_ * dz_, part_);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_,
                       grid.x * grid.y * grid.z, 1, dres_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    total += hres_[0];
  }
  return total;
}

void Solute::interface_exchange(const double* q, double dt, double area) {
  if (!swe_ || !re_) return;
  const dim3 grid(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4)), blk(NT);
  hipLaunchKernelGGL(k_sol_exchange, grid, blk, 0, s_, gs_, gg_, q, dt, area, dz_, (const double*)swe_->d("eta"),
                     (const double*)swe_->d("bottom"), (const double*)re_->d_wc(), cs_, cg_);
  FREHG_LAUNCH_CHECK();
}

// Largest Courant number of a substep of length dt over the domains present.
double Solute::max_courant(double dt) {
  double cfl = 0.0;
  const dim3 blk(NT);
  auto run = [&](const Geo& g, int three, const double* fx, const double* fy, const double* fz) {
    const dim3 grid(ceil_div(g.nx, 64), ceil_div(g.ny, 4), three ? g.nz : 1);
    const unsigned int nb = grid.x * grid.y * grid.z;
    hipLaunchKernelGGL(k_sol_courant, grid, blk, 0, s_, g, three, fx, fy, fz, dt, dx_, dy_, dz_, part_, nb);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_maxmin, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_, nb, dres_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    cfl = std::max(cfl, hres_[0]);
  };
  if (swe_) run(gs_, 0, swe_->d("uu"), swe_->d("vv"), nullptr);
  if (re_) run(gg_, 1, re_->d_qx(), re_->d_qy(), re_->d_qz());
  return cfl;
}

void Solute::decay(double fac) {
  const dim3 blk(NT);
  if (swe_) hipLaunchKernelGGL(k_sol_scale, dim3(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4)), blk, 0, s_, gs_, 1, fac, cs_);
  if (re_)
    hipLaunchKernelGGL(k_sol_scale, dim3(ceil_div(gg_.nx, 64), ceil_div(gg_.ny, 4), gg_.nz), blk, 0, s_, gg_,
                       gg_.nz + 2, fac, cg_);
  FREHG_LAUNCH_CHECK();
}

void Solute::advect() {
  const double dt = cur_dt_;
  const dim3 blk(NT);
  const int muscl = p_.muscl ? 1 : 0;
  if (swe_) {
    FREHG_HIP(hipMemcpyAsync(cs_old_, cs_, (size_t)ns_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    hipLaunchKernelGGL(k_sol_advect, dim3(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4)), blk, 0, s_, gs_, 0, muscl,
                       (const double*)swe_->d("uu"), (const double*)swe_->d("vv"), (const double*)nullptr, dt, dx_,
                       dy_, dz_, (const double*)cs_old_, cs_);
    FREHG_LAUNCH_CHECK();
  }
  if (re_) {
    FREHG_HIP(hipMemcpyAsync(cg_old_, cg_, (size_t)ng_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    hipLaunchKernelGGL(k_sol_advect, dim3(ceil_div(gg_.nx, 64), ceil_div(gg_.ny, 4), gg_.nz), blk, 0, s_, gg_, 1,
                       muscl, (const double*)re_->d_qx(), (const double*)re_->d_qy(), (const double*)re_->d_qz(), dt,
                       dx_, dy_, dz_, (const double*)cg_old_, cg_);
    FREHG_LAUNCH_CHECK();
  }
}

// DiffusionSolver::solve: (I - dt D L) C' = C, x0 = 0, C' into the concentration.
void Solute::diffuse(double dt) {
  if (!p_.implicit_diffusion || !(p_.D > 0.0) || !(dt > 0.0)) return;
  const KrylovConfig kc = krylov_config(p_.solver, "solute diffusion");
  const dim3 blk(NT);
  const double cx = p_.D * dt / (dx_ * dx_), cy = p_.D * dt / (dy_ * dy_), cz = p_.D * dt / (dz_ * dz_);
  if (swe_) {
    hipLaunchKernelGGL(k_sol_diff_op, dim3(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4)), blk, 0, s_, gs_, 0, cx, cy, cz,
                       Ds_, axs_, ays_, (double*)nullptr);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(cs_old_, cs_, (size_t)ns_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    const StencilView<double> A{Ds_, axs_, ays_, nullptr};
    const SolveResult r = kry_s_->solve(A, nullptr, cs_old_, nullptr, cs_, nullptr, true, false, kc);
    tot_.diffusion_iterations += r.iterations;
  }
  if (re_) {
    hipLaunchKernelGGL(k_sol_diff_op, dim3(ceil_div(gg_.nx, 64), ceil_div(gg_.ny, 4), gg_.nz), blk, 0, s_, gg_, 1, cx,
                       cy, cz, Dg_, axg_, ayg_, azg_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(cg_old_, cg_, (size_t)ng_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    const StencilView<double> A{Dg_, axg_, ayg_, azg_};
    const SolveResult r = kry_g_->solve(A, nullptr, cg_old_, nullptr, cg_, nullptr, true, false, kc);
    tot_.diffusion_iterations += r.iterations;
  }
}

// SoluteStepper::step with rain 0 (SoluteStepper.cpp:24-72): decay, advection, diffusion.
void Solute::substep(double dt) {
  cur_dt_ = dt;
  if (p_.k_decay > 0.0) decay(std::exp(-p_.k_decay * dt));
  advect();
  diffuse(dt);
}

// Orchestrator::applySolute (:938-1056).
void Solute::apply(double t, double dt, double rain) {
  if (!(dt > 0.0)) return;
  if (t + dt <= p_.start_time) return;
  if (re_) re_->materialize_fluxes();
  const double area = dx_ * dy_;
  if (swe_ && rain > 0.0 && p_.c_rain != 0.0) {
    const dim3 grid(ceil_div(gs_.nx, 64), ceil_div(gs_.ny, 4)), blk(NT);
    hipLaunchKernelGGL(k_sol_rain, grid, blk, 0, s_, gs_, (const double*)swe_->d("eta"),
                       (const double*)swe_->d("bottom"), rain * dt, p_.c_rain, p_.min_depth, area, cs_, part_);
    FREHG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_, grid.x * grid.y, 1,
                       dres_);
    FREHG_LAUNCH_CHECK();
    FREHG_HIP(hipMemcpyAsync(hres_, dres_, sizeof(double), hipMemcpyDeviceToHost, s_));
    FREHG_HIP(hipStreamSynchronize(s_));
    tot_.rain_added += hres_[0];
  }
  const double mass_before = p_.k_decay > 0.0 ? mass() : 0.0;
  // The flow is frozen over the step, so every substep has the Courant number of the first: a
  // refused count refuses at its first substep and leaves the solute as it was.
  int nsub = p_.substeps;
  const int cap = 4096;
  for (;;) {
    const double sub = dt / (double)nsub;
    const double cfl = max_courant(sub);
    if (!(cfl > p_.cfl_max)) {
      for (int q = 0; q < nsub; ++q) substep(sub);
      tot_.substeps += nsub;
      tot_.max_courant = std::max(tot_.max_courant, cfl);
      break;
    }
    if (nsub < cap) {
      nsub *= 2;
    } else {
      tot_.refused_steps += 1;
      break;
    }
  }
  if (p_.k_decay > 0.0) tot_.decay_loss += std::max(0.0, mass_before - mass());
}

}  // namespace frehg
