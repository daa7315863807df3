// SW <-> GW exchange on the GPU: Coupling::exchange (src/coupling/Coupling.cpp:145-164) as one
// per-column kernel (computeExchangeRates :49-57, limitExchange :78-98, applyExchangeToSurface
// :113-118, applyExchangeToGroundwater :134-142) with a fixed-order FP64 volume reduction.
// Soil is class 0 (re.params().soil, Coupling.cpp:33,68,127), as in the reference.
#include "coupling.hpp"

namespace frehg {

namespace {

constexpr int NT = 256;

struct CpK {
  Geo gs;  // surface
  Geo gg;  // subsurface
  SoilClass soil;
  double Az, dz, visc, min_depth, dt;
};

// hmod (optional) receives 1 at the top cells whose head the exchange changes.
__global__ __launch_bounds__(NT) void k_exchange(CpK P, double* eta, const double* bottom, double* dept, double* wc,
                                                 double* h, unsigned char* hmod, double* q_out, double* part) {
  __shared__ double lds[NT / kWave];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  double qdt = 0.0;
  if (i < P.gs.nx && j < P.gs.ny) {
    const long long si = P.gs.sidx(i, j);
    const long long ti = P.gg.idx(i, j, 0);
    const SoilClass& s = P.soil;
    const double e = eta[si], b = bottom[si];
    const double dp = fmax(e - b, 0.0);
    const double htop = h[ti];
    double qv = P.Az * s.Ks_z * P.visc * (htop - dp) / (0.5 * P.dz);
    if (dp <= P.min_depth && qv < 0.0) qv = 0.0;
    if (qv != 0.0) {
      const double V = P.Az * P.dz;
      const double wct = wc[ti];
      if (qv < 0.0) {
        const double q_min_surf = -dp * P.Az / P.dt;
        const double q_min_gw = -(s.theta_s - wct) * V / P.dt;
        qv = fmax(qv, fmax(q_min_surf, q_min_gw));
      } else {
        const double q_max_gw = (wct - s.theta_r) * V / P.dt;
        qv = fmin(qv, q_max_gw);
      }
    }
    if (q_out) q_out[(size_t)i + (size_t)j * P.gs.nx] = qv;
    const double en = e + qv * P.dt / P.Az;
    eta[si] = en;
    dept[si] = fmax(en - b, 0.0);
    const double V = P.Az * P.dz;
    double wn = wc[ti] - qv * P.dt / V;
    wn = fmin(fmax(wn, s.theta_r), s.theta_s);
    wc[ti] = wn;
    const double hnew = vg_head(s, wn);
    if (hmod && hnew != htop) hmod[ti] = 1;
    h[ti] = hnew;
    qdt = qv * P.dt;
  }
  const double sblk = block_sum<NT>(qdt, lds);
  const unsigned int bid = blockIdx.x + gridDim.x * blockIdx.y;
  if (threadIdx.x == 0) part[bid] = sblk;
}

// Orchestrator::computeWaterVolume (src/driver/Orchestrator.cpp:898-936).
__global__ __launch_bounds__(NT) void k_volume(Geo gs, Geo gg, const double* eta, const double* bottom,
                                               const double* wc, double area, double vol, double* part,
                                               unsigned int nblocks) {
  __shared__ double lds[NT / kWave];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  double sw = 0.0, gw = 0.0;
  if (i < gs.nx && j < gs.ny) {
    if (eta) {
      const long long si = gs.sidx(i, j);
      sw = fmax(0.0, eta[si] - bottom[si]) * area;
    }
    if (wc) {
      for (int k = 0; k < gg.nz; ++k) gw += wc[gg.idx(i, j, k)] * vol;
    }
  }
  const double a = block_sum<NT>(sw, lds);
  const double b = block_sum<NT>(gw, lds);
  const unsigned int bid = blockIdx.x + gridDim.x * blockIdx.y;
  if (threadIdx.x == 0) {
    part[bid] = a;
    part[nblocks + bid] = b;
  }
}

}  // namespace

Coupler::Coupler(hipStream_t s) : s_(s) {
  FREHG_HIP(hipMalloc(&dres_, 4 * sizeof(double)));
  FREHG_HIP(hipHostMalloc(&hres_, 4 * sizeof(double), hipHostMallocDefault));
  FREHG_HIP(hipMalloc(&counter_, 16));
  FREHG_HIP(hipMemsetAsync(counter_, 0, 16, s_));
}

Coupler::~Coupler() {
  for (void* p : {(void*)dres_, (void*)counter_, (void*)part_, (void*)qdev_})
    if (p) (void)hipFree(p);
  if (hres_) (void)hipHostFree(hres_);
}

void Coupler::ensure(unsigned int nblocks, long long ncol) {
  if (nblocks > nblk_) {
    if (part_) FREHG_HIP(hipFree(part_));
    FREHG_HIP(hipMalloc(&part_, 2 * (size_t)nblocks * sizeof(double)));
    nblk_ = nblocks;
  }
  if (ncol > nq_) {
    if (qdev_) FREHG_HIP(hipFree(qdev_));
    FREHG_HIP(hipMalloc(&qdev_, (size_t)ncol * sizeof(double)));
    nq_ = ncol;
  }
}

void Coupler::exchange_async(SweState& swe, ReState& re, const frehg_coupling_params& cp, double dt) {
  const Geo& gs = swe.geo();
  const Geo& gg = re.geo();
  if (gs.nx != gg.nx || gs.ny != gg.ny) throw ArgError("frehg_couple_exchange: surface and subsurface grids differ");
  if (!(dt > 0)) throw ArgError("frehg_couple_exchange: dt must be > 0");
  CpK P;
  P.gs = gs;
  P.gg = gg;
  const frehg_soil_params& sp = re.params().soil;
  P.soil = make_soil_class(sp.alpha, sp.n, sp.theta_s, sp.theta_r, sp.Ks_x, sp.Ks_y, sp.Ks_z, sp.Ss, sp.use_vg,
                           sp.use_mvg, sp.air_entry);
  P.Az = cp.dx * cp.dy;
  P.dz = re.params().dz;
  P.visc = cp.visc;
  P.min_depth = cp.min_depth;
  P.dt = dt;
  const dim3 grid(ceil_div(gs.nx, 64), ceil_div(gs.ny, 4)), blk(NT);
  const unsigned int nb = grid.x * grid.y;
  ensure(nb, (long long)gs.nx * gs.ny);
  hipLaunchKernelGGL(k_exchange, grid, blk, 0, s_, P, swe.d("eta"), swe.d("bottom"), swe.d("dept"), re.d_wc(),
                     re.d_h(), re.d_hmod(), (want_q_ || capture_q_) ? qdev_ : nullptr, part_);
  FREHG_LAUNCH_CHECK();
  // slot 2: volume() uses slots 0 and 1 and may run before this result is read
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_, nb, 1, dres_ + 2);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(hres_ + 2, dres_ + 2, sizeof(double), hipMemcpyDeviceToHost, s_));
}

double Coupler::exchanged_volume() {
  FREHG_HIP(hipStreamSynchronize(s_));  // returns at once when the stream has already drained
  return hres_[2];
}

double Coupler::exchange(SweState& swe, ReState& re, const frehg_coupling_params& cp, double dt, double* q_out) {
  want_q_ = q_out != nullptr;
  try {
    exchange_async(swe, re, cp, dt);
  } catch (...) {
    want_q_ = false;
    throw;
  }
  want_q_ = false;
  const Geo& gs = swe.geo();
  if (q_out)
    FREHG_HIP(hipMemcpyAsync(q_out, qdev_, (size_t)gs.nx * gs.ny * sizeof(double), hipMemcpyDeviceToHost, s_));
  return exchanged_volume();
}

void Coupler::volume(SweState* swe, ReState* re, double* surface, double* subsurface) {
  const Geo& g = swe ? swe->geo() : re->geo();
  const double dx = swe ? swe->dx() : re->params().dx;
  const double dy = swe ? swe->dy() : re->params().dy;
  Geo gg = re ? re->geo() : g;
  const dim3 grid(ceil_div(g.nx, 64), ceil_div(g.ny, 4)), blk(NT);
  const unsigned int nb = grid.x * grid.y;
  ensure(nb, 1);
  hipLaunchKernelGGL(k_volume, grid, blk, 0, s_, g, gg, swe ? swe->d("eta") : nullptr,
                     swe ? swe->d("bottom") : nullptr, re ? re->d_wc() : nullptr, dx * dy,
                     re ? dx * dy * re->params().dz : 0.0, part_, nb);
  FREHG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s_, (const double*)part_, nb, 2, dres_);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipMemcpyAsync(hres_, dres_, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
  FREHG_HIP(hipStreamSynchronize(s_));
  if (surface) *surface = hres_[0];
  if (subsurface) *subsurface = hres_[1];
}

}  // namespace frehg
