// Shared device/host helpers of the gfx950 backend: error handling, wavefront/block reductions,
// Van Genuchten constitutive relations and the per-cell soil view.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace frehg {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ConvergenceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define FREHG_HIP(call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::frehg::HipError(std::string(#call) + " failed: " + hipGetErrorString(e_) + " (" + \
                              __FILE__ + ":" + std::to_string(__LINE__) + ")");                \
  } while (0)

#define FREHG_LAUNCH_CHECK() FREHG_HIP(hipGetLastError())

// Owning device allocation of n elements of T, freed on scope exit (also when an exception leaves
// the scope).
template <typename T>
class DeviceBuffer {
 public:
  explicit DeviceBuffer(size_t n) { FREHG_HIP(hipMalloc(&p_, n * sizeof(T))); }
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  T* get() const { return p_; }

 private:
  T* p_ = nullptr;
};

constexpr int kWave = 64;

// ---------------------------------------------------------------------------------------------
// Reductions. Block sizes are multiples of 64; every partial sum is formed in a fixed order so
// repeated runs are bit-identical.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}

// Sum over the block; result valid in every thread. `lds` holds >= blockDim/64 doubles.
template <int NT>
__device__ inline double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += lds[i];
  return s;
}

// N sums over the block with one round of LDS traffic; each result equals block_sum's bit for bit
// (the same wave sums, added in the same order). `lds` holds >= N * blockDim/64 doubles.
template <int NT, int N>
__device__ inline void block_sums(double (&v)[N], double* lds) {
  constexpr int NW = NT / kWave;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) lds[q * NW + w] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += lds[q * NW + i];
    v[q] = s;
  }
}

// Deterministic sum of n partials by one block of NT threads: each thread keeps eight
// accumulators over a fixed stride pattern (so eight loads are in flight, not one), then a tree.
template <int NT>
__device__ inline double block_sum_partials(const double* part, int n, double* lds) {
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 7 * NT < n; i += 8 * NT) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += part[i + u * NT];
  }
  for (; i < n; i += NT) a[0] += part[i];
  const double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  return block_sum<NT>(s, lds);
}

// One-block reductions of per-block partials (launched as <<<1, kRedThreads>>> after the
// kernel that wrote the partials). Cross-block completion is ordered by the stream, not by
// device-scope atomics, whose serialisation across the XCDs would cost ~0.1 us per block.
constexpr int kRedThreads = 1024;

__global__ __launch_bounds__(kRedThreads) inline void k_reduce_sum(const double* part, unsigned int n, int nsums,
                                                                     double* out) {
  __shared__ double lds[kRedThreads / kWave];
  for (int q = 0; q < nsums; ++q) {
    const double v = block_sum_partials<kRedThreads>(part + (size_t)q * n, (int)n, lds);
    if (threadIdx.x == 0) out[q] = v;
    __syncthreads();
  }
}

// Two-stage form for long rows: kRedStage blocks per sum each reduce one contiguous chunk, then
// one block reduces the chunk sums; both stages use fixed orders, so the result is reproducible.
constexpr unsigned int kRedStage = 64;

__global__ __launch_bounds__(256) inline void k_reduce_sum_stage(const double* part, unsigned int n, double* out) {
  __shared__ double lds[256 / kWave];
  const unsigned int q = blockIdx.y, chunk = (n + gridDim.x - 1) / gridDim.x;
  const unsigned int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double v = 0.0;
  for (unsigned int i = lo + threadIdx.x; i < hi; i += 256) v += part[(size_t)q * n + i];
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) out[q * gridDim.x + blockIdx.x] = v;
}

// out[q] = sum of part[q * n .. q * n + n) for q < nsums; tmp holds nsums * kRedStage doubles.
inline void reduce_sums(const double* part, unsigned int n, int nsums, double* tmp, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_sum_stage, dim3(kRedStage, nsums), dim3(256), 0, s, part, n, tmp);
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s, (const double*)tmp, kRedStage, nsums, out);
}

__global__ __launch_bounds__(kRedThreads) inline void k_reduce_maxmin(const double* part, unsigned int n,
                                                                        double* out) {
  __shared__ double lds[kRedThreads / kWave];
  double a = 0.0, b = 1.0e300;
  for (unsigned int i = threadIdx.x; i < n; i += kRedThreads) {
    a = fmax(a, part[i]);
    b = fmin(b, part[n + i]);
  }
  a = wave_max(a);
  b = wave_min(b);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0];
    for (int i = 1; i < kRedThreads / kWave; ++i) m = fmax(m, lds[i]);
    out[0] = m;
  }
  __syncthreads();
  if (lane == 0) lds[w] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0];
    for (int i = 1; i < kRedThreads / kWave; ++i) m = fmin(m, lds[i]);
    out[1] = m;
  }
}

// Two-stage form of k_reduce_maxmin for long rows: kRedStage blocks each take the max of one
// chunk of part[0..n) and the min of the same chunk of part[n..2n) (to tmp[0..kRedStage) and
// tmp[kRedStage..2 kRedStage)), then one block reduces those. Max and min are exact, so the
// result equals the one-block form bit for bit.
__global__ __launch_bounds__(256) inline void k_reduce_maxmin_stage(const double* part, unsigned int n,
                                                                     double* out) {
  __shared__ double lds[2 * (256 / kWave)];
  const unsigned int chunk = (n + gridDim.x - 1) / gridDim.x;
  const unsigned int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double a = 0.0, b = 1.0e300;
  for (unsigned int i = lo + threadIdx.x; i < hi; i += 256) {
    a = fmax(a, part[i]);
    b = fmin(b, part[n + i]);
  }
  a = wave_max(a);
  b = wave_min(b);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) {
    lds[w] = a;
    lds[256 / kWave + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0], l = lds[256 / kWave];
    for (int i = 1; i < 256 / kWave; ++i) {
      m = fmax(m, lds[i]);
      l = fmin(l, lds[256 / kWave + i]);
    }
    out[blockIdx.x] = m;
    out[gridDim.x + blockIdx.x] = l;
  }
}

// out[0] = max of part[0..n) (floored at 0), out[1] = min of part[n..2n) (capped at 1e300);
// tmp holds 2 * kRedStage doubles.
inline void reduce_maxmin(const double* part, unsigned int n, double* tmp, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_maxmin_stage, dim3(kRedStage), dim3(256), 0, s, part, n, tmp);
  hipLaunchKernelGGL(k_reduce_maxmin, dim3(1), dim3(kRedThreads), 0, s, (const double*)tmp, kRedStage, out);
}

// ---------------------------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs of the device (linear workgroup id mod 8),
// and each XCD has its own L2. xcd_remap is a bijection of [0, n) that hands every XCD one
// contiguous run of tile indices, so stencil neighbours (adjacent tiles, and the cache lines that
// straddle tile edges) are read through the same L2 at about the same time.
__device__ __forceinline__ unsigned int xcd_remap(unsigned int id, unsigned int n) {
  constexpr unsigned int kXcd = 8;
  const unsigned int q = n / kXcd, r = n % kXcd, x = id % kXcd;
  return x * q + (x < r ? x : r) + id / kXcd;
}

struct TileId {
  unsigned int x, y, z, linear;  // tile coordinates (x fastest) and the launch's linear block id
};

// Tile coordinates of this workgroup under xcd_remap, for a grid launched as gridDim.
__device__ __forceinline__ TileId xcd_tile() {
  const unsigned int gx = gridDim.x, gy = gridDim.y;
  const unsigned int id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned int t = xcd_remap(id, gx * gy * gridDim.z);
  TileId r;
  r.x = t % gx;
  r.y = (t / gx) % gy;
  r.z = t / (gx * gy);
  r.linear = id;
  return r;
}

// As xcd_tile, with the layer index varying fastest along each XCD's run of tiles: an XCD
// processes the layers of one column tile one after the other, so the layer above and below a
// tile (its z-neighbours) were read through the same L2 just before.
__device__ __forceinline__ TileId zfast_tile() {
  const unsigned int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const unsigned int id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned int t = xcd_remap(id, gx * gy * gz);
  TileId r;
  r.z = t % gz;
  r.x = (t / gz) % gx;
  r.y = t / (gz * gx);
  r.linear = id;
  return r;
}

// Soil class with derived constants (include/frehg2/re/SoilParams.hpp:13-25).
struct SoilClass {
  double alpha, n, theta_s, theta_r, Ks_x, Ks_y, Ks_z, Ss, air_entry;
  int use_vg, use_mvg;
  double m;      // 1 - 1/n
  double wcm;    // VanGenuchten::wcm (VanGenuchten.hpp:21-28)
  double capc;   // alpha * n * m * (wcm - theta_r)
  double inv_m, inv_n, inv_alpha;
  double ss_ts;  // Ss / theta_s
};

inline SoilClass make_soil_class(double alpha, double n, double theta_s, double theta_r, double Ks_x,
                                 double Ks_y, double Ks_z, double Ss, int use_vg, int use_mvg,
                                 double air_entry);

// Per-cell soil lookup on the halo-padded grid. cls / ks are padded arrays whose ghost entries
// hold the clamped interior value (ReSolver::soilAt clamps ghost indices, ReSolver.cpp:87-96).
// c0 repeats class 0 by value, so that with cls == nullptr (one class) the fixed variants read the
// class parameters as kernel arguments (scalar registers) instead of per-lane loads. The template
// arguments fix, per
// kernel instantiation, whether cls / ks are absent (kAbsent), present (kPresent) or checked at
// each access (kRuntime): fixed variants have no per-access branches, so the loads of a cell's
// whole stencil can issue together.
enum SoilArr { kAbsent = 0, kPresent = 1, kRuntime = 2 };

template <int CLS, int KS>
struct SoilViewT {
  const SoilClass* classes;
  const int32_t* cls;  // nullptr -> every cell is class 0
  const double* ks;    // nullptr -> Ks from the class; else isotropic per-cell Ks
  SoilClass c0;        // classes[0]
  // (a reference either way: a struct chosen at run time between two places would be copied to
  // per-lane stack memory)
  __device__ inline const SoilClass& at(long long c) const {
    if constexpr (CLS == kAbsent) return c0;
    else if constexpr (CLS == kPresent) return classes[cls[c]];
    else return classes[cls ? cls[c] : 0];
  }
  __device__ inline double Ksx(long long c) const {
    if constexpr (KS == kPresent) return ks[c];
    else if constexpr (KS == kAbsent) return at(c).Ks_x;
    else return ks ? ks[c] : at(c).Ks_x;
  }
  __device__ inline double Ksy(long long c) const {
    if constexpr (KS == kPresent) return ks[c];
    else if constexpr (KS == kAbsent) return at(c).Ks_y;
    else return ks ? ks[c] : at(c).Ks_y;
  }
  __device__ inline double Ksz(long long c) const {
    if constexpr (KS == kPresent) return ks[c];
    else if constexpr (KS == kAbsent) return at(c).Ks_z;
    else return ks ? ks[c] : at(c).Ks_z;
  }
};
using SoilView = SoilViewT<kRuntime, kRuntime>;

// ---------------------------------------------------------------------------------------------
// Van Genuchten relations (src/re/VanGenuchten.hpp:18-137). The plain-VG path evaluates
// L = log|alpha h|, lp = n L = log p (p = |alpha h|^n) and l1p = log1p(p) once and derives
//   sqrt(Se) = exp(-m l1p / 2), Se = sqrt(Se)^2           (Se = (1+p)^-m),
//   C = alpha n m (wcm - theta_r) p Se / (|alpha h| (1+p))  (|alpha h|^(n-1) = p/|alpha h|),
//   kr = sqrt(Se) t^2, t = 1 - (p/(1+p))^m = -expm1(m (lp - l1p))   (Se^(1/m) = 1/(1+p)),
// i.e. log, exp, log1p, exp, expm1 and one division instead of ~10 pow, and without the
// cancellation of the pow form near saturation (p << 1). Modified-VG and non-VG soils use the
// reference expressions. kr is K/Ks: K(h, Ks) = Ks kr(h), kr <= 1 and kr = 1 above the cut-off.
struct VGState {
  double theta, cap, kr;
};

__device__ inline double vg_theta_ref(const SoilClass& s, double h) {  // VanGenuchten.hpp:31-47
  double wc;
  if (s.use_vg) {
    const double se = pow(1.0 + pow(fabs(s.alpha * h), s.n), -s.m);
    wc = (h > s.air_entry) ? s.theta_s : s.theta_r + (s.wcm - s.theta_r) * se;
  } else {
    wc = s.theta_r + (s.theta_s - s.theta_r) * exp(0.1634 * h);
  }
  if (wc > s.theta_s) wc = s.theta_s;
  if (wc < s.theta_r) wc = s.theta_r;
  return wc;
}

__device__ inline double vg_kr_ref(const SoilClass& s, double h) {  // VanGenuchten.hpp:84-111 / Ks
  if (!s.use_vg) return exp(0.1634 * h);
  const double se = pow(1.0 + pow(fabs(s.alpha * h), s.n), -s.m);
  double k;
  if (s.use_mvg) {
    const double nume =
        1.0 - pow(1.0 - pow(se * (s.theta_s - s.theta_r) / (s.wcm - s.theta_r), 1.0 / s.m), s.m);
    const double deno =
        1.0 - pow(1.0 - pow((s.theta_s - s.theta_r) / (s.wcm - s.theta_r), 1.0 / s.m), s.m);
    k = (deno == 0.0) ? 1.0 : sqrt(se) * (nume / deno) * (nume / deno);
  } else {
    const double t = 1.0 - pow(1.0 - pow(se, 1.0 / s.m), s.m);
    k = sqrt(se) * t * t;
  }
  if (k > 1.0) k = 1.0;
  if (s.use_mvg) {
    if (h > s.air_entry) k = 1.0;
  } else if (h > 0.0) {
    k = 1.0;
  }
  return k;
}

__device__ inline double vg_cap_ref(const SoilClass& s, double h) {  // VanGenuchten.hpp:68-81
  const double a = fabs(s.alpha * h);
  double c = s.capc * pow(a, s.n - 1.0) / pow(1.0 + pow(a, s.n), s.m + 1.0);
  if (s.use_mvg) {
    if (h > s.air_entry) c = 0.0;
  } else if (h > 0.0) {
    c = 0.0;
  }
  return c;
}

// Out of line (called only for non-VG and modified-VG classes). The class parameters are passed
// as scalars, so that a caller's class (often a kernel argument) never has to be copied to
// per-lane stack memory to have its address taken.
__device__ __noinline__ VGState vg_state_ref(double alpha, double n, double m, double theta_s, double theta_r,
                                             double wcm, double capc, double air_entry, int use_vg, int use_mvg,
                                             double h) {
  SoilClass s{};
  s.alpha = alpha;
  s.n = n;
  s.m = m;
  s.theta_s = theta_s;
  s.theta_r = theta_r;
  s.wcm = wcm;
  s.capc = capc;
  s.air_entry = air_entry;
  s.use_vg = use_vg;
  s.use_mvg = use_mvg;
  VGState o;
  o.theta = vg_theta_ref(s, h);
  o.cap = vg_cap_ref(s, h);
  o.kr = vg_kr_ref(s, h);
  return o;
}

__device__ __forceinline__ VGState vg_state(const SoilClass& s, double h) {
  if (!s.use_vg || s.use_mvg)
    return vg_state_ref(s.alpha, s.n, s.m, s.theta_s, s.theta_r, s.wcm, s.capc, s.air_entry, s.use_vg, s.use_mvg, h);
  VGState o;
  const double a = fabs(s.alpha * h);
  double se, cap, kr;
  if (a == 0.0) {
    se = 1.0;
    cap = 0.0;
    kr = 1.0;
  } else {
    const double lp = s.n * log(a);
    const double p = exp(lp);
    const double l1p = log1p(p);
    const double sq = exp(-0.5 * s.m * l1p);
    se = sq * sq;
    cap = s.capc * p * se / (a * (1.0 + p));
    const double t = -expm1(s.m * (lp - l1p));
    kr = sq * t * t;
    if (kr > 1.0) kr = 1.0;
  }
  double th = (h > s.air_entry) ? s.theta_s : s.theta_r + (s.wcm - s.theta_r) * se;
  if (th > s.theta_s) th = s.theta_s;
  if (th < s.theta_r) th = s.theta_r;
  o.theta = th;
  o.cap = (h > 0.0) ? 0.0 : cap;
  o.kr = (h > 0.0) ? 1.0 : kr;
  return o;
}

// h(theta) (VanGenuchten.hpp:50-65): x = (wcm - theta_r)/(theta - theta_r) >= 1,
// h = -(1/alpha) (x^(1/m) - 1)^(1/n) with x^(1/m) - 1 = expm1(log(x)/m).
__device__ __forceinline__ double vg_head(const SoilClass& s, double wc) {
  if (s.use_vg) {
    double th = wc;
    if (th - s.theta_r < 1.0e-7) th = s.theta_r + 1.0e-7;
    if (th < s.theta_s) {
      const double y = expm1(log((s.wcm - s.theta_r) / (th - s.theta_r)) * s.inv_m);
      return -s.inv_alpha * exp(log(y) * s.inv_n);
    }
    return 0.0;
  }
  return log((wc - s.theta_r) / (s.theta_s - s.theta_r)) / 0.1634;
}

__device__ inline double vg_dKdwc(const SoilClass& s, double wc, double Ks) {  // :114-136
  double th = wc;
  if (th > 0.9999 * s.theta_s && th < s.theta_s) th = 0.9999 * s.theta_s;
  const double mm = s.m;
  const double se = (th - s.theta_r) / (s.theta_s - s.theta_r);
  if (s.use_mvg) {
    const double c2 = (s.theta_s - s.theta_r) / (s.wcm - s.theta_r);
    const double c1 = 1.0 / pow(1.0 - pow(1.0 - pow(c2, 1.0 / mm), mm), 2.0);
    const double term0 = pow(1.0 - pow(c2 * se, 1.0 / mm), mm);
    const double term1 = 0.5 * c1 * Ks * pow(se, -0.5) * (1.0 - term0) * (1.0 - term0);
    const double term2 = 2.0 * c1 * Ks * c2 * pow(se, 0.5) * pow(c2 * se, 1.0 / mm - 1.0) * (1.0 - term0) *
                         pow(1.0 - pow(c2 * se, 1.0 / mm), mm - 1.0);
    return (term1 + term2) / (s.theta_s - s.theta_r);
  }
  const double term0 = pow(1.0 - pow(se, 1.0 / mm), mm);
  const double term1 = 0.5 * Ks * pow(se, -0.5) * (1.0 - term0) * (1.0 - term0);
  const double term2 = 2.0 * Ks * pow(se, (2.0 - mm) / (2.0 * mm)) * (1.0 - term0) *
                       pow(1.0 - pow(se, 1.0 / mm), mm - 1.0);
  return (term1 + term2) / (s.theta_s - s.theta_r);
}

inline SoilClass make_soil_class(double alpha, double n, double theta_s, double theta_r, double Ks_x,
                                 double Ks_y, double Ks_z, double Ss, int use_vg, int use_mvg,
                                 double air_entry) {
  SoilClass s{};
  s.alpha = alpha;
  s.n = n;
  s.theta_s = theta_s;
  s.theta_r = theta_r;
  s.Ks_x = Ks_x;
  s.Ks_y = Ks_y;
  s.Ks_z = Ks_z;
  s.Ss = Ss;
  s.use_vg = use_vg;
  s.use_mvg = use_mvg;
  s.air_entry = air_entry;
  s.m = 1.0 - 1.0 / n;
  s.wcm = use_mvg ? theta_r + (theta_s - theta_r) * std::pow(1.0 + std::pow(std::fabs(air_entry) * alpha, n), s.m)
                  : theta_s;
  s.capc = alpha * n * s.m * (s.wcm - theta_r);
  s.ss_ts = Ss / theta_s;
  s.inv_m = 1.0 / s.m;
  s.inv_n = 1.0 / n;
  s.inv_alpha = 1.0 / alpha;
  return s;
}

// Halo-padded grid geometry. The cells are those of the reference's halo-padded arrays
// (src/core/Grid.hpp:42-60: i in [-1, nx], j in [-1, ny], k in [-1, nz], i fastest), stored with
// a row pitch `sy` rounded up to 16 doubles and cell (0, j, k) at a multiple of 16, so every
// interior row starts on a 128-byte boundary and pairs of cells (i even) are 16-byte aligned.
// Row j of layer k holds cells i = -1 .. nx at kGeoOff - 1 .. kGeoOff + nx of its pitch; the
// rest of the pitch is padding that no kernel writes. The public field interface keeps the
// reference's packed (nx + 2)-pitch layout (copy_field_in / copy_field_out).
constexpr int kGeoOff = 16;

struct Geo {
  int nx, ny, nz;
  long long sy, sz;  // +1 in j, +1 in k
  __host__ __device__ inline long long idx(int i, int j, int k) const {
    return (long long)(kGeoOff + i) + (long long)(j + 1) * sy + (long long)(k + 1) * sz;
  }
  __host__ __device__ inline long long sidx(int i, int j) const {
    return (long long)(kGeoOff + i) + (long long)(j + 1) * sy;
  }
  // storage of a 3D field (nz + 2 layers) / a 2D field (one layer); the tail of kGeoOff doubles
  // keeps the last row's padding (read by paired loads) inside the allocation
  __host__ __device__ inline long long ncell() const { return sz * (nz + 2) + kGeoOff; }
  __host__ __device__ inline long long nsurf() const { return sz + kGeoOff; }
};

inline Geo make_geo(int nx, int ny, int nz) {
  Geo g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.sy = ((long long)nx + 2 + 15) / 16 * 16;
  g.sz = g.sy * (ny + 2);
  return g;
}

// Transfers between a device field and the reference's packed halo-padded layout
// ((nx + 2) x (ny + 2) x planes doubles, planes = nz + 2 for 3D fields, 1 for 2D fields).
inline void copy_field_in(const Geo& g, int planes, double* dev, const double* host, hipStream_t s) {
  const size_t w = (size_t)(g.nx + 2) * sizeof(double);
  const hipError_t e = hipMemcpy2DAsync(dev + kGeoOff - 1, (size_t)g.sy * sizeof(double), host, w, w,
                                        (size_t)(g.ny + 2) * planes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) throw HipError(std::string("copy_field_in: ") + hipGetErrorString(e));
}
inline void copy_field_out(const Geo& g, int planes, double* host, const double* dev, hipStream_t s) {
  const size_t w = (size_t)(g.nx + 2) * sizeof(double);
  const hipError_t e = hipMemcpy2DAsync(host, w, dev + kGeoOff - 1, (size_t)g.sy * sizeof(double), w,
                                        (size_t)(g.ny + 2) * planes, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) throw HipError(std::string("copy_field_out: ") + hipGetErrorString(e));
}

inline unsigned int ceil_div(long long a, long long b) { return (unsigned int)((a + b - 1) / b); }

}  // namespace frehg
