// Stencil solve through the LinearSolver seam without any allocation: frehg_ls_solve_noalloc.
//
// frehg_ls_solve copies a host operator into device arrays it allocates and builds a Krylov
// solver that allocates its own work vectors. This variant is for callers that cannot allocate
// device memory between set-up and the solve (tight serving loops, harnesses that forbid it): the
// caller passes the operator, right-hand side and solution as device arrays in the reference's
// owned row ordering, and every other device buffer of the solve (the halo-padded operator, the
// CG vectors, block partials and control blocks) lives in one file-scope device array, `scratch`,
// of kScratchBytes per GPU. The library resolves its address once per device when a context is
// created on that device; kernels receive that address as an argument.
//
// Scope: FP64 Jacobi-CG on a symmetric 7-point (3D) or 5-point (2D) operator, x0 = 0 and PETSc's
// default convergence test, as frehg_ls_solve with ksp "cg" and FP64 precision. Every call on a
// GPU shares that GPU's scratch array and leaves its contents there: each call first clears the
// CG vectors and writes every entry of the padded operator, and calls that overlap on one GPU
// (two host threads, or two processes) race. Host threads of one process are serialised by a
// per-device lock.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/frehg_gpu.h"
#include "krylov.hpp"
#include "noalloc.hpp"

// The scratch array of the allocation-free solve (zero-initialised at load, held until the
// process ends). Host code never passes its name to a kernel or a copy: the name is a host-side
// stand-in, and the device address comes from hipGetSymbolAddress (scratch_ptr_).
static __device__ __align__(256) unsigned char scratch[frehg::kScratchBytes];

namespace frehg {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_setup_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_call_mu[kMaxDevices];      // serialises this process's calls per device

constexpr int NT = 256;

// Fill kernel (no hipMemsetAsync, so a caller may also capture the call's launches).
__global__ __launch_bounds__(NT) void k_na_zero(double* a, long long n) {
  for (long long c = (long long)blockIdx.x * NT + threadIdx.x; c < n; c += (long long)gridDim.x * NT) a[c] = 0.0;
}

struct NaOperator {
  const double* diag;
  const double* off[6];  // x-, x+, y-, y+, z-, z+ (2D: the first four)
  const double* b;
};

// Every entry of the padded operator, right-hand side and iterate: interior rows from the owned
// arrays (3D rows z-fastest, 2D rows x-fastest), zero elsewhere (ghosts, pitch padding, faces on
// the domain edge). flags: bit 0 = a +face coupling differs from its neighbour's -face coupling
// (CG needs a symmetric operator), bit 1 = a zero or non-finite diagonal.
__global__ __launch_bounds__(NT) void k_na_pack(Geo g, int three, NaOperator op, long long n, double* D, double* ax,
                                                double* ay, double* az, double* b, double* x, int* flags) {
  const long long c = (long long)blockIdx.x * NT + threadIdx.x;
  if (c >= n) return;
  const int nz = three ? g.nz : 1;
  // c = kGeoOff + i + (j+1) sy (+ (k+1) sz): counted from cell i = -1 of row 0, a row's cells
  // i = -1 .. nx lie at 0 .. nx + 1 < sy of its pitch, so the decomposition below is unique
  const long long cc = c - (kGeoOff - 1);
  const long long k1 = three ? cc / g.sz : 1;
  const long long rem = three ? cc - k1 * g.sz : cc;
  const long long j1 = rem / g.sy;
  const long long i = rem - j1 * g.sy - 1;
  const bool interior =
      cc >= 0 && i >= 0 && i < g.nx && j1 >= 1 && j1 <= g.ny && (!three || (k1 >= 1 && k1 <= nz));
  double d = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, vb = 0.0;
  if (interior) {
    const int ii = (int)i, j = (int)j1 - 1, k = (int)k1 - 1;
    const long long r = three ? ((long long)j * g.nx + ii) * nz + k : (long long)j * g.nx + ii;
    d = op.diag[r];
    vb = op.b[r];
    int bad = (d != 0.0 && isfinite(d)) ? 0 : 2;
    // +x / +y / +z couplings inside the domain, each checked against the neighbour's -face value
    auto plus = [&](int dir, bool inside, long long rn) {
      if (!inside) return 0.0;
      const double a = op.off[2 * dir + 1][r], back = op.off[2 * dir][rn];
      if (fabs(a - back) > 1e-12 * (fabs(a) + fabs(back))) bad |= 1;
      return a;
    };
    vx = plus(0, ii < g.nx - 1, three ? r + nz : r + 1);
    vy = plus(1, j < g.ny - 1, three ? r + (long long)g.nx * nz : r + g.nx);
    if (three) vz = plus(2, k < nz - 1, r + 1);
    if (bad) atomicOr(flags, bad);
  }
  D[c] = d;
  ax[c] = vx;
  ay[c] = vy;
  if (az) az[c] = vz;
  b[c] = vb;
  x[c] = 0.0;
}

__global__ __launch_bounds__(NT) void k_na_unpack(Geo g, int three, const double* x, double* x_owned) {
  const int nz = three ? g.nz : 1;
  const long long r = (long long)blockIdx.x * NT + threadIdx.x;
  if (r >= (long long)g.nx * g.ny * nz) return;
  int i, j, k;
  if (three) {
    k = (int)(r % nz);
    const long long col = r / nz;
    i = (int)(col % g.nx);
    j = (int)(col / g.nx);
  } else {
    k = 0;
    i = (int)(r % g.nx);
    j = (int)(r / g.nx);
  }
  x_owned[r] = x[three ? g.idx(i, j, k) : g.sidx(i, j)];
}

inline size_t al256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw ArgError("frehg: device index out of range");
  std::lock_guard<std::mutex> lock(g_setup_mu);
  if (g_scratch_ptr[device]) return;
  // the caller has made `device` current
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    throw HipError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
}

void ls_solve_noalloc(int device, hipStream_t s, int ndim, int nx, int ny, int nz, const double* diag,
                      const double* const* offdiag, const double* b_owned, double* x_owned,
                      const frehg_solver_config& cfg, frehg_solve_stats* st) {
  if (ndim != 2 && ndim != 3) throw ArgError("frehg_ls_solve_noalloc: ndim must be 2 or 3");
  if (nx < 1 || ny < 1 || (ndim == 3 && nz < 1)) throw ArgError("frehg_ls_solve_noalloc: bad sizes");
  if (!diag || !offdiag || !b_owned || !x_owned) throw ArgError("frehg_ls_solve_noalloc: null argument");
  const bool three = ndim == 3;
  const int ndir = three ? 6 : 4;
  for (int d = 0; d < ndir; ++d)
    if (!offdiag[d]) throw ArgError("frehg_ls_solve_noalloc: null off-diagonal");
  const KrylovConfig kc = krylov_config(cfg, "frehg_ls_solve_noalloc");
  if (kc.ksp != FREHG_KSP_CG || kc.precision != 0)
    throw ArgError("frehg_ls_solve_noalloc: FP64 CG only (ksp_type FREHG_KSP_CG, precision FREHG_PREC_FP64)");
  if (device < 0 || device >= kMaxDevices || !g_scratch_ptr[device])
    throw ArgError("frehg_ls_solve_noalloc: no scratch address for this device (create a context on it first)");
  const Geo g = make_geo(nx, ny, three ? nz : 1);
  const long long n = three ? g.ncell() : g.nsurf();
  const size_t vec = al256((size_t)n * sizeof(double));
  const size_t need = (three ? 6 : 5) * vec + al256(sizeof(int)) + Krylov::workspace_bytes(g, three);
  if (need > kScratchBytes)
    throw ArgError("frehg_ls_solve_noalloc: the operator needs " + std::to_string(need >> 20) +
                   " MiB of scratch; the limit is " + std::to_string(kScratchBytes >> 20) +
                   " MiB (about " + std::to_string(kScratchBytes / (8 * 11 + 4)) +
                   " padded cells); use frehg_ls_solve");
  std::lock_guard<std::mutex> lock(g_call_mu[device]);
  char* base = static_cast<char*>(g_scratch_ptr[device]);
  auto take = [&](size_t bytes) {
    char* p = base;
    base += al256(bytes);
    return p;
  };
  double* D = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  double* ax = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  double* ay = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  double* az = three ? reinterpret_cast<double*>(take((size_t)n * sizeof(double))) : nullptr;
  double* b = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  double* x = reinterpret_cast<double*>(take((size_t)n * sizeof(double)));
  int* flags = reinterpret_cast<int*>(take(sizeof(int)));
  const KrylovWorkspace ws = Krylov::carve_workspace(base, g, three);
  // CG vectors from earlier calls hold other geometries: clear them (their ghost entries meet
  // zero coefficients and must be finite), and the flag word.
  const unsigned zb = (unsigned)std::min<long long>(4096, (n + NT - 1) / NT);
  for (double* v : {ws.r, ws.z, ws.p0, ws.p1, ws.q}) {
    hipLaunchKernelGGL(k_na_zero, dim3(zb), dim3(NT), 0, s, v, n);
    FREHG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(NT), 0, s, reinterpret_cast<double*>(flags), 1LL);
  FREHG_LAUNCH_CHECK();
  NaOperator op{};
  op.diag = diag;
  for (int d = 0; d < ndir; ++d) op.off[d] = offdiag[d];
  op.b = b_owned;
  hipLaunchKernelGGL(k_na_pack, dim3(ceil_div(n, NT)), dim3(NT), 0, s, g, three ? 1 : 0, op, n, D, ax, ay, az, b, x,
                     flags);
  FREHG_LAUNCH_CHECK();
  int hflags = 0;
  FREHG_HIP(hipMemcpyAsync(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost, s));
  FREHG_HIP(hipStreamSynchronize(s));
  if (hflags & 2) throw ArgError("frehg_ls_solve_noalloc: zero diagonal (Jacobi needs D != 0)");
  if (hflags & 1) throw ArgError("frehg_ls_solve_noalloc: operator is not symmetric (CG requires A = A^T)");
  Profiler off;  // no timing events (they would be created on first use)
  Krylov kry(g, three, s, &off, ws);
  const StencilView<double> A{D, ax, ay, az};
  const SolveResult sr = kry.solve(A, nullptr, b, nullptr, x, nullptr, true, false, kc);
  const long long rows = (long long)nx * ny * (three ? nz : 1);
  hipLaunchKernelGGL(k_na_unpack, dim3(ceil_div(rows, NT)), dim3(NT), 0, s, g, three ? 1 : 0, (const double*)x,
                     x_owned);
  FREHG_LAUNCH_CHECK();
  FREHG_HIP(hipStreamSynchronize(s));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = sr.iterations;
    st->residual_norm = sr.residual_norm;
    st->rhs_norm = sr.rhs_norm;
    st->converged = sr.converged ? 1 : 0;
  }
}

}  // namespace frehg
