// Streaming probe: the HBM access mix of the groundwater CG SpMV (six FP64 reads and two FP64
// writes per cell, 16-byte loads and stores) with no neighbours, no halo re-reads and no
// reduction. It measures what that mix streams at on this device, the practical ceiling of the
// SpMV beside the 8 TB/s peak. A measurement aid: nothing in the solvers calls it.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/frehg_gpu.h"
#include "common.hpp"

namespace frehg {
namespace {

constexpr int PT = 256;

// p = a + b * c, q = d * p + e * f, pairs of cells per lane, grid-stride over the pairs.
__global__ __launch_bounds__(PT) void k_stream_6r2w(long long npairs, const double2* __restrict__ a,
                                                    const double2* __restrict__ b, const double2* __restrict__ c,
                                                    const double2* __restrict__ d, const double2* __restrict__ e,
                                                    const double2* __restrict__ f, double2* __restrict__ p,
                                                    double2* __restrict__ q) {
  const long long stride = (long long)gridDim.x * PT;
  for (long long i = (long long)blockIdx.x * PT + threadIdx.x; i < npairs; i += stride) {
    const double2 va = a[i], vb = b[i], vc = c[i], vd = d[i], ve = e[i], vf = f[i];
    double2 vp, vq;
    vp.x = va.x + vb.x * vc.x;
    vp.y = va.y + vb.y * vc.y;
    vq.x = vd.x * vp.x + ve.x * vf.x;
    vq.y = vd.y * vp.y + ve.y * vf.y;
    p[i] = vp;
    q[i] = vq;
  }
}

__global__ __launch_bounds__(PT) void k_probe_fill(long long n, double v, double* __restrict__ x) {
  const long long stride = (long long)gridDim.x * PT;
  for (long long i = (long long)blockIdx.x * PT + threadIdx.x; i < n; i += stride) x[i] = v + 1e-9 * (double)(i & 1023);
}

}  // namespace
}  // namespace frehg

extern "C" int frehg_stream_probe(int device, int64_t cells, int reps, double* gbytes_per_s, double* ms_per_launch) {
  using namespace frehg;
  if (cells < 2 || cells > (int64_t)1 << 31 || reps < 1 || !gbytes_per_s || !ms_per_launch) return FREHG_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return FREHG_ERR_HIP;
  const long long npairs = cells / 2;
  const size_t bytes = (size_t)npairs * sizeof(double2);
  std::vector<double*> arr(8, nullptr);
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FREHG_OK;
  auto ok = [&](hipError_t err) {
    if (err != hipSuccess && rc == FREHG_OK) rc = err == hipErrorOutOfMemory ? FREHG_ERR_OOM : FREHG_ERR_HIP;
    return rc == FREHG_OK;
  };
  if (ok(hipStreamCreate(&s)) && ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1))) {
    for (int k = 0; k < 8 && rc == FREHG_OK; ++k) ok(hipMalloc(&arr[k], bytes));
  }
  if (rc == FREHG_OK) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const long long full = (long long)ceil_div(npairs, PT);  // one pair per lane
    for (int k = 0; k < 8; ++k)
      hipLaunchKernelGGL(k_probe_fill, dim3((unsigned int)std::min<long long>(full, (long long)cus * 16)), dim3(PT), 0,
                         s, 2 * npairs, 0.5 + 0.1 * k, arr[k]);
    ok(hipGetLastError());
    // the fastest of a few grid sizes: one pair per lane, or grid-stride over 8, 16 or 32 blocks per CU
    const long long grids[4] = {full, (long long)cus * 8, (long long)cus * 16, (long long)cus * 32};
    double best_ms = 0.0;
    for (long long gsz : grids) {
      const unsigned int blocks = (unsigned int)std::max<long long>(1, std::min<long long>(gsz, full));
      auto launch = [&] {
        hipLaunchKernelGGL(k_stream_6r2w, dim3(blocks), dim3(PT), 0, s, npairs, (const double2*)arr[0],
                           (const double2*)arr[1], (const double2*)arr[2], (const double2*)arr[3],
                           (const double2*)arr[4], (const double2*)arr[5], (double2*)arr[6], (double2*)arr[7]);
      };
      launch();  // warm-up
      ok(hipGetLastError());
      ok(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) launch();
      ok(hipGetLastError());
      ok(hipEventRecord(e1, s));
      if (!ok(hipEventSynchronize(e1))) break;
      float ms = 0.0f;
      if (!ok(hipEventElapsedTime(&ms, e0, e1))) break;
      const double per = (double)ms / reps;
      if (best_ms == 0.0 || per < best_ms) best_ms = per;
    }
    if (rc == FREHG_OK && best_ms > 0.0) {
      *ms_per_launch = best_ms;
      *gbytes_per_s = 64.0 * (double)(2 * npairs) / (best_ms * 1.0e-3) / 1.0e9;
    }
  }
  if (s) (void)hipStreamSynchronize(s);
  for (double* p : arr)
    if (p) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
