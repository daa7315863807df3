// Kernel timing on the context's stream (enabled on request; off by default). The events are
// passed to hipExtLaunchKernelGGL as its start/stop events, so they carry the kernel's own start
// and end times rather than the times at which markers around the launch were processed.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace frehg {

constexpr int kKernelTypes = 9;  // FREHG_KT_COUNT

class Profiler {
 public:
  bool enabled = false;
  unsigned int mask = ~0u;  // kernel types timed while enabled (bit kt)
  hipStream_t stream = nullptr;

  ~Profiler() {
    for (auto& e : pool_) (void)hipEventDestroy(e);
  }

  // Time one launch, or a run of consecutive launches, of kernel type `kt`: pass start(kt) as the
  // start event of the first launch and stop() as the stop event of the last one (both are null,
  // i.e. an ordinary launch, while profiling is off). Call start before stop.
  hipEvent_t start(int kt) {
    if (!enabled || !((mask >> kt) & 1u)) return nullptr;
    pending_kt_ = kt;
    pending_begin_ = take();
    return pending_begin_;
  }
  hipEvent_t stop() {
    if (!enabled || pending_kt_ < 0) return nullptr;
    hipEvent_t e = take();
    recs_.push_back({pending_kt_, pending_begin_, e});
    pending_kt_ = -1;
    return e;
  }
  // Drop the `n` most recent records of type `kt` (launches that returned immediately because the
  // solve had already converged), so averages cover only launches that did work.
  void discard_last(int kt, size_t n) {
    if (!enabled) return;
    for (size_t q = recs_.size(); q > 0 && n > 0; --q) {
      if (recs_[q - 1].kt != kt) continue;
      free_.push_back(recs_[q - 1].e0);
      free_.push_back(recs_[q - 1].e1);
      recs_.erase(recs_.begin() + (q - 1));
      --n;
    }
  }
  // Sum recorded durations into the per-type totals (synchronises on the recorded events).
  void fold() {
    for (auto& r : recs_) {
      FREHG_HIP(hipEventSynchronize(r.e1));
      float ms = 0.f;
      FREHG_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
      total_ms_[r.kt] += ms;
      launches_[r.kt] += 1;
      free_.push_back(r.e0);
      free_.push_back(r.e1);
    }
    recs_.clear();
  }
  void read(int kt, double* ms, long long* n) {
    fold();
    *ms = total_ms_[kt];
    *n = launches_[kt];
  }
  void reset() {
    fold();
    for (int i = 0; i < kKernelTypes; ++i) {
      total_ms_[i] = 0.0;
      launches_[i] = 0;
    }
  }

 private:
  struct Rec {
    int kt;
    hipEvent_t e0, e1;
  };
  hipEvent_t take() {
    if (free_.empty()) {
      hipEvent_t e;
      FREHG_HIP(hipEventCreate(&e));
      pool_.push_back(e);
      return e;
    }
    hipEvent_t e = free_.back();
    free_.pop_back();
    return e;
  }
  std::vector<hipEvent_t> pool_, free_;
  std::vector<Rec> recs_;
  int pending_kt_ = -1;
  hipEvent_t pending_begin_ = nullptr;
  double total_ms_[kKernelTypes] = {};
  long long launches_[kKernelTypes] = {};
};

}  // namespace frehg
