// itsca_kernels.h -- host-side interface of the kernel launch sequence.
#pragma once

#include <hip/hip_runtime.h>

#include "itsca_device.h"

namespace itsca {

constexpr int kNumTimedStages = 5;

// Optional per-stage HIP events; mark(k) records event k (0..kNumTimedStages).
struct StageTimer {
  hipEvent_t ev[kNumTimedStages + 1];
  void mark(int k, hipStream_t stream) { (void)hipEventRecord(ev[k], stream); }
};

struct LaunchConfig {
  int persistentBlocks;
  int* planScratch;  // device ints, at least maxSegs
};

void launchPipeline(const DevState& s, const LaunchConfig& cfg, hipStream_t stream, StageTimer* timer);

}  // namespace itsca
