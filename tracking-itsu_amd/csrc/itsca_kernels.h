// itsca_kernels.h -- host-side interface of the kernel launch sequence.
#pragma once

#include <hip/hip_runtime.h>

#include "itsca_device.h"

namespace itsca {

constexpr int kNumTimedStages = 5;

// Optional HIP events around every launch of a run, on the launch stream:
// event 0 before the first kernel, event k+1 after kernel k.
struct KernelTimer {
  static constexpr int kMax = 48;
  hipEvent_t ev[kMax + 1];
  const char* name[kMax];
  int stage[kMax];
  int n = 0;
  void start(hipStream_t stream)
  {
    n = 0;
    (void)hipEventRecord(ev[0], stream);
  }
  void after(const char* kernel, int stageIndex, hipStream_t stream)
  {
    if (n >= kMax) return;
    name[n] = kernel;
    stage[n] = stageIndex;
    (void)hipEventRecord(ev[n + 1], stream);
    ++n;
  }
};

struct LaunchConfig {
  int persistentBlocks;
  int* planScratch;       // device ints, at least maxSegs
  int maxLayerClusters;   // largest cluster count of any (item, layer)
  bool filter;            // learned tracklet filter after the tracklet count
  float threshold;        // its logit threshold
  bool fusedFilter;       // the filter inside the tracklet search (kTrackletsSearchScore), one launch per layer
  int cus;                // compute units of the device (0: unknown)
};

void launchPipeline(const DevState& s, const LaunchConfig& cfg, hipStream_t stream, KernelTimer* timer);

// Roads of work item `item` as rows of 8 int (ITSCA_ROADS format), at most cap rows.
void launchRoadsOut(const DevState& s, int item, int* rows, int cap, hipStream_t stream);

}  // namespace itsca
