// itsca_mlp_kernels.h -- host-side interface of the learned-filter kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "itsca_device.h"

namespace itsca {

struct MlpBuffers {
  float *master, *mom1, *mom2, *wq;  // FP32 master, AdamW moments, BF16-rounded FP32 copy
  uint16_t* wbf;                     // BF16 bits (MFMA operands)
  float* partial;                    // gradBlocks x (params + 2)
  int gradBlocks;
};

// Scores the tracklet candidates left by the count kernel and keeps those
// whose logit is >= threshold (B = block size of the stage kernels).
void launchTrackletFilter(const DevState& s, int block, int blocks, float threshold, hipStream_t stream);
// Moves the kept counts of the filter (trkKept) into trkCnt after the last layer.
void launchCommitKept(const DevState& s, hipStream_t stream);
// Labels every tracklet of the last run (on a final road or not) and appends
// every stride-th one, from `phase`, to a training set at row `base`.
void launchLabelAndCollect(const DevState& s, int totalTracklets, uint8_t* label, int stride, int phase,
                           uint16_t* outX, uint8_t* outY, int base, int cap, hipStream_t stream);
void launchScoreRows(const uint16_t* wbf, const uint16_t* X, int n, float* out, hipStream_t stream);
// One training step on rows (a*k + off) mod n, k in [k0, k0 + count): the
// gradient and, when lr > 0, an AdamW update with step number `step`.
void launchTrainStep(const MlpBuffers& b, const uint16_t* X, const uint8_t* Y, int n, long a, long off, int k0,
                     int count, float posWeight, int step, float lr, float wd, float* gradOut, float* lossOut,
                     hipStream_t stream);

}  // namespace itsca
