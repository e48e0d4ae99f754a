// itsca_mlp.h -- the learned tracklet filter: a 10 -> 32 -> 32 -> 1 ReLU
// network with BF16 weights and activations and FP32 accumulation.
//
// Parameters are kept as FP32 master copies (for AdamW) plus BF16-rounded
// copies used by every forward pass.  Layout of the flat parameter vector
// (kMlpParams floats):
//   W1[32][10]  b1[32]  W2[32][32]  b2[32]  w3[32]  b3
#pragma once

#include <stdint.h>

#include "itsca_math.h"

namespace itsca {

constexpr int kMlpIn = 10;
constexpr int kMlpInPad = 16;  // K of the first MFMA
constexpr int kMlpHidden = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHidden * kMlpIn;       // 320
constexpr int kOffW2 = kOffB1 + kMlpHidden;                 // 352
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;    // 1376
constexpr int kOffW3 = kOffB2 + kMlpHidden;                 // 1408
constexpr int kOffB3 = kOffW3 + kMlpHidden;                 // 1440
constexpr int kMlpParams = kOffB3 + 1;                      // 1441

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet), and back.
ITSCA_HD uint16_t toBf16(float f)
{
  const uint32_t u = (uint32_t)floatBits(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
ITSCA_HD float fromBf16(uint16_t h) { return bitsFloat((int32_t)((uint32_t)h << 16)); }
ITSCA_HD float roundBf16(float f) { return fromBf16(toBf16(f)); }

// Features of one tracklet candidate: cluster i of layer l (x, y, z, r, phi),
// cluster j of layer l+1, the vertex, and the numbers of accepted tracklets
// starting at i and at j (the latter from layer l+1; 0 for the last layer).
// Every feature is scaled to O(1).
struct TrackletFeatureInput {
  float xi, yi, zi, ri, phii;
  float xj, yj, zj, rj, phij;
  float vx, vy, vz;
  int layer, countI, countJ;
};

ITSCA_HD void trackletFeatures(const TrackletFeatureInput& in, float f[kMlpIn])
{
  const float tanL = (in.zi - in.zj) / (in.ri - in.rj);
  float dphi = in.phij - in.phii;
  if (dphi > kPi) dphi -= kTwoPi;
  if (dphi < -kPi) dphi += kTwoPi;
  const float ax = in.xi - in.vx, ay = in.yi - in.vy, bx = in.xj - in.vx, by = in.yj - in.vy;
  const float cross = ax * by - ay * bx;
  const float dx = bx - ax, dy = by - ay;
  const float denom = sqrtf((ax * ax + ay * ay) * (bx * bx + by * by) * (dx * dx + dy * dy));
  const float curvature = denom > 0.0f ? 2.0f * cross / denom : 0.0f;  // circle through vertex, i, j
  const float pointing = (in.zi - in.vz) / in.ri;
  f[0] = (float)in.layer * 0.2f;
  f[1] = dphi * (1.0f / 0.3f);
  f[2] = (tanL - pointing) * 20.0f;
  f[3] = curvature * 50.0f;
  f[4] = in.ri * 0.025f;
  f[5] = in.rj * 0.025f;
  f[6] = (in.zi - tanL * in.ri - in.vz) * (1.0f / 0.3f);
  f[7] = tanL * (1.0f / 1.5f);
  f[8] = (float)(in.countJ < 32 ? in.countJ : 32) * (1.0f / 16.0f);
  f[9] = (float)(in.countI < 32 ? in.countI : 32) * (1.0f / 16.0f);
  for (int k = 0; k < kMlpIn; ++k)
    if (!(f[k] == f[k])) f[k] = 0.0f;  // NaN guard (degenerate geometry)
}

}  // namespace itsca
