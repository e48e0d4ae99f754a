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

// The part of the features that depends on cluster i only (shared by all of
// its candidates).
struct TrackletSeed {
  float ax, ay, a2;  // cluster i relative to the vertex, and its squared length
  float zi, ri, phii;
  float pointing;    // (z_i - v_z) / r_i
  float countI;      // min(count_i, 32) / 16
};

ITSCA_HD TrackletSeed trackletSeed(float xi, float yi, float zi, float ri, float phii, float vx, float vy, float vz,
                                   int countI)
{
  TrackletSeed t;
  t.ax = xi - vx;
  t.ay = yi - vy;
  t.a2 = t.ax * t.ax + t.ay * t.ay;
  t.zi = zi;
  t.ri = ri;
  t.phii = phii;
  t.pointing = (zi - vz) / ri;
  t.countI = (float)(countI < 32 ? countI : 32) * (1.0f / 16.0f);
  return t;
}

// Features of candidate j of layer l+1 (bx, by relative to the vertex).
ITSCA_HD void candidateFeatures(const TrackletSeed& t, int layer, float bx, float by, float zj, float rj, float phij,
                                float vz, int countJ, float f[kMlpIn])
{
  const float tanL = (t.zi - zj) / (t.ri - rj);
  float dphi = phij - t.phii;
  if (dphi > kPi) dphi -= kTwoPi;
  if (dphi < -kPi) dphi += kTwoPi;
  const float cross = t.ax * by - t.ay * bx;
  const float dx = bx - t.ax, dy = by - t.ay;
  const float denom = sqrtf(t.a2 * (bx * bx + by * by) * (dx * dx + dy * dy));
  const float curvature = denom > 0.0f ? 2.0f * cross / denom : 0.0f;  // circle through vertex, i, j
  f[0] = (float)layer * 0.2f;
  f[1] = dphi * (1.0f / 0.3f);
  f[2] = (tanL - t.pointing) * 20.0f;
  f[3] = curvature * 50.0f;
  f[4] = t.ri * 0.025f;
  f[5] = rj * 0.025f;
  f[6] = (t.zi - tanL * t.ri - vz) * (1.0f / 0.3f);
  f[7] = tanL * (1.0f / 1.5f);
  f[8] = (float)(countJ < 32 ? countJ : 32) * (1.0f / 16.0f);
  f[9] = t.countI;
  // NaN guard (degenerate geometry: r_i = r_j, or a cluster at the vertex's x
  // and y).  With finite coordinates only the features through tanL or the
  // pointing can be NaN: the curvature is guarded above, and every other
  // feature is a finite difference, product or count.
  if (!(f[2] == f[2])) f[2] = 0.0f;
  if (!(f[6] == f[6])) f[6] = 0.0f;
  if (!(f[7] == f[7])) f[7] = 0.0f;
}

ITSCA_HD void trackletFeatures(const TrackletFeatureInput& in, float f[kMlpIn])
{
  const TrackletSeed t = trackletSeed(in.xi, in.yi, in.zi, in.ri, in.phii, in.vx, in.vy, in.vz, in.countI);
  candidateFeatures(t, in.layer, in.xj - in.vx, in.yj - in.vy, in.zj, in.rj, in.phij, in.vz, in.countJ, f);
}

}  // namespace itsca
