// itsca_mlp_device.h -- device side of the learned tracklet filter's
// inference, shared by the filter kernels and by the fused tracklet search:
// the BF16 MFMA network of one wavefront (64 samples per call), the LDS
// feature tile and the feature functions over the sorted-cluster records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "itsca_device.h"
#include "itsca_mlp.h"

namespace itsca {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x16_bf16 (wave64): lane i holds A[i%16][4*(i/16)+r],
// B[4*(i/16)+r][i%16] and C/D[4*(i/16)+r][i%16], r = 0..3.
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16 (gfx950, wave64): lane i holds A[i%16][8*(i/16)+r]
// and B[8*(i/16)+r][i%16], r = 0..7; C/D as above.
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Orders LDS traffic among the lanes of one wavefront.
__device__ __forceinline__ void waveSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Two FP32 values as packed BF16, round to nearest even (one
// v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint32_t packBf16x2(float lo, float hi)
{
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// ReLU of an MFMA result tile, packed as four BF16.  Rounding to nearest even
// keeps the sign, so ReLU is applied after the conversion: a BF16 value read as
// a 16-bit integer is negative exactly when the value is negative (or -0), so
// one v_pk_max_i16 against 0 clamps two values.
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t reluPair(uint32_t packed)
{
  const i16x2 v = __builtin_bit_cast(i16x2, packed);
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, i16x2{0, 0}));
}
__device__ __forceinline__ uint2 reluBf16(f32x4 d)
{
  return make_uint2(reluPair(packBf16x2(d[0], d[1])), reluPair(packBf16x2(d[2], d[3])));
}

__device__ __forceinline__ bf16x8 concat8(uint2 lo, uint2 hi)
{
  return __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
}

// Hidden unit that K index k of a 32-wide layer input stands for.  Layer 1
// (two 16x16x16 MFMAs, one per block of 16 units) leaves units 4*(i/16)+r and
// 16+4*(i/16)+r in lane i; read as the B operand of a 16x16x32 MFMA these are
// K indices 8*(i/16)+r and 8*(i/16)+4+r.  Weights are loaded in this order, so
// no data moves between layers.
__device__ __forceinline__ int hiddenUnit(int k) { return ((k & 4) ? 16 : 0) + 4 * (k >> 3) + (k & 3); }

// Weight operands of one wavefront, loaded once per kernel.  Layer 1 computes
// H1^T = W1 * X^T (K = 16 features, 10 used); layers 2 and 3 take the previous
// D tiles as B operands.  Biases are the accumulator inputs.  The last layer
// (32 -> 1) holds w3 in every row of A, so every row of D is the logit row and
// lane i receives the logit of sample i%16 without a cross-lane reduction.
struct MlpWave {
  bf16x4 a1[2];  // W1 rows 16*ob + i%16, features 4*(i/16)+r (zero past 10)
  bf16x8 a2[2];  // W2 rows 16*ob + i%16, units hiddenUnit(8*(i/16)+r)
  bf16x8 a3;     // w3[hiddenUnit(8*(i/16)+r)], the same for every row
  f32x4 c1[2], c2[2], c3;  // b1, b2 of the D rows 16*ob + 4*(i/16)+r; b3
};

__device__ __forceinline__ MlpWave loadMlp(const uint16_t* w)
{
  MlpWave m;
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  uint16_t a2[2][8], a3[8];
  for (int ob = 0; ob < 2; ++ob) {
    const int row = ob * 16 + col;
    for (int r = 0; r < 4; ++r) {
      const int k = grp * 4 + r;
      m.a1[ob][r] = (short)(k < kMlpIn ? w[kOffW1 + row * kMlpIn + k] : 0);
      const int drow = ob * 16 + grp * 4 + r;
      m.c1[ob][r] = fromBf16(w[kOffB1 + drow]);
      m.c2[ob][r] = fromBf16(w[kOffB2 + drow]);
    }
    for (int r = 0; r < 8; ++r) a2[ob][r] = w[kOffW2 + row * kMlpHidden + hiddenUnit(grp * 8 + r)];
  }
  for (int r = 0; r < 8; ++r) a3[r] = w[kOffW3 + hiddenUnit(grp * 8 + r)];
  for (int ob = 0; ob < 2; ++ob)
    m.a2[ob] = __builtin_bit_cast(bf16x8, make_uint4(a2[ob][0] | (uint32_t)a2[ob][1] << 16,
                                                     a2[ob][2] | (uint32_t)a2[ob][3] << 16,
                                                     a2[ob][4] | (uint32_t)a2[ob][5] << 16,
                                                     a2[ob][6] | (uint32_t)a2[ob][7] << 16));
  m.a3 = __builtin_bit_cast(bf16x8, make_uint4(a3[0] | (uint32_t)a3[1] << 16, a3[2] | (uint32_t)a3[3] << 16,
                                               a3[4] | (uint32_t)a3[5] << 16, a3[6] | (uint32_t)a3[7] << 16));
  const float b3 = fromBf16(w[kOffB3]);
  m.c3 = f32x4{b3, b3, b3, b3};
  return m;
}

// LDS feature tile: 64 sample rows of 32 bytes, each four 8-byte units of 4
// BF16 features.  Unit u of row r is stored at unit u ^ tileSwizzle(r).  The
// MFMA operand read takes unit grp of rows col = 0..15 (16 lanes of 8 bytes);
// unswizzled, rows c, c+4, c+8, c+12 (c, c+8 with 64 banks) fall on the same
// banks.  Swizzled by row bits 2 and 3, those rows use four different units,
// and the 16-byte row stores of rows r, r+4 and r, r+8 go to different halves,
// so both the reads and the stores are free of bank conflicts with 32 or 64
// four-byte banks.
__device__ __forceinline__ int tileSwizzle(int row)
{
  const int b2 = (row >> 2) & 1, b3 = (row >> 3) & 1;
  return ((b2 ^ b3) << 1) | b3;
}

// Row `row` of a tile from its two 16-byte halves (units 0-1 and 2-3).
__device__ __forceinline__ void storeTileRow(uint16_t* tile, int row, uint4 lo, uint4 hi)
{
  const int s = tileSwizzle(row);
  if (s & 1) {  // unit 2k+1 before unit 2k
    lo = make_uint4(lo.z, lo.w, lo.x, lo.y);
    hi = make_uint4(hi.z, hi.w, hi.x, hi.y);
  }
  uint4* r = reinterpret_cast<uint4*>(tile + row * kMlpInPad);
  r[s >> 1] = lo;
  r[(s >> 1) ^ 1] = hi;
}

// Logits of 64 samples whose BF16 features (16 per sample, 32 bytes) sit in
// `tile` (LDS, [64][16], swizzled as above); lane i gets the logit of sample
// i.  The four groups of 16 samples go through the network layer by layer, so
// that independent MFMAs cover each other's latency: 8 + 8 + 4 MFMAs per 64
// samples.
__device__ __forceinline__ float mlpForward64(const MlpWave& m, const uint16_t* tile)
{
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  const uint2* rows = reinterpret_cast<const uint2*>(tile);
  const int unit = grp ^ tileSwizzle(col);  // rows g*16 + col share bits 2 and 3 with col
  bf16x4 x[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) x[g] = __builtin_bit_cast(bf16x4, rows[(g * 16 + col) * 4 + unit]);
  uint2 h1[4][2], h2[4][2];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int ob = 0; ob < 2; ++ob) h1[g][ob] = reluBf16(mfma16(m.a1[ob], x[g], m.c1[ob]));
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int ob = 0; ob < 2; ++ob) h2[g][ob] = reluBf16(mfma32(m.a2[ob], concat8(h1[g][0], h1[g][1]), m.c2[ob]));
  float mine = 0.0f;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const f32x4 z = mfma32(m.a3, concat8(h2[g][0], h2[g][1]), m.c3);  // logit of sample 16*g + col
    if (grp == g) mine = z[0];
  }
  return mine;
}

// 16 BF16 features of one sample as two 16-byte halves (features 10..15 zero).
__device__ __forceinline__ void packFeatureRow(const float f[kMlpIn], uint4& lo, uint4& hi)
{
  static_assert(kMlpIn == 10 && kMlpInPad == 16, "feature row layout");
  lo = make_uint4(packBf16x2(f[0], f[1]), packBf16x2(f[2], f[3]), packBf16x2(f[4], f[5]), packBf16x2(f[6], f[7]));
  hi = make_uint4(packBf16x2(f[8], f[9]), 0u, 0u, 0u);
}

// A training row in global memory (unswizzled, 32 bytes).
__device__ __forceinline__ void packFeatures(const float f[kMlpIn], uint16_t* row)
{
  uint4* r = reinterpret_cast<uint4*>(row);
  packFeatureRow(f, r[0], r[1]);
}

// The cluster-i part of the features: sorted cluster i of layer l, whose
// record is at ri and whose tracklet count (of this work item) is at ci.
__device__ __forceinline__ TrackletSeed seedOf(const DevState& s, float vx, float vy, float vz, int ri, int ci)
{
  const float4 a = s.cxyrz[ri];  // (x, y, r, z)
  return trackletSeed(a.x, a.y, a.w, a.z, s.cphi[ri], vx, vy, vz, s.trkCnt[ci]);
}

// Features of candidate j (sorted cluster of layer l+1; record at rj, count at
// cj) of seed t.
__device__ __forceinline__ void featuresOf(const DevState& s, const TrackletSeed& t, float vx, float vy, float vz, int l,
                                           int rj, int cj, float f[kMlpIn])
{
  const float4 b = s.cxyrz[rj];  // (x, y, r, z)
  candidateFeatures(t, l, b.x - vx, b.y - vy, b.w, b.z, s.cphi[rj], vz,
                    l + 1 < kTrackletLayers ? s.trkCnt[cj] : 0, f);
}

// Record of one candidate cluster j: (x, y, r, z), phi and its tracklet count.
struct CandRec {
  float4 b;
  float phi;
  int count;
};

// The filtered form's kept candidates, as one list per wavefront of a chunk:
// the list of wavefront w (lanes = clusters first + 64w ...) starts at
// trkKeep[(chunk's first cluster + 64 w) * kTrkSlots] and holds the kept
// positions of its clusters, cluster after cluster.  Clusters whose
// candidates overflowed their slots have none (the fill recomputes them).
// The tracklet fill finds cluster c's at the exclusive scan of the capped
// kept counts over its wavefront.
__device__ __forceinline__ int* keptListOf(int* trkKeep, int chunkFirst, int wave)
{
  return trkKeep + (size_t)(chunkFirst + 64 * wave) * kTrkSlots;
}

// Appends a cluster's `capped` kept positions (the start of its slot row
// `row`) to its wavefront's list, at the exclusive scan of `capped` over the
// wavefront.  Called by every lane of the wavefront.
__device__ __forceinline__ void writeKeptList(int* waveList, const int* row, int capped)
{
  const int lane = threadIdx.x & 63;
  int inc = capped;
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  int* dst = waveList + inc - capped;
  for (int n = 0; n < capped; ++n) dst[n] = row[n];
}

// Candidate slots a wavefront scores per round (an LDS map slot -> owning lane).
constexpr int kWaveRound = 1024;
constexpr int kSeedFloats = 8;  // TrackletSeed, staged in LDS per cluster

}  // namespace itsca
