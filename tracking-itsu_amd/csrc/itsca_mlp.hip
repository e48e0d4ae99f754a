// itsca_mlp.hip -- gfx950 kernels of the learned tracklet filter: features,
// labels, BF16 MFMA inference (v_mfma_f32_16x16x16_bf16, FP32 accumulation),
// the FP32 backward pass and AdamW.  All reductions are per-workgroup partials
// summed in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include "itsca_device.h"
#include "itsca_mlp.h"
#include "itsca_mlp_device.h"
#include "itsca_mlp_kernels.h"

namespace itsca {
namespace {

// ---------------------------------------------------------------------------
// Pipeline filter.  Runs after the tracklet count.  Every wavefront handles
// the 64 clusters of its lanes on its own (wavefront barriers only): a
// shuffle scan of the candidate counts numbers their candidate slots (index
// o); the slots are taken in rounds of kWaveRound: (1) an LDS map o -> owning
// lane; (2) 64 candidates at a time are scored (features, BF16 MFMA network),
// the candidate values kept in LDS and the keep decisions as ballot words;
// (3) each lane counts the kept candidates of its own cluster, and the round's
// kept candidates are appended, in slot order (cluster after cluster), to the
// wavefront's kept list (keptListOf), 64 slots per ballot: coalesced stores,
// read in order by the tracklet fill.  Clusters whose candidates overflowed
// their slots are kept whole.  Kept counts go to trkKept (trkCnt stays the raw
// count that the features read) and the chunk sums are rewritten.  The LDS arrays
// allow four workgroups of 256 per CU; the waves-per-SIMD hint holds the
// registers to that occupancy.
// ---------------------------------------------------------------------------


template <int B>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) kTrackletsScore(DevState s, float threshold)
{
  constexpr int kWaves = B / 64;
  __shared__ uint8_t owner[kWaves][kWaveRound];
  __shared__ int candv[kWaves][kWaveRound];
  __shared__ uint32_t keep[kWaves][kWaveRound / 32];
  __shared__ float seeds[kSeedFloats][B];
  __shared__ __attribute__((aligned(16))) uint16_t tiles[kWaves][64 * kMlpInPad];
  __shared__ int waveKept[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const MlpWave m = loadMlp(s.mlpW);
  const int S = s.W * kTrackletLayers;
  const int* chunkStart = s.chunkStart + kStTracklets * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  const int4* chunkInfo = s.chunkInfo + (size_t)kStTracklets * s.maxChunks;
  // this chunk's entry (segment, first cluster, its global index, clusters)
  // and its clusters' raw tracklet counts; the next chunk's are loaded ahead
  auto rawOf = [&](const int4& e) { return (int)threadIdx.x < e.w ? s.trkCnt[e.z + threadIdx.x] : 0; };
  const int G = gridDim.x;
  int4 curInfo = (int)blockIdx.x < nChunks ? chunkInfo[blockIdx.x] : make_int4(0, 0, 0, 0);
  int rawCur = (int)blockIdx.x < nChunks ? rawOf(curInfo) : 0;
  int4 nextInfo = (int)blockIdx.x + G < nChunks ? chunkInfo[blockIdx.x + G] : make_int4(0, 0, 0, 0);
  for (int chunk = blockIdx.x; chunk < nChunks; chunk += G) {
    const int4 ci = curInfo;
    const int raw = rawCur;
    if (chunk + G < nChunks) {
      curInfo = nextInfo;
      rawCur = rawOf(nextInfo);
      if (chunk + 2 * G < nChunks) nextInfo = chunkInfo[chunk + 2 * G];
    }
    const int sg = ci.x;
    const int w = sg / kTrackletLayers, l = sg % kTrackletLayers;
    const Seg cur = s.clSeg[w * kLayers + l], next = s.clSeg[w * kLayers + l + 1];
    const int recCur = recSeg(s, w, l).base, recNext = recSeg(s, w, l + 1).base;
    const int first = ci.y;
    const int i = first + threadIdx.x;
    const int p = cur.base + i;
    const int scored = raw <= kTrkSlots ? raw : 0;
    int inc = scored;  // wavefront inclusive scan
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(inc, d, 64);
      if (lane >= d) inc += t;
    }
    const int ex = inc - scored;
    const int total = __shfl(inc, 63, 64);
    const float vx = s.itemVertex[3 * w], vy = s.itemVertex[3 * w + 1], vz = s.itemVertex[3 * w + 2];
    const int rowBase = cur.base + first + wave * 64;  // slot row of the cluster of lane c: rowBase + c
    if (scored > 0) {  // the cluster-i features, once per cluster
      static_assert(sizeof(TrackletSeed) == kSeedFloats * sizeof(float), "seed layout");
      const TrackletSeed t = seedOf(s, vx, vy, vz, recCur + i, p);
      const float* tv = reinterpret_cast<const float*>(&t);
      for (int k = 0; k < kSeedFloats; ++k) seeds[k][threadIdx.x] = tv[k];
    }
    int kept = 0, listed = 0;
    int* waveList = keptListOf(s.trkKeep, cur.base + first, wave);
    for (int r0 = 0; r0 < total; r0 += kWaveRound) {
      const int rn = min(kWaveRound, total - r0);
      for (int k = max(0, r0 - ex); k < scored && ex + k < r0 + rn; ++k) owner[wave][ex + k - r0] = (uint8_t)lane;
      waveSync();  // owners and seeds visible to the wavefront
      // Loads run two tiles ahead: the slot of tile t+2 and the records of
      // tile t+1 are in flight while tile t is scored.  All of them lie in
      // this round's slots, which the compaction below rewrites only after
      // the last tile.
      const int nTiles = (rn + 63) >> 6;
      auto slotOf = [&](int tile) {  // candidate j of this lane's slot in the tile, -1 past the round
        const int o = tile * 64 + lane;
        const int c = o < rn ? owner[wave][o] : 0;
        const int exc = __shfl(ex, c, 64);
        return o < rn ? s.trkCand[(size_t)(rowBase + c) * kTrkSlots + (r0 + o - exc)] : -1;
      };
      auto recOf = [&](int tile, int j) {  // also keeps j for the compaction
        CandRec q{make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, 0};
        if (j >= 0) {
          candv[wave][tile * 64 + lane] = j;
          q.b = s.cxyrz[recNext + j];
          q.phi = s.cphi[recNext + j];
          q.count = l + 1 < kTrackletLayers ? s.trkCnt[next.base + j] : 0;
        }
        return q;
      };
      CandRec qCur = recOf(0, slotOf(0));
      int jNext = nTiles > 1 ? slotOf(1) : -1;
      for (int tile = 0; tile < nTiles; ++tile) {
        const int jAfter = tile + 2 < nTiles ? slotOf(tile + 2) : -1;
        const CandRec qNext = recOf(tile + 1, jNext);
        const int o = tile * 64 + lane;  // slot index within the round
        float f[kMlpIn];
        if (o < rn) {
          const int c = owner[wave][o];
          TrackletSeed t;
          float* tv = reinterpret_cast<float*>(&t);
          for (int k = 0; k < kSeedFloats; ++k) tv[k] = seeds[k][wave * 64 + c];
          candidateFeatures(t, l, qCur.b.x - vx, qCur.b.y - vy, qCur.b.w, qCur.b.z, qCur.phi, vz, qCur.count, f);
        } else {
          for (int k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
        }
        uint4 lo, hi;
        packFeatureRow(f, lo, hi);
        storeTileRow(tiles[wave], lane, lo, hi);
        waveSync();
        const float z = mlpForward64(m, tiles[wave]);
        const unsigned long long kb = __ballot(o < rn && z >= threshold);
        if (lane == 0) {
          keep[wave][2 * tile] = (uint32_t)kb;
          keep[wave][2 * tile + 1] = (uint32_t)(kb >> 32);
        }
        waveSync();
        qCur = qNext;
        jNext = jAfter;
      }
      // this lane's kept candidates in the round: the set bits of its slots [a, b)
      const int a = max(ex - r0, 0), b = min(ex + scored - r0, rn);
      for (int wd = a >> 5; a < b && wd <= (b - 1) >> 5; ++wd) {
        uint32_t bits = keep[wave][wd];
        const int lo = wd * 32 < a ? a - wd * 32 : 0, hi = (wd + 1) * 32 > b ? b - wd * 32 : 32;
        if (lo > 0) bits &= ~((1u << lo) - 1u);
        if (hi < 32) bits &= (1u << hi) - 1u;
        kept += __builtin_popcount(bits);
      }
      // the round's kept candidates, in slot order (cluster after cluster),
      // appended to the wavefront's list: 64 slots at a time, ranked by a ballot
      for (int t0 = 0; t0 < rn; t0 += 64) {
        const int o = t0 + lane;
        const bool k = o < rn && ((keep[wave][o >> 5] >> (o & 31)) & 1u);
        const unsigned long long kb = __ballot(k);
        if (k) waveList[listed + __builtin_popcountll(kb & ((1ull << lane) - 1ull))] = candv[wave][o];
        listed += __builtin_popcountll(kb);
      }
      waveSync();  // owner, candv and keep are rebuilt by the next round
    }
    const int keptCount = scored == raw ? kept : raw;
    if (i < cur.count) s.trkKept[p] = keptCount;
    int sum = i < cur.count ? keptCount : 0;
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) waveKept[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int k = 0; k < kWaves; ++k) t += waveKept[k];
      s.chunkSum[(size_t)kStTracklets * s.maxChunks + chunk] = t;
    }
    __syncthreads();  // waveKept and the seeds are reused by the next chunk
  }
}

// After scoring every layer: the kept counts become the counts the rest of the
// pipeline reads.  Only the clusters of the tracklet layers that were searched
// change: layer l of an item is searched while its layers 0..l+1 all have
// clusters (kPlanTracklets); the others keep the zero counts of kPrep.
__global__ void kCommitKept(DevState s)
{
  const int S = s.W * kTrackletLayers;
  for (int sg = blockIdx.y; sg < S; sg += gridDim.y) {
    const int w = sg / kTrackletLayers, l = sg % kTrackletLayers;
    bool searched = true;
    for (int k = 0; k <= l + 1; ++k) searched = searched && s.clSeg[w * kLayers + k].count > 0;
    if (!searched) continue;
    const Seg g = s.clSeg[w * kLayers + l];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.count; i += gridDim.x * blockDim.x)
      s.trkCnt[g.base + i] = s.trkKept[g.base + i];
  }
}

// ---------------------------------------------------------------------------
// Training data.  Labels: a tracklet is positive when it is the first or the
// second tracklet of a cell of some final road.
// ---------------------------------------------------------------------------
__global__ void kLabelClear(uint8_t* label, int n)
{
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) label[t] = 0;
}

__global__ void kLabelRoads(DevState s, uint8_t* label)
{
  for (int w = blockIdx.x; w < s.W; w += gridDim.x) {
    const Seg roads = s.roadSeg[w];
    for (int r = threadIdx.x; r < roads.count; r += blockDim.x) {
      for (int k = 0; k < kCellLayers; ++k) {
        const int cell = s.rCells[(size_t)(roads.base + r) * kCellLayers + k];
        if (cell < 0) continue;
        const int c = s.celSeg[w * kCellLayers + k].base + cell;
        label[s.trkSeg[w * kTrackletLayers + k].base + s.ct1[c]] = 1;      // same value from every writer
        label[s.trkSeg[w * kTrackletLayers + k + 1].base + s.ct2[c]] = 1;
      }
    }
  }
}

// Every stride-th tracklet (global index t = phase + stride * row) becomes row
// `base + row` of the training set.
__global__ void kCollect(DevState s, const uint8_t* label, int total, int stride, int phase, uint16_t* outX,
                         uint8_t* outY, int base, int cap)
{
  const int S = s.W * kTrackletLayers;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; phase + (long)row * stride < total;
       row += gridDim.x * blockDim.x) {
    const int t = phase + row * stride;
    if (base + row >= cap) break;
    int lo = 0, hi = S;  // segment of tracklet t
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s.trkSeg[mid].base <= t)
        lo = mid;
      else
        hi = mid;
    }
    const int w = lo / kTrackletLayers, l = lo % kTrackletLayers;
    const int i = s.tFS[t].x, j = s.tFS[t].y;
    float f[kMlpIn];
    const float vx = s.itemVertex[3 * w], vy = s.itemVertex[3 * w + 1], vz = s.itemVertex[3 * w + 2];
    featuresOf(s, seedOf(s, vx, vy, vz, recSeg(s, w, l).base + i, s.clSeg[w * kLayers + l].base + i), vx, vy, vz,
               l, recSeg(s, w, l + 1).base + j, s.clSeg[w * kLayers + l + 1].base + j, f);
    packFeatures(f, outX + (size_t)(base + row) * kMlpInPad);
    outY[base + row] = label[t];
  }
}

// Logits of n samples of a training set (MFMA inference).
__global__ void __launch_bounds__(256) kScoreRows(const uint16_t* w, const uint16_t* X, int n, float* out)
{
  __shared__ __attribute__((aligned(16))) uint16_t tiles[4][64 * kMlpInPad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const MlpWave m = loadMlp(w);
  const int waves = gridDim.x * 4;
  for (int tile = blockIdx.x * 4 + wave; tile * 64 < n; tile += waves) {
    const int o = tile * 64 + lane;
    uint4 lo = make_uint4(0u, 0u, 0u, 0u), hi = lo;
    if (o < n) {
      const uint4* src = reinterpret_cast<const uint4*>(X + (size_t)o * kMlpInPad);
      lo = src[0];
      hi = src[1];
    }
    storeTileRow(tiles[wave], lane, lo, hi);
    waveSync();
    const float z = mlpForward64(m, tiles[wave]);
    if (o < n) out[o] = z;
    waveSync();
  }
}

// ---------------------------------------------------------------------------
// Training step: FP32 forward (BF16-rounded weights, inputs and activations,
// as in inference) and backward (the rounding passed straight through), with
// the weighted binary cross-entropy loss
//   L = sum_s c_s [ softplus(z_s) - y_s z_s ] / sum_s c_s,  c_s = posWeight if y_s else 1.
// Samples of a step are rows (a*k + b) mod n for k in [k0, k0+count).  Each
// workgroup writes its partial gradient (kMlpParams floats) and partial loss
// and weight sum; kAdamW sums them in workgroup order.
// ---------------------------------------------------------------------------
constexpr int kTile = 64;
constexpr int kGradSlots = kMlpParams + 2;  // gradients, loss, weight

__global__ void __launch_bounds__(256) kTrainGrad(const float* wq, const uint16_t* X, const uint8_t* Y, int n,
                                                  long a, long b, int k0, int count, float posWeight, float* partial)
{
  __shared__ float W[kMlpParams];
  __shared__ float xs[kTile][kMlpIn];
  __shared__ float h1[kTile][kMlpHidden], h2[kTile][kMlpHidden], g1[kTile][kMlpHidden], g2[kTile][kMlpHidden];
  __shared__ float dz[kTile];
  __shared__ float lossPart[kTile], weightPart[kTile];
  for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) W[p] = wq[p];
  __syncthreads();
  float acc[(kMlpParams + 255) / 256];
  for (int q = 0; q < (kMlpParams + 255) / 256; ++q) acc[q] = 0.0f;
  float lossAcc = 0.0f, weightAcc = 0.0f;
  const int per = (count + gridDim.x - 1) / gridDim.x;
  const int begin = blockIdx.x * per, end = min(count, begin + per);
  // Per tile of kTile samples, the forward and backward passes: sample
  // s = threadIdx / 4, its four threads each computing a quarter of every
  // hidden layer's units (unit o by thread (s, o / 8)); the logit, loss and
  // output gradient by the first of the four.  Every unit's sum runs in the
  // same order as a one-thread-per-sample pass would run it.
  constexpr int kParts = 256 / kTile, kUnits = kMlpHidden / kParts;
  static_assert(kParts * kTile == 256 && kUnits * kParts == kMlpHidden, "four threads per sample");
  const int sIdx = threadIdx.x / kParts, part = threadIdx.x % kParts;
  for (int t0 = begin; t0 < end; t0 += kTile) {
    const int k = t0 + sIdx;
    const bool live = k < end;
    const long row = live ? (a * (long)(k0 + k) + b) % n : 0;
    for (int f = part; f < kMlpIn; f += kParts) xs[sIdx][f] = live ? fromBf16(X[row * kMlpInPad + f]) : 0.0f;
    __syncthreads();
    for (int o = part * kUnits; o < (part + 1) * kUnits; ++o) {
      float v = W[kOffB1 + o];
      for (int f = 0; f < kMlpIn; ++f) v += W[kOffW1 + o * kMlpIn + f] * xs[sIdx][f];
      h1[sIdx][o] = live ? roundBf16(v > 0.0f ? v : 0.0f) : 0.0f;
    }
    __syncthreads();
    for (int o = part * kUnits; o < (part + 1) * kUnits; ++o) {
      float v = W[kOffB2 + o];
      for (int q = 0; q < kMlpHidden; ++q) v += W[kOffW2 + o * kMlpHidden + q] * h1[sIdx][q];
      h2[sIdx][o] = live ? roundBf16(v > 0.0f ? v : 0.0f) : 0.0f;
    }
    __syncthreads();
    if (part == 0) {
      float loss = 0.0f, weight = 0.0f, d = 0.0f;
      if (live) {
        float z = W[kOffB3];
        for (int o = 0; o < kMlpHidden; ++o) z += W[kOffW3 + o] * h2[sIdx][o];
        const float y = Y[row] ? 1.0f : 0.0f;
        weight = Y[row] ? posWeight : 1.0f;
        const float sp = z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z));  // softplus
        loss = weight * (sp - y * z);
        const float prob = 1.0f / (1.0f + expf(-z));
        d = weight * (prob - y);
      }
      dz[sIdx] = d;
      lossPart[sIdx] = loss;
      weightPart[sIdx] = weight;
    }
    __syncthreads();
    const float d = dz[sIdx];
    for (int o = part * kUnits; o < (part + 1) * kUnits; ++o) g2[sIdx][o] = h2[sIdx][o] > 0.0f ? d * W[kOffW3 + o] : 0.0f;
    __syncthreads();
    for (int q = part * kUnits; q < (part + 1) * kUnits; ++q) {
      float v = 0.0f;
      for (int o = 0; o < kMlpHidden; ++o) v += g2[sIdx][o] * W[kOffW2 + o * kMlpHidden + q];
      g1[sIdx][q] = h1[sIdx][q] > 0.0f ? v : 0.0f;
    }
    __syncthreads();
    for (int q = 0; q < (kMlpParams + 255) / 256; ++q) {
      const int p = threadIdx.x + 256 * q;
      if (p >= kMlpParams) break;
      float v = 0.0f;
      if (p < kOffB1) {
        const int o = p / kMlpIn, f = p % kMlpIn;
        for (int t = 0; t < kTile; ++t) v += g1[t][o] * xs[t][f];
      } else if (p < kOffW2) {
        for (int t = 0; t < kTile; ++t) v += g1[t][p - kOffB1];
      } else if (p < kOffB2) {
        const int o = (p - kOffW2) / kMlpHidden, q2 = (p - kOffW2) % kMlpHidden;
        for (int t = 0; t < kTile; ++t) v += g2[t][o] * h1[t][q2];
      } else if (p < kOffW3) {
        for (int t = 0; t < kTile; ++t) v += g2[t][p - kOffB2];
      } else if (p < kOffB3) {
        for (int t = 0; t < kTile; ++t) v += dz[t] * h2[t][p - kOffW3];
      } else {
        for (int t = 0; t < kTile; ++t) v += dz[t];
      }
      acc[q] += v;
    }
    if (threadIdx.x == 0)
      for (int t = 0; t < kTile; ++t) {
        lossAcc += lossPart[t];
        weightAcc += weightPart[t];
      }
    __syncthreads();
  }
  float* out = partial + (size_t)blockIdx.x * kGradSlots;
  for (int q = 0; q < (kMlpParams + 255) / 256; ++q) {
    const int p = threadIdx.x + 256 * q;
    if (p < kMlpParams) out[p] = acc[q];
  }
  if (threadIdx.x == 0) {
    out[kMlpParams] = lossAcc;
    out[kMlpParams + 1] = weightAcc;
  }
}

// Sums the partials in workgroup order, normalises by the total sample
// weight, and (when lr > 0) applies one AdamW step (decoupled weight decay on
// the weight matrices, none on biases).  Writes the gradient and the loss.
__global__ void __launch_bounds__(256) kAdamW(float* master, float* mom1, float* mom2, float* wq, uint16_t* wbf,
                                              const float* partial, int blocks, int step, float lr, float wd,
                                              float* gradOut, float* lossOut)
{
  __shared__ float total[2];
  if (threadIdx.x == 0) {
    float l = 0.0f, wsum = 0.0f;
    for (int b = 0; b < blocks; ++b) {
      l += partial[(size_t)b * kGradSlots + kMlpParams];
      wsum += partial[(size_t)b * kGradSlots + kMlpParams + 1];
    }
    total[0] = l;
    total[1] = wsum > 0.0f ? wsum : 1.0f;
    if (lossOut) *lossOut = l / total[1];
  }
  __syncthreads();
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) {
    float g = 0.0f;
    for (int b = 0; b < blocks; ++b) g += partial[(size_t)b * kGradSlots + p];
    g /= total[1];
    if (gradOut) gradOut[p] = g;
    if (lr > 0.0f) {
      const float m = beta1 * mom1[p] + (1.0f - beta1) * g;
      const float v = beta2 * mom2[p] + (1.0f - beta2) * g * g;
      mom1[p] = m;
      mom2[p] = v;
      const bool decay = p < kOffB1 || (p >= kOffW2 && p < kOffB2) || (p >= kOffW3 && p < kOffB3);
      float w = master[p];
      if (decay) w -= lr * wd * w;
      w -= lr * (m / c1) / (sqrtf(v / c2) + eps);
      master[p] = w;
      wq[p] = roundBf16(w);
      wbf[p] = toBf16(w);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
void launchTrackletFilter(const DevState& s, int block, int blocks, float threshold, hipStream_t stream)
{
  switch (block) {
    case 64: hipLaunchKernelGGL(kTrackletsScore<64>, dim3(blocks), dim3(64), 0, stream, s, threshold); break;
    case 128: hipLaunchKernelGGL(kTrackletsScore<128>, dim3(blocks), dim3(128), 0, stream, s, threshold); break;
    default: hipLaunchKernelGGL(kTrackletsScore<256>, dim3(blocks), dim3(256), 0, stream, s, threshold); break;
  }
  launchCommitKept(s, stream);
}

void launchCommitKept(const DevState& s, hipStream_t stream)
{
  const int segs = s.W * kTrackletLayers;
  hipLaunchKernelGGL(kCommitKept, dim3(8, segs < 1 ? 1 : (segs < 4096 ? segs : 4096)), dim3(256), 0, stream, s);
}

void launchLabelAndCollect(const DevState& s, int totalTracklets, uint8_t* label, int stride, int phase,
                           uint16_t* outX, uint8_t* outY, int base, int cap, hipStream_t stream)
{
  hipLaunchKernelGGL(kLabelClear, dim3(256), dim3(256), 0, stream, label, totalTracklets);
  hipLaunchKernelGGL(kLabelRoads, dim3(s.W < 1024 ? (s.W > 0 ? s.W : 1) : 1024), dim3(256), 0, stream, s, label);
  hipLaunchKernelGGL(kCollect, dim3(512), dim3(256), 0, stream, s, label, totalTracklets, stride, phase, outX, outY,
                     base, cap);
}

void launchScoreRows(const uint16_t* wbf, const uint16_t* X, int n, float* out, hipStream_t stream)
{
  hipLaunchKernelGGL(kScoreRows, dim3(512), dim3(256), 0, stream, wbf, X, n, out);
}

void launchTrainStep(const MlpBuffers& b, const uint16_t* X, const uint8_t* Y, int n, long a, long off, int k0,
                     int count, float posWeight, int step, float lr, float wd, float* gradOut, float* lossOut,
                     hipStream_t stream)
{
  hipLaunchKernelGGL(kTrainGrad, dim3(b.gradBlocks), dim3(256), 0, stream, b.wq, X, Y, n, a, off, k0, count,
                     posWeight, b.partial);
  hipLaunchKernelGGL(kAdamW, dim3(1), dim3(256), 0, stream, b.master, b.mom1, b.mom2, b.wq, b.wbf, b.partial,
                     b.gradBlocks, step, lr, wd, gradOut, lossOut);
}

}  // namespace itsca
