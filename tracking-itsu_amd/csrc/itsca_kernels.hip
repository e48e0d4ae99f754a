// itsca_kernels.hip -- gfx950 kernels of the plain cellular-automaton track
// finder.  Every stage is count -> scan -> fill: a count kernel computes each
// element's number of outputs and a per-chunk sum, a single-workgroup scan
// turns chunk sums into global offsets, and a fill kernel recomputes the
// outputs and writes them in the reference's emission order.  There are no
// global atomics; LDS atomics are used only where the result is made
// order-independent afterwards.
//
// Compiled with -ffp-contract=off and correctly rounded fp32 divide and sqrt,
// so each float expression rounds as the reference's x86-64 SSE build does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "itsca_device.h"
#include "itsca_kernels.h"
#include "itsca_mlp_device.h"
#include "itsca_mlp_kernels.h"

namespace itsca {
namespace {

// ---------------------------------------------------------------------------
// Block-level helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int waveInclusiveScan(int v)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan over the block; `scratch` holds blockDim/64 + 1 ints.
__device__ __forceinline__ int blockExclusiveScan(int v, int* scratch, int& total)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nWaves = blockDim.x >> 6;
  const int inc = waveInclusiveScan(v);
  if (lane == 63) scratch[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < nWaves; ++w) {
      const int t = scratch[w];
      scratch[w] = s;
      s += t;
    }
    scratch[nWaves] = s;
  }
  __syncthreads();
  const int ex = inc - v + scratch[wave];
  total = scratch[nWaves];
  __syncthreads();
  return ex;
}

// The same in 64 bits (single-workgroup scans of chunk sums whose total may
// pass 2^31); `scratch` holds blockDim/64 + 1 values.
__device__ __forceinline__ long long blockExclusiveScan64(long long v, long long* scratch, long long& total)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nWaves = blockDim.x >> 6;
  long long inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const long long t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  if (lane == 63) scratch[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int w = 0; w < nWaves; ++w) {
      const long long t = scratch[w];
      scratch[w] = s;
      s += t;
    }
    scratch[nWaves] = s;
  }
  __syncthreads();
  const long long ex = inc - v + scratch[wave];
  total = scratch[nWaves];
  __syncthreads();
  return ex;
}

__device__ __forceinline__ int blockSum(int v, int* scratch)
{
  int total;
  blockExclusiveScan(v, scratch, total);
  return total;
}

// Chunks of a stage are dealt so that each XCD (8 per device, blocks dealt
// round-robin) walks a contiguous range, which keeps neighbouring chunks --
// which read the same next-layer data -- in the same L2.
__device__ __forceinline__ int xcdBlockIndex()
{
  const int g = gridDim.x;
  if ((g & 7) != 0) return blockIdx.x;
  return (blockIdx.x & 7) * (g >> 3) + (blockIdx.x >> 3);
}

__device__ __forceinline__ int segmentOfChunk(const int* chunkStart, int S, int chunk)
{
  int lo = 0, hi = S;  // chunkStart[lo] <= chunk < chunkStart[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chunkStart[mid] <= chunk)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Stage (a): per-layer cluster transform and stable counting sort by bin.
// PrimaryVertexContext.cxx:31-142, Cluster.cxx:35-45.
// One workgroup per (item, layer).
// ---------------------------------------------------------------------------
// kPrepPer: clusters per thread on the LDS path.  Layers up to kPrepPer * 1024
// clusters are sorted through an LDS array of that many ints: 64 KB for
// kPrepPer = 16 (two workgroups per CU), 128 KB for kPrepPer = 32 (batches
// with larger layers, such as pile-up events).  Positions are packed as 16
// bits, so kPrepPer * 1024 <= 65536.  The waves-per-SIMD hint keeps the
// scalar registers to what eight waves allow: at seven, only one workgroup of
// 16 waves fits a CU.
constexpr int kPrepThreads = 1024;

template <int kPrepPer>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) kPrep(DevState s)
{
  constexpr int kPrepLds = kPrepPer * kPrepThreads;
  static_assert(kPrepLds <= 65536, "sorted positions are packed as 16 bits");
  static_assert(kBins * kMaxZSub <= kPrepLds, "sub-bin counters live in the bucket array");
  __shared__ int lBucket[kPrepLds];
  __shared__ int hist[kTableSize];
  __shared__ int tab[kTableSize];
  __shared__ int scratch[1024 / 64 + 1];
  __shared__ int geo[4];  // min r, max r (float bits; r >= 0 orders like ints), min and max z bin index
  __shared__ int negCount;  // clusters with a negative bin (z below the layer in phi row 0)
  const int w = blockIdx.x / kLayers, l = blockIdx.x % kLayers;
  if (blockIdx.x == 0 && threadIdx.x < kNumStatus) s.status[threadIdx.x] = 0;
  const Seg seg = s.clSeg[w * kLayers + l];
  if (s.recItem[w] != w) {  // reads the records of an earlier vertex: only its own counts are reset
    for (int i = threadIdx.x; i < seg.count; i += blockDim.x) s.trkCnt[seg.base + i] = 0;
    return;
  }
  const int e = s.itemEvent[w];
  const int rawBase = s.evLayerOff[e * 8 + l];
  const float vx = s.itemVertex[3 * w], vy = s.itemVertex[3 * w + 1];
  const int n = seg.count, base = seg.base;
  for (int b = threadIdx.x; b < kTableSize; b += blockDim.x) hist[b] = 0;
  if (threadIdx.x == 0) {
    geo[0] = __float_as_int(3.0e38f);
    geo[1] = 0;
    geo[2] = kZBins;
    geo[3] = -1;
    negCount = 0;
  }
  __syncthreads();
  int rMinBits = __float_as_int(3.0e38f), rMaxBits = 0, zMin = kZBins, zMax = -1;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float x = s.rx[rawBase + i], y = s.ry[rawBase + i], z = s.rz[rawBase + i];
    const float dx = x - vx, dy = y - vy;
    const float phi = clusterPhi(dx, dy);
    const float r = sqrtf(dx * dx + dy * dy);
    const int zb = zBinIndex(l, z);
    const int b = binIndex(zb, phiBinIndex(phi));
    s.sPhi[base + i] = phi;
    s.sR[base + i] = r;
    s.sBin[base + i] = b;
    s.trkCnt[base + i] = 0;
    atomicAdd(b >= 0 ? &hist[b] : &negCount, 1);
    rMinBits = min(rMinBits, __float_as_int(r));  // a NaN r has bits above every finite r
    rMaxBits = max(rMaxBits, __float_as_int(r));
    zMin = min(zMin, zb);
    zMax = max(zMax, zb);
  }
  for (int d = 32; d >= 1; d >>= 1) {
    rMinBits = min(rMinBits, __shfl_xor(rMinBits, d, 64));
    rMaxBits = max(rMaxBits, __shfl_xor(rMaxBits, d, 64));
    zMin = min(zMin, __shfl_xor(zMin, d, 64));
    zMax = max(zMax, __shfl_xor(zMax, d, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&geo[0], rMinBits);
    atomicMax(&geo[1], rMaxBits);
    atomicMin(&geo[2], zMin);
    atomicMax(&geo[3], zMax);
  }
  __syncthreads();
  if (threadIdx.x == 0)
    s.segGeo[w * kLayers + l] = make_float4(__int_as_float(geo[0]), __int_as_float(geo[1]),
                                            geo[2] >= 0 && geo[3] < kZBins ? 1.0f : 0.0f, 0.0f);
  // Exclusive scan of the 401 bin counts: table[b] = #clusters with bin < b
  // for b >= 1, and table[0] = 0 (PrimaryVertexContext.cxx:90-119).  Clusters
  // with a negative bin sort first, at positions [0, nNeg), so they fall in
  // bin 0's range [table[0], table[1]) as in the reference; tab[b] is where the
  // clusters of bin b itself start.
  int* table = s.table + (size_t)(w * kLayers + l) * kTableSize;
  const int nNeg = negCount;
  {
    const int v = threadIdx.x < kTableSize ? hist[threadIdx.x] : 0;
    int total;
    const int ex = blockExclusiveScan(v, scratch, total) + nNeg;
    if (threadIdx.x < kTableSize) {
      hist[threadIdx.x] = ex;  // becomes the insertion cursor
      tab[threadIdx.x] = ex;
      table[threadIdx.x] = threadIdx.x == 0 ? 0 : ex;
    }
  }
  __syncthreads();
  // Stable sorted positions, one tile of 1024 clusters (in index order) at a
  // time: clusters of equal bin keep their index order.  Inside a wavefront,
  // a cluster's rank among those of its bin comes from a bin match (one
  // ballot per bit of the bin); the wavefronts' bin counts of the tile are
  // scanned across wavefronts on top of each bin's running cursor (hist).
  // Every thread calls it with the same t0; it returns the position of
  // cluster t0 + threadIdx.x, or -1 past the end.
  constexpr int kPrepWaves = kPrepThreads / 64;
  static_assert(kPrepWaves * kTableSize <= kPrepLds, "per-wavefront bin counts live in the bucket array");
  int* waveHist = lBucket;  // [kPrepWaves][kTableSize]
  // A cluster with a negative bin takes its rank among those (ordered by bin,
  // then index, as a stable sort orders them) from a scan of the layer's bins;
  // such clusters lie outside the detector and are normally absent.
  auto negativeRank = [&](int i, int b) {
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const int bj = s.sBin[base + j];
      rank += (bj < b || (bj == b && j < i)) ? 1 : 0;
    }
    return rank;
  };
  auto tilePositions = [&](int t0) -> int {
    const int i = t0 + threadIdx.x;
    const int bRaw = i < n ? s.sBin[base + i] : 511;
    const bool neg = bRaw < 0;
    const bool valid = i < n && !neg;
    const int b = valid ? bRaw : 511;  // bins of the match are 0..400 (9 bits)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < 9; ++bit) {
      const unsigned long long set = __ballot((b >> bit) & 1);
      peers &= ((b >> bit) & 1) ? set : ~set;
    }
    const unsigned long long below = peers & ((1ull << lane) - 1ull);
    for (int k = threadIdx.x; k < kPrepWaves * kTableSize; k += blockDim.x) waveHist[k] = 0;
    __syncthreads();
    if (valid && below == 0ull) waveHist[wave * kTableSize + b] = __popcll(peers);
    __syncthreads();
    for (int bb = threadIdx.x; bb < kTableSize; bb += blockDim.x) {
      int run = hist[bb];
      for (int wv = 0; wv < kPrepWaves; ++wv) {
        const int c = waveHist[wv * kTableSize + bb];
        waveHist[wv * kTableSize + bb] = run;
        run += c;
      }
      hist[bb] = run;
    }
    __syncthreads();
    int pos = valid ? waveHist[wave * kTableSize + b] + __popcll(below) : -1;
    __syncthreads();  // the counts are cleared by the next tile
    if (nNeg > 0 && neg) pos = negativeRank(i, bRaw);
    return pos;
  };
  // When the layer fits in LDS the positions stay in registers and the
  // records are permuted through LDS; else the source of every sorted
  // position goes to global scratch and the records are gathered.
  if (n <= kPrepLds) {
    // The sorted position of each of this thread's clusters, i = threadIdx.x + k * 1024,
    // two 16-bit positions per register.
    uint32_t pos2[kPrepPer / 2];
#pragma unroll
    for (int k = 0; k < kPrepPer / 2; ++k) pos2[k] = 0;
#pragma unroll
    for (int k = 0; k < kPrepPer; ++k) {
      if (k * kPrepThreads < n) {
        const int pos = tilePositions(k * kPrepThreads);
        if (pos >= 0) pos2[k >> 1] |= (uint32_t)pos << (16 * (k & 1));
      }
    }
    // Each field goes through LDS: read in original order (coalesced), stored
    // at its sorted position in LDS, written out in sorted order (coalesced).
    // Addresses are a uniform base pointer plus an unsigned 32-bit byte offset.
    auto ld = [](const void* b, int i) {
      return *reinterpret_cast<const int*>(static_cast<const char*>(b) + 4u * (unsigned)i);
    };
    auto st = [](void* b, int i, int v) { *reinterpret_cast<int*>(static_cast<char*>(b) + 4u * (unsigned)i) = v; };
    // The loads of a field are issued together (indices past the layer
    // clamped to its last cluster), then stored: a conditional load per
    // element would wait for each one before its store.
    auto stageField = [&](const void* src) {
      if (n > 0) {
        int v[kPrepPer];
#pragma unroll
        for (int k = 0; k < kPrepPer; ++k) v[k] = ld(src, min((int)threadIdx.x + k * kPrepThreads, n - 1));
#pragma unroll
        for (int k = 0; k < kPrepPer; ++k)
          if ((int)threadIdx.x + k * kPrepThreads < n) lBucket[(pos2[k >> 1] >> (16 * (k & 1))) & 0xffffu] = v[k];
      }
      __syncthreads();
    };
    auto field = [&](const void* src, void* dst) {
      stageField(src);
      for (int p = threadIdx.x; p < n; p += kPrepThreads) st(dst, p, lBucket[p]);
      __syncthreads();
    };
    field(s.rid + rawBase, s.cid + base);
    field(s.rmc + rawBase, s.cmc + base);
    field(s.sBin + base, s.cbin + base);
    field(s.sR + base, s.cr + base);
    field(s.rz + rawBase, s.cz + base);
    // x and y, in sorted order, go to the bin and r scratch (read above, free
    // now) until the packed records are written.
    field(s.rx + rawBase, s.sBin + base);
    field(s.ry + rawBase, s.sR + base);
    // A sorted position is always written by the same thread, which reads back
    // its own x, y, r and z for the packed records.
    stageField(s.sPhi + base);
    for (int p = threadIdx.x; p < n; p += kPrepThreads) {
      const float phi = __int_as_float(lBucket[p]);
      const float r = __int_as_float(ld(s.cr + base, p)), z = __int_as_float(ld(s.cz + base, p));
      st(s.cphi + base, p, __float_as_int(phi));
      s.crzp[base + p] = make_float4(r, z, phi, 0.0f);
      s.cxyrz[base + p] = make_float4(__int_as_float(ld(s.sBin + base, p)), __int_as_float(ld(s.sR + base, p)), r, z);
    }
  } else {
    for (int t0 = 0; t0 < n; t0 += blockDim.x) {
      const int pos = tilePositions(t0);
      if (pos >= 0) s.sInv[base + pos] = t0 + threadIdx.x;
    }
    __threadfence_block();
    __syncthreads();
    // The sorted records are written position by position (coalesced stores,
    // gathered loads).
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
      const int i = s.sInv[base + p];
      const int q = base + p;
      const float z = s.rz[rawBase + i], phi = s.sPhi[base + i], r = s.sR[base + i];
      s.cxyrz[q] = make_float4(s.rx[rawBase + i], s.ry[rawBase + i], r, z);
      s.cz[q] = z;
      s.cphi[q] = phi;
      s.cr[q] = r;
      s.crzp[q] = make_float4(r, z, phi, 0.0f);
      s.cid[q] = s.rid[rawBase + i];
      s.cmc[q] = s.rmc[rawBase + i];
      s.cbin[q] = s.sBin[base + i];
    }
  }
  // The tracklet search index (no counterpart in the reference; only with
  // more than one sub-bin per bin).  Inside every bin of at most kSubMaxBin
  // clusters the clusters are grouped by sub-bin (phi sub-row, then z
  // sub-bin; a counting sort with LDS counters; the order inside a sub-bin is
  // arbitrary and never matters) and written to crzpS with their offset in the
  // bin; subTab gets the start of every sub-bin.  Larger bins and the
  // phi == 2pi slot keep the sorted order.  Every sorted position p is visited
  // by the thread that wrote its record.
  const int S = s.zSub, P = s.phiSub, NS = S * P;
  if (NS <= 1) return;
  __threadfence_block();
  __syncthreads();
  int* subCnt = lBucket;  // kBins * NS counters, then cursors
  for (int k = threadIdx.x; k < kBins * NS; k += blockDim.x) subCnt[k] = 0;
  __syncthreads();
  auto smallBin = [&](int b) { return b >= 0 && b < kBins && tab[b + 1] - tab[b] <= kSubMaxBin; };
  auto subOf = [&](int b, int p) {  // sub-bin of sorted position p in bin b (row b / kPhiBins, z bin b % kPhiBins)
    return phiSubIndex(s.cphi[base + p], b / kPhiBins, P) * S + zSubIndex(l, s.cz[base + p], b % kPhiBins, S);
  };
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const int b = s.cbin[base + p];
    if (smallBin(b)) atomicAdd(&subCnt[b * NS + subOf(b, p)], 1);
  }
  __syncthreads();
  // Byte t of a bin's entry: members with sub-bin < t (t = NS and above: all).
  uint4* subTab = s.subTab + (size_t)(w * kLayers + l) * kBins;
  for (int b = threadIdx.x; b < kBins; b += blockDim.x) {
    uint32_t e[4] = {0u, 0u, 0u, 0u};
    if (smallBin(b)) {
      int acc = 0;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        e[t >> 2] |= (uint32_t)acc << (8 * (t & 3));
        if (t < NS) {
          const int c = subCnt[b * NS + t];
          subCnt[b * NS + t] = acc;
          acc += c;
        }
      }
    }
    subTab[b] = make_uint4(e[0], e[1], e[2], e[3]);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const int b = s.cbin[base + p];
    const int lo = b < 0 ? 0 : (b < kBins ? tab[b] : tab[kBins]);
    int q = p;
    if (smallBin(b)) q = lo + atomicAdd(&subCnt[b * NS + subOf(b, p)], 1);
    const float4 rec = s.crzp[base + p];
    s.crzpS[base + q] = make_float4(rec.x, rec.y, rec.z, __int_as_float(p - lo));
  }
}

// ---------------------------------------------------------------------------
// Plans.  A segment is one (item, layer); plan kernels turn per-segment element
// counts into chunk ranges.  Single workgroup.
// ---------------------------------------------------------------------------
__device__ void planChunks(const int* counts, int S, int chunk, int* chunkStart, int* scratch)
{
  int carry = 0;
  for (int s0 = 0; s0 < S; s0 += blockDim.x) {
    const int sIdx = s0 + threadIdx.x;
    const int c = sIdx < S ? (counts[sIdx] + chunk - 1) / chunk : 0;
    int total;
    const int ex = blockExclusiveScan(c, scratch, total);
    if (sIdx < S) chunkStart[sIdx] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) chunkStart[S] = carry;
}

// Per chunk of a stage: its segment, its first element within the segment,
// the global index of that element and the number of elements it holds, so
// that a stage kernel finds them with one load instead of a binary search of
// dependent loads per chunk.  Grid-wide; launched after the stage's plan,
// when the segments of the stage's elements are final.  elements: where the
// element segment of segment sg is -- 0: clusters of (item, layer) with
// sg = w*6+l; 1: tracklets, sg = w*5+l; 2: cells, sg = w*4+l; 3: cells of
// cell layer `layer`, sg = w.
__global__ void __launch_bounds__(256) kChunkSegs(DevState s, int stage, int S, int elements, int layer)
{
  const int* chunkStart = s.chunkStart + stage * (s.maxSegs + 1);
  int4* info = s.chunkInfo + (size_t)stage * s.maxChunks;
  const int nChunks = chunkStart[S];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nChunks; c += gridDim.x * blockDim.x) {
    const int sg = segmentOfChunk(chunkStart, S, c);
    const int first = (c - chunkStart[sg]) * s.chunk;
    Seg e;
    if (elements == 0)
      e = s.clSeg[(sg / kTrackletLayers) * kLayers + sg % kTrackletLayers];
    else if (elements == 1)
      e = s.trkSeg[(sg / kCellLayers) * kTrackletLayers + sg % kCellLayers];
    else if (elements == 2)
      e = s.celSeg[(sg / (kCellLayers - 1)) * kCellLayers + sg % (kCellLayers - 1)];
    else
      e = s.celSeg[sg * kCellLayers + layer];
    info[c] = make_int4(sg, first, e.base + first, min(s.chunk, e.count - first));
  }
}

// The chunks of a stage that a workgroup walks (dealt XCD by XCD, as
// xcdBlockIndex), each with its entry (segment, first element, global index
// of the first element, elements) from the stage's chunk map.  Entries are
// loaded two chunks ahead; after take(), next() is the following chunk's
// entry, so a kernel can issue that chunk's first loads while it works on the
// current one.
struct ChunkWalk {
  const int4* info;
  int n, step, chunk;
  int4 cur, nxt;
  __device__ __forceinline__ ChunkWalk(const DevState& s, int stage, int nChunks, int start)
      : info(s.chunkInfo + (size_t)stage * s.maxChunks), n(nChunks), step(gridDim.x), chunk(start)
  {
    cur = chunk < n ? info[chunk] : make_int4(0, 0, 0, 0);
    nxt = chunk + step < n ? info[chunk + step] : make_int4(0, 0, 0, 0);
  }
  __device__ __forceinline__ bool valid() const { return chunk < n; }
  __device__ __forceinline__ int4 take()
  {
    const int4 c = cur;
    cur = nxt;
    nxt = chunk + 2 * step < n ? info[chunk + 2 * step] : make_int4(0, 0, 0, 0);
    return c;
  }
  __device__ __forceinline__ bool hasNext() const { return chunk + step < n; }
  __device__ __forceinline__ const int4& next() const { return cur; }
  __device__ __forceinline__ void advance() { chunk += step; }
};

// Tracklet layer l is computed only while every layer 0..l+1 has clusters
// (early return, Tracker.cxx:49-52).
__global__ void __launch_bounds__(1024) kPlanTracklets(DevState s, int* counts)
{
  __shared__ int scratch[1024 / 64 + 1];
  const int S = s.W * kTrackletLayers;
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const int w = i / kTrackletLayers, l = i % kTrackletLayers;
    bool ok = true;
    for (int k = 0; k <= l + 1; ++k) ok = ok && s.clSeg[w * kLayers + k].count > 0;
    counts[i] = ok ? s.clSeg[w * kLayers + l].count : 0;
  }
  __syncthreads();
  planChunks(counts, S, s.chunk, s.chunkStart + kStTracklets * (s.maxSegs + 1), scratch);
}

// Single-workgroup scan of a stage's chunk sums into global output offsets and
// per-segment (base, count); flags overflow of the output capacity.
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 16;

__global__ void __launch_bounds__(kScanThreads) kScanChunks(DevState s, int stage, int S, Seg* outSeg, int cap,
                                                            int statusSlot)
{
  __shared__ long long scratch[kScanThreads / 64 + 1];
  __shared__ int tile[kScanThreads * (kScanPer + 1)];
  const int* chunkStart = s.chunkStart + stage * (s.maxSegs + 1);
  const int* sums = s.chunkSum + (size_t)stage * s.maxChunks;
  int* offs = s.chunkOff + (size_t)stage * s.maxChunks;
  const int nChunks = chunkStart[S];
  // Tiles of 1024 * kScanPer sums go through LDS (coalesced loads and
  // stores); each thread scans kScanPer consecutive sums of the tile (an LDS
  // row padded by one against bank conflicts), then one block scan of the
  // thread totals.  The running sums are 64-bit (a stage's total may pass
  // 2^31, e.g. roads); offsets are stored clamped to the capacity, which is
  // where every fill stops writing.
  // The sums of the next tile are loaded into registers while the current
  // tile is scanned.
  constexpr int kTile = kScanThreads * kScanPer;
  long long carry = 0;
  auto slot = [](int idx) { return (idx / kScanPer) * (kScanPer + 1) + idx % kScanPer; };
  int next[kScanPer];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int idx = k * kScanThreads + threadIdx.x;
    next[k] = idx < nChunks ? sums[idx] : 0;
  }
  for (int c0 = 0; c0 < nChunks; c0 += kTile) {
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) tile[slot(k * kScanThreads + threadIdx.x)] = next[k];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      const int idx = c0 + kTile + k * kScanThreads + threadIdx.x;
      next[k] = idx < nChunks ? sums[idx] : 0;
    }
    __syncthreads();
    int* row = tile + threadIdx.x * (kScanPer + 1);
    long long pre[kScanPer];
    long long mine = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      pre[k] = mine;
      mine += row[k];
    }
    long long total;
    const long long ex = carry + blockExclusiveScan64(mine, scratch, total);  // its barriers order the row reads
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) row[k] = (int)min(ex + pre[k], (long long)cap);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      const int idx = k * kScanThreads + threadIdx.x;
      if (c0 + idx < nChunks) offs[c0 + idx] = tile[slot(idx)];
    }
    carry += total;
    __syncthreads();  // the tile is refilled
  }
  // On overflow every segment of the stage is emptied, so no later stage reads
  // past the capacity; the host grows the buffers and runs again.
  const bool overflow = carry > (long long)cap;
  const int grand = (int)min(carry, (long long)cap);
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const int a = chunkStart[i], b = chunkStart[i + 1];
    const int base = a < nChunks ? offs[a] : grand;
    const int end = b < nChunks ? offs[b] : grand;
    outSeg[i].base = overflow ? 0 : base;
    outSeg[i].count = overflow ? 0 : end - base;
  }
  if (threadIdx.x == 0 && overflow) s.status[statusSlot] = (int)min(carry, 0x7fffffffLL);
}

// ---------------------------------------------------------------------------
// Stage (b): tracklets.  Tracker.cxx:44-117, TrackingUtils.cxx:30-48,
// Tracklet.cxx:32-40.  One thread per cluster of layer l.
// ---------------------------------------------------------------------------
struct TrackletWindow {
  int x, y, z, w;
  int rows;
  bool empty;
};

__device__ __forceinline__ TrackletWindow trackletWindow(float z, float r, float phi, float vz, int l)
{
  TrackletWindow win;
  const float tanLambda = (z - vz) / r;
  const float dirZ = tanLambda * (layerR(l + 1) - r) + z;
  const float zMin = dirZ - 2 * kZCut;
  const float phiMin = phi - kPhiCut;
  const float zMax = dirZ + 2 * kZCut;
  const float phiMax = phi + kPhiCut;
  const float hz = layerHalfZ(l + 1);
  if (zMax < -hz || zMin > hz || zMin > zMax) {
    win.x = win.y = win.z = win.w = 0;
  } else {
    const int zlo = zBinIndex(l + 1, zMin), zhi = zBinIndex(l + 1, zMax);
    win.x = zlo > 0 ? zlo : 0;
    win.y = phiBinIndex(normalizedPhi(phiMin));
    win.z = zhi < kZBins - 1 ? zhi : kZBins - 1;
    win.w = phiBinIndex(normalizedPhi(phiMax));
  }
  // (0,0,0,0) doubles as the empty rectangle (Tracker.cxx:68-71).
  win.empty = win.x == 0 && win.y == 0 && win.z == 0 && win.w == 0;
  int rows = win.w - win.y + 1;
  if (rows < 0) rows += kPhiBins;
  win.rows = rows;
  return win;
}

// Visits every accepted candidate of cluster i in the reference's order
// (window row order, then ascending candidate position).
// Skipping z bins.  Within a window row the reference tests every cluster of
// the z bins [win.x, win.z] (a +-1 cm window) although the cut keeps only
// |tanL*(r_j - r) + z - z_j| < maxDeltaZ (0.1-0.3 cm).  A cluster sits in the z
// bin slot of its row that zBinIndex (monotone in z) gives it, as long as that
// index lies in [0, kZBins) -- otherwise getBinIndex spills it into another
// row, and nothing is skipped.  With r_j in the next layer's [min r, max r],
// a cluster can pass only if z_j lies in [zLo, zHi] = the range of
// tanL*(r_j - r) + z, widened by maxDeltaZ and a margin (kZSkipMargin, far
// above float rounding at these magnitudes); bins wholly below zBinIndex(zLo)
// or above zBinIndex(zHi) are skipped.  The reference's extra candidate (the
// inclusive upper bound) is still tested, after the others, so the accepted
// candidates and their order are unchanged.
constexpr float kZSkipMargin = 0.01f;


// The search of one cluster: its window, and the z bins of each window row
// whose clusters are tested.
//
// phi gate (the learned form only, no counterpart in the reference): with
// phiCut < kPhiCut the phi cut becomes |dphi| < phiCut (the same comparisons,
// the smaller bound), so the accepted candidates are the plain ones that pass
// the gate, in the plain order.  The window rows are the plain ones, but the
// bins of a row are scanned only when the row can hold a candidate inside the
// gate: the rows of [phi - phiCut, phi + phiCut] widened by kPhiSkipMargin
// (phiBinIndex is monotone), a cyclic sub-range of the plain rows.  The
// clusters of the other rows' bins lie in those rows (every z bin index of the
// next layer in [0, kZBins), else nothing is narrowed), so outside the gate.
// Every row's extra candidate (the first cluster after its bins, which can sit
// in any later row when bins are empty) is still tested.
constexpr float kPhiSkipMargin = 1.0e-3f;

struct TrackletSearch {
  TrackletWindow win;
  float tanLambda, maxDeltaZ;
  int zFirst, zLast;
  float zLo, zHi;  // every candidate that can pass has z in [zLo, zHi] (when sub)
  bool sub;        // the z range holds: bins, then z sub-bins, outside it are skipped
  float phiCut;    // the phi cut: kPhiCut, or the learned form's gate
  // the window rows whose bins are scanned: scanRows rows from scanFirst
  // (narrow: a gated sub-range of the window's rows, else all of them)
  int scanFirst, scanRows;
  bool narrow;
};

// The learned form's phi gate of tracklet layer l (kPhiCut: none).
__device__ __forceinline__ float trackletPhiCut(const DevState& s, int l) { return s.gated ? s.phiGate[l] : kPhiCut; }

__device__ __forceinline__ TrackletSearch trackletSearch(float4 geoNext, float z, float r, float phi, float vz, int l,
                                                         float phiCut)
{
  TrackletSearch t;
  t.win = trackletWindow(z, r, phi, vz, l);
  t.tanLambda = (z - vz) / r;
  t.maxDeltaZ = trackletMaxDeltaZ(l);
  t.zFirst = t.win.x;
  t.zLast = t.win.z;
  t.zLo = t.zHi = 0.0f;
  t.sub = false;
  t.phiCut = phiCut < kPhiCut ? phiCut : kPhiCut;
  t.scanFirst = t.win.y;
  t.scanRows = t.win.rows;
  t.narrow = false;
  if (t.phiCut < kPhiCut && !t.win.empty && geoNext.z != 0.0f) {
    const int y = phiBinIndex(normalizedPhi(phi - (t.phiCut + kPhiSkipMargin)));
    const int w = phiBinIndex(normalizedPhi(phi + (t.phiCut + kPhiSkipMargin)));
    const bool inRange = y >= 0 && y < kPhiBins && w >= 0 && w < kPhiBins && t.win.y >= 0 && t.win.y < kPhiBins &&
                         t.win.w >= 0 && t.win.w < kPhiBins;
    if (inRange) {
      int rows = w - y + 1;
      if (rows < 0) rows += kPhiBins;
      t.scanFirst = y;
      t.scanRows = rows;
      t.narrow = true;
    }
  }
  if (!t.win.empty && geoNext.z != 0.0f) {
    const float zA = t.tanLambda * (geoNext.x - r) + z, zB = t.tanLambda * (geoNext.y - r) + z;
    const float zLo = fminf(zA, zB) - t.maxDeltaZ - kZSkipMargin, zHi = fmaxf(zA, zB) + t.maxDeltaZ + kZSkipMargin;
    if (zLo > -1.0e4f && zHi < 1.0e4f && zLo <= zHi) {  // finite and ordered; else test every bin
      t.zFirst = max(t.zFirst, zBinIndex(l + 1, zLo));
      t.zLast = min(t.zLast, zBinIndex(l + 1, zHi));
      t.zLo = zLo;
      t.zHi = zHi;
      t.sub = true;
    }
  }
  return t;
}

// Where the candidate records of the next layer are read from: global memory
// (a uniform base plus an unsigned 32-bit byte offset; `global` in sorted
// order, `globalS` in search-index order, `subTab` the layer's sub-bin
// starts), or an LDS copy of some window rows of the search-index order
// (kTrackletsCount): row ri = (row - rowRef) mod kPhiBins of the copy holds
// the positions of its searched bins at LDS index j + off[ri].
struct CandidateSource {
  const char* global;
  const char* globalS;
  const uint4* subTab;
  int zSub, phiSub;
  const float4* lds;
  const int* off;
  int rowRef;
};

// Byte t (0..15) of a sub-bin entry.
__device__ __forceinline__ int subStart(uint4 e, int t)
{
  const uint32_t d = t < 4 ? e.x : (t < 8 ? e.y : (t < 12 ? e.z : e.w));
  return (int)((d >> (8 * (t & 3))) & 0xffu);
}

// The phi sub-rows [lo, hi] of window row `row` (P sub-rows per row) that can
// hold a candidate passing the phi cut phiCut (kPhiCut, or the learned form's
// gate) of a cluster at `phi`; lo > hi: none.  A candidate passes when
// |phi - phi_j| < phiCut or ||phi - phi_j| - 2pi| < phiCut, so phi_j lies
// within phiCut (plus kPhiSkipMargin, far above the rounding of those
// differences at |phi| < 7) of phi + 2pi k for some k in {-1, 0, 1}.  The candidates of row `row` have phi_j in the row's range,
// widened here by 0.01 rad; for every k whose interval meets it, phiSubIndex
// (monotone in phi) bounds the candidate's sub-row by those of the interval's
// ends.

__device__ __forceinline__ void phiSubRange(float phi, float phiCut, int row, int P, int& lo, int& hi)
{
  constexpr float kRowWidth = kTwoPi / (float)kPhiBins;
  const float rowLo = (float)row * kRowWidth - 0.01f, rowHi = (float)(row + 1) * kRowWidth + 0.01f;
  lo = P;
  hi = -1;
#pragma unroll
  for (int k = -1; k <= 1; ++k) {
    const float a = phi - (phiCut + kPhiSkipMargin) + (float)k * kTwoPi;
    const float b = phi + (phiCut + kPhiSkipMargin) + (float)k * kTwoPi;
    if (b >= rowLo && a <= rowHi) {
      lo = min(lo, phiSubIndex(a, row, P));
      hi = max(hi, phiSubIndex(b, row, P));
    }
  }
}

// Visits every accepted candidate of the search, in the reference's order
// (window row order, then ascending candidate position).
//
// With the z range known (t.sub) and more than one sub-bin per bin, each row
// is searched bin by bin over the bins [zFirst, zLast].  In a bin of at most
// kSubMaxBin clusters only the sub-bins whose z sub-bin lies in
// [zSubIndex(zLo), zSubIndex(zHi)] and whose phi sub-row lies in the row's
// phiSubRange are tested: a candidate that passes has z in [zLo, zHi] and phi
// in that range, hence (both indices being monotone) a sub-bin among them.
// The accepted ones are collected as bits of their offsets in the bin and
// emitted in ascending order, which is the sorted order.  Larger bins keep the
// sorted order in the search index and are tested whole; a row whose phi
// range holds no sub-row is skipped.  Otherwise the row's bins [zFirst, zLast]
// are tested in sorted order, as one range.  The reference's extra candidate (the inclusive
// upper bound) is tested last, from the sorted records in global memory.
// kStaged: the records of the tested bins are in the LDS copy (search-index
// order with more than one sub-bin, else sorted order).  kScanBlock: window
// candidates whose loads are issued together (more registers, fewer partial
// blocks).  kSub: the bin-by-bin sub-bin search is compiled in (without it the
// caller must have src.zSub * src.phiSub == 1).
//
// kWrap: the second phi comparison (|dphi - 2pi| < cut, the window across
// phi = 0) is evaluated.  Callers may leave it out for a cluster whose phi
// lies in (kWrapFreeLo, kWrapFreeHi): every candidate phi lies in [0, 2pi]
// (normalized, Cluster.cxx:35-45), so |phi - phi_j| <= 5.79 and
// |dphi - 2pi| >= 0.49 > kPhiCut -- that comparison is false for every
// candidate, and the result is unchanged.
constexpr float kWrapFreeLo = 0.5f, kWrapFreeHi = 5.7f;

__device__ __forceinline__ bool phiWrapFree(float phi) { return phi > kWrapFreeLo && phi < kWrapFreeHi; }

template <bool kStaged, int kScanBlock, bool kSub, bool kWrap = true, typename F>
__device__ __forceinline__ void searchRows(const TrackletSearch& t, const int* table, int nNext, int l, float z,
                                           float r, float phi, const CandidateSource& src, F&& emit)
{
  if (t.win.empty) return;
  const TrackletWindow& win = t.win;
  const float tanLambda = t.tanLambda, maxDeltaZ = t.maxDeltaZ;
  const float phiCut = t.phiCut;
  auto passes = [&](const float4& q) {
    const float deltaZ = fabsf(tanLambda * (q.x - r) + z - q.y);
    const float deltaPhi = fabsf(phi - q.z);
    if (!kWrap) return (deltaZ < maxDeltaZ) & (deltaPhi < phiCut);
    return (deltaZ < maxDeltaZ) & ((deltaPhi < phiCut) | (fabsf(deltaPhi - kTwoPi) < phiCut));
  };
  auto loadGlobal = [](const char* g, int j) { return *reinterpret_cast<const float4*>(g + 16u * (unsigned)j); };
  uint64_t bits0 = 0, bits1 = 0;  // accepted offsets 0..63 and 64..127 of the current bin
  // Tests the candidates [j0, j1) kScanBlock at a time: the block's loads are
  // issued first, the cuts are combined without short-circuit branches, and in
  // the last, partial block the indices past the range are clamped to its last
  // element and masked out.  kBits: accepted candidates set the bit of their offset in
  // the bin (record w); else they are emitted in ascending order.
  auto scan = [&](int j0, int j1, auto load, auto bitsMode) {
    constexpr bool kBits = decltype(bitsMode)::value;
    auto block = [&](int jb, auto full) {
      constexpr bool kFull = decltype(full)::value;
      float4 q[kScanBlock];
#pragma unroll
      for (int k = 0; k < kScanBlock; ++k) q[k] = load(kFull ? jb + k : min(jb + k, j1 - 1));
      uint32_t mask = 0;
#pragma unroll
      for (int k = 0; k < kScanBlock; ++k) {
        const bool ok = passes(q[k]) & (kFull || jb + k < j1);
        if (kBits) {
          const int o = __float_as_int(q[k].w);
          bits0 |= (ok & (o < 64)) ? 1ull << (o & 63) : 0ull;
          bits1 |= (ok & (o >= 64)) ? 1ull << (o & 63) : 0ull;
        } else {
          mask |= (ok ? 1u : 0u) << k;
        }
      }
      if (!kBits) {
        while (mask != 0u) {
          const int k = __builtin_ctz(mask);
          mask &= mask - 1u;
          emit(jb + k);
        }
      }
    };
    int jb = j0;
    for (; jb + kScanBlock <= j1; jb += kScanBlock) block(jb, std::true_type{});
    if (jb < j1) block(jb, std::false_type{});
  };
  int row = win.y;
  for (int count = 0; count < win.rows; ++count) {
    const int firstBin = binIndex(win.x, row);
    int maxBin = firstBin + win.z - win.x + 1;
    // Window rows past the last phi row would read past the table; the
    // reference reads the following memory there (only reachable when phi
    // lies within 2 ulp below 0.3).  Clamp to the table end.
    if (maxBin > kBins) maxBin = kBins;
    const int jLast = table[maxBin];  // the reference's inclusive upper bound
    bool scanned = true;  // the row's bins are scanned (gated: the row meets the gate)
    if (t.narrow) {
      int d = row - t.scanFirst;
      if (d < 0) d += kPhiBins;
      scanned = d < t.scanRows;
    }
    int off = 0;
    if (kStaged && scanned) {  // the LDS copy holds the scanned rows
      int ri = row - src.rowRef;
      if (ri < 0) ri += kPhiBins;
      off = src.off[ri];
    }
    if (!scanned) {
      // only the extra candidate below
    } else if (kSub && t.sub && src.zSub * src.phiSub > 1) {
      auto loadS = [&](int j) { return kStaged ? src.lds[j + off] : loadGlobal(src.globalS, j); };
      const int S = src.zSub, P = src.phiSub;
      int psLo = 0, psHi = 0;
      if (P > 1) phiSubRange(phi, phiCut, row, P, psLo, psHi);
      for (int zb = t.zFirst; psLo <= psHi && zb <= t.zLast; ++zb) {  // table positions clamped like maxBin
        const int b = min(firstBin + (zb - win.x), maxBin);
        const int bS = table[b];
        const int bE = min(table[min(b + 1, maxBin)], nNext);
        if (bE <= bS) continue;
        if (bE - bS <= kSubMaxBin) {
          const uint4 e = src.subTab[b];
          const int zsLo = zSubIndex(l + 1, t.zLo, zb, S), zsHi = zSubIndex(l + 1, t.zHi, zb, S);
          bits0 = bits1 = 0;
          for (int ps = psLo; ps <= psHi; ++ps)
            scan(bS + subStart(e, ps * S + zsLo), bS + subStart(e, ps * S + zsHi + 1), loadS, std::true_type{});
          while (bits0 != 0ull) {
            const int k = __builtin_ctzll(bits0);
            bits0 &= bits0 - 1ull;
            emit(bS + k);
          }
          while (bits1 != 0ull) {
            const int k = __builtin_ctzll(bits1);
            bits1 &= bits1 - 1ull;
            emit(bS + 64 + k);
          }
        } else {
          scan(bS, bE, loadS, std::false_type{});
        }
      }
    } else if (t.zFirst <= t.zLast) {
      const int jFirst = table[min(firstBin + (t.zFirst - win.x), maxBin)];
      const int jEnd = min(table[min(firstBin + (t.zLast - win.x) + 1, maxBin)], nNext);
      scan(jFirst, jEnd, [&](int j) { return kStaged ? src.lds[j + off] : loadGlobal(src.global, j); },
           std::false_type{});
    }
    if (jLast < nNext) {  // the extra candidate of the inclusive bound
      if (passes(loadGlobal(src.global, jLast))) emit(jLast);
    }
    ++row;
    if (row == kPhiBins) row = 0;
  }
}

// The search of one cluster with every candidate read from global memory
// (the fill's path for clusters that overflowed their slots).
// rw: the item whose records are read; baseNext: their next-layer segment.
template <typename F>
__device__ __forceinline__ void forEachTracklet(const DevState& s, int rw, int l, const int* table, int baseNext,
                                                int nNext, float4 geoNext, float z, float r, float phi, float vz,
                                                F&& emit)
{
  const TrackletSearch t = trackletSearch(geoNext, z, r, phi, vz, l, trackletPhiCut(s, l));
  const CandidateSource src{reinterpret_cast<const char*>(s.crzp + baseNext),
                            reinterpret_cast<const char*>(s.crzpS + baseNext),
                            s.subTab + (size_t)(rw * kLayers + l + 1) * kBins,
                            s.zSub,
                            s.phiSub,
                            nullptr,
                            nullptr,
                            0};
  searchRows<false, 2, true>(t, table, nNext, l, z, r, phi, src, emit);  // rare path: few registers
}

// LDS copy of the next layer for the tracklet count.  A chunk's clusters are
// consecutive in (phi row, z bin) order, so the union of their windows is a
// few phi rows times a range of z bins; those clusters of the next layer are
// copied once per chunk and read from LDS by every cluster of the chunk.
constexpr int kStageCap = 1984;  // next-layer clusters per chunk (31 KB: five workgroups of 256 per CU)
constexpr int kTrackletsCountPerCu = 5;  // workgroups of kTrackletsCount resident per CU (its LDS)
constexpr int kStageRows = 8;    // phi rows of the copy

// The LDS copy of one chunk: the union of its windows, its rows in the copy.
struct StageLds {
  int uni[5];  // min row offset, max row end, min z bin, max z bin, no-copy flag
  int rowLo[kStageRows], rowOff[kStageRows], rowStart[kStageRows + 1];
  int copyRows, copyRef;  // rows of the LDS copy (0: none) and the window row of its first row
};

// Copies the next-layer records (`src`, the next layer's segment, in the order
// the search reads) of the union of the chunk's windows to `stage`, at most
// `cap` of them, and returns the number of rows copied (0: none fit; the
// search then reads global memory).  Called by every thread of the block; t
// is the thread's search (empty past the chunk), rowRef the window row before
// the chunk's first cluster.  Rows are taken relative to rowRef: the chunk's
// clusters sit in consecutive rows, and a window starts at most one row
// before its cluster's row.
__device__ __forceinline__ int stageUnion(const TrackletSearch& t, int rowRef, const int* table, int nNext,
                                          const float4* src, float4* stage, int cap, StageLds& L)
{
  if (threadIdx.x == 0) {
    L.uni[0] = kPhiBins;
    L.uni[1] = -kPhiBins;
    L.uni[2] = kZBins;
    L.uni[3] = -1;
    L.uni[4] = 0;
  }
  __syncthreads();
  int dMin = kPhiBins, dMax = -kPhiBins, zMin = kZBins, zMax = -1, noCopy = 0;
  if (!t.win.empty) {  // the rows whose bins the search scans
    if (t.scanFirst >= kPhiBins || t.scanRows > kStageRows) noCopy = 1;  // the phi == 2pi row, or a wide window
    int d = t.scanFirst - rowRef;
    if (d < 0) d += kPhiBins;
    if (d >= kPhiBins / 2) d -= kPhiBins;
    dMin = d;
    dMax = d + t.scanRows - 1;
    if (t.zFirst <= t.zLast) {
      zMin = t.zFirst;
      zMax = t.zLast;
    }
  }
  for (int m = 32; m >= 1; m >>= 1) {
    dMin = min(dMin, __shfl_xor(dMin, m, 64));
    dMax = max(dMax, __shfl_xor(dMax, m, 64));
    zMin = min(zMin, __shfl_xor(zMin, m, 64));
    zMax = max(zMax, __shfl_xor(zMax, m, 64));
    noCopy = max(noCopy, __shfl_xor(noCopy, m, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&L.uni[0], dMin);
    atomicMax(&L.uni[1], dMax);
    atomicMin(&L.uni[2], zMin);
    atomicMax(&L.uni[3], zMax);
    atomicMax(&L.uni[4], noCopy);
  }
  __syncthreads();
  // The copy: rows [dMin, dMax] of the union, z bins [zMin, zMax] of each.
  // The first wavefront reads the rows' table entries in parallel (lane ri:
  // row ri) and numbers them with a shuffle scan.
  const int nr = L.uni[1] - L.uni[0] + 1;
  const bool fitsRows = L.uni[4] == 0 && nr >= 1 && nr <= kStageRows;
  if (threadIdx.x < 64 && fitsRows) {
    const int ri = threadIdx.x;
    const int ref = ((rowRef + L.uni[0]) % kPhiBins + kPhiBins) % kPhiBins;
    int lo = 0, hi = 0;
    if (ri < nr && L.uni[2] <= L.uni[3]) {
      const int row = (ref + ri) % kPhiBins;
      lo = table[binIndex(L.uni[2], row)];
      hi = max(lo, min(table[binIndex(L.uni[3], row) + 1], nNext));
    }
    const int len = hi - lo;
    int inc = len;  // inclusive scan over the rows (at most kStageRows lanes carry a length)
#pragma unroll
    for (int d = 1; d < kStageRows; d <<= 1) {
      const int t = __shfl_up(inc, d, 64);
      if (ri >= d) inc += t;
    }
    const int start = inc - len;
    if (ri < nr) {
      L.rowLo[ri] = lo;
      L.rowStart[ri] = start;
      L.rowOff[ri] = start - lo;
    }
    const int total = __shfl(inc, nr - 1, 64);
    if (ri == 0) {
      L.rowStart[nr] = total;
      L.copyRows = total <= cap ? nr : 0;
      L.copyRef = ref;
    }
  } else if (threadIdx.x == 0 && !fitsRows) {
    L.copyRows = 0;
    L.copyRef = 0;
  }
  __syncthreads();
  const int nrows = L.copyRows;
  if (nrows > 0) {
    // Direct global-to-LDS loads (global_load_lds_dwordx4: lane L of a
    // wavefront writes 16 bytes at the wavefront's LDS base + 16 L): no
    // registers hold the records, and every load of the copy is in flight at
    // once.  A lane past the end issues none.
    const int total = L.rowStart[nrows];
    const int waveFirst = threadIdx.x & ~63;
    int ri = 0;  // k only grows, so the row index only moves forward
    for (int k0 = 0; k0 < total; k0 += blockDim.x) {
      const int k = k0 + threadIdx.x;
      if (k < total) {
        while (ri + 1 < nrows && L.rowStart[ri + 1] <= k) ++ri;
        __builtin_amdgcn_global_load_lds(src + (L.rowLo[ri] + (k - L.rowStart[ri])), stage + (k0 + waveFirst), 16, 0,
                                         0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // the copy is in LDS
    __syncthreads();
  }
  return nrows;
}

// The window row before the chunk's first cluster (sorted position `first`
// of the records at `cbin`).
__device__ __forceinline__ int chunkRowRef(const int* cbin, int first)
{
  return (max(cbin[first], 0) / kPhiBins + kPhiBins - 1) % kPhiBins;
}

// Count: each cluster's accepted candidates, in emission order, are kept in
// kTrkSlots scratch slots so that the fill does not repeat the window search.
// Five waves per SIMD (the LDS copy allows five workgroups of 256 per CU).
// kSub: launched with a search index (more than one sub-bin per bin); the
// instance without one leaves the sub-bin search out and fits five waves
// without spills.
template <int B, bool kSub>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(5))) kTrackletsCount(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  __shared__ float4 stage[kStageCap];
  __shared__ StageLds L;
  const int S = s.W * kTrackletLayers;
  const int* chunkStart = s.chunkStart + kStTracklets * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  for (ChunkWalk cw(s, kStTracklets, nChunks, xcdBlockIndex()); cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int sg = ci.x;
    const int w = sg / kTrackletLayers, l = sg % kTrackletLayers;
    const Seg cur = s.clSeg[w * kLayers + l], next = s.clSeg[w * kLayers + l + 1];
    const int rw = s.recItem[w];
    const Seg rcur = s.clSeg[rw * kLayers + l], rnext = s.clSeg[rw * kLayers + l + 1];
    const int first = ci.y;
    const int i = first + threadIdx.x;
    const int p = cur.base + i;
    const int* table = s.table + (size_t)(rw * kLayers + l + 1) * kTableSize;
    TrackletSearch t;
    float z = 0.0f, r = 0.0f, phi = 0.0f;
    t.win.empty = true;
    if (i < cur.count) {
      z = s.cz[rcur.base + i];
      r = s.cr[rcur.base + i];
      phi = s.cphi[rcur.base + i];
      t = trackletSearch(s.segGeo[rw * kLayers + l + 1], z, r, phi, s.itemVertex[3 * w + 2], l, trackletPhiCut(s, l));
    }
    const bool indexed = s.zSub * s.phiSub > 1;  // the search index exists (crzpS)
    const float4* copySrc = indexed ? s.crzpS : s.crzp;
    const int nrows = stageUnion(t, chunkRowRef(s.cbin + rcur.base, first), table, next.count, copySrc + rnext.base,
                                 stage, kStageCap, L);
    // every cluster of the wavefront away from phi = 0: the wrap comparison is left out
    const bool waveWrapFree = __ballot(i < cur.count && !phiWrapFree(phi)) == 0ull;
    int cnt = 0;
    if (i < cur.count) {
      int* slots = s.trkCand + (size_t)p * kTrkSlots;
      auto emit = [&](int j) {
        if (cnt < kTrkSlots) slots[cnt] = j;
        ++cnt;
      };
      const CandidateSource src{reinterpret_cast<const char*>(s.crzp + rnext.base),
                                reinterpret_cast<const char*>(s.crzpS + rnext.base),
                                s.subTab + (size_t)(rw * kLayers + l + 1) * kBins,
                                s.zSub,
                                s.phiSub,
                                stage,
                                L.rowOff,
                                L.copyRef};
      if (nrows > 0 && (t.sub || !indexed)) {  // with sub-bins the copy serves only the bin-by-bin search
        if (waveWrapFree)
          searchRows<true, 8, kSub, false>(t, table, next.count, l, z, r, phi, src, emit);
        else
          searchRows<true, 8, kSub, true>(t, table, next.count, l, z, r, phi, src, emit);
      } else {
        searchRows<false, 8, kSub>(t, table, next.count, l, z, r, phi, src, emit);
      }
      s.trkCnt[p] = cnt;
    }
    const int total = blockSum(cnt, scratch);  // its barriers also end this chunk's use of the copy
    if (threadIdx.x == 0) s.chunkSum[(size_t)kStTracklets * s.maxChunks + chunk] = total;
  }
}

// ---------------------------------------------------------------------------
// Learned form, second variant (itsca_set_filter mode 2): the tracklet search
// and the filter fused, one tracklet layer per launch, from the last layer
// down (a candidate's features read the raw tracklet count of its next-layer
// cluster, which the previous launch wrote).  Per chunk: the search of
// kTrackletsCount, with each cluster's accepted candidates kept in LDS as
// 16-bit positions (the first kCandLds of them; the rest up to kTrkSlots go to
// the cluster's slot row); then, lane by lane, step k scores the k-th
// candidate of every lane's cluster (features, BF16 MFMA network) and appends
// a kept one to the lane's slot row.  The raw count goes to trkCnt, the kept
// count to trkKept (kCommitKept moves it after the last layer), the chunk's
// kept sum to the chunk sums.  Clusters with more than kTrkSlots candidates
// are kept whole (the fill recomputes them).  Needs one sub-bin per bin (no
// search index) and next layers of at most 65536 clusters.  The same tracklets as kTrackletsCount +
// kTrackletsScore, with no candidate round trip through memory, but slower on
// MI355X: the search and the scoring alternate within a workgroup at four
// waves per SIMD, and the six launches each pay the chunk set-up.
//
// The workgroup walks the chunks of its stride sequence (xcdBlockIndex + k *
// gridDim.x, as every stage kernel) that belong to the launch's layer, listed
// in LDS B at a time.  The MFMA tiles alias the copy of the next layer, which
// the search no longer reads by then; the LDS of a workgroup of 256 (27.6 KB
// copy, 8 KB candidates) lets four of them share a CU.
// ---------------------------------------------------------------------------
constexpr int kCandLds = 16;           // candidates per cluster kept in LDS
constexpr int kFusedStageCap = 1728;   // next-layer clusters of the copy (central events need at most ~1720)

template <int B>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4)))
kTrackletsSearchScore(DevState s, int layer, float threshold)
{
  constexpr int kWaves = B / 64;
  struct ScoreLds {
    uint16_t tiles[kWaves][64 * kMlpInPad];  // 16-byte aligned rows
  };
  static_assert(sizeof(ScoreLds) <= kFusedStageCap * sizeof(float4), "the scoring arrays alias the copy");
  __shared__ float4 stage[kFusedStageCap];
  __shared__ uint16_t cand[kCandLds * B];
  __shared__ StageLds L;
  __shared__ int scratch[B / 64 + 1];
  __shared__ int list[B];
  __shared__ int waveKept[kWaves];
  ScoreLds& sc = *reinterpret_cast<ScoreLds*>(stage);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l = layer;
  const int S = s.W * kTrackletLayers;
  const int nChunks = s.chunkStart[kStTracklets * (s.maxSegs + 1) + S];
  const int4* info = s.chunkInfo + (size_t)kStTracklets * s.maxChunks;
  const int G = gridDim.x, b0 = xcdBlockIndex();
  const int mine = b0 < nChunks ? (nChunks - 1 - b0) / G + 1 : 0;
  for (int r0 = 0; r0 < mine; r0 += B) {
    const int kk = r0 + (int)threadIdx.x;
    const int c = kk < mine ? b0 + kk * G : -1;
    const int on = c >= 0 && info[c].x % kTrackletLayers == l ? 1 : 0;
    int n;
    const int at = blockExclusiveScan(on, scratch, n);
    if (on) list[at] = c;
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int chunk = list[q];
      const int4 ci = info[chunk];
      const int w = ci.x / kTrackletLayers;
      const Seg cur = s.clSeg[w * kLayers + l], next = s.clSeg[w * kLayers + l + 1];
      const int rw = s.recItem[w];
      const Seg rcur = s.clSeg[rw * kLayers + l], rnext = s.clSeg[rw * kLayers + l + 1];
      const int first = ci.y;
      const int i = first + threadIdx.x;
      const int p = cur.base + i;
      const int* table = s.table + (size_t)(rw * kLayers + l + 1) * kTableSize;
      const float vx = s.itemVertex[3 * w], vy = s.itemVertex[3 * w + 1], vz = s.itemVertex[3 * w + 2];
      TrackletSearch t;
      float z = 0.0f, r = 0.0f, phi = 0.0f;
      t.win.empty = true;
      if (i < cur.count) {
        z = s.cz[rcur.base + i];
        r = s.cr[rcur.base + i];
        phi = s.cphi[rcur.base + i];
        t = trackletSearch(s.segGeo[rw * kLayers + l + 1], z, r, phi, vz, l, trackletPhiCut(s, l));
      }
      const int nrows = stageUnion(t, chunkRowRef(s.cbin + rcur.base, first), table, next.count,
                                   s.crzp + rnext.base, stage, kFusedStageCap, L);
      int* myRow = s.trkCand + (size_t)p * kTrkSlots;
      int raw = 0;
      if (i < cur.count) {
        auto emit = [&](int j) {
          if (raw < kCandLds)
            cand[raw * B + threadIdx.x] = (uint16_t)j;
          else if (raw < kTrkSlots)
            myRow[raw] = j;
          ++raw;
        };
        const CandidateSource src{reinterpret_cast<const char*>(s.crzp + rnext.base), nullptr, nullptr, 1, 1, stage,
                                  L.rowOff, L.copyRef};
        if (nrows > 0)
          searchRows<true, 8, false>(t, table, next.count, l, z, r, phi, src, emit);
        else
          searchRows<false, 8, false>(t, table, next.count, l, z, r, phi, src, emit);
        s.trkCnt[p] = raw;
      }
      __syncthreads();  // the copy is no longer read: its LDS holds the scoring arrays from here

      // Scoring, lane by lane: step k scores the k-th candidate of every
      // lane's cluster (one 64-sample tile per step; lanes whose cluster has
      // fewer candidates idle), and a kept candidate is appended to the
      // lane's slot row at once.  The weights are loaded per chunk (cache
      // hits) so that they hold no registers during the search.
      const MlpWave m = loadMlp(s.mlpW);
      const int scored = raw <= kTrkSlots ? raw : 0;
      int steps = scored;
      for (int d = 32; d >= 1; d >>= 1) steps = max(steps, __shfl_xor(steps, d, 64));
      TrackletSeed ts{};
      if (scored > 0) {
        const float4 a = s.cxyrz[rcur.base + i];  // (x, y, r, z)
        ts = trackletSeed(a.x, a.y, a.w, a.z, phi, vx, vy, vz, raw);
      }
      // candidate k of this lane's cluster, -1 past its last; a slot-row entry
      // is read two steps before any kept candidate can be written over it
      auto candAt = [&](int k) {
        if (k >= scored) return -1;
        return k < kCandLds ? (int)cand[k * B + threadIdx.x] : myRow[k];
      };
      auto recOf = [&](int j) {
        CandRec rec{make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, 0};
        if (j >= 0) {
          rec.b = s.cxyrz[rnext.base + j];
          rec.phi = s.cphi[rnext.base + j];
          rec.count = l + 1 < kTrackletLayers ? s.trkCnt[next.base + j] : 0;
        }
        return rec;
      };
      int kept = 0;
      int jCur = candAt(0), jNext = candAt(1);
      CandRec qCur = recOf(jCur);
      for (int k = 0; k < steps; ++k) {
        const int jAfter = candAt(k + 2);
        const CandRec qNext = recOf(jNext);
        float f[kMlpIn];
        if (jCur >= 0)
          candidateFeatures(ts, l, qCur.b.x - vx, qCur.b.y - vy, qCur.b.w, qCur.b.z, qCur.phi, vz, qCur.count, f);
        else
          for (int e = 0; e < kMlpIn; ++e) f[e] = 0.0f;
        uint4 lo, hi;
        packFeatureRow(f, lo, hi);
        storeTileRow(sc.tiles[wave], lane, lo, hi);
        waveSync();
        const float zl = mlpForward64(m, sc.tiles[wave]);
        waveSync();  // the tile is rewritten by the next step
        if (jCur >= 0 && zl >= threshold) myRow[kept++] = jCur;
        jCur = jNext;
        jNext = jAfter;
        qCur = qNext;
      }
      const int keptCount = scored == raw ? kept : raw;
      if (i < cur.count) s.trkKept[p] = keptCount;
      writeKeptList(keptListOf(s.trkKeep, cur.base + first, wave), myRow, (i < cur.count && raw <= kTrkSlots) ? kept : 0);
      int sum = i < cur.count ? keptCount : 0;
      for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
      if (lane == 0) waveKept[wave] = sum;
      __syncthreads();
      if (threadIdx.x == 0) {
        int tk = 0;
        for (int k = 0; k < kWaves; ++k) tk += waveKept[k];
        s.chunkSum[(size_t)kStTracklets * s.maxChunks + chunk] = tk;
      }
      __syncthreads();  // the next chunk's copy overwrites the scoring arrays
    }
    __syncthreads();  // the list is rebuilt by the next round
  }
}

// Tracklet ctor, Tracklet.cxx:32-40 (absolute x, y for the direction).
__device__ __forceinline__ void writeTracklet(const DevState& s, int k, int i, int j, int curBase, int nextBase)
{
  if (k >= s.tCap) return;
  const int pi = curBase + i, q = nextBase + j;
  s.tFS[k] = make_int2(i, j);
  const float4 a = s.cxyrz[pi], b = s.cxyrz[q];  // (x, y, r, z)
  s.tTP[k] = make_float2((a.w - b.w) / (a.z - b.z), fdlibmAtan2f(a.y - b.y, a.x - b.x));
  s.cellCnt[k] = 0;  // stays 0 where the cell stage does not run
}

// Enumerates the slot outputs of a chunk, one per thread: element c (thread c)
// owns `capped` outputs (0 for an element whose outputs overflowed its slots),
// numbered from its exclusive scan `exc`; exCap holds every element's exc and
// must be written before the call.  The owner of each output comes from an LDS
// map built in rounds of kOwnerRound outputs, so every output costs one LDS
// lookup.  visit(c, k) handles the k-th output of element c.  B <= 256 (the
// map stores element indices as bytes).
constexpr int kOwnerRound = 4096;

template <int B, typename F>
__device__ __forceinline__ void forEachSlotOutput(int capped, int exc, int totalCap, const int* exCap,
                                                  uint8_t* owner, F&& visit)
{
  static_assert(B <= 256, "owner indices are bytes");
  for (int r0 = 0; r0 < totalCap; r0 += kOwnerRound) {
    const int rn = min(kOwnerRound, totalCap - r0);
    for (int k = max(0, r0 - exc); k < capped && exc + k < r0 + rn; ++k) owner[exc + k - r0] = (uint8_t)threadIdx.x;
    __syncthreads();
    for (int q = threadIdx.x; q < rn; q += B) {
      const int c = owner[q];
      visit(c, r0 + q - exCap[c]);
    }
    __syncthreads();  // the map is rebuilt by the next round
  }
}

// Fill: one output record per thread of the chunk (coalesced writes, no
// divergent window search); clusters with more than kTrkSlots tracklets are
// recomputed by their own thread.  Six waves per SIMD: the overflow search
// tests its candidates two at a time so that the kernel fits that many
// without spills.
template <int B, bool kKeptList>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(6))) kTrackletsFill(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  __shared__ int exFull[B], exCap[B];
  __shared__ uint8_t owner[kOwnerRound];
  const int S = s.W * kTrackletLayers;
  const int* chunkStart = s.chunkStart + kStTracklets * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  ChunkWalk cw(s, kStTracklets, nChunks, xcdBlockIndex());
  // tracklet count of this thread's cluster in a chunk; the next chunk's is loaded ahead
  auto elemOf = [&](const int4& e) { return (int)threadIdx.x < e.w ? s.trkCnt[e.z + threadIdx.x] : 0; };
  int cntNext = cw.valid() ? elemOf(cw.next()) : int{};
  for (; cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int cnt = cntNext;
    if (cw.hasNext()) cntNext = elemOf(cw.next());
    const int sg = ci.x;
    const int w = sg / kTrackletLayers, l = sg % kTrackletLayers;
    const Seg cur = s.clSeg[w * kLayers + l], next = s.clSeg[w * kLayers + l + 1];
    const int first = ci.y;
    const int i = first + threadIdx.x;
    const int p = cur.base + i;
    const int capped = cnt <= kTrkSlots ? cnt : 0;
    int total, totalCap;
    const int ex = blockExclusiveScan(cnt, scratch, total);
    const int exc = blockExclusiveScan(capped, scratch, totalCap);
    exFull[threadIdx.x] = ex;
    exCap[threadIdx.x] = exc;
    const int chunkBase = s.chunkOff[(size_t)kStTracklets * s.maxChunks + chunk];
    if (i < cur.count) s.trkOff[p] = chunkBase + ex - s.trkSeg[sg].base;
    const int rw = s.recItem[w];
    const Seg rcur = s.clSeg[rw * kLayers + l], rnext = s.clSeg[rw * kLayers + l + 1];
    forEachSlotOutput<B>(capped, exc, totalCap, exCap, owner, [&](int c, int k) {
      // the filtered form reads its wavefront's kept list (the scoring wrote it in
      // this order: offset = the capped counts' scan over the wavefront), the plain
      // form the slot rows
      const int j = kKeptList ? keptListOf(s.trkKeep, cur.base + first, c >> 6)[exCap[c] - exCap[c & ~63] + k]
                              : s.trkCand[(size_t)(cur.base + first + c) * kTrkSlots + k];
      writeTracklet(s, chunkBase + exFull[c] + k, first + c, j, rcur.base, rnext.base);
    });
    if (cnt > kTrkSlots) {
      const int* table = s.table + (size_t)(rw * kLayers + l + 1) * kTableSize;
      const int q = rcur.base + i;
      int k = chunkBase + ex;
      forEachTracklet(s, rw, l, table, rnext.base, next.count, s.segGeo[rw * kLayers + l + 1], s.cz[q], s.cr[q],
                      s.cphi[q], s.itemVertex[3 * w + 2],
                      [&](int j) { writeTracklet(s, k++, i, j, rcur.base, rnext.base); });
    }
  }
}

// ---------------------------------------------------------------------------
// Stage (c): cells.  Tracker.cxx:119-235, Cell.cxx:24-32, MathUtils.h:53-60.
// One thread per tracklet of layer l; candidates are the tracklets of layer
// l+1 that start at its second cluster (a contiguous range).
// ---------------------------------------------------------------------------
// Cell layer l is computed only while every tracklet layer 0..l+1 is
// non-empty (early return, Tracker.cxx:124-128).
__global__ void __launch_bounds__(1024) kPlanCells(DevState s, int* counts)
{
  __shared__ int scratch[1024 / 64 + 1];
  const int S = s.W * kCellLayers;
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const int w = i / kCellLayers, l = i % kCellLayers;
    bool ok = true;
    for (int k = 0; k <= l + 1; ++k) ok = ok && s.trkSeg[w * kTrackletLayers + k].count > 0;
    counts[i] = ok ? s.trkSeg[w * kTrackletLayers + l].count : 0;
  }
  __syncthreads();
  planChunks(counts, S, s.chunk, s.chunkStart + kStCells * (s.maxSegs + 1), scratch);
}

struct CellOut {
  float nx, ny, nz, curvature;
};

// Geometry shared by every candidate of tracklet t (layer l, global index).
struct CellSeed {
  int l, tnBase, cl2Base, uFirst, nCand;
  float x1, y1, z1, r1, r1sq, x2, y2, r2sq, d1x, d1y, d1z, vx, vy, vz, tTanL, tPhi;
  float maxDeltaZ, maxDCA;  // the layer's cuts, read once per tracklet
};

__device__ __forceinline__ CellSeed cellSeedOf(const DevState& s, int w, int l, int2 fs, float2 tp);

__device__ __forceinline__ CellSeed cellSeed(const DevState& s, int w, int l, int t)
{
  return cellSeedOf(s, w, l, s.tFS[t], s.tTP[t]);
}

// The same from the tracklet's record (tFS[t], tTP[t]) already loaded.
__device__ __forceinline__ CellSeed cellSeedOf(const DevState& s, int w, int l, int2 fs, float2 tp)
{
  CellSeed c;
  c.l = l;
  c.tnBase = s.trkSeg[w * kTrackletLayers + l + 1].base;
  const Seg cl0 = recSeg(s, w, l), cl1 = recSeg(s, w, l + 1);  // records
  c.cl2Base = recSeg(s, w, l + 2).base;
  const int item1 = s.clSeg[w * kLayers + l + 1].base;  // the item's own counts and offsets
  const int k = fs.y;
  c.nCand = s.trkCnt[item1 + k];  // tracklets of layer l+1 starting at this cluster
  c.uFirst = c.nCand > 0 ? s.trkOff[item1 + k] : 0;
  c.vx = s.itemVertex[3 * w];
  c.vy = s.itemVertex[3 * w + 1];
  c.vz = s.itemVertex[3 * w + 2];
  const int p1 = cl0.base + fs.x, p2 = cl1.base + k;
  const float4 q1 = s.cxyrz[p1], q2 = s.cxyrz[p2];  // (x, y, r, z)
  c.x1 = q1.x;
  c.y1 = q1.y;
  c.z1 = q1.w;
  c.r1 = q1.z;
  c.x2 = q2.x;
  c.y2 = q2.y;
  const float r2 = q2.z;
  c.r1sq = c.r1 * c.r1;
  c.r2sq = r2 * r2;
  c.d1x = c.x2 - c.x1;
  c.d1y = c.y2 - c.y1;
  c.d1z = c.r2sq - c.r1sq;
  c.tTanL = tp.x;
  c.tPhi = tp.y;
  c.maxDeltaZ = cellMaxDeltaZ(l);
  c.maxDCA = cellMaxDCA(l);
  return c;
}

// Cheap cuts of candidate iu (Tracker.cxx:161-173): tanLambda and phi
// continuity, then the z of the line through the first cluster at the vertex.
// The reference evaluates the z cut only after the first two pass; the cuts
// have no side effects, so all three are evaluated and combined without
// branches (same comparisons, same result).
//
// kWrap: the second phi comparison (|dphi - 2pi| < cut) is evaluated.
// Tracklet phis are atan2f values in [-pi, pi] (Tracklet.cxx:36), so for a
// tracklet with |phi| <= kCellWrapFree, |dphi| <= 5.95 and |dphi - 2pi| >=
// 0.34 > kCellMaxDeltaPhi: that comparison is false, and callers whose
// tracklets all lie there may leave it out.
constexpr float kCellWrapFree = 2.8f;

template <bool kWrap = true>
__device__ __forceinline__ bool cellCheapCuts(const CellSeed& c, float2 up)
{
  const float uTanL = up.x, uPhi = up.y;
  const float deltaTanLambda = fabsf(c.tTanL - uTanL);
  const float deltaPhi = fabsf(c.tPhi - uPhi);
  const bool phiOk = kWrap ? ((deltaPhi < kCellMaxDeltaPhi) | (fabsf(deltaPhi - kTwoPi) < kCellMaxDeltaPhi))
                           : (deltaPhi < kCellMaxDeltaPhi);
  const bool continuity = (deltaTanLambda < kCellMaxDeltaTanLambda) & phiOk;
  const float avgTanLambda = 0.5f * (c.tTanL + uTanL);
  const float dirZ = -avgTanLambda * c.r1 + c.z1;
  const float deltaZ = fabsf(dirZ - c.vz);
  return continuity & (deltaZ < c.maxDeltaZ);
}

// Conformal-plane fit and DCA cut (Tracker.cxx:175-218, MathUtils.h:53-60)
// of the candidate whose third cluster has the record q3 = (x, y, r, z).
__device__ __forceinline__ bool cellFitAt(const CellSeed& c, float4 q3, CellOut& o)
{
  const float x3 = q3.x, y3 = q3.y, r3 = q3.z;
  const float r3sq = r3 * r3;
  const float d2x = x3 - c.x1, d2y = y3 - c.y1, d2z = r3sq - c.r1sq;
  const float nx = (c.d1y * d2z) - (c.d1z * d2y);
  const float ny = (c.d1z * d2x) - (c.d1x * d2z);
  const float nz = (c.d1x * d2y) - (c.d1y * d2x);
  const float norm = sqrtf(nx * nx + ny * ny + nz * nz);
  if (norm < kFloatMin || fabsf(nz) < kFloatMin) return false;
  const float inv = 1.0f / norm;
  const float hx = nx * inv, hy = ny * inv, hz = nz * inv;
  const float planeDistance = -hx * (c.x2 - c.vx) - (hy * c.y2 - c.vy) - hz * c.r2sq;
  const float hz2 = hz * hz;
  const float radius = sqrtf((1.0f - hz2 - 4.0f * planeDistance * hz) / (4.0f * hz2));
  const float cx = -0.5f * hx / hz, cy = -0.5f * hy / hz;
  const float dca = fabsf(radius - sqrtf(cx * cx + cy * cy));
  if (dca > c.maxDCA) return false;  // NaN passes, as in the reference CPU path
  o.nx = hx;
  o.ny = hy;
  o.nz = hz;
  o.curvature = 1.0f / radius;
  return true;
}

__device__ __forceinline__ float4 cellThird(const DevState& s, const CellSeed& c, int iu)
{
  return s.cxyrz[c.cl2Base + s.tFS[c.tnBase + iu].y];
}

__device__ __forceinline__ bool cellFit(const DevState& s, const CellSeed& c, int iu, CellOut& o)
{
  return cellFitAt(c, cellThird(s, c, iu), o);
}

// Cheap-cut survivors kept per thread before the fit (16 KB of LDS for 256
// threads: eight workgroups per CU).  A thread with more fits the rest one at
// a time inside the candidate loop while its workgroup waits; in central
// events 0.2 % of the layer-1 tracklets have 9-13 survivors.
constexpr int kCellList = 16;
constexpr int kFitBatch = 4;  // survivors whose fit loads are issued together
constexpr int kCandBlockMax = 8;  // candidates whose loads are issued together (up to four 16-byte loads)
static_assert(kCandBlockMax - 1 <= kTrackletPad, "the last block reads past the range");

// Two consecutive (tanLambda, phi) records in one 16-byte load (8-byte
// aligned).  The last block of a candidate range reads up to kCandBlockMax - 1
// records past it -- masked out, and inside the kTrackletPad records that
// every tracklet array is allocated beyond its capacity.
__device__ __forceinline__ void loadPair(const float2* p, float2& a, float2& b)
{
  typedef float Float4A8 __attribute__((ext_vector_type(4), aligned(8)));
  const Float4A8 v = *reinterpret_cast<const Float4A8*>(p);
  a = make_float2(v.x, v.y);
  b = make_float2(v.z, v.w);
}

// Count: the cheap cuts run over every candidate; survivors (usually 0-2) are
// fitted in a second loop, so the expensive fit does not sit inside the
// divergent candidate loop.  Accepted candidates are kept in kCellSlots slots.
// kCandBlock: candidates whose loads are issued together; kWrapFree: wavefronts
// whose tracklets are all away from phi = +-pi run the cuts without the wrap
// comparison.  The plain form (10 candidates per tracklet on central events)
// runs <8, true>; the filtered form (about 2) runs <4, false>, whose smaller
// register file keeps eight waves per SIMD.
template <int B, int kCandBlock, bool kWrapFree>
__global__ void __launch_bounds__(B) kCellsCount(DevState s)
{
  static_assert(kCandBlock % 2 == 0 && kCandBlock <= kCandBlockMax, "pairs of records, within the padding");
  __shared__ int scratch[B / 64 + 1];
  __shared__ int list[kCellList][B];
  const int S = s.W * kCellLayers;
  const int* chunkStart = s.chunkStart + kStCells * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  ChunkWalk cw(s, kStCells, nChunks, xcdBlockIndex());
  // the record of this thread's tracklet in a chunk; the next chunk's is loaded ahead
  struct Rec {
    int2 fs;
    float2 tp;
  };
  auto recOf = [&](const int4& e) {
    Rec r{make_int2(0, 0), make_float2(0.0f, 0.0f)};
    if ((int)threadIdx.x < e.w) {
      r.fs = s.tFS[e.z + threadIdx.x];
      r.tp = s.tTP[e.z + threadIdx.x];
    }
    return r;
  };
  Rec recNext = cw.valid() ? recOf(cw.next()) : Rec{make_int2(0, 0), make_float2(0.0f, 0.0f)};
  for (; cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const Rec rec = recNext;
    if (cw.hasNext()) recNext = recOf(cw.next());
    const int sg = ci.x;
    const int w = sg / kCellLayers, l = sg % kCellLayers;
    const Seg tl = s.trkSeg[w * kTrackletLayers + l];
    const int it = ci.y + threadIdx.x;
    // every tracklet of the wavefront away from phi = +-pi: the wrap comparison is left out
    const bool waveWrapFree = kWrapFree && __ballot(it < tl.count && !(fabsf(rec.tp.y) <= kCellWrapFree)) == 0ull;
    int cnt = 0;
    if (it < tl.count) {
      const int t = tl.base + it;
      const CellSeed c = cellSeedOf(s, w, l, rec.fs, rec.tp);
      int* slots = s.cellCand + (size_t)t * kCellSlots;
      int m = 0;
      const int uEnd = c.uFirst + c.nCand;
      auto survivor = [&](int iu) {
        if (m < kCellList) {
          list[m][threadIdx.x] = iu;
        } else {  // long survivor list: fit inline
          CellOut o;
          if (m == kCellList) {
            for (int q = 0; q < kCellList; ++q)
              if (cellFit(s, c, list[q][threadIdx.x], o)) {
                if (cnt < kCellSlots) slots[cnt] = list[q][threadIdx.x];
                ++cnt;
              }
          }
          if (cellFit(s, c, iu, o)) {
            if (cnt < kCellSlots) slots[cnt] = iu;
            ++cnt;
          }
        }
        ++m;
      };
      // Blocks of kCandBlock candidates: loads first (indices past the range
      // clamped and masked out), cuts without branches, survivors in
      // ascending order.
      // (full blocks need no range test per candidate; the last one masks
      // the candidates past the range)
      auto candidates = [&](auto wrap) {
        constexpr bool kWrap = decltype(wrap)::value;
        auto block = [&](int ib, auto full) {
          constexpr bool kFull = decltype(full)::value;
          float2 up[kCandBlock];
#pragma unroll
          for (int k = 0; k < kCandBlock; k += 2) loadPair(s.tTP + c.tnBase + ib + k, up[k], up[k + 1]);
          uint32_t mask = 0;
#pragma unroll
          for (int k = 0; k < kCandBlock; ++k)
            mask |= ((cellCheapCuts<kWrap>(c, up[k]) & (kFull || ib + k < uEnd)) ? 1u : 0u) << k;
          while (mask != 0u) {
            const int k = __builtin_ctz(mask);
            mask &= mask - 1u;
            survivor(ib + k);
          }
        };
        int ib = c.uFirst;
        for (; ib + kCandBlock <= uEnd; ib += kCandBlock) block(ib, std::true_type{});
        if (ib < uEnd) block(ib, std::false_type{});
      };
      if constexpr (kWrapFree) {
        if (waveWrapFree)
          candidates(std::false_type{});
        else
          candidates(std::true_type{});
      } else {
        candidates(std::true_type{});
      }
      if (m <= kCellList) {  // survivors fitted kFitBatch at a time, their loads issued together
        for (int q0 = 0; q0 < m; q0 += kFitBatch) {
          int iu[kFitBatch];
          float4 q3[kFitBatch];
#pragma unroll
          for (int b = 0; b < kFitBatch; ++b) iu[b] = list[min(q0 + b, m - 1)][threadIdx.x];
#pragma unroll
          for (int b = 0; b < kFitBatch; ++b) q3[b] = cellThird(s, c, iu[b]);
#pragma unroll
          for (int b = 0; b < kFitBatch; ++b) {
            CellOut o;
            if (q0 + b < m && cellFitAt(c, q3[b], o)) {
              if (cnt < kCellSlots) slots[cnt] = iu[b];
              ++cnt;
            }
          }
        }
      }
      s.cellCnt[t] = cnt;
    }
    const int total = blockSum(cnt, scratch);
    if (threadIdx.x == 0) s.chunkSum[(size_t)kStCells * s.maxChunks + chunk] = total;
  }
}

__device__ __forceinline__ void writeCell(const DevState& s, int tnBase, int k, int it, int iu, const CellOut& o,
                                          int first, int second)
{
  if (k >= s.cCap) return;
  s.c1[k] = first;
  s.c2[k] = second;
  s.c3[k] = s.tFS[tnBase + iu].y;
  s.ct1[k] = it;
  s.ct2[k] = iu;
  s.cnc[k] = make_float4(o.nx, o.ny, o.nz, o.curvature);
  s.clevel[k] = 1;
  s.cpaths[k] = 1;
  s.edgeCnt[k] = 0;
}

// Fill: one cell per thread of the chunk, refitted from the kept candidate;
// tracklets with more than kCellSlots cells are redone by their own thread.
// Eight waves per SIMD (the registers fit without spills).
template <int B>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(8))) kCellsFill(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  __shared__ int exFull[B], exCap[B];
  __shared__ uint8_t owner[kOwnerRound];
  const int S = s.W * kCellLayers;
  const int* chunkStart = s.chunkStart + kStCells * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  ChunkWalk cw(s, kStCells, nChunks, xcdBlockIndex());
  // cell count of this thread's tracklet in a chunk (0 past its end); the next
  // chunk's is loaded while the current one is filled
  auto countOf = [&](const int4& e) { return (int)threadIdx.x < e.w ? s.cellCnt[e.z + threadIdx.x] : 0; };
  int cntNext = cw.valid() ? countOf(cw.next()) : 0;
  for (; cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int cnt = cntNext;
    if (cw.hasNext()) cntNext = countOf(cw.next());
    const int sg = ci.x;
    const int w = sg / kCellLayers, l = sg % kCellLayers;
    const Seg tl = s.trkSeg[w * kTrackletLayers + l];
    const int firstT = ci.y;
    const int it = firstT + threadIdx.x;
    const int t = tl.base + it;
    const int capped = cnt <= kCellSlots ? cnt : 0;
    int total, totalCap;
    const int ex = blockExclusiveScan(cnt, scratch, total);
    const int exc = blockExclusiveScan(capped, scratch, totalCap);
    exFull[threadIdx.x] = ex;
    exCap[threadIdx.x] = exc;
    const int chunkBase = s.chunkOff[(size_t)kStCells * s.maxChunks + chunk];
    if (it < tl.count) s.cellOff[t] = chunkBase + ex - s.celSeg[sg].base;
    forEachSlotOutput<B>(capped, exc, totalCap, exCap, owner, [&](int c, int n) {
      const int tt = tl.base + firstT + c;
      const int iu = s.cellCand[(size_t)tt * kCellSlots + n];
      const CellSeed cs = cellSeed(s, w, l, tt);
      CellOut out;
      cellFit(s, cs, iu, out);  // accepted in the count pass; same inputs, same result
      const int2 fs = s.tFS[tt];
      writeCell(s, cs.tnBase, chunkBase + exFull[c] + n, firstT + c, iu, out, fs.x, fs.y);
    });
    if (cnt > kCellSlots) {
      const CellSeed cs = cellSeed(s, w, l, t);
      int k = chunkBase + ex;
      for (int iu = cs.uFirst; iu < cs.uFirst + cs.nCand; ++iu) {
        CellOut out;
        if (cellCheapCuts(cs, s.tTP[cs.tnBase + iu]) && cellFit(s, cs, iu, out))
          writeCell(s, cs.tnBase, k++, it, iu, out, s.tFS[t].x, s.tFS[t].y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Stage (c): neighbours.  Tracker.cxx:416-474.  The successors of cell c of
// layer l are the cells of layer l+1 whose first tracklet is c's second
// tracklet: a contiguous range.  Edges are produced in (c, n) order by
// count -> scan -> fill, then transposed into each n's predecessor list in
// ascending c -- the order in which the reference pushes them.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) kPlanEdges(DevState s, int* counts)
{
  __shared__ int scratch[1024 / 64 + 1];
  const int S = s.W * (kCellLayers - 1);
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const int w = i / (kCellLayers - 1), l = i % (kCellLayers - 1);
    const int cNext = s.celSeg[w * kCellLayers + l + 1].count;
    counts[i] = cNext > 0 ? s.celSeg[w * kCellLayers + l].count : 0;
  }
  __syncthreads();
  planChunks(counts, S, s.chunk, s.chunkStart + kStEdges * (s.maxSegs + 1), scratch);
}

template <typename F>
__device__ __forceinline__ void forEachSuccessor(const DevState& s, int w, int l, int c /*global cell of l*/,
                                                 F&& emit)
{
  const int tnBase = s.trkSeg[w * kTrackletLayers + l + 1].base;
  const int cnBase = s.celSeg[w * kCellLayers + l + 1].base;
  const int u = tnBase + s.ct2[c];
  const int cnt = s.cellCnt[u];
  if (cnt == 0) return;
  const int first = s.cellOff[u];
  const float4 a = s.cnc[c];  // (nx, ny, nz, curvature)
  const float maxNormal = neighbourMaxNormalDelta(l), maxCurv = neighbourMaxCurvatureDelta(l);
  const int end = first + cnt;
  for (int ib = first; ib < end; ib += 4) {  // four candidates at a time, loads first
    float4 b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = s.cnc[cnBase + min(ib + k, end - 1)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float dx = a.x - b[k].x, dy = a.y - b[k].y, dz = a.z - b[k].z;
      const float deltaNormal = (dx * dx) + (dy * dy) + (dz * dz);
      const float deltaCurvature = fabsf(a.w - b[k].w);
      if (ib + k < end && deltaNormal < maxNormal && deltaCurvature < maxCurv) emit(ib + k);
    }
  }
}

template <int B>
__global__ void __launch_bounds__(B) kEdgesCount(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  const int S = s.W * (kCellLayers - 1);
  const int* chunkStart = s.chunkStart + kStEdges * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  for (ChunkWalk cw(s, kStEdges, nChunks, xcdBlockIndex()); cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int sg = ci.x;
    const int w = sg / (kCellLayers - 1), l = sg % (kCellLayers - 1);
    const Seg cl = s.celSeg[w * kCellLayers + l];
    const int ic = ci.y + threadIdx.x;
    int cnt = 0;
    if (ic < cl.count) {
      const int c = cl.base + ic;
      int* slots = s.succCand + (size_t)c * kEdgeSlots;
      forEachSuccessor(s, w, l, c, [&](int in) {
        if (cnt < kEdgeSlots) slots[cnt] = in;
        ++cnt;
      });
      s.succCnt[c] = cnt;
    }
    const int total = blockSum(cnt, scratch);
    if (threadIdx.x == 0) s.chunkSum[(size_t)kStEdges * s.maxChunks + chunk] = total;
  }
}

template <int B>
__global__ void __launch_bounds__(B) kEdgesFill(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  __shared__ int exFull[B], exCap[B];
  __shared__ uint8_t owner[kOwnerRound];
  const int S = s.W * (kCellLayers - 1);
  const int* chunkStart = s.chunkStart + kStEdges * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  ChunkWalk cw(s, kStEdges, nChunks, xcdBlockIndex());
  // successor count of this thread's cell in a chunk; the next chunk's is loaded ahead
  auto elemOf = [&](const int4& e) { return (int)threadIdx.x < e.w ? s.succCnt[e.z + threadIdx.x] : 0; };
  int cntNext = cw.valid() ? elemOf(cw.next()) : int{};
  for (; cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int cnt = cntNext;
    if (cw.hasNext()) cntNext = elemOf(cw.next());
    const int sg = ci.x;
    const int w = sg / (kCellLayers - 1), l = sg % (kCellLayers - 1);
    const Seg cl = s.celSeg[w * kCellLayers + l];
    const int firstC = ci.y;
    const int ic = firstC + threadIdx.x;
    const int c = cl.base + ic;
    const int capped = cnt <= kEdgeSlots ? cnt : 0;
    int total, totalCap;
    const int ex = blockExclusiveScan(cnt, scratch, total);
    const int exc = blockExclusiveScan(capped, scratch, totalCap);
    exFull[threadIdx.x] = ex;
    exCap[threadIdx.x] = exc;
    const int chunkBase = s.chunkOff[(size_t)kStEdges * s.maxChunks + chunk];
    forEachSlotOutput<B>(capped, exc, totalCap, exCap, owner, [&](int o, int n) {
      const int k = chunkBase + exFull[o] + n;
      if (k < s.eCap) {
        s.eFrom[k] = firstC + o;
        s.eTo[k] = s.succCand[(size_t)(cl.base + firstC + o) * kEdgeSlots + n];
      }
    });
    if (cnt > kEdgeSlots) {
      int k = chunkBase + ex;
      forEachSuccessor(s, w, l, c, [&](int in) {
        if (k < s.eCap) {
          s.eFrom[k] = ic;
          s.eTo[k] = in;
        }
        ++k;
      });
    }
  }
}

// Predecessor lists: a stable counting sort of the (c, n) edges of one
// (item, layer) by n.  One workgroup per (item, layer); keys in passes of
// kTransposeKeys.  Bucket placement by LDS atomics, then each edge is written
// at bucket start + (number of bucket members with a smaller edge index).  The
// fill writes the edges in ascending c, so that rank keeps the predecessors
// ascending whatever order the atomics took.  One LDS array serves as counts,
// bucket starts and insertion cursors: after the placement the cursor of
// bucket k stands at its end, which is the start of bucket k + 1.
// 76 KB of LDS, two workgroups of 1024 per CU: the largest cell layer of a
// central event (about 18k cells) takes one pass over its edges.
constexpr int kTransposeKeys = 19456;

__global__ void __launch_bounds__(1024) kEdgesTranspose(DevState s)
{
  __shared__ int hist[kTransposeKeys];
  __shared__ int scratch[1024 / 64 + 1];
  const int w = blockIdx.x / (kCellLayers - 1), l = blockIdx.x % (kCellLayers - 1);
  const Seg cn = s.celSeg[w * kCellLayers + l + 1];
  const Seg e = s.edgSeg[w * (kCellLayers - 1) + l];
  const int* to = s.eTo + e.base;
  const int* from = s.eFrom + e.base;
  int* tmp = s.eTmp + e.base;
  int carry = 0;
  for (int k0 = 0; k0 < cn.count; k0 += kTransposeKeys) {
    const int nk = min(kTransposeKeys, cn.count - k0);
    const int passStart = carry;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    for (int a = threadIdx.x; a < e.count; a += blockDim.x) {
      const int kk = to[a] - k0;
      if (kk >= 0 && kk < nk) atomicAdd(&hist[kk], 1);
    }
    __syncthreads();
    for (int k1 = 0; k1 < nk; k1 += blockDim.x) {
      const int k = k1 + threadIdx.x;
      const int v = k < nk ? hist[k] : 0;
      int total;
      const int ex = blockExclusiveScan(v, scratch, total);
      if (k < nk) {
        s.edgeOff[cn.base + k0 + k] = carry + ex;
        s.edgeCnt[cn.base + k0 + k] = v;
        hist[k] = carry + ex;  // insertion cursor
      }
      carry += total;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < e.count; a += blockDim.x) {
      const int kk = to[a] - k0;
      if (kk >= 0 && kk < nk) tmp[atomicAdd(&hist[kk], 1)] = a;
    }
    __threadfence_block();
    __syncthreads();
    for (int a = threadIdx.x; a < e.count; a += blockDim.x) {
      const int kk = to[a] - k0;
      if (kk < 0 || kk >= nk) continue;
      const int lo = kk > 0 ? hist[kk - 1] : passStart, hi = hist[kk];  // cursors ended at the bucket ends
      int rank = 0;
      for (int b = lo; b < hi; ++b) rank += tmp[b] < a ? 1 : 0;
      s.pred[e.base + lo + rank] = from[a];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage (d): CA levels and road counts.  Tracker.cxx:461-468 computes, layer
// by layer, level(n) = max(level(n), level(c) + 1) over accepted neighbours c;
// since layer l is final before layer l+1 is visited this equals
// 1 + max(0, max level of predecessors).  paths(n) counts the predecessor
// chains with levels stepping down by exactly one (the chains findTracks
// follows, Tracker.cxx:476-571).  One launch per cell layer 1..4.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) kPlanLayer(DevState s, int* counts, int layer, int stage)
{
  __shared__ int scratch[1024 / 64 + 1];
  const int S = s.W;
  for (int w = threadIdx.x; w < S; w += blockDim.x) counts[w] = s.celSeg[w * kCellLayers + layer].count;
  __syncthreads();
  planChunks(counts, S, s.chunk, s.chunkStart + stage * (s.maxSegs + 1), scratch);
}

template <int B>
__global__ void __launch_bounds__(B) kLevels(DevState s, int layer, int stage)
{
  __shared__ int scratch[B / 64 + 1];
  const int S = s.W;
  const int* chunkStart = s.chunkStart + stage * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  ChunkWalk cw(s, stage, nChunks, xcdBlockIndex());
  // predecessor count and offset of this thread's cell in a chunk; the next chunk's are loaded ahead
  auto predOf = [&](const int4& e) {
    return (int)threadIdx.x < e.w ? make_int2(s.edgeCnt[e.z + threadIdx.x], s.edgeOff[e.z + threadIdx.x])
                                  : make_int2(0, 0);
  };
  int2 predNext = cw.valid() ? predOf(cw.next()) : make_int2(0, 0);
  for (; cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int2 pr = predNext;
    if (cw.hasNext()) predNext = predOf(cw.next());
    const int w = ci.x;
    const Seg cn = s.celSeg[w * kCellLayers + layer];
    const Seg cp = s.celSeg[w * kCellLayers + layer - 1];
    const int in = ci.y + threadIdx.x;
    int roots = 0;
    if (in < cn.count) {
      const int n = cn.base + in;
      const int cnt = pr.x;
      int level = 1;
      uint32_t paths = 1;
      int only = -1;  // the only predecessor at the highest level, if there is one
      if (cnt > 0) {
        // One pass over the predecessors: the highest level so far and the
        // saturating sum of the path counts of the predecessors at that level
        // (a higher level restarts the sum), each predecessor's level and
        // count loaded together.  The sum saturates at 2^32 - 1, so its order
        // does not matter.
        const int off = s.edgSeg[w * (kCellLayers - 1) + layer - 1].base + pr.y;
        int best = 0, atBest = 0;
        paths = 0;
        for (int a = 0; a < cnt; ++a) {
          const int pi = s.pred[off + a];
          const int c = cp.base + pi;
          const int lv = s.clevel[c];
          const uint32_t p = s.cpaths[c];
          if (lv > best) {
            best = lv;
            paths = p;
            only = pi;
            atBest = 1;
          } else if (lv == best) {
            paths = paths + p < paths ? 0xffffffffu : paths + p;  // saturate
            ++atBest;
          }
        }
        level = best + 1;
        if (atBest != 1) only = -1;
      }
      s.clevel[n] = level;
      s.cpaths[n] = paths;
      s.cFirst[n] = only;
      if (layer == kCellLayers - 1 && level == kCellsMinLevel) {
        // A saturated count stays saturated through every later sum, so a
        // root reached by one exceeds the bound too.
        if (paths > kMaxRootRoads) s.status[kStatusRootRoads] = 1;
        roots = (int)min(paths, kMaxRootRoads);
      }
    }
    if (layer == kCellLayers - 1) {
      const int total = blockSum(roots, scratch);
      if (threadIdx.x == 0) s.chunkSum[(size_t)kStRoads * s.maxChunks + chunk] = total;
    }
  }
}

// Roads: every root (layer-4 cell of level 5) emits its chains in the order of
// the reference's depth-first traversal -- ascending predecessor at every
// layer -- with Moore's majority vote of the MC labels (Tracker.cxx:573-652).
// The roads of a chunk of roots are dealt one per thread (a root may have many
// chains, most have one).  Road q of a root is its q-th chain in that order:
// at every layer, the eligible predecessors (level one less) are taken in
// ascending order, each standing for paths(p) chains, and q falls in one of
// them.
template <int B>
__global__ void __launch_bounds__(B) kRoadsFill(DevState s)
{
  __shared__ int scratch[B / 64 + 1];
  __shared__ int exAll[B];
  __shared__ uint8_t owner[kOwnerRound];
  constexpr int L4 = kCellLayers - 1;
  const int S = s.W;
  const int* chunkStart = s.chunkStart + kStRoads * (s.maxSegs + 1);
  const int nChunks = chunkStart[S];
  for (ChunkWalk cw(s, kStRoads, nChunks, xcdBlockIndex()); cw.valid(); cw.advance()) {
    const int chunk = cw.chunk;
    const int4 ci = cw.take();
    const int w = ci.x;
    const Seg c4 = s.celSeg[w * kCellLayers + L4];
    const int first = ci.y;
    const int in = first + threadIdx.x;
    int cnt = 0;
    if (in < c4.count && s.clevel[c4.base + in] == kCellsMinLevel) cnt = (int)min(s.cpaths[c4.base + in], kMaxRootRoads);
    int total;
    const int ex = blockExclusiveScan(cnt, scratch, total);
    exAll[threadIdx.x] = ex;
    const int out0 = s.chunkOff[(size_t)kStRoads * s.maxChunks + chunk];
    if (cnt > 0) s.roadOff[c4.base + in] = out0 + ex - s.roadSeg[w].base;
    Seg cs[kCellLayers];
    for (int k = 0; k < kCellLayers; ++k) cs[k] = s.celSeg[w * kCellLayers + k];
    const int rw = s.recItem[w];
    const Seg cl[kLayers] = {s.clSeg[rw * kLayers + 0], s.clSeg[rw * kLayers + 1], s.clSeg[rw * kLayers + 2],
                             s.clSeg[rw * kLayers + 3], s.clSeg[rw * kLayers + 4], s.clSeg[rw * kLayers + 5],
                             s.clSeg[rw * kLayers + 6]};
    forEachSlotOutput<B>(cnt, ex, total, exAll, owner, [&](int c, int q0) {
      const long long outL = (long long)out0 + exAll[c] + q0;  // out0 <= rCap, the rest < 2^28
      if (outL >= s.rCap) return;
      const int out = (int)outL;
      int cell[kCellLayers];
      cell[L4] = first + c;
      uint32_t q = (uint32_t)q0;
      for (int layer = L4; layer > 0; --layer) {
        const int n = cs[layer].base + cell[layer];
        // A cell with one predecessor at the level below its own: every chain
        // through it continues there, whatever q.
        const int only = s.cFirst[n];
        if (only >= 0) {
          cell[layer - 1] = only;
          continue;
        }
        const int lvl = s.clevel[n];
        const int off = s.edgSeg[w * (kCellLayers - 1) + layer - 1].base + s.edgeOff[n];
        const int cnt2 = s.edgeCnt[n];
        int pick = -1;
        for (int a = 0; a < cnt2; ++a) {
          const int p = s.pred[off + a];
          const int pc = cs[layer - 1].base + p;
          if (s.clevel[pc] != lvl - 1) continue;
          const uint32_t paths = s.cpaths[pc];
          if (q < paths) {
            pick = p;
            break;
          }
          q -= paths;
        }
        // paths(n) is the sum over the eligible predecessors, and q < paths(root),
        // so some predecessor is always picked; anything else fails the run.
        if (pick < 0) {
          s.status[kStatusUndecodable] = 1;
          pick = 0;
        }
        cell[layer - 1] = pick;
      }
      for (int k = 0; k < kCellLayers; ++k) s.rCells[(size_t)out * kCellLayers + k] = cell[k];
      // Moore's voting over the 7 clusters of the chain.
      const int base0 = cs[0].base + cell[0];
      int value = s.cmc[cl[0].base + s.c1[base0]];
      int count = 1;
      bool fake = false;
      const int second = s.cmc[cl[1].base + s.c2[base0]];
      if (second == value) {
        ++count;
      } else {
        value = second;
        count = 1;
        fake = true;
      }
      for (int k = 0; k < kCellLayers; ++k) {
        const int current = s.cmc[cl[k + 2].base + s.c3[cs[k].base + cell[k]]];
        if (current == value) {
          ++count;
        } else {
          --count;
          fake = true;
        }
        if (count == 0) {
          value = current;
          count = 1;
        }
      }
      s.rLabel[out] = value;
      s.rFake[out] = fake ? 1 : 0;
    });
  }
}

// Road rows of one work item in the download format (cell of layers 0..4,
// size, label, fake: 8 int per road) into a caller's device buffer of `cap`
// rows; roads past cap are not written.
__global__ void __launch_bounds__(256) kRoadsOut(DevState s, int item, int* rows, int cap)
{
  const Seg g = s.roadSeg[item];
  const int n = min(g.count, cap);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = g.base + i;
    int size = 0;
    for (int k = 0; k < kCellLayers; ++k) {
      const int c = s.rCells[(size_t)r * kCellLayers + k];
      rows[(size_t)i * 8 + k] = c;
      size += c >= 0 ? 1 : 0;
    }
    rows[(size_t)i * 8 + 5] = size;
    rows[(size_t)i * 8 + 6] = s.rLabel[r];
    rows[(size_t)i * 8 + 7] = s.rFake[r];
  }
}

}  // namespace

void launchRoadsOut(const DevState& s, int item, int* rows, int cap, hipStream_t stream)
{
  hipLaunchKernelGGL(kRoadsOut, dim3(256), dim3(256), 0, stream, s, item, rows, cap);
}

// ---------------------------------------------------------------------------
// Launch sequence of one plain-CA run over the whole batch.
// ---------------------------------------------------------------------------
template <int B>
void launchPipelineB(const DevState& s, const LaunchConfig& cfg, hipStream_t stream, KernelTimer* timer)
{
  // workgroups of kChunkSegs for a stage of at most `elements` elements: one
  // thread per chunk, at most 512 workgroups
  auto segBlocks = [&](long elements) {
    const long chunks = elements / B + (long)s.W * kLayers + 1;
    return (unsigned)std::min<long>(512, (chunks + 255) / 256);
  };
  const int G = cfg.persistentBlocks;
  // Some kernels run faster with at most the workgroups that are resident at
  // once (perCu per CU; measured kernel by kernel): a grid of several rounds
  // runs its last, partial round on a partly idle chip.  The others keep the
  // tuned grid.
  auto resident = [&](int perCu) { return cfg.cus > 0 ? std::min(G, perCu * cfg.cus) : G; };
  const int G8 = resident(2048 / B);  // eight wavefronts per SIMD
  int* counts = cfg.planScratch;
  auto done = [&](const char* name, int stage) {
    if (timer) timer->after(name, stage, stream);
  };
  if (timer) timer->start(stream);
  if (cfg.maxLayerClusters <= 16 * kPrepThreads)
    hipLaunchKernelGGL(kPrep<16>, dim3(s.W * kLayers), dim3(kPrepThreads), 0, stream, s);
  else
    hipLaunchKernelGGL(kPrep<32>, dim3(s.W * kLayers), dim3(kPrepThreads), 0, stream, s);
  done("kPrep", 0);
  hipLaunchKernelGGL(kPlanTracklets, dim3(1), dim3(1024), 0, stream, s, counts);
  hipLaunchKernelGGL(kChunkSegs, dim3(segBlocks((long)s.W * kTrackletLayers * cfg.maxLayerClusters)), dim3(256), 0, stream, s,
                     (int)kStTracklets, s.W * kTrackletLayers, 0, 0);
  done("kPlanTracklets", 1);
  // the fused search and filter keeps candidates as 16-bit positions and reads the copy in sorted order
  if (cfg.filter && cfg.fusedFilter && s.zSub * s.phiSub == 1 && cfg.maxLayerClusters <= 65536) {
    for (int l = kTrackletLayers - 1; l >= 0; --l) {
      hipLaunchKernelGGL(kTrackletsSearchScore<B>, dim3(G), dim3(B), 0, stream, s, l, cfg.threshold);
      done("kTrackletsSearchScore", 1);
    }
    launchCommitKept(s, stream);
    done("kCommitKept", 1);
  } else {
    const int Gt = resident(kTrackletsCountPerCu);
    if (s.zSub * s.phiSub > 1)
      hipLaunchKernelGGL((kTrackletsCount<B, true>), dim3(Gt), dim3(B), 0, stream, s);
    else
      hipLaunchKernelGGL((kTrackletsCount<B, false>), dim3(Gt), dim3(B), 0, stream, s);
    done("kTrackletsCount", 1);
    if (cfg.filter) {
      launchTrackletFilter(s, B, resident(1024 / B), cfg.threshold, stream);  // four wavefronts per SIMD (its LDS)
      done("kTrackletsScore", 1);
    }
  }
  hipLaunchKernelGGL(kScanChunks, dim3(1), dim3(1024), 0, stream, s, (int)kStTracklets, s.W * kTrackletLayers,
                     s.trkSeg, s.tCap, (int)kStatusOverflowTracklets);
  done("kScanChunks", 1);
  if (cfg.filter)
    hipLaunchKernelGGL((kTrackletsFill<B, true>), dim3(G), dim3(B), 0, stream, s);
  else
    hipLaunchKernelGGL((kTrackletsFill<B, false>), dim3(G), dim3(B), 0, stream, s);
  done("kTrackletsFill", 1);
  hipLaunchKernelGGL(kPlanCells, dim3(1), dim3(1024), 0, stream, s, counts);
  hipLaunchKernelGGL(kChunkSegs, dim3(segBlocks(s.tCap)), dim3(256), 0, stream, s, (int)kStCells, s.W * kCellLayers, 1, 0);
  done("kPlanCells", 2);
  if (cfg.filter)
    hipLaunchKernelGGL((kCellsCount<B, 4, false>), dim3(G), dim3(B), 0, stream, s);
  else
    hipLaunchKernelGGL((kCellsCount<B, 8, true>), dim3(G), dim3(B), 0, stream, s);
  done("kCellsCount", 2);
  hipLaunchKernelGGL(kScanChunks, dim3(1), dim3(1024), 0, stream, s, (int)kStCells, s.W * kCellLayers, s.celSeg,
                     s.cCap, (int)kStatusOverflowCells);
  done("kScanChunks", 2);
  hipLaunchKernelGGL(kCellsFill<B>, dim3(G8), dim3(B), 0, stream, s);
  done("kCellsFill", 2);
  hipLaunchKernelGGL(kPlanEdges, dim3(1), dim3(1024), 0, stream, s, counts);
  hipLaunchKernelGGL(kChunkSegs, dim3(segBlocks(s.cCap)), dim3(256), 0, stream, s, (int)kStEdges, s.W * (kCellLayers - 1), 2,
                     0);
  done("kPlanEdges", 3);
  hipLaunchKernelGGL(kEdgesCount<B>, dim3(G8), dim3(B), 0, stream, s);
  done("kEdgesCount", 3);
  hipLaunchKernelGGL(kScanChunks, dim3(1), dim3(1024), 0, stream, s, (int)kStEdges, s.W * (kCellLayers - 1),
                     s.edgSeg, s.eCap, (int)kStatusOverflowEdges);
  done("kScanChunks", 3);
  hipLaunchKernelGGL(kEdgesFill<B>, dim3(2 * G), dim3(B), 0, stream, s);  // measured faster with twice the grid
  done("kEdgesFill", 3);
  hipLaunchKernelGGL(kEdgesTranspose, dim3(s.W * (kCellLayers - 1)), dim3(1024), 0, stream, s);
  done("kEdgesTranspose", 3);
  for (int layer = 1; layer < kCellLayers; ++layer) {
    // The last level launch shares its chunk plan with the roads stage: it also
    // writes each chunk's number of roads.
    const int stage = layer == kCellLayers - 1 ? (int)kStRoads : kStLevels + layer - 1;
    hipLaunchKernelGGL(kPlanLayer, dim3(1), dim3(1024), 0, stream, s, counts, layer, stage);
    hipLaunchKernelGGL(kChunkSegs, dim3(segBlocks(s.cCap)), dim3(256), 0, stream, s, stage, s.W, 3, layer);
    done("kPlanLayer", 4);
    hipLaunchKernelGGL(kLevels<B>, dim3(G8), dim3(B), 0, stream, s, layer, stage);
    done("kLevels", 4);
  }
  hipLaunchKernelGGL(kScanChunks, dim3(1), dim3(1024), 0, stream, s, (int)kStRoads, s.W, s.roadSeg, s.rCap,
                     (int)kStatusOverflowRoads);
  done("kScanChunks", 4);
  hipLaunchKernelGGL(kRoadsFill<B>, dim3(2 * G), dim3(B), 0, stream, s);  // measured faster with twice the grid
  done("kRoadsFill", 4);
}

void launchPipeline(const DevState& s, const LaunchConfig& cfg, hipStream_t stream, KernelTimer* timer)
{
  switch (s.chunk) {
    case 64: launchPipelineB<64>(s, cfg, stream, timer); break;
    case 128: launchPipelineB<128>(s, cfg, stream, timer); break;
    default: launchPipelineB<256>(s, cfg, stream, timer); break;
  }
}

}  // namespace itsca
