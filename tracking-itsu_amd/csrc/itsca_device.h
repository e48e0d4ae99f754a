// itsca_device.h -- device-side state of one batch of (event, vertex) work
// items, shared by the kernels and the runtime.
//
// HBM layout (all structure-of-arrays, 4-byte elements):
//   raw clusters      per event, layer after layer, file order (upload once)
//   sorted clusters   per work item and layer at clSeg[w*7+l].base; a work item
//                     is one primary vertex of one event (Tracker.cxx:251-262)
//   index tables      401 int per (item, layer): first sorted position of each
//                     z-phi bin (PrimaryVertexContext.cxx:90-119)
//   tracklets         compact over all (item, layer) segments, trkSeg[w*6+l]
//   cells             compact, celSeg[w*5+l]
//   predecessor lists compact, edgSeg[w*4+l]: for every cell of layer l+1 the
//                     ascending list of neighbouring cells of layer l
//   roads             compact, roadSeg[w]
// Indices stored inside records are local to their (item, layer), exactly as
// the reference stores positions in its per-layer vectors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "itsca_math.h"

namespace itsca {

struct Seg {
  int base;
  int count;
};

// Stage plans: each stage's persistent kernels walk chunks of `chunk`
// elements (= their workgroup size: 64, 128 or 256, chosen at run time);
// chunkStart[s] is the first chunk of segment s.
constexpr int kMinChunk = 64;
// Accepted candidates kept per element between a count and its fill.
constexpr int kTrkSlots = 64;  // per cluster (next-layer cluster positions)
// per tracklet (next-layer tracklet indices).  A tracklet with more cells is
// redone serially by one thread of the fill while its workgroup waits; in
// central events 0.2 % of the layer-1 tracklets have 5-9 cells, enough to put
// one in a third of the layer-1 chunks with 4 slots, and none has more than 8.
constexpr int kCellSlots = 8;
constexpr int kEdgeSlots = 8;  // per cell (next-layer cell indices)
// Records allocated past the tracklet capacity (tTP): the cell count reads its
// candidates in blocks whose last one may run past the end of the range.
constexpr int kTrackletPad = 8;

// Tracklet search index: inside every z-phi bin of at most kSubMaxBin
// clusters, the clusters ordered by sub-bin -- phi sub-row (phiSubIndex,
// P = DevState::phiSub parts per row, at most kMaxPhiSub), then z sub-bin
// (zSubIndex, S = DevState::zSub parts per bin, at most kMaxZSub): sub-bin
// phiSub * S + zSub -- with the start of each sub-bin as a byte of a 16-byte
// entry, so S * P <= kMaxSubBins.
constexpr int kMaxZSub = 8;
constexpr int kMaxPhiSub = 4;
constexpr int kMaxSubBins = 15;
constexpr int kSubMaxBin = 128;
constexpr int kDefaultZSub = 8;

enum Stage { kStTracklets = 0, kStCells = 1, kStEdges = 2, kStLevels = 3, kStRoads = 7, kNumStages = 8 };
// kStLevels + (l-1) for cell layers 1..4 (levels of layer 0 are all 1).

enum StatusSlot {
  kStatusOverflowTracklets = 0,  // capacity exceeded: the number of outputs needed
  kStatusOverflowCells = 1,
  kStatusOverflowEdges = 2,
  kStatusOverflowRoads = 3,
  kStatusRootRoads = 4,    // a root cell has more than kMaxRootRoads roads (not retried)
  kStatusUndecodable = 5,  // a road chain could not be decoded from the path counts (not retried)
  kNumStatus = 8
};

// Roads of one root cell (chains of level-descending cells ending at it).  A
// root with more sets kStatusRootRoads and the run fails with a capacity
// error instead of dropping roads.  The bound keeps every per-chunk road sum
// (256 roots) inside 32 bits.
#ifndef ITSCA_MAX_ROOT_ROADS
#define ITSCA_MAX_ROOT_ROADS (1 << 20)
#endif
constexpr uint32_t kMaxRootRoads = ITSCA_MAX_ROOT_ROADS;

struct DevState {
  int W;  // work items
  int maxSegs;
  int chunk;  // elements per chunk = workgroup size of the stage kernels
  // work items
  const int* itemEvent;
  const float* itemVertex;  // 3 per item
  // raw clusters of the uploaded events
  const float *rx, *ry, *rz;
  const int *rid, *rmc;
  const int* evLayerOff;  // 8 per event, absolute into the raw arrays
  // sorted clusters.  clSeg indexes the per-item arrays aligned with the
  // sorted clusters (trkCnt, trkOff, trkKept, trkCand); the records
  // themselves (cx..cbin, crzp, crzpS, table, subTab, segGeo) are those of
  // item recItem[w]: the item itself, or an earlier vertex of the same event
  // with the same x and y, whose preparation is identical (phi and r are
  // taken about (x, y), the bins use the absolute z).
  const Seg* clSeg;     // W*7
  const int* recItem;   // W
  float4* cxyrz;  // (x, y, r, z) per sorted cluster: one load for the tracklet fill and the cell fit
  float *cz, *cphi, *cr;
  float4* crzp;  // (r, z, phi, 0) per sorted cluster: one load per tracklet candidate
  float4* crzpS;  // the same records, z-sub-bin order inside each bin; w = offset in the bin (int bits)
  uint4* subTab;  // per (item, layer) and bin: byte t = offset of sub-bin t in crzpS (t <= zSub * phiSub)
  int zSub;       // z sub-bins per bin of the search index (1..kMaxZSub)
  int phiSub;     // phi sub-rows per row of the search index (1..kMaxPhiSub)
  float4* segGeo;  // per (item, layer): (min r, max r, 1 if every z bin index lies in [0, kZBins) else 0, 0)
  int *cid, *cmc, *cbin;
  float *sPhi, *sR;  // prep scratch, original order
  int *sBin, *sInv;  // per-cluster scratch of the sort: bin, source of each sorted position
  int* table;  // W*7*401
  // tracklets
  int *trkCnt, *trkOff;  // per sorted cluster
  int* trkKept;          // per sorted cluster: candidates kept by the learned filter
  const uint16_t* mlpW;  // learned filter parameters, BF16 bits
  // learned form: per tracklet layer, the phi gate of the tracklet search (the
  // phi cut becomes |dphi| < phiGate[l] <= kPhiCut); gated = 0: plain search
  float phiGate[kTrackletLayers];
  int gated;
  int* trkCand;          // kTrkSlots per sorted cluster
  int* trkKeep;          // filtered form: per chunk of clusters, their kept candidates as one list (chunk's first cluster * kTrkSlots)
  Seg* trkSeg;           // W*6
  int2* tFS;    // (first cluster, second cluster) positions
  float2* tTP;  // (tanLambda, phi): one load per cell candidate
  int tCap;
  // cells
  int *cellCnt, *cellOff;  // per tracklet
  int* cellCand;           // kCellSlots per tracklet
  Seg* celSeg;             // W*5
  int *c1, *c2, *c3, *ct1, *ct2, *clevel;
  float4* cnc;  // (nx, ny, nz, curvature) of the cell
  uint32_t* cpaths;  // number of level-descending paths from layer 0
  int cCap;
  // neighbour edges (c of layer l -> n of layer l+1), first in (c, n) order,
  // then transposed into predecessor lists
  int *succCnt, *succCand;  // per cell: accepted successors, kEdgeSlots slots
  int* cFirst;  // per cell: its only predecessor at the highest level, else -1 (succCnt's memory, free after the edges)
  int *eFrom, *eTo, *eTmp;  // compact, edgSeg order
  int *edgeCnt, *edgeOff;   // per cell of layer l+1: its predecessor list
  Seg* edgSeg;              // W*4
  int* pred;
  int eCap;
  // roads
  int* roadOff;  // per cell (roots in layer 4)
  Seg* roadSeg;  // W
  int *rCells, *rLabel, *rFake;
  int rCap;
  // plans
  int* chunkStart;  // kNumStages * (maxSegs + 1)
  int *chunkSum, *chunkOff;  // kNumStages * maxChunks
  int4* chunkInfo;           // kNumStages * maxChunks: (segment, first element, its global index, elements) per chunk
  int maxChunks;
  int* status;  // kNumStatus
};

// Segment of the sorted-cluster records that work item w reads for layer l.
__device__ __forceinline__ Seg recSeg(const DevState& s, int w, int l) { return s.clSeg[s.recItem[w] * kLayers + l]; }

}  // namespace itsca
