// ca_oracle.hpp -- CPU restatement of the reference plain cellular-automaton
// track finder (one primary vertex at a time).
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker for the gfx950 path: only
// the tests, the smoke check and the CPU-baseline leg of the benchmark may
// load it.  It is never part of the product library.
//
// It restates, in the reference's evaluation order and with the reference's
// libm (std::atan2 on floats, i.e. glibc atan2f), the functions
//   PrimaryVertexContext::initialize      src/CA/PrimaryVertexContext.cxx:31-142
//   Cluster vertex ctor                   src/CA/Cluster.cxx:35-45
//   TrackingUtils::getBinsRect            src/CA/TrackingUtils.cxx:30-48
//   TrackerTraits<false>::computeLayerTracklets  src/CA/Tracker.cxx:44-117
//   Tracklet ctor                         src/CA/Tracklet.cxx:32-40
//   TrackerTraits<false>::computeLayerCells      src/CA/Tracker.cxx:119-235
//   Tracker::findCellsNeighbours          src/CA/Tracker.cxx:416-474
//   Tracker::findTracks / traverseCellsTree      src/CA/Tracker.cxx:476-571
//   Tracker::computeMontecarloLabels      src/CA/Tracker.cxx:573-652
// Two deliberate differences, both documented at their site: the candidate
// loop of the tracklet search is guarded by "< N" (the reference may read one
// element past the end of the next layer), and the lookup tables are sized
// from exact counts instead of the fitted memory coefficients.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace itsca_oracle {

constexpr int kLayers = 7;
constexpr int kTrackletLayers = 6;
constexpr int kCellLayers = 5;
constexpr int kTableSize = 401;

// Same 36-byte layout and member order as the reference Cluster
// (include/ITSReconstruction/CA/Cluster.h:36-44) so that std::sort sees the
// same sequence of comparisons.
struct Cluster {
  float x, y, z, phi, r;
  int id;
  float alpha;
  int mc;
  int bin;
};

struct Tracklet {
  int first, second;
  float tanLambda, phi;
};

struct Cell {
  int c1, c2, c3, t1, t2;
  float nx, ny, nz, curvature;
  int level;
};

struct Road {
  int cells[kCellLayers];
  int size;
  int label;
  bool fake;
};

struct EventView {
  const float *x, *y, *z;  // all layers concatenated, layer by layer, file order
  const int *id, *mc;
  const int* layerOffsets;  // kLayers + 1 entries
  float vx, vy, vz;
};

struct VertexResult {
  float vx = 0, vy = 0, vz = 0;
  std::array<std::vector<Cluster>, kLayers> clusters;
  std::array<std::array<int, kTableSize>, kTrackletLayers> tables{};  // [l-1] indexes layer l
  std::array<std::vector<Tracklet>, kTrackletLayers> tracklets;
  std::array<std::vector<int>, kCellLayers> trackletLookup;  // [l] indexed by clusters of layer l+1
  std::array<std::vector<Cell>, kCellLayers> cells;
  std::array<std::vector<int>, kCellLayers - 1> cellLookup;  // [l] indexed by tracklets of layer l+1
  std::array<std::vector<std::vector<int>>, kCellLayers - 1> neighbours;  // [l] indexed by cells of layer l+1
  std::vector<Road> roads;
};

// Runs the plain CA for one vertex.  stableSort selects std::stable_sort for
// the per-layer bin ordering (the reference uses std::sort).  exclusiveBound
// is not the reference: it drops the extra candidate of the reference's
// inclusive upper bound (Tracker.cxx:84-89), which leaves every tracklet the
// window search can find exactly once, whatever the sort order; the tests use
// it to show that the two sort orders differ only in the duplicates that
// extra candidate adds.
void runVertex(const EventView& ev, bool stableSort, VertexResult& out, bool exclusiveBound = false);

// Canonical text dump; floats are written as their IEEE bit patterns.
std::string dump(const VertexResult& r);

}  // namespace itsca_oracle
