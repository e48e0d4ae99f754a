// Tracker.h -- the reference tracker entry point, Tracker<IsGPU>
// (include/ITSReconstruction/CA/Tracker.h:50-79), on the gfx950 back end.
//
// Tracker<true> runs the whole plain CA (index tables, tracklets, cells,
// neighbours, levels, roads and MC labels) on the device through the C ABI of
// libitsca_hip.so.  Errors are thrown as std::runtime_error carrying the
// library's message, like the reference's CUDA path (gpu/cuda/Utils.cu:27-38).
// The reference's CPU back end, Tracker<false>, is not part of this library.
#pragma once

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "Event.h"
#include "PrimaryVertexContext.h"
#include "Road.h"

struct itsca_ctx;

namespace o2 {
namespace ITS {
namespace CA {

template <bool IsGPU>
class Tracker;

template <>
class Tracker<true> {
 public:
  explicit Tracker(int device = 0);
  ~Tracker();
  Tracker(const Tracker&) = delete;
  Tracker& operator=(const Tracker&) = delete;

  // One vector of roads per primary vertex (Tracker.cxx:244-265).
  std::vector<std::vector<Road>> clustersToTracks(const Event& event);
  // Same for a batch of events in one device run: result[e][v].
  std::vector<std::vector<std::vector<Road>>> clustersToTracks(const std::vector<Event>& events);
  // Per-vertex device time per stage, tab-separated like the reference's
  // TimeBenchmark output (Tracker.cxx:368-402): prep, tracklets, cells,
  // neighbours, levels+roads, total.
  std::vector<std::vector<Road>> clustersToTracksTimeBenchmark(const Event& event, std::ofstream& out);
  // The same device stage times printed to std::cout in the reference's
  // DEBUG format (Tracker.cxx:267-301, 654-684).  The MC labels are computed
  // inside the roads kernel, so their own line reports 0 ms.
  std::vector<std::vector<Road>> clustersToTracksVerbose(const Event& event);
  // Per vertex, the reference's GPU-mode MemoryOccupancy lines
  // (Tracker.cxx:303-366 with TRACKINGITSU_GPU_MODE): clusters per layer;
  // cell capacity per layer (the cell buffer is one compact array shared by
  // the layers, so each layer may use the whole of it); cells per layer;
  // roads.
  std::vector<std::vector<Road>> clustersToTracksMemoryBenchmark(const Event& event, std::ofstream& out);

  // Host snapshot of the device state of vertex `vertex` of the last run
  // (the batch's work item order: event by event, vertex by vertex).
  PrimaryVertexContext getPrimaryVertexContext(int item) const;

  // Learned tracklet filter (no counterpart in the reference): tracklet
  // candidates whose network logit is below the threshold are dropped before
  // cell building; every later cut is the exact one, so the roads found are a
  // subset of the plain roads.  The weights file is the one written by
  // `python -m itsca.filter_training` (1441 parameters, a threshold and the
  // per-layer phi gates of the tracklet search; no gates: an empty vector).
  void useLearnedFilter(const std::string& weightsFile);
  void useLearnedFilter(const std::vector<float>& params, float threshold,
                        const std::vector<float>& phiGates = std::vector<float>());
  void disableLearnedFilter();

 private:
  void runBatch(const std::vector<const Event*>& events);
  // Runs vertex v of the event as its own work item with per-stage timing.
  void runVertexProfiled(const Event& event, int v, float ms[5]);
  std::vector<Road> downloadRoads(int item) const;
  // The roads of every work item of the last run, in one copy per array.
  std::vector<std::vector<Road>> allRoads() const;
  itsca_ctx* mCtx = nullptr;
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
