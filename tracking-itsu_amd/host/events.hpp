// events.hpp -- in-memory event batches, the reference text formats, and the
// seeded synthetic generator for the benchmark workloads.
//
// Text formats follow src/CA/IOUtils.cxx:38-78 (events) and :80-123 (labels).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace itsca {

constexpr int kNumLayers = 7;

struct RawEvent {
  std::vector<float> vertices;  // 3 per primary vertex (x, y, z)
  // Per layer, clusters in file order.  clusterId is the running index in file
  // order over the whole event (IOUtils.cxx:63,71).
  std::vector<float> x[kNumLayers], y[kNumLayers], z[kNumLayers], alpha[kNumLayers];
  std::vector<int32_t> id[kNumLayers], mc[kNumLayers];

  int numVertices() const { return (int)vertices.size() / 3; }
  int layerSize(int l) const { return (int)x[l].size(); }
};

struct McLabel {
  int mc;
  float pt, phi, eta;
  int pdg;
  int nClusters;
};

struct EventBatch {
  std::vector<RawEvent> events;
  std::vector<std::vector<McLabel>> labels;  // one list per event (may be empty)
};

// Workload presets of the generator.
enum GenPreset {
  kGenSmall = 1,    // 100 primaries, 1% noise, one vertex at z ~ N(0, 5 cm)
  kGenCentral = 2,  // 8000 primaries + 2% secondaries (r < 2 cm) + 2% noise
  kGenMixed = 3,    // like kGenCentral with multiplicity log-uniform in [50, 8000]
  kGenPileUp = 5,   // 4 collisions at z ~ U(-10, 10), 2000-8000 primaries each, 3% noise
};

// Generates `numEvents` events; event i uses seed `seed + i`.
void generateEvents(int preset, uint64_t seed, int numEvents, EventBatch& out);

// Reference-format text I/O.  loadEvents skips unparseable lines like the
// reference reader and starts a new event when a vertex line follows clusters.
// writeEvents emits each event's clusters in cluster-id order (the order of
// the file they were read from).
bool writeEvents(const EventBatch& batch, const std::string& path, std::string& err);
bool writeLabels(const EventBatch& batch, const std::string& path, std::string& err);
// loadEvents also reads the binary side-car below (recognised by its magic).
bool loadEvents(const std::string& path, EventBatch& out, std::string& err);
// Labels file (IOUtils.cxx:80-123 format: header line, events separated by
// "-1" lines); every row is kept (the reference reader's pdg/cluster filter is
// applied by IOUtils::loadLabels).
bool loadLabelRows(const std::string& path, EventBatch& batch, std::string& err);

// Binary side-car of an event batch (no counterpart in the reference): the
// content of the text format, little endian, read back through mmap.
//   "ITSCAEV1", uint32 events; per event: uint32 vertices, uint32 clusters of
//   each of the 7 layers, float vertices[3 * vertices]; then per layer
//   float x[n], y[n], z[n], alpha[n], int32 clusterId[n], mc[n].
bool writeEventsBinary(const EventBatch& batch, const std::string& path, std::string& err);
bool loadEventsBinary(const std::string& path, EventBatch& out, std::string& err);
bool isEventsBinary(const std::string& path);

// Pile-up merger (benchmarks/scripts/root/generatePileUpData.cxx:44-207):
// consecutive events are merged until a merged event holds `pileUp` vertices;
// it lists their vertices, then their clusters in file order, and the MC
// labels of clusters (and label rows) that follow the k-th vertex of the
// merged event are offset by (k - 1) * 100000.  Cluster ids follow the merged
// file order: a collision's clusters come after those of the earlier ones.
constexpr int kPileUpMcOffset = 100000;
bool mergePileUp(const EventBatch& in, int pileUp, EventBatch& out, std::string& err);

}  // namespace itsca
