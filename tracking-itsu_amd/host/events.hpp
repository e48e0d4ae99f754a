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
bool writeEvents(const EventBatch& batch, const std::string& path, std::string& err);
bool writeLabels(const EventBatch& batch, const std::string& path, std::string& err);
bool loadEvents(const std::string& path, EventBatch& out, std::string& err);

}  // namespace itsca
