// host_capi.cpp -- C entry points of libitsca.so for event batches
// (include/itsca_host.h).
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <unordered_map>

#include "events.hpp"
#include "itsca_host.h"
#include "roads_report.hpp"

struct itsca_events {
  itsca::EventBatch batch;
};

extern "C" {

itsca_events* itsca_events_generate(int preset, uint64_t seed, int num_events)
{
  if (num_events < 0 || (preset != 1 && preset != 2 && preset != 3 && preset != 5)) return nullptr;
  itsca_events* ev = new (std::nothrow) itsca_events();
  if (!ev) return nullptr;
  itsca::generateEvents(preset, seed, num_events, ev->batch);
  return ev;
}

itsca_events* itsca_events_load(const char* path)
{
  if (!path) return nullptr;
  itsca_events* ev = new (std::nothrow) itsca_events();
  if (!ev) return nullptr;
  std::string err;
  if (!itsca::loadEvents(path, ev->batch, err)) {
    delete ev;
    return nullptr;
  }
  return ev;
}

int itsca_events_write(const itsca_events* ev, const char* events_path, const char* labels_path)
{
  if (!ev || !events_path) return -1;
  std::string err;
  if (!itsca::writeEvents(ev->batch, events_path, err)) return -1;
  if (labels_path && !itsca::writeLabels(ev->batch, labels_path, err)) return -1;
  return 0;
}

int itsca_events_write_binary(const itsca_events* ev, const char* path)
{
  if (!ev || !path) return -1;
  std::string err;
  return itsca::writeEventsBinary(ev->batch, path, err) ? 0 : -1;
}

int itsca_events_load_labels(itsca_events* ev, const char* labels_path)
{
  if (!ev || !labels_path) return -1;
  std::string err;
  return itsca::loadLabelRows(labels_path, ev->batch, err) ? 0 : -1;
}

itsca_events* itsca_events_merge_pileup(const itsca_events* ev, int pile_up)
{
  if (!ev || pile_up < 1) return nullptr;
  itsca_events* out = new (std::nothrow) itsca_events();
  if (!out) return nullptr;
  std::string err;
  if (!itsca::mergePileUp(ev->batch, pile_up, out->batch, err)) {
    delete out;
    return nullptr;
  }
  return out;
}

void itsca_events_free(itsca_events* ev) { delete ev; }

int itsca_events_roads_report(const itsca_events* ev, int event, const int32_t* const* roads,
                              const int32_t* num_roads, int num_vertices, long counts[3])
{
  if (!ev || !counts || num_vertices < 0 || (num_vertices > 0 && (!roads || !num_roads))) return -1;
  if (event < 0 || event >= (int)ev->batch.labels.size()) return -1;
  // IOUtils.cxx:109: only charged pions that left 7 clusters are labels
  std::unordered_map<int, int> labels;
  for (const itsca::McLabel& m : ev->batch.labels[event])
    if (std::abs(m.pdg) == 211 && m.nClusters == 7) labels.emplace(m.mc, m.mc);
  counts[0] = counts[1] = counts[2] = 0;
  itsca::RoadClassifier classify;
  for (int v = 0; v < num_vertices; ++v)
    for (int i = 0; i < num_roads[v]; ++i) {
      const int32_t* row = roads[v] + (size_t)i * 8;  // 5 cells, size, label, fake
      const int cls = classify.next(row[6], row[7] != 0, [&](int label, int& mc) {
        const auto it = labels.find(label);
        if (it == labels.end()) return false;
        mc = it->second;
        return true;
      });
      if (cls >= 0) ++counts[cls];
    }
  return 0;
}

int itsca_events_count(const itsca_events* ev) { return ev ? (int)ev->batch.events.size() : -1; }

int itsca_events_num_vertices(const itsca_events* ev, int event)
{
  if (!ev || event < 0 || event >= (int)ev->batch.events.size()) return -1;
  return ev->batch.events[event].numVertices();
}

int itsca_events_layer_offsets(const itsca_events* ev, int event, int32_t offsets[8])
{
  if (!ev || !offsets || event < 0 || event >= (int)ev->batch.events.size()) return -1;
  const itsca::RawEvent& e = ev->batch.events[event];
  offsets[0] = 0;
  for (int l = 0; l < itsca::kNumLayers; ++l) offsets[l + 1] = offsets[l] + e.layerSize(l);
  return 0;
}

int itsca_events_copy(const itsca_events* ev, int event, float* x, float* y, float* z, int32_t* cluster_id,
                      int32_t* mc_id, float* vertices)
{
  if (!ev || event < 0 || event >= (int)ev->batch.events.size()) return -1;
  const itsca::RawEvent& e = ev->batch.events[event];
  size_t o = 0;
  for (int l = 0; l < itsca::kNumLayers; ++l) {
    const size_t n = (size_t)e.layerSize(l);
    if (n) {
      if (x) std::memcpy(x + o, e.x[l].data(), n * 4);
      if (y) std::memcpy(y + o, e.y[l].data(), n * 4);
      if (z) std::memcpy(z + o, e.z[l].data(), n * 4);
      if (cluster_id) std::memcpy(cluster_id + o, e.id[l].data(), n * 4);
      if (mc_id) std::memcpy(mc_id + o, e.mc[l].data(), n * 4);
    }
    o += n;
  }
  if (vertices && !e.vertices.empty()) std::memcpy(vertices, e.vertices.data(), e.vertices.size() * 4);
  return 0;
}

}  // extern "C"

// C linkage from the declaration in itsca_host.h.
int itsca_filter_load(const char* path, float* params, int n_params, float* threshold)
{
  if (!path || !params || !threshold || n_params < 1) return -1;
  std::ifstream in(path);
  if (!in) return -1;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  size_t pos = text.find("\"params_f32_bits\"");
  if (pos == std::string::npos) return -1;
  pos = text.find('[', pos);
  const size_t end = pos == std::string::npos ? std::string::npos : text.find(']', pos);
  if (end == std::string::npos) return -1;
  int n = 0;
  for (size_t q = text.find('"', pos); q != std::string::npos && q < end; q = text.find('"', q)) {
    const size_t r = text.find('"', q + 1);
    if (r == std::string::npos || r > end || n >= n_params) return -1;
    const uint32_t bits = (uint32_t)std::stoul(text.substr(q + 1, r - q - 1), nullptr, 16);
    std::memcpy(&params[n++], &bits, 4);
    q = r + 1;
  }
  if (n != n_params) return -1;
  pos = text.find("\"threshold\"");
  if (pos == std::string::npos) return -1;
  pos = text.find(':', pos);
  if (pos == std::string::npos) return -1;
  *threshold = std::strtof(text.c_str() + pos + 1, nullptr);
  return 0;
}

// C linkage from the declaration in itsca_host.h.
int itsca_filter_load_gates(const char* path, float* gates, int n_gates)
{
  if (!path || !gates || n_gates < 1) return -1;
  std::ifstream in(path);
  if (!in) return -1;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  for (int l = 0; l < n_gates; ++l) gates[l] = 0.3f;  // no gate
  size_t pos = text.find("\"phi_gates\"");
  if (pos == std::string::npos) return 0;
  pos = text.find('[', pos);
  const size_t end = pos == std::string::npos ? std::string::npos : text.find(']', pos);
  if (end == std::string::npos) return -1;
  const char* c = text.c_str() + pos + 1;
  for (int l = 0; l < n_gates; ++l) {
    char* next = nullptr;
    const float v = std::strtof(c, &next);
    if (next == c || (size_t)(next - text.c_str()) > end) return -1;
    gates[l] = v;
    c = next;
    while (*c == ',' || *c == ' ' || *c == '\n') ++c;
  }
  return 0;
}
