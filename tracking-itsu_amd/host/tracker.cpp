// tracker.cpp -- Tracker<true> and IOUtils of the C++ API, on top of the C ABI
// of libitsca_hip.so (include/itsca.h).
#include <cstdint>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <tuple>
#include <unordered_set>

#include "events.hpp"
#include "roads_report.hpp"
#include "ITSReconstruction/CA/IOUtils.h"
#include "ITSReconstruction/CA/Tracker.h"
#include "itsca.h"
#include "itsca_host.h"

namespace o2 {
namespace ITS {
namespace CA {

namespace {

void check(itsca_ctx* ctx, itsca_status st, const char* what)
{
  if (st != ITSCA_OK) throw std::runtime_error(std::string(what) + ": " + itsca_last_error(ctx));
}

template <typename T>
std::vector<T> download(itsca_ctx* ctx, int item, int what, int layer, size_t rows, size_t fields)
{
  std::vector<int32_t> buf(rows * fields + 1);
  size_t written = 0;
  check(ctx, itsca_download(ctx, item, what, layer, buf.data(), buf.size() * 4, &written), "itsca_download");
  buf.resize(written / 4);
  std::vector<T> out;
  (void)out;
  return std::vector<T>(buf.begin(), buf.end());
}

float asFloat(int32_t v)
{
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

}  // namespace

void Event::printPrimaryVertices() const
{
  for (const float3& v : mPrimaryVertices) std::cout << "-1\t" << v.x << "\t" << v.y << "\t" << v.z << std::endl;
}

Tracker<true>::Tracker(int device)
{
  itsca_config cfg{};
  cfg.device = device;
  const itsca_status st = itsca_create(&mCtx, &cfg);
  if (st != ITSCA_OK) {
    const std::string msg = mCtx ? itsca_last_error(mCtx) : "itsca_create failed";
    itsca_destroy(mCtx);
    mCtx = nullptr;
    throw std::runtime_error(msg);
  }
}

Tracker<true>::~Tracker() { itsca_destroy(mCtx); }

void Tracker<true>::useLearnedFilter(const std::string& weightsFile)
{
  std::vector<float> params(ITSCA_MLP_PARAMS);
  float threshold = 0.0f;
  std::vector<float> gates(ITSCA_TRACKLET_LAYERS);
  if (itsca_filter_load(weightsFile.c_str(), params.data(), ITSCA_MLP_PARAMS, &threshold) != 0 ||
      itsca_filter_load_gates(weightsFile.c_str(), gates.data(), ITSCA_TRACKLET_LAYERS) != 0)
    throw std::runtime_error("cannot read learned filter weights from " + weightsFile);
  useLearnedFilter(params, threshold, gates);
}

void Tracker<true>::useLearnedFilter(const std::vector<float>& params, float threshold,
                                     const std::vector<float>& phiGates)
{
  if (params.size() != ITSCA_MLP_PARAMS) throw std::runtime_error("learned filter: expected 1441 parameters");
  if (!phiGates.empty() && phiGates.size() != ITSCA_TRACKLET_LAYERS)
    throw std::runtime_error("learned filter: expected one phi gate per tracklet layer");
  check(mCtx, itsca_mlp_set_params(mCtx, params.data()), "itsca_mlp_set_params");
  check(mCtx, itsca_set_filter_gates(mCtx, phiGates.empty() ? nullptr : phiGates.data()), "itsca_set_filter_gates");
  check(mCtx, itsca_set_filter(mCtx, 1, threshold), "itsca_set_filter");
}

void Tracker<true>::disableLearnedFilter() { check(mCtx, itsca_set_filter(mCtx, 0, 0.0f), "itsca_set_filter"); }

void Tracker<true>::runBatch(const std::vector<const Event*>& events)
{
  check(mCtx, itsca_clear_events(mCtx), "itsca_clear_events");
  std::vector<float> x, y, z, v;
  std::vector<int32_t> id, mc;
  for (const Event* ev : events) {
    x.clear(), y.clear(), z.clear(), id.clear(), mc.clear(), v.clear();
    int32_t offs[8];
    offs[0] = 0;
    for (int l = 0; l < Constants::ITS::LayersNumber; ++l) {
      for (const Cluster& c : ev->getLayer(l).getClusters()) {
        x.push_back(c.xCoordinate);
        y.push_back(c.yCoordinate);
        z.push_back(c.zCoordinate);
        id.push_back(c.clusterId);
        mc.push_back(c.monteCarloId);
      }
      offs[l + 1] = (int32_t)x.size();
    }
    for (int k = 0; k < ev->getPrimaryVerticesNum(); ++k) {
      const float3& p = ev->getPrimaryVertex(k);
      v.insert(v.end(), {p.x, p.y, p.z});
    }
    check(mCtx,
          itsca_add_event(mCtx, x.data(), y.data(), z.data(), id.data(), mc.data(), offs, v.data(),
                          ev->getPrimaryVerticesNum(), nullptr),
          "itsca_add_event");
  }
  check(mCtx, itsca_upload(mCtx), "itsca_upload");
  check(mCtx, itsca_run(mCtx), "itsca_run");
}

std::vector<Road> Tracker<true>::downloadRoads(int item) const
{
  itsca_counts n;
  check(mCtx, itsca_get_counts(mCtx, item, &n), "itsca_get_counts");
  const std::vector<int32_t> rows = download<int32_t>(mCtx, item, ITSCA_ROADS, 0, n.roads, 8);
  std::vector<Road> roads;
  roads.reserve(n.roads);
  for (int i = 0; i < n.roads; ++i) {
    Road r;
    for (int k = 0; k < Constants::ITS::CellsPerRoad; ++k)
      if (rows[(size_t)i * 8 + k] >= 0) r.addCell(k, rows[(size_t)i * 8 + k]);
    r.setLabel(rows[(size_t)i * 8 + 6]);
    r.setFakeRoad(rows[(size_t)i * 8 + 7] != 0);
    roads.push_back(r);
  }
  return roads;
}

std::vector<std::vector<Road>> Tracker<true>::allRoads() const
{
  const int items = itsca_num_items(mCtx);
  std::vector<int32_t> first(items + 1);
  size_t total = 0;
  const itsca_status st = itsca_download_roads(mCtx, nullptr, 0, first.data(), &total);
  if (st != ITSCA_OK && total == 0) check(mCtx, st, "itsca_download_roads");
  std::vector<int32_t> rows(total * 8 + 1);
  check(mCtx, itsca_download_roads(mCtx, rows.data(), total, first.data(), &total), "itsca_download_roads");
  std::vector<std::vector<Road>> out(items);
  for (int w = 0; w < items; ++w) {
    out[w].reserve(first[w + 1] - first[w]);
    for (int i = first[w]; i < first[w + 1]; ++i) {
      const int32_t* r = &rows[(size_t)i * 8];
      Road road;
      for (int k = 0; k < Constants::ITS::CellsPerRoad; ++k)
        if (r[k] >= 0) road.addCell(k, r[k]);
      road.setLabel(r[6]);
      road.setFakeRoad(r[7] != 0);
      out[w].push_back(road);
    }
  }
  return out;
}

std::vector<std::vector<Road>> Tracker<true>::clustersToTracks(const Event& event)
{
  runBatch({&event});
  return allRoads();
}

std::vector<std::vector<std::vector<Road>>> Tracker<true>::clustersToTracks(const std::vector<Event>& events)
{
  std::vector<const Event*> ptrs;
  for (const Event& e : events) ptrs.push_back(&e);
  runBatch(ptrs);
  std::vector<std::vector<Road>> items = allRoads();
  std::vector<std::vector<std::vector<Road>>> out(events.size());
  size_t item = 0;
  for (size_t e = 0; e < events.size(); ++e)
    for (int v = 0; v < events[e].getPrimaryVerticesNum(); ++v) out[e].push_back(std::move(items[item++]));
  return out;
}

void Tracker<true>::runVertexProfiled(const Event& event, int v, float ms[5])
{
  Event single(event.getEventId());
  single.addPrimaryVertex(event.getPrimaryVertex(v).x, event.getPrimaryVertex(v).y, event.getPrimaryVertex(v).z);
  for (int l = 0; l < Constants::ITS::LayersNumber; ++l)
    for (const Cluster& c : event.getLayer(l).getClusters())
      single.pushClusterToLayer(l, c.clusterId, c.xCoordinate, c.yCoordinate, c.zCoordinate, c.alphaAngle,
                                c.monteCarloId);
  check(mCtx, itsca_set_profiling(mCtx, 1), "itsca_set_profiling");
  try {
    runBatch({&single});
    check(mCtx, itsca_stage_times(mCtx, ms), "itsca_stage_times");
  } catch (...) {
    (void)itsca_set_profiling(mCtx, 0);
    throw;
  }
  check(mCtx, itsca_set_profiling(mCtx, 0), "itsca_set_profiling");
}

std::vector<std::vector<Road>> Tracker<true>::clustersToTracksTimeBenchmark(const Event& event, std::ofstream& out)
{
  std::vector<std::vector<Road>> roads;
  for (int v = 0; v < event.getPrimaryVerticesNum(); ++v) {
    float ms[5];
    runVertexProfiled(event, v, ms);
    float total = 0;
    for (float t : ms) {
      out << t << "\t";
      total += t;
    }
    out << total << std::endl;
    roads.push_back(downloadRoads(0));
  }
  return roads;
}

std::vector<std::vector<Road>> Tracker<true>::clustersToTracksVerbose(const Event& event)
{
  static const char* const kTasks[] = {"Tracklets Finding", "Cells Finding", "Neighbours Finding", "Tracks Finding",
                                       "Computing Montecarlo Labels"};
  std::vector<std::vector<Road>> roads;
  for (int v = 0; v < event.getPrimaryVerticesNum(); ++v) {
    float ms[5];
    runVertexProfiled(event, v, ms);
    std::cout << std::setw(2) << " - Context initialized in: " << ms[0] << "ms" << std::endl;
    const float task[5] = {ms[1], ms[2], ms[3], ms[4], 0.0f};  // labels: inside the roads kernel
    float total = ms[0];
    for (int k = 0; k < 5; ++k) {
      std::cout << std::setw(2) << " - " << kTasks[k] << " completed in: " << task[k] << "ms" << std::endl;
      total += task[k];
    }
    std::cout << std::setw(2) << " - Vertex " << v + 1 << " completed in: " << total << "ms" << std::endl;
    roads.push_back(downloadRoads(0));
  }
  return roads;
}

std::vector<std::vector<Road>> Tracker<true>::clustersToTracksMemoryBenchmark(const Event& event,
                                                                              std::ofstream& out)
{
  std::vector<std::vector<Road>> roads;
  for (int v = 0; v < event.getPrimaryVerticesNum(); ++v) {
    float ms[5];
    runVertexProfiled(event, v, ms);
    itsca_counts n;
    check(mCtx, itsca_get_counts(mCtx, 0, &n), "itsca_get_counts");
    itsca_capacities cap;
    check(mCtx, itsca_get_capacities(mCtx, &cap), "itsca_get_capacities");
    for (int l = 0; l < Constants::ITS::LayersNumber; ++l) out << n.clusters[l] << "\t";
    out << std::endl;
    for (int l = 0; l < Constants::ITS::CellsPerRoad; ++l) out << cap.cells << "\t";
    out << std::endl;
    for (int l = 0; l < Constants::ITS::CellsPerRoad; ++l) out << n.cells[l] << "\t";
    out << std::endl;
    out << n.roads << std::endl;
    roads.push_back(downloadRoads(0));
  }
  return roads;
}

PrimaryVertexContext Tracker<true>::getPrimaryVertexContext(int item) const
{
  itsca_counts n;
  check(mCtx, itsca_get_counts(mCtx, item, &n), "itsca_get_counts");
  PrimaryVertexContext ctx;
  for (int l = 0; l < Constants::ITS::LayersNumber; ++l) {
    const auto rows = download<int32_t>(mCtx, item, ITSCA_CLUSTERS, l, n.clusters[l], 8);
    for (int i = 0; i < n.clusters[l]; ++i) {
      const int32_t* r = &rows[(size_t)i * 8];
      Cluster c(r[0], l, asFloat(r[4]), asFloat(r[5]), asFloat(r[6]), 0.0f, r[7]);
      c.phiCoordinate = asFloat(r[1]);
      c.rCoordinate = asFloat(r[2]);
      c.indexTableBinIndex = r[3];
      ctx.mClusters[l].push_back(c);
    }
  }
  for (int l = 1; l < Constants::ITS::LayersNumber; ++l) {
    const auto t = download<int32_t>(mCtx, item, ITSCA_INDEX_TABLE, l, PrimaryVertexContext::TableSize, 1);
    for (int b = 0; b < PrimaryVertexContext::TableSize; ++b) ctx.mIndexTables[l - 1][b] = t[b];
  }
  for (int l = 0; l < Constants::ITS::TrackletsPerRoad; ++l) {
    const auto rows = download<int32_t>(mCtx, item, ITSCA_TRACKLETS, l, n.tracklets[l], 4);
    for (int i = 0; i < n.tracklets[l]; ++i)
      ctx.mTracklets[l].push_back(
          {rows[(size_t)i * 4], rows[(size_t)i * 4 + 1], asFloat(rows[(size_t)i * 4 + 2]), asFloat(rows[(size_t)i * 4 + 3])});
  }
  for (int l = 0; l < Constants::ITS::CellsPerRoad; ++l)
    ctx.mTrackletsLookupTable[l] = download<int>(mCtx, item, ITSCA_TRACKLET_LOOKUP, l, n.clusters[l + 1], 1);
  for (int l = 0; l < Constants::ITS::CellsPerRoad; ++l) {
    const auto rows = download<int32_t>(mCtx, item, ITSCA_CELLS, l, n.cells[l], 10);
    for (int i = 0; i < n.cells[l]; ++i) {
      const int32_t* r = &rows[(size_t)i * 10];
      ctx.mCells[l].emplace_back(r[0], r[1], r[2], r[3], r[4], float3{asFloat(r[5]), asFloat(r[6]), asFloat(r[7])},
                                 asFloat(r[8]), r[9]);
    }
  }
  for (int l = 0; l < Constants::ITS::CellsPerRoad - 1; ++l) {
    ctx.mCellsLookupTable[l] = download<int>(mCtx, item, ITSCA_CELL_LOOKUP, l, n.tracklets[l + 1], 1);
    const auto off = download<int>(mCtx, item, ITSCA_NEIGHBOUR_OFFSETS, l, n.cells[l + 1] + 1, 1);
    const auto lst = download<int>(mCtx, item, ITSCA_NEIGHBOURS, l, n.neighbours[l], 1);
    ctx.mCellsNeighbours[l].resize(n.cells[l + 1]);
    for (int c = 0; c < n.cells[l + 1]; ++c)
      ctx.mCellsNeighbours[l][c].assign(lst.begin() + off[c], lst.begin() + off[c + 1]);
  }
  ctx.mRoads = downloadRoads(item);
  return ctx;
}

namespace IOUtils {

// IOUtils.cxx:38-78; also reads the binary side-car (itsca::writeEventsBinary).
std::vector<Event> loadEventData(const std::string& fileName)
{
  std::vector<Event> events;
  if (itsca::isEventsBinary(fileName)) {
    itsca::EventBatch batch;
    std::string err;
    if (!itsca::loadEventsBinary(fileName, batch, err)) return events;
    for (const itsca::RawEvent& raw : batch.events) {
      events.emplace_back((int)events.size());
      Event& ev = events.back();
      for (int v = 0; v < raw.numVertices(); ++v)
        ev.addPrimaryVertex(raw.vertices[3 * v], raw.vertices[3 * v + 1], raw.vertices[3 * v + 2]);
      for (int l = 0; l < Constants::ITS::LayersNumber; ++l)
        for (int i = 0; i < raw.layerSize(l); ++i)
          ev.pushClusterToLayer(l, raw.id[l][i], raw.x[l][i], raw.y[l][i], raw.z[l][i], raw.alpha[l][i], raw.mc[l][i]);
    }
    return events;
  }
  std::ifstream in(fileName);
  std::string line, unused;
  int clusterId = -1;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    int layer;
    float x, y, z, alpha;
    int mc;
    if (!(ss >> layer >> x >> y >> z)) continue;
    if (layer == -1) {
      if (clusterId != 0) events.emplace_back((int)events.size());
      events.back().addPrimaryVertex(x, y, z);
      clusterId = 0;
    } else if (ss >> unused >> unused >> unused >> alpha >> mc) {
      if (events.empty() || layer < 0 || layer >= Constants::ITS::LayersNumber) continue;
      events.back().pushClusterToLayer(layer, clusterId, x, y, z, alpha, mc);
      ++clusterId;
    }
  }
  return events;
}

// IOUtils.cxx:80-123 (pions with 7 clusters only)
std::vector<std::unordered_map<int, Label>> loadLabels(int eventsNum, const std::string& fileName)
{
  std::vector<std::unordered_map<int, Label>> labels;
  labels.reserve(eventsNum);
  std::unordered_map<int, Label> current;
  std::ifstream in(fileName);
  std::string line;
  std::getline(in, line);
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    int mc, pdg, n;
    float pt, phi, eta;
    if (!(ss >> mc)) continue;
    if (mc == -1) {
      labels.push_back(current);
      current.clear();
    } else if (ss >> pt >> phi >> eta >> pdg >> n) {
      if (std::abs(pdg) == Constants::PDGCodes::PionCode && n == 7)
        current.emplace(std::piecewise_construct, std::forward_as_tuple(mc),
                        std::forward_as_tuple(mc, pt, phi, eta, pdg, n));
    }
  }
  labels.push_back(current);
  return labels;
}

// IOUtils.cxx:125-170 (correct / duplicate / fake classification)
void writeRoadsReport(std::ofstream& correct, std::ofstream& duplicate, std::ofstream& fake,
                      const std::vector<std::vector<Road>>& roads, const std::unordered_map<int, Label>& labels)
{
  itsca::RoadClassifier classify;
  std::ofstream* out[3] = {&correct, &duplicate, &fake};
  correct << -1 << std::endl;
  fake << -1 << std::endl;
  for (const std::vector<Road>& vertexRoads : roads) {
    for (const Road& road : vertexRoads) {
      const auto it = labels.find(road.getLabel());
      const int cls = classify.next(road.getLabel(), road.isFakeRoad(), [&](int, int& mc) {
        if (it == labels.end()) return false;
        mc = it->second.monteCarloId;
        return true;
      });
      if (cls >= 0) *out[cls] << it->second << std::endl;
    }
  }
}

}  // namespace IOUtils

}  // namespace CA
}  // namespace ITS
}  // namespace o2
