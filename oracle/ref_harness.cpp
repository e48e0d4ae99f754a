// ref_harness.cpp -- drives the reference CPU tracker, compiled from its own
// sources under the reference checkout, and writes its per-vertex state in the
// oracle's canonical dump format.
//
// TEST INFRASTRUCTURE ONLY: used to pin the CPU oracle (and through it the
// gfx950 path) to the reference.  Built by oracle/Makefile into oracle/_ref/.
// The translation unit is compiled with -Dprivate=public -Dprotected=public so
// that it can step through Tracker's protected stages and read its context,
// exactly as Tracker::clustersToTracks does (src/CA/Tracker.cxx:244-265).
//
// usage: ref_harness dump <events.txt> <out_prefix>
//          writes <out_prefix>_e<E>_v<V>.txt per (event, vertex) and prints
//          per-vertex diagnostics (over-read of the next layer, lookup-table
//          estimate vs actual)
//        ref_harness time <events.txt> <repeats>
//          times the same per-vertex sequence per event (after one warm-up run)
//        ref_harness report <events.txt> <labels.txt> <dir>
//          writes the reference's road report files into <dir>
//
// The reference's tracklet search reads one element past the end of the next
// layer's cluster vector whenever the window's last bin is at or after the
// layer's last occupied bin (inclusive bound, src/CA/Tracker.cxx:85-91).  That
// read is undefined and crashes the optimised build on ordinary events.  After
// PrimaryVertexContext::initialize the harness therefore reserves one more
// element per layer and constructs a sentinel cluster there whose z is NaN:
// every cut against it is false, so the read becomes defined and behaves
// exactly like a "< N" guard on the candidate loop.
//
// The reference also sizes each cell lookup table from fitted memory
// coefficients (src/CA/PrimaryVertexContext.cxx:68-74) and then writes it at
// every tracklet index (src/CA/Tracker.cxx:220-225); on small events the
// estimate is far below the tracklet count and the write is out of bounds.
// After the tracklet stage the harness grows each table to the actual
// tracklet count, filled with UnusedIndex, which is what an exactly sized
// table holds.  Each such growth is reported.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <unordered_map>
#include <string>
#include <vector>

#include "ITSReconstruction/CA/Constants.h"
#include "ITSReconstruction/CA/Event.h"
#include "ITSReconstruction/CA/IOUtils.h"
#include "ITSReconstruction/CA/Tracker.h"
#include "ca_oracle.hpp"

using namespace o2::ITS::CA;

namespace {

// PrimaryVertexContext::initialize followed by the sentinel described above,
// then the stages of Tracker::clustersToTracks (src/CA/Tracker.cxx:253-259).
void trackVertex(Tracker<false>& tracker, const Event& ev, int v, std::string* notes = nullptr)
{
  PrimaryVertexContext& ctx = tracker.mPrimaryVertexContext;
  ctx.initialize(ev, v);
  for (int l = 0; l < 7; ++l) {
    std::vector<Cluster>& cl = ctx.getClusters()[l];
    cl.reserve(cl.size() + 1);
    new (cl.data() + cl.size()) Cluster(-1, l, NAN, NAN, NAN, 0.0f, -1);
  }
  tracker.computeTracklets();
  for (int l = 0; l < 4; ++l) {
    std::vector<int>& lk = ctx.getCellsLookupTable()[l];
    const size_t need = ctx.getTracklets()[l + 1].size();
    if (lk.size() < need) {
      if (notes) {
        char buf[160];
        std::snprintf(buf, sizeof buf, " cell-lookup[%d] estimate %zu < tracklets %zu;", l, lk.size(), need);
        *notes += buf;
      }
      lk.resize(need, Constants::ITS::UnusedIndex);
    }
  }
  tracker.computeCells();
  tracker.findCellsNeighbours();
  tracker.findTracks();
  tracker.computeMontecarloLabels();
}

void convert(Tracker<false>& tracker, itsca_oracle::VertexResult& out, int& overread)
{
  PrimaryVertexContext& ctx = tracker.mPrimaryVertexContext;
  const float3& v = ctx.getPrimaryVertex();
  out.vx = v.x;
  out.vy = v.y;
  out.vz = v.z;
  for (int l = 0; l < 7; ++l) {
    out.clusters[l].clear();
    for (const Cluster& c : ctx.getClusters()[l]) {
      itsca_oracle::Cluster o;
      o.x = c.xCoordinate;
      o.y = c.yCoordinate;
      o.z = c.zCoordinate;
      o.phi = c.phiCoordinate;
      o.r = c.rCoordinate;
      o.id = c.clusterId;
      o.alpha = c.alphaAngle;
      o.mc = c.monteCarloId;
      o.bin = c.indexTableBinIndex;
      out.clusters[l].push_back(o);
    }
  }
  for (int l = 0; l < 6; ++l)
    for (int b = 0; b < 401; ++b) out.tables[l][b] = ctx.getIndexTables()[l][b];
  overread = 0;
  for (int l = 0; l < 6; ++l) {
    out.tracklets[l].clear();
    const int nNext = (int)ctx.getClusters()[l + 1].size();
    for (const Tracklet& t : ctx.getTracklets()[l]) {
      if (t.secondClusterIndex >= nNext) ++overread;
      out.tracklets[l].push_back({t.firstClusterIndex, t.secondClusterIndex, t.tanLambda, t.phiCoordinate});
    }
  }
  for (int l = 0; l < 5; ++l) out.trackletLookup[l] = ctx.getTrackletsLookupTable()[l];
  for (int l = 0; l < 5; ++l) {
    out.cells[l].clear();
    for (const Cell& c : ctx.getCells()[l]) {
      const float3& n = c.getNormalVectorCoordinates();
      out.cells[l].push_back({c.getFirstClusterIndex(), c.getSecondClusterIndex(), c.getThirdClusterIndex(),
                              c.getFirstTrackletIndex(), c.getSecondTrackletIndex(), n.x, n.y, n.z,
                              c.getCurvature(), c.getLevel()});
    }
  }
  for (int l = 0; l < 4; ++l) {
    const std::vector<int>& lk = ctx.getCellsLookupTable()[l];
    const size_t need = ctx.getTracklets()[l + 1].size();
    out.cellLookup[l].assign(lk.begin(), lk.begin() + std::min(lk.size(), need));
    out.neighbours[l] = ctx.getCellsNeighbours()[l];
  }
  out.roads.clear();
  for (Road& rd : ctx.getRoads()) {
    itsca_oracle::Road o;
    for (int k = 0; k < 5; ++k) o.cells[k] = rd[k];
    o.size = rd.getRoadSize();
    o.label = rd.getLabel();
    o.fake = rd.isFakeRoad();
    out.roads.push_back(o);
  }
}

int dumpMode(const char* events, const char* prefix)
{
  std::vector<Event> evs = IOUtils::loadEventData(events);
  Tracker<false> tracker{};
  itsca_oracle::VertexResult res;
  for (size_t e = 0; e < evs.size(); ++e) {
    for (int v = 0; v < evs[e].getPrimaryVerticesNum(); ++v) {
      std::string notes;
      trackVertex(tracker, evs[e], v, &notes);
      int overread = 0;
      convert(tracker, res, overread);
      char path[4096];
      std::snprintf(path, sizeof path, "%s_e%zu_v%d.txt", prefix, e, v);
      FILE* f = std::fopen(path, "w");
      if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path);
        return 1;
      }
      const std::string s = itsca_oracle::dump(res);
      std::fwrite(s.data(), 1, s.size(), f);
      std::fclose(f);
      std::printf("event %zu vertex %d: tracklets", e, v);
      for (int l = 0; l < 6; ++l) std::printf(" %zu", res.tracklets[l].size());
      std::printf(" cells");
      for (int l = 0; l < 5; ++l) std::printf(" %zu", res.cells[l].size());
      std::printf(" roads %zu overread %d%s\n", res.roads.size(), overread, notes.c_str());
    }
  }
  return 0;
}

int timeMode(const char* events, int repeats)
{
  std::vector<Event> evs = IOUtils::loadEventData(events);
  if (evs.empty()) return 1;
  Tracker<false> tracker{};
  for (int v = 0; v < evs[0].getPrimaryVerticesNum(); ++v) trackVertex(tracker, evs[0], v);  // warm-up
  long vertices = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < repeats; ++r)
    for (Event& ev : evs) {
      for (int v = 0; v < ev.getPrimaryVerticesNum(); ++v) trackVertex(tracker, ev, v);
      vertices += ev.getPrimaryVerticesNum();
    }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("{\"events\": %ld, \"vertices\": %ld, \"seconds\": %.6f, \"ms_per_event\": %.4f}\n",
              (long)evs.size() * repeats, vertices, s, 1e3 * s / ((double)evs.size() * repeats));
  return 0;
}

// The reference's road report (main.cpp:49-61, 130-134): every event's roads
// (Tracker.cxx:244-265, through trackVertex) classified against the labels
// file by IOUtils::loadLabels and IOUtils::writeRoadsReport
// (IOUtils.cxx:80-170), into <dir>/CorrectRoads.txt, DuplicateRoads.txt and
// FakeRoads.txt.
int reportMode(const char* events, const char* labels, const std::string& dir)
{
  std::vector<Event> evs = IOUtils::loadEventData(events);
  std::vector<std::unordered_map<int, Label>> labelsMap = IOUtils::loadLabels((int)evs.size(), labels);
  std::ofstream correct(dir + "/CorrectRoads.txt"), duplicate(dir + "/DuplicateRoads.txt"),
      fake(dir + "/FakeRoads.txt");
  Tracker<false> tracker{};
  for (size_t e = 0; e < evs.size(); ++e) {
    std::vector<std::vector<Road>> roads;
    for (int v = 0; v < evs[e].getPrimaryVerticesNum(); ++v) {
      trackVertex(tracker, evs[e], v);
      roads.emplace_back(tracker.mPrimaryVertexContext.getRoads());
    }
    if (e < labelsMap.size()) IOUtils::writeRoadsReport(correct, duplicate, fake, roads, labelsMap[e]);
    std::printf("event %zu:", e);
    for (const auto& r : roads) std::printf(" %zu", r.size());
    std::printf(" roads\n");
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc >= 5 && !std::strcmp(argv[1], "report")) return reportMode(argv[2], argv[3], argv[4]);
  if (argc >= 4 && !std::strcmp(argv[1], "dump")) return dumpMode(argv[2], argv[3]);
  if (argc >= 4 && !std::strcmp(argv[1], "time")) return timeMode(argv[2], std::atoi(argv[3]));
  std::fprintf(stderr, "usage: ref_harness dump <events> <prefix> | time <events> <repeats> | "
                       "report <events> <labels> <dir>\n");
  return 2;
}
