// ca_oracle.cpp -- CPU restatement of the reference plain CA.
//
// TEST INFRASTRUCTURE ONLY (see ca_oracle.hpp).  Build flags must match the
// reference CPU build: -O3, no -march (SSE2, no FMA), no fast-math, so every
// float operation below rounds exactly like the reference's.
#include "ca_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace itsca_oracle {
namespace {

// Constants.h:31-33
constexpr float Pi = 3.14159265359f;
constexpr float TwoPi = 2.0f * Pi;
constexpr float FloatMinThreshold = 1e-20f;
// Constants.h:43-50
constexpr float LayersZ[kLayers] = {16.333f, 16.333f, 16.333f, 42.140f, 42.140f, 73.745f, 73.745f};
constexpr float LayersR[kLayers] = {2.33959f, 3.14076f, 3.91924f, 19.6213f, 24.5597f, 34.388f, 39.3329f};
// Constants.h:54-72
constexpr float TrackletMaxDeltaZ[kTrackletLayers] = {0.1f, 0.1f, 0.3f, 0.3f, 0.3f, 0.3f};
constexpr float CellMaxDeltaTanLambda = 0.025f;
constexpr float CellMaxDeltaZ[kCellLayers] = {0.2f, 0.4f, 0.5f, 0.6f, 3.0f};
constexpr float CellMaxDCA[kCellLayers] = {0.05f, 0.04f, 0.05f, 0.2f, 0.4f};
constexpr float CellMaxDeltaPhi = 0.14f;
constexpr float ZCoordinateCut = 0.5f;
constexpr float PhiCoordinateCut = 0.3f;
constexpr float NeighbourMaxNormalDelta[kCellLayers - 1] = {0.002f, 0.009f, 0.002f, 0.005f};
constexpr float NeighbourMaxCurvatureDelta[kCellLayers - 1] = {0.008f, 0.0025f, 0.003f, 0.0035f};
constexpr int CellsMinLevel = 5;
// Constants.h:77-85
constexpr int ZBins = 20;
constexpr int PhiBins = 20;
constexpr float InversePhiBinSize = PhiBins / TwoPi;
const float InverseZBinSize[kLayers] = {
    (float)(0.5 * ZBins / 16.333f), (float)(0.5 * ZBins / 16.333f), (float)(0.5 * ZBins / 16.333f),
    (float)(0.5 * ZBins / 42.140f), (float)(0.5 * ZBins / 42.140f), (float)(0.5 * ZBins / 73.745f),
    (float)(0.5 * ZBins / 73.745f)};
constexpr int UnusedIndex = -1;

// MathUtils.h:47-51
inline float normalizedPhi(float phi)
{
  return (phi < 0) ? phi + TwoPi : (phi > TwoPi) ? phi - TwoPi : phi;
}
// IndexTableUtils.h:47-62
inline int zBin(int layer, float z) { return (int)((z + LayersZ[layer]) * InverseZBinSize[layer]); }
inline int phiBin(float phi) { return (int)(phi * InversePhiBinSize); }
inline int bin(int zIndex, int phiIndex) { return std::min(phiIndex * PhiBins + zIndex, ZBins * PhiBins); }

struct Rect {
  int x, y, z, w;
};

// TrackingUtils.cxx:30-48
Rect binsRect(const Cluster& c, int layer, float dirZ)
{
  const float zMin = dirZ - 2 * ZCoordinateCut;
  const float phiMin = c.phi - PhiCoordinateCut;
  const float zMax = dirZ + 2 * ZCoordinateCut;
  const float phiMax = c.phi + PhiCoordinateCut;
  if (zMax < -LayersZ[layer + 1] || zMin > LayersZ[layer + 1] || zMin > zMax) return Rect{0, 0, 0, 0};
  return Rect{std::max(0, zBin(layer + 1, zMin)), phiBin(normalizedPhi(phiMin)),
              std::min(ZBins - 1, zBin(layer + 1, zMax)), phiBin(normalizedPhi(phiMax))};
}

// PrimaryVertexContext.cxx:31-142 (cluster transform, sort, index tables)
void initialize(const EventView& ev, bool stableSort, VertexResult& out)
{
  for (int l = 0; l < kLayers; ++l) {
    std::vector<Cluster>& cl = out.clusters[l];
    cl.clear();
    for (int i = ev.layerOffsets[l]; i < ev.layerOffsets[l + 1]; ++i) {
      Cluster c;
      c.x = ev.x[i];
      c.y = ev.y[i];
      c.z = ev.z[i];
      const float dx = c.x - ev.vx, dy = c.y - ev.vy;
      // Cluster.cxx:37-42; MathUtils.h:37-45
      c.phi = normalizedPhi(std::atan2(-dy, -dx) + Pi);
      c.r = std::sqrt(dx * dx + dy * dy);
      c.id = ev.id[i];
      c.alpha = 0.0f;
      c.mc = ev.mc[i];
      c.bin = bin(zBin(l, c.z), phiBin(c.phi));  // absolute z, not vertex-relative
      cl.push_back(c);
    }
    auto byBin = [](const Cluster& a, const Cluster& b) { return a.bin < b.bin; };
    if (stableSort)
      std::stable_sort(cl.begin(), cl.end(), byBin);
    else
      std::sort(cl.begin(), cl.end(), byBin);
  }
  for (int l = 1; l < kLayers; ++l) {
    std::array<int, kTableSize>& table = out.tables[l - 1];
    const int n = (int)out.clusters[l].size();
    int previous = 0;
    table[0] = 0;
    for (int i = 0; i < n; ++i) {
      const int b = out.clusters[l][i].bin;
      if (b > previous) {
        for (int k = previous + 1; k <= b; ++k) table[k] = i;
        previous = b;
      }
    }
    for (int k = previous + 1; k <= ZBins * PhiBins; ++k) table[k] = n;
  }
  for (int l = 0; l < kCellLayers; ++l) out.trackletLookup[l].assign(out.clusters[l + 1].size(), UnusedIndex);
  for (int l = 0; l < kTrackletLayers; ++l) out.tracklets[l].clear();
  for (int l = 0; l < kCellLayers; ++l) out.cells[l].clear();
  for (int l = 0; l < kCellLayers - 1; ++l) {
    out.cellLookup[l].clear();
    out.neighbours[l].clear();
  }
  out.roads.clear();
}

// Tracker.cxx:44-117 and Tracklet.cxx:32-40
void computeTracklets(VertexResult& ctx, bool exclusiveBound)
{
  for (int l = 0; l < kTrackletLayers; ++l) {
    if (ctx.clusters[l].empty() || ctx.clusters[l + 1].empty()) return;  // early return, not continue
    const std::vector<Cluster>& cur = ctx.clusters[l];
    const std::vector<Cluster>& next = ctx.clusters[l + 1];
    const std::array<int, kTableSize>& table = ctx.tables[l];
    const int nNext = (int)next.size();
    for (int i = 0; i < (int)cur.size(); ++i) {
      const Cluster& c = cur[i];
      const float tanLambda = (c.z - ctx.vz) / c.r;
      const float dirZ = tanLambda * (LayersR[l + 1] - c.r) + c.z;
      const Rect rect = binsRect(c, l, dirZ);
      if (rect.x == 0 && rect.y == 0 && rect.z == 0 && rect.w == 0) continue;
      int phiBinsNum = rect.w - rect.y + 1;
      if (phiBinsNum < 0) phiBinsNum += PhiBins;
      for (int iPhi = rect.y, count = 0; count < phiBinsNum; ++count) {
        const int firstBin = bin(rect.x, iPhi);
        const int maxBin = firstBin + rect.z - rect.x + 1;
        const int firstRow = table[firstBin];
        const int maxRow = table[maxBin];
        // Inclusive upper bound as in the reference (Tracker.cxx:88); the
        // "< nNext" guard stops the one read past the end of the layer.
        const int lastRow = exclusiveBound ? maxRow - 1 : maxRow;
        for (int j = firstRow; j <= lastRow && j < nNext; ++j) {
          const Cluster& n = next[j];
          const float deltaZ = std::abs(tanLambda * (n.r - c.r) + c.z - n.z);
          const float deltaPhi = std::abs(c.phi - n.phi);
          if (deltaZ < TrackletMaxDeltaZ[l] &&
              (deltaPhi < PhiCoordinateCut || std::abs(deltaPhi - TwoPi) < PhiCoordinateCut)) {
            if (l > 0 && ctx.trackletLookup[l - 1][i] == UnusedIndex)
              ctx.trackletLookup[l - 1][i] = (int)ctx.tracklets[l].size();
            Tracklet t;
            t.first = i;
            t.second = j;
            t.tanLambda = (c.z - n.z) / (c.r - n.r);
            t.phi = std::atan2(c.y - n.y, c.x - n.x);
            ctx.tracklets[l].push_back(t);
          }
        }
        ++iPhi;
        if (iPhi == PhiBins) iPhi = 0;
      }
    }
  }
}

// Tracker.cxx:119-235, Cell.cxx:24-32, MathUtils.h:53-60
void computeCells(VertexResult& ctx)
{
  for (int l = 0; l < kCellLayers; ++l) {
    if (ctx.tracklets[l + 1].empty() || ctx.tracklets[l].empty()) return;  // early return
    if (l > 0) ctx.cellLookup[l - 1].assign(ctx.tracklets[l].size(), UnusedIndex);
    const std::vector<Tracklet>& cur = ctx.tracklets[l];
    const std::vector<Tracklet>& next = ctx.tracklets[l + 1];
    const int nNext = (int)next.size();
    for (int it = 0; it < (int)cur.size(); ++it) {
      const Tracklet& t = cur[it];
      const int k = t.second;
      const int firstNext = ctx.trackletLookup[l][k];
      if (firstNext == UnusedIndex) continue;
      const Cluster& c1 = ctx.clusters[l][t.first];
      const Cluster& c2 = ctx.clusters[l + 1][t.second];
      const float r1sq = c1.r * c1.r;
      const float r2sq = c2.r * c2.r;
      const float d1x = c2.x - c1.x, d1y = c2.y - c1.y, d1z = r2sq - r1sq;
      for (int iu = firstNext; iu < nNext && next[iu].first == k; ++iu) {
        const Tracklet& u = next[iu];
        const float deltaTanLambda = std::abs(t.tanLambda - u.tanLambda);
        const float deltaPhi = std::abs(t.phi - u.phi);
        if (!(deltaTanLambda < CellMaxDeltaTanLambda &&
              (deltaPhi < CellMaxDeltaPhi || std::abs(deltaPhi - TwoPi) < CellMaxDeltaPhi)))
          continue;
        const float avgTanLambda = 0.5f * (t.tanLambda + u.tanLambda);
        const float dirZ = -avgTanLambda * c1.r + c1.z;
        const float deltaZ = std::abs(dirZ - ctx.vz);
        if (!(deltaZ < CellMaxDeltaZ[l])) continue;
        const Cluster& c3 = ctx.clusters[l + 2][u.second];
        const float r3sq = c3.r * c3.r;
        const float d2x = c3.x - c1.x, d2y = c3.y - c1.y, d2z = r3sq - r1sq;
        const float nx = (d1y * d2z) - (d1z * d2y);
        const float ny = (d1z * d2x) - (d1x * d2z);
        const float nz = (d1x * d2y) - (d1y * d2x);
        const float norm = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (norm < FloatMinThreshold || std::abs(nz) < FloatMinThreshold) continue;
        const float inv = 1.0f / norm;
        const float hx = nx * inv, hy = ny * inv, hz = nz * inv;
        const float planeDistance = -hx * (c2.x - ctx.vx) - (hy * c2.y - ctx.vy) - hz * r2sq;
        const float hz2 = hz * hz;
        const float radius = std::sqrt((1.0f - hz2 - 4.0f * planeDistance * hz) / (4.0f * hz2));
        const float cx = -0.5f * hx / hz, cy = -0.5f * hy / hz;
        const float dca = std::abs(radius - std::sqrt(cx * cx + cy * cy));
        if (dca > CellMaxDCA[l]) continue;  // a NaN DCA is accepted, as in the reference
        const float curvature = 1.0f / radius;
        if (l > 0 && ctx.cellLookup[l - 1][it] == UnusedIndex) ctx.cellLookup[l - 1][it] = (int)ctx.cells[l].size();
        Cell cell;
        cell.c1 = t.first;
        cell.c2 = u.first;
        cell.c3 = u.second;
        cell.t1 = it;
        cell.t2 = iu;
        cell.nx = hx;
        cell.ny = hy;
        cell.nz = hz;
        cell.curvature = curvature;
        cell.level = 1;
        ctx.cells[l].push_back(cell);
      }
    }
  }
}

// Tracker.cxx:416-474
void findNeighbours(VertexResult& ctx)
{
  for (int l = 0; l < kCellLayers - 1; ++l) {
    ctx.neighbours[l].clear();
    if (ctx.cells[l + 1].empty() || ctx.cellLookup[l].empty()) continue;
    const int nNext = (int)ctx.cells[l + 1].size();
    for (int ic = 0; ic < (int)ctx.cells[l].size(); ++ic) {
      const Cell& c = ctx.cells[l][ic];
      const int u = c.t2;
      const int first = ctx.cellLookup[l][u];
      if (first == UnusedIndex || ctx.cells[l + 1][first].t1 != u) continue;
      ctx.neighbours[l].resize(nNext);
      for (int in = first; in < nNext && ctx.cells[l + 1][in].t1 == u; ++in) {
        Cell& n = ctx.cells[l + 1][in];
        const float dx = c.nx - n.nx, dy = c.ny - n.ny, dz = c.nz - n.nz;
        const float deltaNormal = (dx * dx) + (dy * dy) + (dz * dz);
        const float deltaCurvature = std::abs(c.curvature - n.curvature);
        if (deltaNormal < NeighbourMaxNormalDelta[l] && deltaCurvature < NeighbourMaxCurvatureDelta[l]) {
          ctx.neighbours[l][in].push_back(ic);
          if (c.level >= n.level) n.level = c.level + 1;
        }
      }
    }
  }
}

void newRoad(VertexResult& ctx, int layer, int cell)
{
  Road r;
  for (int k = 0; k < kCellLayers; ++k) r.cells[k] = UnusedIndex;
  r.size = 0;
  r.label = UnusedIndex;
  r.fake = false;
  ctx.roads.push_back(r);
  // Road::addCell, Road.cxx:42-50
  ctx.roads.back().cells[layer] = cell;
  ctx.roads.back().size = 1;
}

void addCell(Road& r, int layer, int cell)
{
  if (r.cells[layer] == UnusedIndex) ++r.size;
  r.cells[layer] = cell;
}

// Tracker.cxx:531-571
void traverse(VertexResult& ctx, int cellId, int layer)
{
  const int level = ctx.cells[layer][cellId].level;
  addCell(ctx.roads.back(), layer, cellId);
  if (layer > 0) {
    const std::vector<int>& preds = ctx.neighbours[layer - 1][cellId];
    bool first = true;
    for (int p : preds) {
      if (level - 1 != ctx.cells[layer - 1][p].level) continue;
      if (first)
        first = false;
      else
        ctx.roads.push_back(ctx.roads.back());
      traverse(ctx, p, layer - 1);
    }
  }
}

// Tracker.cxx:476-529
void findTracks(VertexResult& ctx)
{
  for (int level = kCellLayers; level >= CellsMinLevel; --level) {
    const int minimumLevel = level - 1;
    for (int layer = kCellLayers - 1; layer >= minimumLevel; --layer) {
      for (int ic = 0; ic < (int)ctx.cells[layer].size(); ++ic) {
        if (ctx.cells[layer][ic].level != level) continue;
        newRoad(ctx, layer, ic);
        const std::vector<int>& preds = ctx.neighbours[layer - 1][ic];
        bool first = true;
        for (int p : preds) {
          if (level - 1 != ctx.cells[layer - 1][p].level) continue;
          if (first)
            first = false;
          else
            newRoad(ctx, layer, ic);
          traverse(ctx, p, layer - 1);
        }
      }
    }
  }
}

// Tracker.cxx:573-652 (Moore's majority vote; a mismatch anywhere marks fake)
void computeLabels(VertexResult& ctx)
{
  for (Road& road : ctx.roads) {
    int value = UnusedIndex;
    int count = 0;
    bool fake = false;
    bool firstCell = true;
    for (int k = 0; k < kCellLayers; ++k) {
      const int ci = road.cells[k];
      if (ci == UnusedIndex) {
        if (firstCell) continue;
        break;
      }
      const Cell& cell = ctx.cells[k][ci];
      if (firstCell) {
        value = ctx.clusters[k][cell.c1].mc;
        count = 1;
        const int second = ctx.clusters[k + 1][cell.c2].mc;
        if (second == value) {
          ++count;
        } else {
          value = second;
          count = 1;
          fake = true;
        }
        firstCell = false;
      }
      const int current = ctx.clusters[k + 2][cell.c3].mc;
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
    road.label = value;
    road.fake = fake;
  }
}

void hexf(std::string& s, float f)
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  char buf[16];
  std::snprintf(buf, sizeof buf, "%08x", u);
  s += buf;
}

void ints(std::string& s, const std::vector<int>& v)
{
  char buf[16];
  for (size_t i = 0; i < v.size(); ++i) {
    std::snprintf(buf, sizeof buf, i ? " %d" : "%d", v[i]);
    s += buf;
  }
  s += "\n";
}

}  // namespace

void runVertex(const EventView& ev, bool stableSort, VertexResult& out, bool exclusiveBound)
{
  out.vx = ev.vx;
  out.vy = ev.vy;
  out.vz = ev.vz;
  initialize(ev, stableSort, out);
  computeTracklets(out, exclusiveBound);
  computeCells(out);
  findNeighbours(out);
  findTracks(out);
  computeLabels(out);
}

std::string dump(const VertexResult& r)
{
  std::string s;
  char buf[160];
  s += "vertex ";
  hexf(s, r.vx);
  s += " ";
  hexf(s, r.vy);
  s += " ";
  hexf(s, r.vz);
  s += "\n";
  for (int l = 0; l < kLayers; ++l) {
    std::snprintf(buf, sizeof buf, "clusters %d %zu\n", l, r.clusters[l].size());
    s += buf;
    for (const Cluster& c : r.clusters[l]) {
      std::snprintf(buf, sizeof buf, "%d ", c.id);
      s += buf;
      hexf(s, c.phi);
      s += " ";
      hexf(s, c.r);
      std::snprintf(buf, sizeof buf, " %d\n", c.bin);
      s += buf;
    }
  }
  for (int l = 1; l < kLayers; ++l) {
    std::snprintf(buf, sizeof buf, "table %d %d\n", l, kTableSize);
    s += buf;
    ints(s, std::vector<int>(r.tables[l - 1].begin(), r.tables[l - 1].end()));
  }
  for (int l = 0; l < kTrackletLayers; ++l) {
    std::snprintf(buf, sizeof buf, "tracklets %d %zu\n", l, r.tracklets[l].size());
    s += buf;
    for (const Tracklet& t : r.tracklets[l]) {
      std::snprintf(buf, sizeof buf, "%d %d ", t.first, t.second);
      s += buf;
      hexf(s, t.tanLambda);
      s += " ";
      hexf(s, t.phi);
      s += "\n";
    }
  }
  for (int l = 0; l < kCellLayers; ++l) {
    std::snprintf(buf, sizeof buf, "tlookup %d %zu\n", l, r.trackletLookup[l].size());
    s += buf;
    ints(s, r.trackletLookup[l]);
  }
  for (int l = 0; l < kCellLayers; ++l) {
    std::snprintf(buf, sizeof buf, "cells %d %zu\n", l, r.cells[l].size());
    s += buf;
    for (const Cell& c : r.cells[l]) {
      std::snprintf(buf, sizeof buf, "%d %d %d %d %d ", c.c1, c.c2, c.c3, c.t1, c.t2);
      s += buf;
      hexf(s, c.nx);
      s += " ";
      hexf(s, c.ny);
      s += " ";
      hexf(s, c.nz);
      s += " ";
      hexf(s, c.curvature);
      std::snprintf(buf, sizeof buf, " %d\n", c.level);
      s += buf;
    }
  }
  for (int l = 0; l < kCellLayers - 1; ++l) {
    // Exact size: the tracklets of layer l+1.  The reference array is sized by
    // an estimate; only its first T_{l+1} entries are meaningful.
    std::vector<int> lk = r.cellLookup[l];
    lk.resize(r.tracklets[l + 1].size(), -1);
    std::snprintf(buf, sizeof buf, "clookup %d %zu\n", l, lk.size());
    s += buf;
    ints(s, lk);
  }
  for (int l = 0; l < kCellLayers - 1; ++l) {
    const size_t n = r.cells[l + 1].size();
    std::snprintf(buf, sizeof buf, "neighbours %d %zu\n", l, n);
    s += buf;
    for (size_t i = 0; i < n; ++i) {
      std::vector<int> lst;
      if (i < r.neighbours[l].size()) lst = r.neighbours[l][i];
      std::snprintf(buf, sizeof buf, "%zu", lst.size());
      s += buf;
      for (int p : lst) {
        std::snprintf(buf, sizeof buf, " %d", p);
        s += buf;
      }
      s += "\n";
    }
  }
  std::snprintf(buf, sizeof buf, "roads %zu\n", r.roads.size());
  s += buf;
  for (const Road& rd : r.roads) {
    std::snprintf(buf, sizeof buf, "%d %d %d %d %d %d %d %d\n", rd.cells[0], rd.cells[1], rd.cells[2], rd.cells[3],
                  rd.cells[4], rd.size, rd.label, rd.fake ? 1 : 0);
    s += buf;
  }
  return s;
}

}  // namespace itsca_oracle
