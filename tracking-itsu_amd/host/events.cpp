// events.cpp -- seeded generator and reference text-format I/O.
//
// The generator places one cluster per crossed layer for helices in a uniform
// Bz = 0.5 T field, smears it by 5 um in r-phi and z, and adds uniform noise.
// All sampling is done in double from a xoshiro256** stream seeded by
// splitmix64, in a fixed operation order, so a (preset, seed) pair always
// yields the same float32 event on the same libm.
#include "events.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace itsca {
namespace {

// Generator geometry (cylinder radii and half-lengths, cm).
constexpr double kGenR[kNumLayers] = {2.34, 3.15, 3.93, 19.6, 24.55, 34.39, 39.34};
constexpr double kGenHalfZ[kNumLayers] = {13.5, 13.5, 13.5, 42.1, 42.1, 73.7, 73.7};
constexpr double kSmear = 5e-4;  // 5 um
constexpr double kPiD = 3.14159265358979323846;

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed)
  {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t t = z;
      t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
      t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
      s[i] = t ^ (t >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next()
  {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal()
  {
    const double u1 = 1.0 - uniform();  // (0, 1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPiD * u2);
  }
  uint64_t below(uint64_t n) { return next() % n; }
};

struct Track {
  double x0, y0, z0;  // production point
  double phi0, eta, pt;
  int charge;
  int mc;
};

// First outgoing crossing of the helix with the cylinder of radius rho.
// Returns false if the circle never reaches rho.
bool crossCylinder(const Track& t, double rho, double& px, double& py, double& pz)
{
  const double R = 100.0 * t.pt / 0.15;  // cm: R[m] = pT / (0.3 * 0.5)
  const double q = t.charge;
  const double cx = t.x0 + q * R * std::sin(t.phi0);
  const double cy = t.y0 - q * R * std::cos(t.phi0);
  const double d = std::sqrt(cx * cx + cy * cy);
  if (d < 1e-12 || d > R + rho || d < std::fabs(R - rho)) return false;
  const double a = (rho * rho - R * R + d * d) / (2.0 * d);
  const double h2 = rho * rho - a * a;
  const double h = h2 > 0 ? std::sqrt(h2) : 0.0;
  const double ux = cx / d, uy = cy / d;
  double best = -1;
  for (int sgn = -1; sgn <= 1; sgn += 2) {
    const double qx = a * ux - sgn * h * uy;
    const double qy = a * uy + sgn * h * ux;
    // position = c + qR(-sin phi, cos phi), phi = phi0 - q s / R
    const double sphi = -(qx - cx) / (q * R);
    const double cphi = (qy - cy) / (q * R);
    const double phi = std::atan2(sphi, cphi);
    double turn = q * (t.phi0 - phi);
    turn = std::fmod(turn, 2.0 * kPiD);
    if (turn <= 1e-12) turn += 2.0 * kPiD;
    const double s = turn * R;
    if (best < 0 || s < best) {
      best = s;
      px = qx;
      py = qy;
    }
  }
  pz = t.z0 + best * std::sinh(t.eta);
  return true;
}

double samplePt(Rng& rng)
{
  for (;;) {
    const double pt = -0.6 * std::log(1.0 - rng.uniform());
    if (pt >= 0.1 && pt <= 20.0) return pt;
  }
}

Track sampleTrack(Rng& rng, double x0, double y0, double z0, int mc)
{
  Track t;
  t.x0 = x0;
  t.y0 = y0;
  t.z0 = z0;
  t.eta = rng.uniform(-1.2, 1.2);
  t.phi0 = rng.uniform(0.0, 2.0 * kPiD);
  t.pt = samplePt(rng);
  t.charge = rng.uniform() < 0.5 ? -1 : 1;
  t.mc = mc;
  return t;
}

struct LayerHits {
  std::vector<float> x, y, z, alpha;
  std::vector<int32_t> mc;
  void add(double px, double py, double pz, int m)
  {
    x.push_back((float)px);
    y.push_back((float)py);
    z.push_back((float)pz);
    alpha.push_back((float)std::atan2(py, px));
    mc.push_back(m);
  }
};

void addTrackHits(Rng& rng, const Track& t, LayerHits* hits, std::vector<McLabel>& labels)
{
  int n = 0;
  for (int l = 0; l < kNumLayers; ++l) {
    double px = 0, py = 0, pz = 0;
    if (!crossCylinder(t, kGenR[l], px, py, pz)) continue;
    if (std::fabs(pz) > kGenHalfZ[l]) continue;
    const double dphi = rng.normal() * kSmear / kGenR[l];
    const double c = std::cos(dphi), s = std::sin(dphi);
    const double sx = c * px - s * py, sy = s * px + c * py;
    const double sz = pz + rng.normal() * kSmear;
    hits[l].add(sx, sy, sz, t.mc);
    ++n;
  }
  McLabel lab;
  lab.mc = t.mc;
  lab.pt = (float)t.pt;
  lab.phi = (float)t.phi0;
  lab.eta = (float)t.eta;
  lab.pdg = t.charge > 0 ? 211 : -211;
  lab.nClusters = n;
  labels.push_back(lab);
}

void generateOne(int preset, uint64_t seed, RawEvent& ev, std::vector<McLabel>& labels)
{
  Rng rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)preset);
  int collisions = preset == kGenPileUp ? 4 : 1;
  double noiseFrac = preset == kGenSmall ? 0.01 : (preset == kGenPileUp ? 0.03 : 0.02);
  double secondaryFrac = preset == kGenSmall ? 0.0 : 0.02;
  LayerHits hits[kNumLayers];
  ev.vertices.clear();
  for (int c = 0; c < collisions; ++c) {
    double vz;
    int primaries;
    if (preset == kGenPileUp) {
      vz = rng.uniform(-10.0, 10.0);
      primaries = 2000 + (int)rng.below(6001);
    } else {
      vz = 5.0 * rng.normal();
      if (preset == kGenSmall) {
        primaries = 100;
      } else if (preset == kGenMixed) {
        primaries = (int)std::lround(std::exp(std::log(50.0) + rng.uniform() * (std::log(8000.0) - std::log(50.0))));
      } else {
        primaries = 8000;
      }
    }
    const float fvz = (float)vz;
    ev.vertices.push_back(0.0f);
    ev.vertices.push_back(0.0f);
    ev.vertices.push_back(fvz);
    const int mcBase = c * 100000;  // pile-up merge offset per collision
    int mc = mcBase;
    for (int i = 0; i < primaries; ++i) addTrackHits(rng, sampleTrack(rng, 0.0, 0.0, fvz, mc++), hits, labels);
    const int secondaries = (int)std::lround(secondaryFrac * primaries);
    for (int i = 0; i < secondaries; ++i) {
      const double r = rng.uniform(0.0, 2.0);
      const double a = rng.uniform(0.0, 2.0 * kPiD);
      addTrackHits(rng, sampleTrack(rng, r * std::cos(a), r * std::sin(a), fvz, mc++), hits, labels);
    }
  }
  for (int l = 0; l < kNumLayers; ++l) {
    const int noise = (int)std::lround(noiseFrac * (double)hits[l].x.size());
    for (int i = 0; i < noise; ++i) {
      const double a = rng.uniform(0.0, 2.0 * kPiD);
      const double zz = rng.uniform(-kGenHalfZ[l], kGenHalfZ[l]);
      hits[l].add(kGenR[l] * std::cos(a), kGenR[l] * std::sin(a), zz, -1);
    }
  }
  // Deterministic shuffle within each layer, then number clusters in file
  // order (layer by layer).
  int32_t nextId = 0;
  for (int l = 0; l < kNumLayers; ++l) {
    const int n = (int)hits[l].x.size();
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)rng.below((uint64_t)i + 1)]);
    ev.x[l].resize(n);
    ev.y[l].resize(n);
    ev.z[l].resize(n);
    ev.alpha[l].resize(n);
    ev.id[l].resize(n);
    ev.mc[l].resize(n);
    for (int i = 0; i < n; ++i) {
      const int p = perm[i];
      ev.x[l][i] = hits[l].x[p];
      ev.y[l][i] = hits[l].y[p];
      ev.z[l][i] = hits[l].z[p];
      ev.alpha[l][i] = hits[l].alpha[p];
      ev.mc[l][i] = hits[l].mc[p];
      ev.id[l][i] = nextId++;
    }
  }
}

}  // namespace

void generateEvents(int preset, uint64_t seed, int numEvents, EventBatch& out)
{
  out.events.assign(numEvents, RawEvent());
  out.labels.assign(numEvents, {});
  for (int e = 0; e < numEvents; ++e) generateOne(preset, seed + (uint64_t)e, out.events[e], out.labels[e]);
}

bool writeEvents(const EventBatch& batch, const std::string& path, std::string& err)
{
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::vector<std::pair<int32_t, int32_t>> order;  // (cluster id, layer * 2^24 + index)
  for (const RawEvent& ev : batch.events) {
    for (int v = 0; v < ev.numVertices(); ++v)
      std::fprintf(f, "-1\t%.9g\t%.9g\t%.9g\n", ev.vertices[3 * v], ev.vertices[3 * v + 1], ev.vertices[3 * v + 2]);
    // Clusters in cluster-id order, which is the order of the file they were
    // read from (layer by layer for generated events).
    order.clear();
    for (int l = 0; l < kNumLayers; ++l)
      for (int i = 0; i < ev.layerSize(l); ++i) order.emplace_back(ev.id[l][i], (l << 24) | i);
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
                       return a.first < b.first;
                     });
    for (const auto& o : order) {
      const int l = o.second >> 24, i = o.second & 0xffffff;
      std::fprintf(f, "%d\t%.9g\t%.9g\t%.9g\t0\t0\t0\t%.9g\t%d\n", l, ev.x[l][i], ev.y[l][i], ev.z[l][i],
                   ev.alpha[l][i], ev.mc[l][i]);
    }
  }
  std::fclose(f);
  return true;
}

bool writeLabels(const EventBatch& batch, const std::string& path, std::string& err)
{
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::fprintf(f, "mc\tpT\tphi\teta\tpdg\tnClusters\n");
  for (size_t e = 0; e < batch.labels.size(); ++e) {
    if (e > 0) std::fprintf(f, "-1\n");
    for (const McLabel& m : batch.labels[e])
      std::fprintf(f, "%d\t%.9g\t%.9g\t%.9g\t%d\t%d\n", m.mc, m.pt, m.phi, m.eta, m.pdg, m.nClusters);
  }
  std::fclose(f);
  return true;
}

bool loadEvents(const std::string& path, EventBatch& out, std::string& err)
{
  if (isEventsBinary(path)) return loadEventsBinary(path, out, err);
  std::ifstream in(path);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  out.events.clear();
  out.labels.clear();
  std::string line, unused;
  int32_t clusterId = -1;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    int layer;
    float x, y, z, alpha;
    int mc;
    if (!(ss >> layer >> x >> y >> z)) continue;
    if (layer == -1) {
      if (clusterId != 0) out.events.emplace_back();
      out.events.back().vertices.insert(out.events.back().vertices.end(), {x, y, z});
      clusterId = 0;
    } else if (ss >> unused >> unused >> unused >> alpha >> mc) {
      if (out.events.empty() || layer < 0 || layer >= kNumLayers) continue;
      RawEvent& ev = out.events.back();
      ev.x[layer].push_back(x);
      ev.y[layer].push_back(y);
      ev.z[layer].push_back(z);
      ev.alpha[layer].push_back(alpha);
      ev.id[layer].push_back(clusterId);
      ev.mc[layer].push_back(mc);
      ++clusterId;
    }
  }
  out.labels.assign(out.events.size(), {});
  return true;
}

bool loadLabelRows(const std::string& path, EventBatch& batch, std::string& err)
{
  std::ifstream in(path);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  batch.labels.assign(batch.events.size(), {});
  std::string line;
  std::getline(in, line);  // header (IOUtils.cxx:92)
  size_t e = 0;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    McLabel m;
    if (!(ss >> m.mc)) continue;
    if (m.mc == -1) {
      ++e;
      continue;
    }
    if (!(ss >> m.pt >> m.phi >> m.eta >> m.pdg >> m.nClusters)) continue;
    if (e < batch.labels.size()) batch.labels[e].push_back(m);
  }
  return true;
}

namespace {
constexpr char kMagic[8] = {'I', 'T', 'S', 'C', 'A', 'E', 'V', '1'};

template <typename T>
void put(std::vector<char>& b, const T* p, size_t n)
{
  const char* c = reinterpret_cast<const char*>(p);
  b.insert(b.end(), c, c + n * sizeof(T));
}
}  // namespace

bool writeEventsBinary(const EventBatch& batch, const std::string& path, std::string& err)
{
  std::vector<char> b;
  put(b, kMagic, 8);
  const uint32_t ne = (uint32_t)batch.events.size();
  put(b, &ne, 1);
  for (const RawEvent& ev : batch.events) {
    uint32_t head[1 + kNumLayers];
    head[0] = (uint32_t)ev.numVertices();
    for (int l = 0; l < kNumLayers; ++l) head[1 + l] = (uint32_t)ev.layerSize(l);
    put(b, head, 1 + kNumLayers);
    put(b, ev.vertices.data(), ev.vertices.size());
    for (int l = 0; l < kNumLayers; ++l) {
      const size_t n = (size_t)ev.layerSize(l);
      put(b, ev.x[l].data(), n);
      put(b, ev.y[l].data(), n);
      put(b, ev.z[l].data(), n);
      put(b, ev.alpha[l].data(), n);
      put(b, ev.id[l].data(), n);
      put(b, ev.mc[l].data(), n);
    }
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  std::fclose(f);
  if (!ok) err = "cannot write " + path;
  return ok;
}

bool isEventsBinary(const std::string& path)
{
  char m[8];
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, kMagic, 8) == 0;
  std::fclose(f);
  return ok;
}

bool loadEventsBinary(const std::string& path, EventBatch& out, std::string& err)
{
  const int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    err = "cannot open " + path;
    return false;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < 12) {
    close(fd);
    err = "not an event file: " + path;
    return false;
  }
  const size_t size = (size_t)st.st_size;
  void* map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (map == MAP_FAILED) {
    err = "cannot map " + path;
    return false;
  }
  const char* p = static_cast<const char*>(map);
  size_t at = 0;
  bool ok = std::memcmp(p, kMagic, 8) == 0;
  at = 8;
  auto take = [&](void* dst, size_t bytes) {
    if (!ok || at + bytes > size) {
      ok = false;
      return;
    }
    if (bytes) std::memcpy(dst, p + at, bytes);
    at += bytes;
  };
  uint32_t ne = 0;
  take(&ne, 4);
  out.events.clear();
  out.labels.clear();
  if (ok && ne > size / (4 * (1 + kNumLayers))) ok = false;  // more events than the file can hold
  if (ok) out.events.resize(ne);
  for (uint32_t e = 0; ok && e < ne; ++e) {
    RawEvent& ev = out.events[e];
    uint32_t head[1 + kNumLayers];
    take(head, sizeof(head));
    if (!ok) break;
    for (int k = 0; k < 1 + kNumLayers; ++k)
      if (head[k] > size) ok = false;
    if (!ok) break;
    ev.vertices.resize(3 * (size_t)head[0]);
    take(ev.vertices.data(), ev.vertices.size() * 4);
    for (int l = 0; ok && l < kNumLayers; ++l) {
      const size_t n = head[1 + l];
      ev.x[l].resize(n);
      ev.y[l].resize(n);
      ev.z[l].resize(n);
      ev.alpha[l].resize(n);
      ev.id[l].resize(n);
      ev.mc[l].resize(n);
      take(ev.x[l].data(), n * 4);
      take(ev.y[l].data(), n * 4);
      take(ev.z[l].data(), n * 4);
      take(ev.alpha[l].data(), n * 4);
      take(ev.id[l].data(), n * 4);
      take(ev.mc[l].data(), n * 4);
    }
  }
  munmap(map, size);
  if (!ok) {
    out.events.clear();
    err = "truncated or malformed event file: " + path;
    return false;
  }
  out.labels.assign(out.events.size(), {});
  return true;
}

bool mergePileUp(const EventBatch& in, int pileUp, EventBatch& out, std::string& err)
{
  if (pileUp < 1) {
    err = "pile-up must be at least 1";
    return false;
  }
  out.events.clear();
  out.labels.clear();
  const bool withLabels = in.labels.size() == in.events.size();
  int verticesNum = 0;  // vertices of the merged event so far
  int32_t idOffset = 0;  // clusters of the merged event so far
  for (size_t e = 0; e < in.events.size(); ++e) {
    const RawEvent& src = in.events[e];
    if (out.events.empty() || verticesNum >= pileUp) {
      out.events.emplace_back();
      out.labels.emplace_back();
      verticesNum = 0;
      idOffset = 0;
    }
    RawEvent& dst = out.events.back();
    dst.vertices.insert(dst.vertices.end(), src.vertices.begin(), src.vertices.end());
    verticesNum += src.numVertices();
    // Clusters follow the event's last vertex line (generatePileUpData.cxx:88-104).
    const int32_t offset = verticesNum > 1 ? (verticesNum - 1) * kPileUpMcOffset : 0;
    for (int l = 0; l < kNumLayers; ++l) {
      for (int i = 0; i < src.layerSize(l); ++i) {
        dst.x[l].push_back(src.x[l][i]);
        dst.y[l].push_back(src.y[l][i]);
        dst.z[l].push_back(src.z[l][i]);
        dst.alpha[l].push_back(src.alpha[l][i]);
        dst.id[l].push_back(idOffset + src.id[l][i]);  // the source's file order, after the earlier collisions
        dst.mc[l].push_back(src.mc[l][i] + offset);
      }
    }
    int32_t n = 0;
    for (int l = 0; l < kNumLayers; ++l) n += src.layerSize(l);
    idOffset += n;
    if (withLabels)
      for (McLabel m : in.labels[e]) {
        m.mc += offset;
        out.labels.back().push_back(m);
      }
  }
  return true;
}

}  // namespace itsca
