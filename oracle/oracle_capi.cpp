// oracle_capi.cpp -- C entry points of the CPU oracle for the Python tests and
// the CPU-baseline leg of the benchmark.
//
// TEST INFRASTRUCTURE ONLY (see ca_oracle.hpp).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "ca_oracle.hpp"

using itsca_oracle::VertexResult;

extern "C" {

void* itsca_oracle_new(void) { return new VertexResult(); }

void itsca_oracle_free(void* h) { delete static_cast<VertexResult*>(h); }

// Runs one vertex.  Arrays hold every layer's clusters, layer after layer, in
// file order; layer_offsets has 8 entries.  mode bit 0: stable sort; bit 1:
// exclusive candidate bound (not the reference; see runVertex).
int itsca_oracle_run(void* h, const float* x, const float* y, const float* z, const int* id, const int* mc,
                     const int* layer_offsets, const float* vertex, int mode)
{
  if (!h || !layer_offsets || !vertex) return -1;
  itsca_oracle::EventView ev{x, y, z, id, mc, layer_offsets, vertex[0], vertex[1], vertex[2]};
  itsca_oracle::runVertex(ev, (mode & 1) != 0, *static_cast<VertexResult*>(h), (mode & 2) != 0);
  return 0;
}

// Writes the canonical dump to path; returns its length or -1.
long itsca_oracle_dump(void* h, const char* path)
{
  const std::string s = itsca_oracle::dump(*static_cast<VertexResult*>(h));
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  std::fwrite(s.data(), 1, s.size(), f);
  std::fclose(f);
  return (long)s.size();
}

// what: 0 clusters(layer), 1 tracklets(layer), 2 cells(layer), 3 roads,
//       4 neighbour edges(layer)
long itsca_oracle_count(void* h, int what, int layer)
{
  const VertexResult& r = *static_cast<VertexResult*>(h);
  switch (what) {
    case 0: return (long)r.clusters[layer].size();
    case 1: return (long)r.tracklets[layer].size();
    case 2: return (long)r.cells[layer].size();
    case 3: return (long)r.roads.size();
    case 4: {
      long e = 0;
      for (const auto& v : r.neighbours[layer]) e += (long)v.size();
      return e;
    }
    default: return -1;
  }
}

// Copies records as flat int32/float32 rows.
//   0 clusters: id, phi, r, bin, x, y, z, mc   (8 x 4 B)
//   1 tracklets: first, second, tanLambda, phi  (4 x 4 B)
//   2 cells: c1, c2, c3, t1, t2, nx, ny, nz, curvature, level (10 x 4 B)
//   3 roads: cells[5], size, label, fake (8 x 4 B)
long itsca_oracle_copy(void* h, int what, int layer, void* dst)
{
  const VertexResult& r = *static_cast<VertexResult*>(h);
  char* out = static_cast<char*>(dst);
  long n = 0;
  auto put = [&](const void* p) {
    std::memcpy(out, p, 4);
    out += 4;
  };
  switch (what) {
    case 0:
      for (const auto& c : r.clusters[layer]) {
        put(&c.id), put(&c.phi), put(&c.r), put(&c.bin), put(&c.x), put(&c.y), put(&c.z), put(&c.mc);
        ++n;
      }
      return n;
    case 1:
      for (const auto& t : r.tracklets[layer]) {
        put(&t.first), put(&t.second), put(&t.tanLambda), put(&t.phi);
        ++n;
      }
      return n;
    case 2:
      for (const auto& c : r.cells[layer]) {
        put(&c.c1), put(&c.c2), put(&c.c3), put(&c.t1), put(&c.t2);
        put(&c.nx), put(&c.ny), put(&c.nz), put(&c.curvature), put(&c.level);
        ++n;
      }
      return n;
    case 3:
      for (const auto& rd : r.roads) {
        for (int k = 0; k < 5; ++k) put(&rd.cells[k]);
        const int fake = rd.fake ? 1 : 0;
        put(&rd.size), put(&rd.label), put(&fake);
        ++n;
      }
      return n;
    default: return -1;
  }
}

// CPU baseline: runs every vertex of a batch of events `repeats` times on
// `threads` threads, one independent tracker state per thread (events are
// independent).  Events are given as concatenated arrays with per-event
// layer offsets (8 per event, absolute into the arrays) and one vertex each.
// Returns the wall-clock seconds of the whole run.
double itsca_oracle_time(const float* x, const float* y, const float* z, const int* id, const int* mc,
                         const int* layer_offsets, const float* vertices, int num_events, int repeats,
                         int threads)
{
  if (threads < 1) threads = 1;
  auto worker = [&](int tid) {
    VertexResult r;
    for (int rep = 0; rep < repeats; ++rep)
      for (int e = tid; e < num_events; e += threads) {
        int offs[8];
        for (int l = 0; l < 8; ++l) offs[l] = layer_offsets[8 * e + l];
        itsca_oracle::EventView ev{x, y, z, id, mc, offs, vertices[3 * e], vertices[3 * e + 1], vertices[3 * e + 2]};
        itsca_oracle::runVertex(ev, false, r);
      }
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
