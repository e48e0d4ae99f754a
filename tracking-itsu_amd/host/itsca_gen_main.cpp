// itsca-gen -- writes seeded synthetic events (and MC labels) in the reference
// text formats (src/CA/IOUtils.cxx:38-123).
//
// usage: itsca-gen <preset 1|2|3|5> <seed> <num_events> <events.txt> [labels.txt]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "events.hpp"

int main(int argc, char** argv)
{
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s <preset 1|2|3|5> <seed> <num_events> <events.txt> [labels.txt]\n", argv[0]);
    return 2;
  }
  const int preset = std::atoi(argv[1]);
  if (preset != 1 && preset != 2 && preset != 3 && preset != 5) {
    std::fprintf(stderr, "unknown preset %d\n", preset);
    return 2;
  }
  itsca::EventBatch batch;
  itsca::generateEvents(preset, std::strtoull(argv[2], nullptr, 10), std::atoi(argv[3]), batch);
  std::string err;
  if (!itsca::writeEvents(batch, argv[4], err) || (argc > 5 && !itsca::writeLabels(batch, argv[5], err))) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  return 0;
}
