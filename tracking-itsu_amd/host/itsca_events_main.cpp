// itsca-events -- event-file tools: the binary side-car and the pile-up
// merger (the counterpart of benchmarks/scripts/root/generatePileUpData.cxx).
//
// usage: itsca-events binary <events.txt> <events.itsb>
//        itsca-events pileup <N> <events> <merged_events.txt> [<labels.txt> <merged_labels.txt>]
// Inputs may be text or binary; the merged events are written as text.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "events.hpp"

int main(int argc, char** argv)
{
  const std::string cmd = argc > 1 ? argv[1] : "";
  std::string err;
  itsca::EventBatch in;
  if (cmd == "binary" && argc == 4) {
    if (!itsca::loadEvents(argv[2], in, err) || !itsca::writeEventsBinary(in, argv[3], err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    return 0;
  }
  if (cmd == "pileup" && (argc == 5 || argc == 7)) {
    const int n = std::atoi(argv[2]);
    itsca::EventBatch out;
    bool ok = itsca::loadEvents(argv[3], in, err) && (argc == 5 || itsca::loadLabelRows(argv[5], in, err)) &&
              itsca::mergePileUp(in, n, out, err) && itsca::writeEvents(out, argv[4], err) &&
              (argc == 5 || itsca::writeLabels(out, argv[6], err));
    if (!ok) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    std::printf("merged %zu events into %zu with pile-up %d\n", in.events.size(), out.events.size(), n);
    return 0;
  }
  std::fprintf(stderr,
               "usage: %s binary <events.txt> <events.itsb>\n"
               "       %s pileup <N> <events> <merged_events.txt> [<labels.txt> <merged_labels.txt>]\n",
               argv[0], argv[0]);
  return 2;
}
