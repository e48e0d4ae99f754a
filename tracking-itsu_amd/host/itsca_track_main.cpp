// itsca-track -- tracker CLI, the counterpart of the reference main.cpp
// (main.cpp:27-153) on the gfx950 back end.
//
// usage: itsca-track <events.txt> [labels.txt] [--filter weights.json]
// Runs a warm-up pass, then every event with its own timing, and prints the
// road count per vertex and the average time per vertex.  With labels it
// writes CorrectRoads.txt, DuplicateRoads.txt and FakeRoads.txt next to the
// events file, as the reference does.  --filter turns on the learned tracklet
// filter with the given weights file.
#include <chrono>
#include <cstdio>
#include <iostream>
#include <limits>
#include <string>

#include "ITSReconstruction/CA/IOUtils.h"
#include "ITSReconstruction/CA/Tracker.h"

using namespace o2::ITS::CA;

int main(int argc, char** argv)
{
  std::vector<std::string> args;
  std::string filterWeights;
  for (int a = 1; a < argc; ++a) {
    const std::string v = argv[a];
    if (v == "--filter" && a + 1 < argc)
      filterWeights = argv[++a];
    else
      args.push_back(v);
  }
  if (args.empty()) {
    std::cerr << "Please, provide a data file." << std::endl;
    return 1;
  }
  const std::string eventsFile = args[0];
  const std::string dir = eventsFile.find('/') == std::string::npos ? "" : eventsFile.substr(0, eventsFile.rfind('/') + 1);
  std::vector<Event> events = IOUtils::loadEventData(eventsFile);
  if (events.empty()) {
    std::cerr << "No events in " << eventsFile << std::endl;
    return 1;
  }
  std::vector<std::unordered_map<int, Label>> labels;
  std::ofstream correct, duplicate, fake;
  const bool report = args.size() > 1;
  if (report) {
    labels = IOUtils::loadLabels((int)events.size(), args[1]);
    correct.open(dir + "CorrectRoads.txt");
    duplicate.open(dir + "DuplicateRoads.txt");
    fake.open(dir + "FakeRoads.txt");
  }
  int vertices = 0;
  for (const Event& e : events) vertices += e.getPrimaryVerticesNum();
  try {
    Tracker<true> tracker;
    if (!filterWeights.empty()) tracker.useLearnedFilter(filterWeights);
    tracker.clustersToTracks(events[0]);  // warm-up
    double total = 0, minT = std::numeric_limits<double>::max(), maxT = 0;
    for (size_t i = 0; i < events.size(); ++i) {
      std::cout << "Processing event " << i + 1 << std::endl;
      const auto t0 = std::chrono::steady_clock::now();
      const std::vector<std::vector<Road>> roads = tracker.clustersToTracks(events[i]);
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      total += ms;
      minT = std::min(minT, ms);
      maxT = std::max(maxT, ms);
      for (size_t v = 0; v < roads.size(); ++v)
        std::cout << "Found " << roads[v].size() << " roads for vertex " << v + 1 << std::endl;
      std::cout << "Event " << i + 1 << " processed in: " << ms << "ms" << std::endl << std::endl;
      if (report && i < labels.size()) IOUtils::writeRoadsReport(correct, duplicate, fake, roads, labels[i]);
    }
    std::cout << "Avg time: " << total / vertices << "ms" << std::endl;
    std::cout << "Min time: " << minT << "ms" << std::endl;
    std::cout << "Max time: " << maxT << "ms" << std::endl;
  } catch (const std::exception& e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  return 0;
}
