// IOUtils.h -- reference text formats (include/ITSReconstruction/CA/IOUtils.h:36-41,
// src/CA/IOUtils.cxx:38-170).
#pragma once

#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "Event.h"
#include "Label.h"
#include "Road.h"

namespace o2 {
namespace ITS {
namespace CA {
namespace IOUtils {
std::vector<Event> loadEventData(const std::string& fileName);
std::vector<std::unordered_map<int, Label>> loadLabels(int eventsNum, const std::string& fileName);
void writeRoadsReport(std::ofstream& correct, std::ofstream& duplicate, std::ofstream& fake,
                      const std::vector<std::vector<Road>>& roads, const std::unordered_map<int, Label>& labels);
}  // namespace IOUtils
}  // namespace CA
}  // namespace ITS
}  // namespace o2
