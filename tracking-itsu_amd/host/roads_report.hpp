// roads_report.hpp -- the road classification of IOUtils::writeRoadsReport
// (src/CA/IOUtils.cxx:125-170), shared by the C++ API's report writer and the
// C entry point that counts the classes.
//
// Per event, in road order over its vertices: a road whose MC label is not in
// the labels map is skipped; a fake road is "fake"; otherwise the first road
// of each MC id is "correct" and every later one "duplicate".
#pragma once

#include <unordered_set>

namespace itsca {

enum RoadClass { kRoadCorrect = 0, kRoadDuplicate = 1, kRoadFake = 2 };

class RoadClassifier {
 public:
  // Returns the class of the next road, or -1 if its label is not in the map
  // (find(label) must give the map entry's MC id, or nothing).
  template <typename Find>
  int next(int label, bool fake, Find&& find)
  {
    int mc;
    if (!find(label, mc)) return -1;
    if (fake) return kRoadFake;
    if (mFound.count(mc)) return kRoadDuplicate;
    mFound.insert(mc);
    return kRoadCorrect;
  }

 private:
  std::unordered_set<int> mFound;
};

}  // namespace itsca
