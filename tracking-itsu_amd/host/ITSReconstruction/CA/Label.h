// Label.h -- MC truth record of the reference API
// (include/ITSReconstruction/CA/Label.h:27-40, Label.cxx:26-40).
#pragma once

#include <ostream>

namespace o2 {
namespace ITS {
namespace CA {

struct Label {
  Label(int mc, float pt, float phi, float eta, int pdg, int nClusters)
      : monteCarloId{mc}, transverseMomentum{pt}, phiCoordinate{phi}, pseudorapidity{eta}, pdgCode{pdg},
        numberOfClusters{nClusters}
  {
  }
  int monteCarloId;
  float transverseMomentum;
  float phiCoordinate;
  float pseudorapidity;
  int pdgCode;
  int numberOfClusters;
  friend std::ostream& operator<<(std::ostream& os, const Label& l)
  {
    return os << l.monteCarloId << "\t" << l.transverseMomentum << "\t" << l.phiCoordinate << "\t"
              << l.pseudorapidity << "\t" << l.pdgCode << "\t" << l.numberOfClusters;
  }
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
