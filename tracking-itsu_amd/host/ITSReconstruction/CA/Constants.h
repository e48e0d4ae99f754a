// Constants.h -- the layer and road constants of the reference API
// (include/ITSReconstruction/CA/Constants.h:36-41, 73, 96).  The cut values
// live with the kernels (csrc/itsca_math.h).
#pragma once

namespace o2 {
namespace ITS {
namespace CA {
namespace Constants {
namespace ITS {
constexpr int LayersNumber{7};
constexpr int TrackletsPerRoad{6};
constexpr int CellsPerRoad{LayersNumber - 2};
constexpr int UnusedIndex{-1};
}  // namespace ITS
namespace Thresholds {
constexpr int CellsMinLevel{5};
}
namespace PDGCodes {
constexpr int PionCode{211};
}
}  // namespace Constants
}  // namespace CA
}  // namespace ITS
}  // namespace o2
