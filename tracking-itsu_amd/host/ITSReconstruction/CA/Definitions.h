// Definitions.h -- plain vector types of the reference API
// (include/ITSReconstruction/CA/Definitions.h:71-75, CPU branch).
#pragma once

namespace o2 {
namespace ITS {
namespace CA {

struct float3 {
  float x, y, z;
};
struct int4 {
  int x, y, z, w;
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
