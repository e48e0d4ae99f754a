// Definitions.h -- back-end selector and plain vector types of the reference
// API (include/ITSReconstruction/CA/Definitions.h:20-24, 71-75).
#pragma once

// The reference selects Tracker<TRACKINGITSU_GPU_MODE> at compile time
// (Definitions.h:20-24: true in its CUDA build).  This library's only back end
// is the gfx950 one, Tracker<true>, so callers written against the macro
// (main.cpp:15, 74) get it.
#ifndef TRACKINGITSU_GPU_MODE
#define TRACKINGITSU_GPU_MODE true
#endif

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
