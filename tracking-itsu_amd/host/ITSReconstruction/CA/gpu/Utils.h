// gpu/Utils.h -- the profiler switches of the reference's GPU utilities
// (include/ITSReconstruction/CA/gpu/Utils.h:51-52, src/CA/gpu/cuda/Utils.cu:161-169),
// which the reference main.cpp calls around its event loop under
// GPU_PROFILING_MODE (main.cpp:77-79, 144-146).
//
// The reference switches the CUDA profiler on and off.  Here they open and
// close a ROCTx range named "itsca profiled region" and resume / pause ROCTx
// collection, so `rocprofv3 --marker-trace` shows the region and
// `rocprofv3 --selected-regions` limits collection to it.  Outside a profiler
// both calls do nothing.  The reference's device helpers and allocation
// wrappers of this header belong to its CUDA runtime layer, which the C ABI of
// libitsca_hip.so replaces.
#pragma once

namespace o2 {
namespace ITS {
namespace CA {
namespace GPU {
namespace Utils {
namespace Host {
void gpuStartProfiler();
void gpuStopProfiler();
}  // namespace Host
}  // namespace Utils
}  // namespace GPU
}  // namespace CA
}  // namespace ITS
}  // namespace o2
