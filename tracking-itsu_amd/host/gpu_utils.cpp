// gpu_utils.cpp -- GPU::Utils::Host::gpuStartProfiler / gpuStopProfiler
// (src/CA/gpu/cuda/Utils.cu:161-169) as a ROCTx range plus resume / pause of
// ROCTx-selected collection.
#include <rocprofiler-sdk-roctx/roctx.h>

#include "ITSReconstruction/CA/gpu/Utils.h"

namespace o2 {
namespace ITS {
namespace CA {
namespace GPU {
namespace Utils {
namespace Host {

void gpuStartProfiler()
{
  roctxProfilerResume(0);
  roctxRangePush("itsca profiled region");
}

void gpuStopProfiler()
{
  roctxRangePop();
  roctxProfilerPause(0);
}

}  // namespace Host
}  // namespace Utils
}  // namespace GPU
}  // namespace CA
}  // namespace ITS
}  // namespace o2
