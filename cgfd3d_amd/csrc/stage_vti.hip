// stage_vti.hip -- the stage kernels instantiated for the VTI elastic medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_vti(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_VTI>(L, P, kind, dx, dy, dz);
}

}  // namespace cgfdh
