// stage_iso.hip -- the stage kernels instantiated for the isotropic elastic medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_iso(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_ISO>(L, P, kind, dx, dy, dz);
}

}  // namespace cgfdh
