// stage_aniso.hip -- the stage kernels instantiated for the general anisotropic elastic medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_aniso(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_ANISO>(L, P, kind, dx, dy, dz);
}

}  // namespace cgfdh
