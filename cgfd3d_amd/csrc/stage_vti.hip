// stage_vti.hip -- the stage kernels instantiated for the VTI elastic medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_vti(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_VTI>(L, P, kind, dx, dy, dz);
}

void launch_rows_multi_vti(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                          int dy, int dz)
{
  launch_rows_multi_mt<MEDIUM_VTI>(st, R, nblocks, kind, dx, dy, dz);
}

}  // namespace cgfdh
