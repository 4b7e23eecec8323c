// stage_visco.hip -- the stage kernels instantiated for the visco-elastic (GMB) medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_visco(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_VISCO>(L, P, kind, dx, dy, dz);
}

void launch_rows_multi_visco(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                          int dy, int dz)
{
  launch_rows_multi_mt<MEDIUM_VISCO>(st, R, nblocks, kind, dx, dy, dz);
}

}  // namespace cgfdh
