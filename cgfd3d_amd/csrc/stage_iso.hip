// stage_iso.hip -- the stage kernels instantiated for the isotropic elastic medium.
#include <hip/hip_runtime.h>

#include "stage_launch.h"

namespace cgfdh {

void launch_stage_iso(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  launch_stage_mt<MEDIUM_ISO>(L, P, kind, dx, dy, dz);
}

void launch_rows_multi_iso(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                          int dy, int dz)
{
  launch_rows_multi_mt<MEDIUM_ISO>(st, R, nblocks, kind, dx, dy, dz);
}

}  // namespace cgfdh
