"""MI355X (gfx950) engine for the time-stepping hot path of a curvilinear
collocated-grid elastic wave solver (the drv_rk_curv_col / sv_curv_col_el_{iso,
vti,aniso} path of CGFD3D), behind the C ABI in include/cgfdh.h."""
from ._lib import CgfdhError, LIB_PATH  # noqa: F401
from .solver import (MEDIUM_ELASTIC_ANISO, MEDIUM_ELASTIC_ISO, MEDIUM_ELASTIC_VTI,  # noqa: F401
                     MEDIUM_VISCO_ISO, SRC_GAUSSIAN, SRC_POINT, Batch, PinnedBuffer, Solver,
                     L2M_EL_ANISO, L2M_EL_ISO, L2M_EL_VTI, L2M_ONECMP, bin2model_grid,
                     depth_to_axis, media_bin2model,
                     estimate_dt, gen_grid_layer, grid2model_grid, layer2model, layer2model_grid,
                     locate_points, media_grid2model, media_layer2model)
