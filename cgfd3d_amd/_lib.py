"""ctypes binding of the C ABI in include/cgfdh.h (libcgfdh.so, built in-tree).

There is no fallback: if the library is missing or fails to load, every entry
point raises, so nothing silently runs off the GPU path.
"""
import ctypes as C
import os

_P = C.c_void_p
_PKG = os.path.dirname(os.path.abspath(__file__))
# CGFDH_LIB may name another build of the same library (A/B timing runs)
LIB_PATH = os.environ.get("CGFDH_LIB") or os.path.join(_PKG, "libcgfdh.so")

_lib = None


class CgfdhDesc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("medium_type", C.c_int), ("free_top", C.c_int),
                ("pml_on", C.c_int * 6), ("pml_layers", C.c_int * 6),
                ("dt", C.c_float), ("nt_total", C.c_int), ("nmaxwell", C.c_int)]


# (name, restype, argtypes) of every symbol include/cgfdh.h declares
SIGNATURES = [
    ("cgfdh_create", C.c_int, [_P, _P]),
    ("cgfdh_destroy", C.c_int, [_P]),
    ("cgfdh_last_error", C.c_char_p, []),
    ("cgfdh_set_metric", C.c_int, [_P, _P]),
    ("cgfdh_set_medium_iso", C.c_int, [_P, _P, _P, _P]),
    ("cgfdh_set_medium_vti", C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("cgfdh_set_medium_aniso", C.c_int, [_P, _P, _P]),
    ("cgfdh_set_medium_visco", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("cgfdh_set_free_surface_matD", C.c_int, [_P, _P]),
    ("cgfdh_get_free_surface_matD", C.c_int, [_P, _P]),
    ("cgfdh_set_pml_face", C.c_int, [_P, C.c_int, _P, _P, _P]),
    ("cgfdh_set_free_surface", C.c_int, [_P, _P, _P, _P]),
    ("cgfdh_set_moment_sources", C.c_int,
     [_P, C.c_int, _P, _P, _P, C.c_int, _P, C.c_int, C.c_int, _P]),
    ("cgfdh_set_sources", C.c_int,
     [_P, C.c_int, _P, _P, _P, C.c_int, _P, _P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P,
      _P]),
    ("cgfdh_set_dd_sources", C.c_int, [_P, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("cgfdh_put_dd_values", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("cgfdh_set_receivers", C.c_int, [_P, C.c_int, _P, _P]),
    ("cgfdh_check_nan", C.c_int, [_P, _P, _P]),
    ("cgfdh_locate_points", C.c_int, [C.c_int] * 3 + [_P, C.c_int, _P, _P, _P, _P]),
    ("cgfdh_depth_to_axis", C.c_int, [C.c_int] * 3 + [_P, C.c_int, _P, _P, _P]),
    ("cgfdh_layer2model", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_long, _P, _P, _P, _P, _P, _P]),
    ("cgfdh_layer2model_grid", C.c_int, [C.c_int] * 6 + [_P, _P] + [C.c_int] * 3 + [_P] * 4),
    ("cgfdh_media_layer2model_onecmp", C.c_int,
     [_P] * 4 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_layer2model_el_iso", C.c_int,
     [_P] * 6 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_layer2model_el_vti", C.c_int,
     [_P] * 9 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_layer2model_el_aniso", C.c_int,
     [_P] * 25 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_bin2model_grid", C.c_int, [C.c_int] + [_P] * 5 + [C.c_long] + [_P] * 4),
    ("cgfdh_media_bin2model_el_iso", C.c_int,
     [_P] * 6 + [C.c_size_t] * 3 + [C.c_float] * 4 + [C.c_int] + [_P] * 4 +
     [C.c_char_p] * 3),
    ("cgfdh_media_bin2model_vis_iso", C.c_int,
     [_P] * 5 + [C.c_size_t] * 3 + [C.c_float] * 4 + [C.c_int] + [_P] * 4 +
     [C.c_char_p] * 2),
    ("cgfdh_grid2model_grid", C.c_int,
     [C.c_int] * 4 + [_P, C.c_int, C.c_int, _P, _P] + [C.c_int] * 3 + [_P] * 4),
    ("cgfdh_media_grid2model_onecmp", C.c_int,
     [_P] * 4 + [C.c_int] * 3 + [C.c_float] * 4 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_grid2model_el_iso", C.c_int,
     [_P] * 6 + [C.c_int] * 3 + [C.c_float] * 4 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_grid2model_el_vti", C.c_int,
     [_P] * 9 + [C.c_int] * 3 + [C.c_float] * 4 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_media_grid2model_el_aniso", C.c_int,
     [_P] * 25 + [C.c_int] * 3 + [C.c_float] * 4 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cgfdh_gen_grid_layer", C.c_int,
     [C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P] + [C.c_int] * 9 + [_P]),
    ("cgfdh_set_wavefield", C.c_int, [_P, _P]),
    ("cgfdh_get_wavefield", C.c_int, [_P, _P]),
    ("cgfdh_pml_aux_len", C.c_size_t, [_P]),
    ("cgfdh_set_pml_aux", C.c_int, [_P, _P]),
    ("cgfdh_get_pml_aux", C.c_int, [_P, _P]),
    ("cgfdh_run", C.c_int, [_P, C.c_int, C.c_int]),
    ("cgfdh_get_seismograms", C.c_int, [_P, _P]),
    ("cgfdh_get_peak_ground_motion", C.c_int, [_P, _P, _P]),
    ("cgfdh_get_subvolume", C.c_int, [_P, _P, C.c_int, _P, _P]),
    ("cgfdh_get_subvolume_noalloc", C.c_int, [_P, _P, C.c_int, _P, _P]),
    ("cgfdh_set_ablexp", C.c_int, [_P, _P, _P, _P, _P]),
    ("cgfdh_set_graves_qs", C.c_int, [_P, _P, C.c_float]),
    ("cgfdh_get_strain_subvolume", C.c_int, [_P, _P, _P]),
    ("cgfdh_run_async", C.c_int, [_P, C.c_int, C.c_int]),
    ("cgfdh_subvolume_async", C.c_int, [_P, _P, C.c_int, _P, _P, _P]),
    ("cgfdh_wait", C.c_int, [_P, C.c_int]),
    ("cgfdh_sync", C.c_int, [_P]),
    ("cgfdh_host_alloc", C.c_int, [C.c_size_t, _P]),
    ("cgfdh_host_free", C.c_int, [_P]),
    ("cgfdh_estimate_dt", C.c_int, [C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int,
                                    C.c_float, _P, _P, _P, _P]),
    ("cgfdh_onestage", C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
    ("cgfdh_last_timing", C.c_int, [_P, _P, _P, _P]),
    ("cgfdh_set_kernel_variant", C.c_int, [_P, C.c_int]),
    ("cgfdh_set_tile_group", C.c_int, [_P, C.c_int, C.c_int]),
    ("cgfdh_get_kernel_choice", C.c_int, [_P, _P, _P, _P]),
    ("cgfdh_batch_create", C.c_int, [C.c_int, _P, _P]),
    ("cgfdh_batch_destroy", C.c_int, [_P]),
    ("cgfdh_batch_autotune", C.c_int, [_P, C.c_int]),
    ("cgfdh_batch_run", C.c_int, [_P, C.c_int, C.c_int]),
    ("cgfdh_batch_last_timing", C.c_int, [_P, _P]),
    ("cgfds_gemm_bf16", C.c_int,
     [C.c_int, C.c_int, C.c_int, _P, C.c_long, C.c_int, _P, C.c_long, C.c_int, _P, C.c_int,
      _P, C.c_long, _P, C.c_long, _P, C.c_long, C.c_int, _P, _P]),
    ("cgfds_colsum_bf16", C.c_int, [C.c_int, C.c_int, _P, C.c_long, _P, C.c_int, _P, _P]),
    ("cgfds_loss_grad", C.c_int, [C.c_int, C.c_long, _P, _P, _P, C.c_float, _P, _P, _P]),
    ("cgfds_adam", C.c_int, [C.c_long, _P, _P, _P, _P, _P] + [C.c_float] * 8 + [_P]),
    ("cgfds_mt_contract", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
    ("cgfds_mt_loss_grad", C.c_int,
     [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, C.c_float, _P, _P, _P]),
    ("cgfds_f32_to_bf16", C.c_int, [C.c_long, _P, _P, _P]),
]


class CgfdhError(RuntimeError):
    pass


def lib():
    """Load libcgfdh.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CgfdhError(f"{LIB_PATH} is missing: build it with "
                         "`python -m cgfd3d_amd.build` (hipcc, gfx950)")
    # One HIP runtime per process: torch ships its own libamdhip64 and loads
    # it by a name that does not match an already loaded /opt/rocm runtime,
    # so a process that loaded this library first and imports torch later
    # ends up with two runtimes, and torch then sees no device.  Loading
    # torch first makes this library bind to torch's runtime (same soname),
    # so device pointers and streams are shared with it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().cgfdh_last_error().decode(errors="replace")
        raise CgfdhError(f"{what}: rc={rc}: {msg}")
    return rc
