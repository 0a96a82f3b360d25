"""Load one build of the imaging library.

Mirrors bullseye/mo/library_loader.py:42-59 (load_library(architecture, precision) with
ctypes.pydll, so the GIL is held during calls as in the reference).  'GPU' and 'MI355X' both
select the gfx950 builds; 'mixed' is the float64-geometry / float32-data build.  The restypes
of the extension entry points are declared here; the reference's 19 symbols keep ctypes'
default int restype except the two walltime getters.
"""
import ctypes
import os

from . import LIB_DIR, LIBRARY_NAMES

ENTRY_POINTS = (
    "initLibrary", "releaseLibrary", "weight_uniformly", "normalize", "repack_input_data",
    "finalize", "finalize_psf", "grid_single_pol", "facet_single_pol", "grid_duel_pol",
    "facet_duel_pol", "grid_4_cor", "facet_4_cor", "facet_4_cor_corrections",
    "grid_sampling_function", "facet_sampling_function",
)
ALL_SYMBOLS = ENTRY_POINTS + ("get_gridding_walltime", "get_inversion_walltime",
                              "gridding_barrier")
EXTENSION_SYMBOLS = ("bullseye_last_error", "bullseye_set_options", "bullseye_get_options",
                     "bullseye_reset_grids", "bullseye_get_stats", "bullseye_set_wplane_support",
                     "bullseye_degrid", "bullseye_form_stokes", "bullseye_dft_image",
                     "bullseye_dft_predict", "bullseye_batch_kernel_set",
                     "bullseye_batch_set_tuning", "bullseye_batch_get_tuning",
                     "bullseye_batch_plan", "bullseye_batch_execute", "bullseye_batch_autotune",
                     "bullseye_batch_release", "bullseye_fp64_atomic_rate",
                     "bullseye_grid_noalloc")


class bullseye_options(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("device_resident", ctypes.c_uint32),
        ("compute_path", ctypes.c_uint32),
        ("apply_taper_correction", ctypes.c_uint32),
        ("channels_per_stream", ctypes.c_uint32),
        ("rows_per_stream", ctypes.c_uint32),
        ("walk_order", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 9),
    ]


class bullseye_stats(ctypes.Structure):
    _fields_ = [
        ("samples_gridded", ctypes.c_uint64),
        ("taps", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("prep_ms", ctypes.c_double),
        ("copy_ms", ctypes.c_double),
        ("time_inner", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class bullseye_batch_facet(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32),
        ("ny", ctypes.c_uint32),
        ("kernel_set", ctypes.c_uint32),
        ("sampling_function", ctypes.c_uint32),
        ("cell_x", ctypes.c_double),
        ("cell_y", ctypes.c_double),
        ("ra", ctypes.c_double),
        ("dec", ctypes.c_double),
        ("output", ctypes.c_void_p),
        ("normalization", ctypes.c_void_p),
    ]


class bullseye_batch_tuning(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("channels_per_stream", ctypes.c_uint32),
        ("base_rows", ctypes.c_uint32),
        ("tuned", ctypes.c_uint32),
        ("shape", ctypes.c_uint32 * 32),
        ("chunk", ctypes.c_uint32 * 32),
        ("facet_major", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("tuned_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
    ]


PATH_DEFAULT, PATH_FP64, PATH_FP32_KAHAN, PATH_FP32 = 0, 1, 2, 3
BATCH_PERSISTENT_GRAPH, BATCH_PERSISTENT, BATCH_PER_FACET = 0, 1, 2
SHAPE_DEFAULT, SHAPE_COLUMN, SHAPE_PAIRS1, SHAPE_PAIRS2, SHAPE_PAIRS4 = 0, 1, 2, 3, 4


def library_path(precision):
    if precision not in LIBRARY_NAMES:
        raise Exception("Invalid precision mode selected, only single, double or mixed allowed")
    return os.path.join(LIB_DIR, LIBRARY_NAMES[precision])


def _share_torch_runtime():
    """PyTorch ships its own HIP runtime whose SONAME (libamdhip64.so.7) equals the system one.
    Loading torch first makes the dynamic linker satisfy this library's dependency with torch's
    copy, so one HIP runtime serves both (two runtimes in one process do not share devices)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(architecture, precision, share_torch_runtime=True):
    """Load the imaging library; raises like the reference on a bad argument and loudly when
    the HIP build is missing (there is no CPU fallback in this package)."""
    if share_torch_runtime:
        _share_torch_runtime()
    if architecture not in ("GPU", "MI355X"):
        raise Exception("Invalid architecture, only GPU or MI355X allowed (this package ships "
                        "no CPU imaging library)")
    path = library_path(precision)
    if not os.path.exists(path):
        raise OSError("HIP imaging library not built: %s (run __graft_entry__.build())" % path)
    lib = ctypes.pydll.LoadLibrary(path)
    lib.get_gridding_walltime.restype = ctypes.c_double
    lib.get_inversion_walltime.restype = ctypes.c_double
    lib.gridding_barrier.restype = None
    lib.bullseye_last_error.restype = ctypes.c_char_p
    lib.bullseye_set_options.argtypes = [ctypes.POINTER(bullseye_options)]
    lib.bullseye_get_options.argtypes = [ctypes.POINTER(bullseye_options)]
    lib.bullseye_get_stats.argtypes = [ctypes.POINTER(bullseye_stats)]
    lib.bullseye_fp64_atomic_rate.argtypes = [ctypes.POINTER(ctypes.c_double)]
    lib.bullseye_degrid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p]
    lib.bullseye_grid_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.bullseye_dft_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.bullseye_dft_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                         ctypes.c_void_p]
    lib.bullseye_batch_kernel_set.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_float]
    lib.bullseye_batch_set_tuning.argtypes = [ctypes.POINTER(bullseye_batch_tuning)]
    lib.bullseye_batch_get_tuning.argtypes = [ctypes.POINTER(bullseye_batch_tuning)]
    lib.bullseye_batch_plan.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(bullseye_batch_facet), ctypes.c_uint32,
                                        ctypes.c_int]
    lib.bullseye_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(bullseye_batch_facet), ctypes.c_uint32,
                                            ctypes.POINTER(bullseye_batch_tuning)]
    return lib
