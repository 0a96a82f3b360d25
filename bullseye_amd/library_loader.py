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
                     "bullseye_dft_predict")


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


PATH_DEFAULT, PATH_FP64, PATH_FP32_KAHAN, PATH_FP32 = 0, 1, 2, 3


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
    lib.bullseye_degrid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p]
    lib.bullseye_dft_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.bullseye_dft_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                         ctypes.c_void_p]
    return lib
