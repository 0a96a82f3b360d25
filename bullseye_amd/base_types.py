"""NumPy dtypes per precision build.

Mirrors bullseye/mo/base_types.py:64-112 ('single', 'double') and adds 'mixed': float64 uvw,
wavelengths and geometry with float32 visibilities, weights, kernels and grids.  Unlike the
reference, the dtypes are returned as a value object instead of module globals, so several
precisions can be used in one process (the parity tests drive all three builds).
"""
import ctypes
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class BaseTypes:
    precision: str
    grid_type: type
    psf_type: type
    fir_type: type
    w_fir_type: type
    detaper_type: type
    visibility_type: type
    visibility_component_type: type
    uvw_type: type
    uvw_ctypes_convert_type: type
    reference_wavelength_type: type
    weight_type: type
    normalization_type: type = np.float64


def force_precision(precision):
    """Return the dtype set of one precision build ('single', 'double' or 'mixed')."""
    if precision == "single":
        return BaseTypes("single", np.complex64, np.complex64, np.float32, np.complex64,
                         np.float32, np.complex64, np.float32, np.float32, ctypes.c_float,
                         np.float32, np.float32)
    if precision == "double":
        return BaseTypes("double", np.complex128, np.complex128, np.float64, np.complex128,
                         np.float64, np.complex128, np.float64, np.float64, ctypes.c_double,
                         np.float64, np.float64)
    if precision == "mixed":
        return BaseTypes("mixed", np.complex64, np.complex64, np.float32, np.complex64,
                         np.float32, np.complex64, np.float32, np.float64, ctypes.c_double,
                         np.float64, np.float32)
    raise Exception("base_types don't understand your argument.. must be 'single'/'double'/'mixed'")
