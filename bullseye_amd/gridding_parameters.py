"""ctypes mirror of the gridding_parameters struct (include/bullseye_amd.h).

Field order and types follow bullseye/mo/gridding_parameters.py:46-112, which mirrors
bullseye/mo/cpu_gpu_common/gridding_parameters.h:44-109.  The only precision-dependent
fields are the uvw_base_type scalars (cell sizes, phase centre, wmax_est).
"""
import ctypes
from ctypes import c_bool, c_size_t, c_uint, c_void_p

from .base_types import force_precision

_CLASSES = {}


def _fields(uvw_ct):
    return [
        ("visibilities", c_void_p),
        ("uvw_coords", c_void_p),
        ("reference_wavelengths", c_void_p),
        ("visibility_weights", c_void_p),
        ("flags", c_void_p),
        ("flagged_rows", c_void_p),
        ("field_array", c_void_p),
        ("spw_index_array", c_void_p),
        ("imaging_field", c_uint),
        ("baseline_count", c_size_t),
        ("row_count", c_size_t),
        ("chunk_max_row_count", c_size_t),
        ("channel_count", c_size_t),
        ("number_of_polarization_terms", c_size_t),
        ("number_of_polarization_terms_being_gridded", c_size_t),
        ("spw_count", c_size_t),
        ("no_timestamps_read", c_size_t),
        ("nx", c_size_t),
        ("ny", c_size_t),
        ("cell_size_x", uvw_ct),
        ("cell_size_y", uvw_ct),
        ("conv", c_void_p),
        ("conv_support", c_size_t),
        ("conv_oversample", c_size_t),
        ("polarization_index", c_size_t),
        ("second_polarization_index", c_size_t),
        ("output_buffer", c_void_p),
        ("phase_centre_ra", uvw_ct),
        ("phase_centre_dec", uvw_ct),
        ("facet_centres", c_void_p),
        ("num_facet_centres", c_size_t),
        ("jones_terms", c_void_p),
        ("should_invert_jones_terms", c_bool),
        ("antenna_1_ids", c_void_p),
        ("antenna_2_ids", c_void_p),
        ("timestamp_ids", c_void_p),
        ("antenna_count", c_size_t),
        ("enabled_channels", c_void_p),
        ("channel_grid_indicies", c_void_p),
        ("cube_channel_dim_size", c_size_t),
        ("should_grid_sampling_function", c_bool),
        ("sampling_function_buffer", c_void_p),
        ("sampling_function_channel_grid_indicies", c_void_p),
        ("sampling_function_channel_count", c_size_t),
        ("is_final_data_chunk", c_bool),
        ("wplanes", c_size_t),
        ("wmax_est", uvw_ct),
        ("baseline_starting_indexes", c_void_p),
        ("antenna_jones_starting_indexes", c_void_p),
        ("jones_time_indicies_per_antenna", c_void_p),
        ("normalization_terms", c_void_p),
    ]


def gridding_parameters_class(precision):
    """The ctypes Structure class for one precision build."""
    if precision not in _CLASSES:
        bt = force_precision(precision)
        cls = type("gridding_parameters_%s" % precision, (ctypes.Structure,),
                   {"_fields_": _fields(bt.uvw_ctypes_convert_type)})
        _CLASSES[precision] = cls
    return _CLASSES[precision]


def gridding_parameters(precision):
    """A zero-initialised struct instance for one precision build."""
    return gridding_parameters_class(precision)()
