"""Fill a gridding_parameters struct from an observation, the way bullseye's pipeliner does
(bullseye/bin/bullseye_pipeliner.py:246-305 for the set-up, :327-350 per chunk).

Arrays are converted to the dtypes of the selected precision build (base_types.force_precision)
and kept alive by the Binding object.  With device=None they stay in host memory (NumPy, the
reference's model); with a torch device they are copied to HBM once and the struct carries
device pointers (use with bullseye_options.device_resident = 1).
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .base_types import force_precision
from .gridding_parameters import gridding_parameters


@dataclass
class ImagingSpec:
    npix: int
    cell: float                     # arcsec
    conv: np.ndarray                # AA [csz] or w-projection [wplanes, csz, csz]
    support: int
    oversample: int
    facet_centres: Optional[np.ndarray] = None   # [F, 2] arcsec (None: grid_* modes)
    n_pol_gridded: int = 1
    pol_index: int = 0
    second_pol_index: int = 1
    cube: int = 1
    channel_grid_index: Optional[np.ndarray] = None  # [spw*chan] (None: all 0)
    enabled_channels: Optional[np.ndarray] = None    # [spw*chan] (None: all True)
    sf_channel_grid_index: Optional[np.ndarray] = None
    sf_count: int = 0
    wplanes: int = 1
    wmax: float = 0.0
    jones: Optional[np.ndarray] = None   # [time, ant, facet, spw, chan, 4]
    imaging_field: int = 0

    @property
    def n_facets(self):
        return 1 if self.facet_centres is None else int(self.facet_centres.shape[0])


def jones_reference_layout(obs, n_facets, dtype):
    """Broadcast the intrinsic (time, antenna) Jones set to the reference's
    [time][antenna][facet][spw][channel][4] layout (correlation_gridding_policies.h:528-532)."""
    j = obs.jones.reshape(obs.n_time, obs.n_ant, 1, 1, 1, 4)
    return np.ascontiguousarray(
        np.broadcast_to(j, (obs.n_time, obs.n_ant, n_facets, 1, obs.n_chan, 4)).astype(dtype))


class Binding:
    def __init__(self, precision, obs, spec, device=None, rows=None, with_psf=False):
        import ctypes
        self.precision = precision
        self.bt = bt = force_precision(precision)
        self.obs, self.spec, self.device = obs, spec, device
        self.keep = {}
        p = self.params = gridding_parameters(precision)
        r0, r1 = (0, obs.rows) if rows is None else rows
        nrow = r1 - r0
        C = obs.n_chan
        F = spec.n_facets
        P = spec.n_pol_gridded
        self._set("visibilities", obs.vis[r0:r1], bt.visibility_type)
        self._set("uvw_coords", obs.uvw[r0:r1], bt.uvw_type)
        self._set("reference_wavelengths", obs.wavelengths, bt.reference_wavelength_type)
        self._set("visibility_weights", obs.weights[r0:r1], bt.weight_type)
        self._set("flags", obs.flags[r0:r1], np.bool_)
        self._set("flagged_rows", obs.flagged_rows[r0:r1], np.bool_)
        self._set("field_array", obs.field[r0:r1], np.uint32)
        self._set("spw_index_array", obs.spw[r0:r1], np.uint32)
        self._set("antenna_1_ids", obs.ant1[r0:r1], np.uint32)
        self._set("antenna_2_ids", obs.ant2[r0:r1], np.uint32)
        self._set("timestamp_ids", obs.time_idx[r0:r1], np.uint64)
        p.imaging_field = spec.imaging_field
        p.baseline_count = obs.baseline_count
        p.row_count = nrow
        p.chunk_max_row_count = nrow
        p.channel_count = C
        p.number_of_polarization_terms = obs.n_corr
        p.number_of_polarization_terms_being_gridded = P
        p.spw_count = 1
        p.no_timestamps_read = obs.n_time
        p.nx = p.ny = spec.npix
        p.cell_size_x = p.cell_size_y = spec.cell
        conv_t = bt.fir_type if spec.wplanes <= 1 else bt.w_fir_type
        self._set("conv", spec.conv, conv_t)
        p.conv_support = spec.support
        p.conv_oversample = spec.oversample
        p.polarization_index = spec.pol_index
        p.second_polarization_index = spec.second_pol_index
        self.output = self._alloc("output_buffer", (F, spec.cube, P, spec.npix, spec.npix),
                                  bt.grid_type)
        p.phase_centre_ra, p.phase_centre_dec = obs.phase_centre
        fc = spec.facet_centres if spec.facet_centres is not None else np.zeros((1, 2))
        self._set("facet_centres", fc, bt.uvw_type)
        p.num_facet_centres = F
        if spec.jones is not None:
            self._set("jones_terms", spec.jones, bt.visibility_type)
            p.should_invert_jones_terms = True
        p.antenna_count = obs.n_ant
        en = spec.enabled_channels if spec.enabled_channels is not None else np.ones(C, bool)
        cgi = (spec.channel_grid_index if spec.channel_grid_index is not None
               else np.zeros(C, np.uint64))
        self._set("enabled_channels", en, np.bool_)
        self._set("channel_grid_indicies", cgi, np.uint64)
        p.cube_channel_dim_size = spec.cube
        if with_psf:
            sfc = spec.sf_count if spec.sf_count else 1
            sfi = (spec.sf_channel_grid_index if spec.sf_channel_grid_index is not None
                   else np.zeros(C, np.uint64))
            p.should_grid_sampling_function = True
            self.psf = self._alloc("sampling_function_buffer",
                                   (F, sfc, 1, spec.npix, spec.npix), bt.psf_type)
            self._set("sampling_function_channel_grid_indicies", sfi, np.uint64)
            p.sampling_function_channel_count = sfc
        p.is_final_data_chunk = True
        p.wplanes = spec.wplanes
        p.wmax_est = spec.wmax
        self._ctypes = ctypes

    # -- helpers -----------------------------------------------------------------------------
    def _to_backend(self, arr):
        if self.device is None:
            return arr, arr.ctypes.data
        import torch
        t = torch.from_numpy(np.ascontiguousarray(arr)).to(self.device)
        return t, t.data_ptr()

    def _set(self, name, arr, dtype):
        a = np.ascontiguousarray(np.asarray(arr).astype(dtype, copy=False))
        obj, ptr = self._to_backend(a)
        self.keep[name] = obj
        setattr(self.params, name, ptr)
        return obj

    def _alloc(self, name, shape, dtype):
        return self._set(name, np.zeros(shape, dtype=dtype), dtype)

    def set_array(self, name, arr, dtype):
        return self._set(name, arr, dtype)

    def host(self, name):
        obj = self.keep[name]
        if self.device is None:
            return obj
        return obj.cpu().numpy()

    def byref(self):
        return self._ctypes.byref(self.params)
