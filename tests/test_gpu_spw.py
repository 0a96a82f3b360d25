"""Two spectral windows against the oracle: the row table's spw, the channel table's spw * chan + c
(staged in LDS, and read from device memory when spw * chan exceeds 1024 entries), per-spw
wavelengths, enabled channels and grid slices, and the Jones offset of (facet, spw, channel).

Rows alternate between the windows by timestep; window 1 observes 1.5-1.6 GHz, has every fifth
channel disabled, grids into slice 1 of a 2-slice cube and carries its own Jones terms (window 0's
rotated by a phase), so a wrong window index in any table moves or rescales samples.
Bars as in test_gpu_parity.
"""
import ctypes

import numpy as np
import pytest

from common import (MODES_PSF, make_spec, rel_err, round_to, set_options, small_obs,
                    wproj_spec)
from bullseye_amd import synth
from bullseye_amd.binding import Binding, force_precision
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu

NSPW = 2


def two_window_obs(mode, precision, n_chan, seed):
    obs = round_to(small_obs(mode, seed=seed, n_ant=6, n_time=6, n_chan=n_chan), precision)
    lam0 = obs.wavelengths[0]
    lam1 = synth.C_LIGHT / np.linspace(1.5e9, 1.6e9, n_chan)
    bt = force_precision(precision)
    obs.wl = np.stack([lam0, lam1]).astype(bt.reference_wavelength_type).astype(np.float64)
    obs.spw = (obs.time_idx % NSPW).astype(np.uint32)
    return obs


def two_window_spec(obs, mode, precision, wproj):
    C = obs.n_chan
    if wproj:
        spec, halves = wproj_spec(obs, mode, precision, S=7, OS=4, planes=8,
                                  halves=[1, 2, 3, 4, 5, 6, 7, 7])
        w = np.abs(obs.uvw[:, 2:3] / obs.wavelengths.min())
        spec.wmax = float(np.float32(w.max() * 0.9))
    else:
        spec, halves = make_spec(obs, mode, precision=precision, cube=2), None
    spec.cube = 2
    spec.channel_grid_index = np.repeat(np.arange(NSPW, dtype=np.uint64), C)
    en = np.ones(NSPW * C, bool)
    en[C::5] = False
    spec.enabled_channels = en
    if mode in MODES_PSF:
        spec.sf_count = 2
        spec.sf_channel_grid_index = spec.channel_grid_index
    if spec.jones is not None:
        F = spec.n_facets
        j = obs.jones.reshape(obs.n_time, obs.n_ant, 1, 1, 1, 4)
        J = np.broadcast_to(j, (obs.n_time, obs.n_ant, F, NSPW, C, 4)).astype(np.complex128)
        J = J.copy()
        J[:, :, :, 1] *= np.exp(0.3j)
        spec.jones = np.ascontiguousarray(J)
    return spec, halves


def bind(precision, obs, spec, mode):
    b = Binding(precision, obs, spec, with_psf=mode in MODES_PSF)
    b.params.spw_count = NSPW
    return b


CASES = [("facet_4_cor_corrections", "mixed", PATH_FP32_KAHAN, 8, False),
         ("facet_4_cor_corrections", "double", PATH_FP64, 8, False),
         ("facet_4_cor_corrections", "mixed", PATH_FP32_KAHAN, 600, False),
         ("grid_single_pol", "mixed", PATH_FP32_KAHAN, 600, False),
         ("facet_sampling_function", "mixed", PATH_FP32_KAHAN, 8, False),
         ("grid_single_pol", "double", PATH_FP64, 8, True)]


@pytest.mark.parametrize("mode,precision,path,n_chan,wproj", CASES)
def test_two_spectral_windows(mode, precision, path, n_chan, wproj):
    print("case two spws", mode, precision, path, n_chan, "wproj" if wproj else "aa", flush=True)
    obs = two_window_obs(mode, precision, n_chan, seed=71)
    spec, halves = two_window_spec(obs, mode, precision, wproj)
    ref = bind("double", obs, spec, mode)
    want_norm = oracle.run(ref, mode, normalize=False)
    want = ref.psf if mode in MODES_PSF else ref.output
    assert np.abs(want[:, 1]).max() > 0 and np.abs(want[:, 0]).max() > 0
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    if halves is not None:
        arr = (ctypes.c_uint32 * len(halves))(*halves)
        assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    b = bind(precision, obs, spec, mode)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        got = b.psf if mode in MODES_PSF else b.output
        if mode not in MODES_PSF:
            n = len(want_norm)
            norm = np.ctypeslib.as_array(ctypes.cast(b.params.normalization_terms,
                                                     ctypes.POINTER(ctypes.c_double)), (n,)).copy()
    finally:
        lib.releaseLibrary()
        if halves is not None:
            lib.bullseye_set_wplane_support(None, 0)
    rms, peak = rel_err(got, want)
    tol = (1e-11, 1e-9) if (precision, path) == ("double", PATH_FP64) else (1e-6, 2e-5)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        # FP64 sums of ~5e4 terms (600 channels) added in another order than the oracle's
        np.testing.assert_allclose(norm, want_norm, rtol=1e-12 if n_chan <= 100 else 1e-11)
