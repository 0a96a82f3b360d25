"""Edge cases through the C ABI, against the oracle: an empty chunk, fully flagged data, rows in
random order with whole baselines missing, and the widest supports each kernel accepts (the
generic AA kernel at W = 21, a w-projection plane of half support 31, i.e. class width 63), plus
the rejections beyond them.  Bars as in test_gpu_parity.
"""
import ctypes

import numpy as np
import pytest

from common import (MODES_PSF, make_spec, rel_err, round_to, set_options, small_obs,
                    wproj_spec)
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, bullseye_stats, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = {("double", PATH_FP64): (1e-11, 1e-9), ("mixed", PATH_FP32_KAHAN): (1e-6, 2e-5)}


def oracle_grid(obs, spec, mode):
    ref = Binding("double", obs, spec, with_psf=mode in MODES_PSF)
    norm = oracle.run(ref, mode, normalize=False)
    return (ref.psf if mode in MODES_PSF else ref.output), norm


def grid_chunks(precision, path, obs, spec, mode, chunks, halves=None):
    """Grid the row ranges in `chunks` as successive chunks of one imaging pass."""
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    if halves is not None:
        arr = (ctypes.c_uint32 * len(halves))(*halves)
        assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    psf = mode in MODES_PSF
    bs = [Binding(precision, obs, spec, rows=r, with_psf=psf) for r in chunks]
    for b in bs[:-1]:
        b.params.is_final_data_chunk = False
    try:
        assert lib.initLibrary(bs[0].byref()) == 0, lib.bullseye_last_error()
        for b in bs:
            b.params.normalization_terms = bs[0].params.normalization_terms
            b.params.output_buffer = bs[0].params.output_buffer
            if psf:
                b.params.sampling_function_buffer = bs[0].params.sampling_function_buffer
            assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
            lib.gridding_barrier()
        p = bs[0].params
        n = p.num_facet_centres * p.cube_channel_dim_size * \
            p.number_of_polarization_terms_being_gridded
        norm = np.ctypeslib.as_array(
            ctypes.cast(p.normalization_terms, ctypes.POINTER(ctypes.c_double)), (n,)).copy()
    finally:
        lib.releaseLibrary()
        if halves is not None:
            lib.bullseye_set_wplane_support(None, 0)
    return (bs[0].psf if psf else bs[0].output), norm


def check(got, want, precision, path):
    rms, peak = rel_err(got, want)
    tol = TOL[(precision, path)]
    assert np.abs(want).max() > 0
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)


@pytest.mark.parametrize("mode", ["facet_single_pol", "facet_4_cor_corrections",
                                  "grid_sampling_function"])
def test_empty_chunks_contribute_nothing(mode):
    """An empty first chunk and an empty last chunk around the data."""
    print("case empty chunks", mode, flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    obs = round_to(small_obs(mode, seed=41), precision)
    spec = make_spec(obs, mode, precision=precision)
    want, want_norm = oracle_grid(obs, spec, mode)
    got, norm = grid_chunks(precision, path, obs, spec, mode,
                            [(0, 0), (0, obs.rows), (obs.rows, obs.rows)])
    check(got, want, precision, path)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


@pytest.mark.parametrize("mode", ["facet_single_pol", "grid_4_cor", "facet_sampling_function"])
def test_fully_flagged_data(mode):
    """Every sample flagged: empty grids and zero sums for the visibility modes; the sampling
    function ignores sample flags (reference quirk) and still matches the oracle."""
    print("case all flagged", mode, flush=True)
    precision, path = "double", PATH_FP64
    obs = round_to(small_obs(mode, seed=43), precision)
    obs.flags[:] = True
    spec = make_spec(obs, mode, precision=precision)
    want, want_norm = oracle_grid(obs, spec, mode)
    got, norm = grid_chunks(precision, path, obs, spec, mode, [(0, obs.rows)])
    if mode in MODES_PSF:
        check(got, want, precision, path)
    else:
        assert np.abs(want).max() == 0 and np.all(want_norm == 0)
        assert np.abs(got).max() == 0 and np.all(norm == 0)


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
def test_shuffled_rows_and_missing_baselines(kernel):
    """Rows in random order (not time- or baseline-sorted), a quarter of the baselines absent
    and a tenth of the remaining rows dropped: the device sort must still grid every row once."""
    print("case shuffled rows", kernel, flush=True)
    precision, path = "double", PATH_FP64
    mode = "facet_single_pol"
    obs = round_to(small_obs(mode, seed=45), precision)
    rng = np.random.default_rng(5)
    bl = obs.ant1.astype(np.int64) * obs.n_ant + obs.ant2
    gone = rng.choice(np.unique(bl), len(np.unique(bl)) // 4, replace=False)
    keep = ~np.isin(bl, gone) & (rng.random(obs.rows) > 0.1)
    order = rng.permutation(np.flatnonzero(keep))
    for name in ("uvw", "vis", "weights", "flags", "flagged_rows", "field", "spw", "ant1",
                 "ant2", "time_idx"):
        setattr(obs, name, getattr(obs, name)[order])
    if kernel == "aa":
        spec, halves = make_spec(obs, mode, precision=precision), None
    else:
        spec, halves = wproj_spec(obs, mode, precision, S=7, OS=4, planes=8,
                                  halves=[1, 2, 3, 4, 5, 6, 7, 7])
    want, want_norm = oracle_grid(obs, spec, mode)
    got, norm = grid_chunks(precision, path, obs, spec, mode, [(0, obs.rows)], halves)
    check(got, want, precision, path)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


@pytest.mark.parametrize("precision,path", [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN)])
def test_widest_aa_support(precision, path):
    """W = 21 (conv_support 10): the generic AA kernel's widest footprint."""
    print("case AA W=21", precision, path, flush=True)
    mode = "facet_duel_pol"
    obs = round_to(small_obs(mode, seed=47), precision)
    spec = make_spec(obs, mode, precision=precision, support=10, oversample=15, npix=192)
    want, want_norm = oracle_grid(obs, spec, mode)
    got, norm = grid_chunks(precision, path, obs, spec, mode, [(0, obs.rows)])
    check(got, want, precision, path)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


@pytest.mark.parametrize("precision,path", [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN)])
def test_widest_wproj_support(precision, path):
    """A plane of half support 31: class width 63, four residue pairs per lane."""
    print("case w-projection M=63", precision, path, flush=True)
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=49), precision)
    spec, halves = wproj_spec(obs, mode, precision, S=31, OS=2, planes=4, halves=[3, 15, 31, 31],
                              npix=256)
    want, want_norm = oracle_grid(obs, spec, mode)
    got, norm = grid_chunks(precision, path, obs, spec, mode, [(0, obs.rows)], halves)
    check(got, want, precision, path)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


def test_supports_beyond_the_kernels_are_rejected():
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=51), "mixed")
    lib = load_library("GPU", "mixed")
    spec = make_spec(obs, mode, precision="mixed", support=11, oversample=7, npix=128)
    b = Binding("mixed", obs, spec)
    assert lib.initLibrary(b.byref()) == 0
    try:
        assert lib.grid_single_pol(b.byref()) != 0
        assert b"too large" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
    spec, _ = wproj_spec(obs, mode, "mixed", S=32, OS=2, planes=2, npix=256)
    b = Binding("mixed", obs, spec)
    try:
        assert lib.initLibrary(b.byref()) != 0
        assert b"exceeds 31" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()


def test_failed_call_leaves_the_library_usable():
    """A call that fails midway (a NULL row array) reports the error; the next call of the same
    pass grids normally and the barrier's timing and counters are unaffected."""
    print("case failed call", flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    mode = "facet_single_pol"
    obs = round_to(small_obs(mode, seed=53), precision)
    spec = make_spec(obs, mode, precision=precision)
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    b = Binding(precision, obs, spec)
    bad = Binding(precision, obs, spec)
    bad.params.uvw_coords = None
    try:
        assert lib.initLibrary(b.byref()) == 0
        bad.params.output_buffer = b.params.output_buffer
        bad.params.is_final_data_chunk = False
        assert lib.facet_single_pol(bad.byref()) != 0
        assert b"NULL" in lib.bullseye_last_error()
        assert lib.facet_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.bullseye_last_error() == b"" or b"NULL" in lib.bullseye_last_error()
        st = bullseye_stats()
        lib.bullseye_get_stats(ctypes.byref(st))
        assert st.kernel_ms > 0 and st.taps > 0
        assert lib.get_gridding_walltime() > 0
        norm = np.ctypeslib.as_array(
            ctypes.cast(b.params.normalization_terms, ctypes.POINTER(ctypes.c_double)),
            (len(want_norm),)).copy()
    finally:
        lib.releaseLibrary()
    check(b.output, want, precision, path)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)
