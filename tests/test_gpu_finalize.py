"""normalize, weight_uniformly, finalize, finalize_psf, the taper correction and
repack_input_data through the C ABI, against the oracle (host loops and NumPy FFT restating
fft_and_repacking_routines.cpp:83-150) on the same inputs."""
import ctypes

import numpy as np
import pytest

from common import make_spec, rel_err, round_to, set_options, small_obs
from bullseye_amd import kernels
from bullseye_amd.binding import Binding, ImagingSpec
from bullseye_amd.library_loader import PATH_FP64, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu


def float_images(buf, F, n, fstride, npix):
    """The float32 images finalize leaves at float offset f*fstride*npix of each facet block."""
    fl = np.asarray(buf).view(np.float32).reshape(-1) if np.asarray(buf).dtype != np.float32 \
        else np.asarray(buf).reshape(-1)
    return np.stack([fl[f * fstride * npix: f * fstride * npix + n * npix] for f in range(F)])


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("taper", [0, 1])
def test_finalize_matches_numpy(precision, taper):
    print("case finalize", precision, taper, flush=True)
    mode = "facet_duel_pol"
    obs = round_to(small_obs(mode, seed=21), precision)
    spec = make_spec(obs, mode, precision=precision, npix=128, cube=2)
    ref = Binding("double", obs, spec)
    oracle.run(ref, mode)
    npix = 128 * 128
    F, P, C = spec.n_facets, 2, 2
    want = oracle.finalize_images(ref.output, C)  # [F, C, ny, nx]
    t = kernels.grid_taper(spec.conv, spec.support, spec.oversample, 128)
    t2 = np.outer(t, t)
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=PATH_FP64, apply_taper_correction=taper)
    b = Binding(precision, obs, spec)
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.normalize(b.byref()) == 0
        assert lib.finalize(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.get_inversion_walltime() > 0
    finally:
        lib.releaseLibrary()
        set_options(lib)
    got = float_images(b.output, F, C, C * P, npix).reshape(want.shape)
    if taper:
        # the division is exact: image * T(x) T(y) is the uncorrected image everywhere ...
        rms, peak = rel_err(got * t2, want)
        assert rms < 1e-6 and peak < 1e-5, (rms, peak)
        # ... and the corrected image itself matches where the taper is not tiny (T >= 0.25;
        # near the edges T falls to ~2e-4 and the division amplifies float32 FFT noise)
        inner = slice(32, 96)
        rms, peak = rel_err(got[..., inner, inner], (want / t2)[..., inner, inner])
        assert rms < 2e-6 and peak < 2e-5, (rms, peak)
    else:
        rms, peak = rel_err(got, want)
        assert rms < 1e-6 and peak < 1e-5, (rms, peak)


def test_finalize_psf_and_weight_uniformly():
    print("case psf + uniform", flush=True)
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=22), "double")
    spec = make_spec(obs, mode, precision="double", npix=128)
    spec.sf_count = 1
    ref = Binding("double", obs, spec, with_psf=True)
    oracle.run(ref, "grid_sampling_function", normalize=False)
    oracle.run(ref, mode, normalize=False, weight_uniformly=True)
    oracle.load("double").oracle_normalize(ref.byref())
    want_img = oracle.finalize_images(ref.output, 1)
    want_psf = oracle.finalize_images(ref.psf, 1)
    lib = load_library("GPU", "double")
    set_options(lib)
    b = Binding("double", obs, spec, with_psf=True)
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert lib.grid_sampling_function(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.grid_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.weight_uniformly(b.byref()) == 0
        assert lib.normalize(b.byref()) == 0
        assert lib.finalize(b.byref()) == 0
        assert lib.finalize_psf(b.byref()) == 0
    finally:
        lib.releaseLibrary()
    npix = 128 * 128
    got = float_images(b.output, 1, 1, 1, npix).reshape(want_img.shape)
    got_psf = float_images(b.psf, 1, 1, 1, npix).reshape(want_psf.shape)
    assert rel_err(got, want_img)[0] < 1e-6
    assert rel_err(got_psf, want_psf)[0] < 1e-6


def test_device_resident_finalize():
    import torch
    print("case device finalize", flush=True)
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=23), "mixed")
    spec = make_spec(obs, mode, precision="mixed", npix=128)
    ref = Binding("double", obs, spec)
    oracle.run(ref, mode)
    want = oracle.finalize_images(ref.output, 1)
    lib = load_library("GPU", "mixed")
    set_options(lib, device_resident=1)
    b = Binding("mixed", obs, spec, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert lib.grid_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.normalize(b.byref()) == 0
        assert lib.finalize(b.byref()) == 0, lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
        set_options(lib)
    got = float_images(b.host("output_buffer"), 1, 1, 1, 128 * 128).reshape(want.shape)
    assert rel_err(got, want)[0] < 2e-6


def test_taper_correction_restores_point_source_flux():
    """With the taper correction an off-centre point source images at its flux; without it the
    kernel's image-plane taper attenuates it (no counterpart in the reference)."""
    print("case taper flux", flush=True)
    from bullseye_amd import synth
    obs = synth.simulate(n_ant=24, n_time=60, n_chan=4, array_size_m=1000.0, seed=3, n_src=1,
                         noise=0, flag_fraction=0)
    npix, cell = 256, 20.0
    a = np.pi / 648000
    l0, m0 = 70 * cell * a, -40 * cell * a
    k = 2 * np.pi * obs.freqs / synth.C_LIGHT
    d = obs.uvw[:, 0] * l0 + obs.uvw[:, 1] * m0 + obs.uvw[:, 2] * (np.sqrt(1 - l0**2 - m0**2) - 1)
    obs.vis[:, :, 0] = np.exp(1j * d[:, None] * k[None, :]).astype(np.complex64)
    conv = kernels.pswf_aa_kernel(3, 63)
    spec = ImagingSpec(npix=npix, cell=cell, conv=conv, support=3, oversample=63)
    lib = load_library("GPU", "double")
    peaks = []
    for taper in (0, 1):
        set_options(lib, apply_taper_correction=taper)
        b = Binding("double", obs, spec)
        try:
            assert lib.initLibrary(b.byref()) == 0
            assert lib.grid_single_pol(b.byref()) == 0
            lib.gridding_barrier()
            assert lib.normalize(b.byref()) == 0
            assert lib.finalize(b.byref()) == 0
        finally:
            lib.releaseLibrary()
        img = float_images(b.output, 1, 1, 1, npix * npix).reshape(npix, npix)
        peaks.append(img[npix // 2 - 40, npix // 2 - 70])
    set_options(lib)
    assert peaks[0] < 0.97
    assert abs(peaks[1] - 1.0) < 0.01, peaks


def test_repack_input_data_matches_reference_order():
    """repack_input_data reorders rows by baseline, time order kept, and fills the prefix scan
    (gpu_algorithm/wrapper.cu:367-454)."""
    import copy
    print("case repack", flush=True)
    obs = small_obs("grid_single_pol", seed=30, n_ant=6, n_time=5, n_chan=2)
    orig = copy.deepcopy(obs)   # the repack reorders the arrays in place
    spec = make_spec(obs, "grid_single_pol", precision="double")
    b = Binding("double", obs, spec)
    starts = np.zeros(obs.baseline_count + 1, dtype=np.uint64)
    b.params.baseline_starting_indexes = starts.ctypes.data
    lib = load_library("GPU", "double")
    set_options(lib)
    assert lib.repack_input_data(b.byref()) == 0, lib.bullseye_last_error()
    want_starts, perm = oracle.baseline_starts(orig.ant1, orig.ant2, orig.n_ant)
    p = perm.astype(int)
    np.testing.assert_array_equal(starts, want_starts)
    for name, arr in [("uvw_coords", orig.uvw), ("visibilities", orig.vis),
                      ("visibility_weights", orig.weights), ("flags", orig.flags),
                      ("antenna_1_ids", orig.ant1), ("antenna_2_ids", orig.ant2)]:
        np.testing.assert_array_equal(b.keep[name], arr[p])


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("device", [False, True])
def test_weight_uniformly_multi_slice(precision, device):
    """Sampling function in four slices (one per channel), visibilities in two grid slices with
    a non-trivial channel map, two facets: weight_uniformly (device kernel with device-resident
    buffers, host loop otherwise) then normalize, against the oracle's chain
    (cpu_algorithm/wrapper.cpp:116-163)."""
    import torch
    print("case uniform multi-slice", precision, device, flush=True)
    mode = "facet_single_pol"
    obs = round_to(small_obs(mode, seed=31, n_chan=4), precision)
    spec = make_spec(obs, mode, precision=precision, npix=128, n_facets=2)
    spec.cube = 2
    spec.channel_grid_index = np.array([1, 0, 1, 1], np.uint64)
    spec.sf_count = 4
    spec.sf_channel_grid_index = np.arange(4, dtype=np.uint64)
    ref = Binding("double", obs, spec, with_psf=True)
    oracle.run(ref, "facet_sampling_function", normalize=False)
    oracle.run(ref, mode, normalize=False, weight_uniformly=True)
    oracle.load("double").oracle_normalize(ref.byref())
    lib = load_library("GPU", precision)
    set_options(lib, device_resident=int(device), compute_path=PATH_FP64)
    dev = torch.device("cuda:0") if device else None
    b = Binding(precision, obs, spec, device=dev, with_psf=True)
    if device:
        torch.cuda.synchronize()
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.facet_sampling_function(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.weight_uniformly(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.normalize(b.byref()) == 0, lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
        set_options(lib)
    got = b.host("output_buffer")
    rms, peak = rel_err(got, ref.output)
    bar = (1e-11, 1e-9) if precision == "double" else (2e-7, 2e-6)
    assert rms < bar[0] and peak < bar[1], (rms, peak)
    # slice 0 is weighted by sampling-function channel 1 only, slice 1 by channels 0, 2 and 3
    assert np.abs(ref.output[:, 0]).max() > 0 and np.abs(ref.output[:, 1]).max() > 0


def test_device_resident_cannot_change_while_initialised():
    """bullseye_set_options refuses a device_resident change between initLibrary and
    releaseLibrary (the grids were seeded from, and publish to, one buffer model) and leaves
    the options as they were; the context keeps gridding correctly."""
    print("case options refusal", flush=True)
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=32), "double")
    spec = make_spec(obs, mode, precision="double", npix=128)
    ref = Binding("double", obs, spec)
    oracle.run(ref, mode)
    lib = load_library("GPU", "double")
    set_options(lib, device_resident=0)
    b = Binding("double", obs, spec)
    from bullseye_amd.library_loader import bullseye_options
    try:
        assert lib.initLibrary(b.byref()) == 0
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1)
        assert lib.bullseye_set_options(ctypes.byref(o)) == 1
        assert b"device_resident" in lib.bullseye_last_error()
        cur = bullseye_options(struct_size=ctypes.sizeof(bullseye_options))
        assert lib.bullseye_get_options(ctypes.byref(cur)) == 0
        assert cur.device_resident == 0
        # a compute-path change is accepted (queued work finishes on its own path first)
        set_options(lib, device_resident=0, compute_path=PATH_FP64)
        assert lib.grid_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.normalize(b.byref()) == 0
    finally:
        lib.releaseLibrary()
        set_options(lib)
    assert rel_err(b.output, ref.output)[0] < 1e-11
