"""HIP library vs the C oracle on identical seeded inputs, through the C ABI.

Bars (relative to the oracle's peak): FP64 path into a float64 grid (DOUBLE build) rms <= 1e-11,
peak <= 1e-9; FP64 path into a float32 grid (MIXED/SINGLE builds) rms <= 2e-7, peak <= 1e-6 (the
output rounding); FP32 paths rms <= 1e-6 (Kahan) / 2e-6 (plain), peak <= 2e-5.  Normalisation
terms to 1e-12 relative (FP64 sums on every path).
"""
import ctypes

import numpy as np
import pytest

from common import (ALL_MODES, MODES_PSF, make_spec, rel_err, round_to, run_library,
                    set_options, small_obs)
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import (PATH_FP32, PATH_FP32_KAHAN, PATH_FP64, bullseye_stats,
                                         load_library)
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = {PATH_FP64: (1e-11, 1e-9), PATH_FP32_KAHAN: (1e-6, 2e-5), PATH_FP32: (2e-6, 2e-5),
       "fp64_float_grid": (2e-7, 1e-6)}


def tolerance(precision, path):
    if path == PATH_FP64 and precision != "double":
        return TOL["fp64_float_grid"]
    return TOL[path]
CASES = [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN), ("mixed", PATH_FP64),
         ("mixed", PATH_FP32), ("single", PATH_FP32_KAHAN)]


def oracle_grid(obs, spec, mode):
    ref = Binding("double", obs, spec, with_psf=mode in MODES_PSF)
    norm = oracle.run(ref, mode)
    out = ref.psf if mode in MODES_PSF else ref.output
    return out, norm


@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", ALL_MODES)
def test_mode_parity(precision, path, mode):
    print("case", precision, path, mode, flush=True)
    obs = round_to(small_obs(mode, seed=3), precision)
    spec = make_spec(obs, mode, precision=precision)
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    b = Binding(precision, obs, spec, with_psf=mode in MODES_PSF)
    norm = run_library(lib, b, mode)
    got = b.psf if mode in MODES_PSF else b.output
    rms, peak = rel_err(got, want)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert np.abs(want).max() > 0
    tol = tolerance(precision, path)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode", ["facet_single_pol", "facet_4_cor_corrections"])
def test_device_resident(mode):
    import torch
    print("case device_resident", mode, flush=True)
    obs = round_to(small_obs(mode, seed=5), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP32_KAHAN, device_resident=1)
    b = Binding("mixed", obs, spec, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    try:
        norm = run_library(lib, b, mode)
    finally:
        set_options(lib)
    rms, peak = rel_err(b.host("output_buffer"), want)
    assert rms <= 1e-6 and peak <= 2e-5, (rms, peak)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


def test_cube_channels_and_chunks():
    """Per-channel grid slices (cube > 1) and two row chunks accumulated into one grid."""
    print("case cube+chunks", flush=True)
    mode = "facet_single_pol"
    obs = round_to(small_obs(mode, seed=9), "mixed")
    spec = make_spec(obs, mode, precision="mixed", cube=4)
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP64)
    half = obs.rows // 2
    b1 = Binding("mixed", obs, spec, rows=(0, half))
    b2 = Binding("mixed", obs, spec, rows=(half, obs.rows))
    b1.params.is_final_data_chunk = False
    assert lib.initLibrary(b1.byref()) == 0
    try:
        assert lib.facet_single_pol(b1.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        b2.params.normalization_terms = b1.params.normalization_terms
        b2.params.output_buffer = b1.params.output_buffer
        assert lib.facet_single_pol(b2.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        assert lib.normalize(b2.byref()) == 0
    finally:
        lib.releaseLibrary()
    rms, peak = rel_err(b1.output, want)
    assert rms <= 2e-7 and peak <= 1e-6, (rms, peak)


def time_chunks(obs, spec, k):
    """k host-buffer chunks of whole timesteps (rows are time-ordered), each with the Jones
    terms and rebased timestamp ids of its own timesteps, as the pipeliner hands them over."""
    t = obs.time_idx.astype(np.int64)
    bounds = np.searchsorted(t, np.linspace(0, obs.n_time, k + 1).round().astype(np.int64))
    out = []
    for i in range(k):
        r0, r1 = int(bounds[i]), int(bounds[i + 1])
        b = Binding("mixed", obs, spec, rows=(r0, r1))
        if spec.jones is not None:
            t0, t1 = int(t[r0]), int(t[r1 - 1]) + 1
            b._set("timestamp_ids", (t[r0:r1] - t0).astype(np.uint64), np.uint64)
            b._set("jones_terms", spec.jones[t0:t1], b.bt.visibility_type)
            b.params.no_timestamps_read = t1 - t0
        b.params.is_final_data_chunk = i == k - 1
        out.append(b)
    return out


@pytest.mark.parametrize("mode", ["facet_4_cor_corrections", "grid_single_pol"])
def test_host_chunks_back_to_back(mode):
    """Host buffers in five chunks called back to back, no barrier between the calls: each
    call stages its inputs on the copy stream into the staging set that the call before the
    previous one read, while the previous call's kernels run.  The accumulated grid equals
    the oracle's over all rows."""
    print("case host chunks back to back", mode, flush=True)
    obs = round_to(small_obs(mode, seed=21, n_time=25), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP64)
    chunks = time_chunks(obs, spec, 5)
    assert lib.initLibrary(chunks[0].byref()) == 0, lib.bullseye_last_error()
    for bc in chunks[1:]:   # one accumulation: the normalisation array initLibrary published
        bc.params.normalization_terms = chunks[0].params.normalization_terms
        bc.params.output_buffer = chunks[0].params.output_buffer
    try:
        for bc in chunks:
            assert getattr(lib, mode)(bc.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        norm = np.ctypeslib.as_array(
            ctypes.cast(chunks[0].params.normalization_terms, ctypes.POINTER(ctypes.c_double)),
            (len(want_norm),)).copy()
        assert lib.normalize(chunks[-1].byref()) == 0
    finally:
        lib.releaseLibrary()
    rms, peak = rel_err(chunks[0].output, want)
    assert rms <= 2e-7 and peak <= 1e-6, (rms, peak)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-9)


def test_singular_jones_reports_error():
    print("case singular jones", flush=True)
    mode = "facet_4_cor_corrections"
    obs = small_obs(mode, seed=2, n_ant=6, n_time=4, n_chan=2)
    spec = make_spec(obs, mode, precision="mixed")
    spec.jones[1, 2] = 0
    lib = load_library("GPU", "mixed")
    set_options(lib)
    b = Binding("mixed", obs, spec)
    assert lib.initLibrary(b.byref()) == 0
    try:
        rc = lib.facet_4_cor_corrections(b.byref())
        assert rc != 0
        assert b"SINGULAR" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()


# A wide band (1.0-1.5 GHz in 8 channels) with short time steps (120 over 8 h): a channel step
# moves the footprint about 3x further than a time step, so the measured order is time inner.
def wideband_obs(mode, seed=4):
    return small_obs(mode, seed=seed, n_ant=10, n_time=120, f_lo=1.0e9, f_hi=1.5e9)


WALK_MODES = ["grid_single_pol", "grid_duel_pol", "facet_4_cor_corrections",
              "facet_sampling_function"]


@pytest.mark.parametrize("generic", [False, True])
@pytest.mark.parametrize("walk", [0, 1, 2])
@pytest.mark.parametrize("precision,path", [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN)])
@pytest.mark.parametrize("mode", WALK_MODES)
def test_walk_orders(mode, precision, path, walk, generic):
    """Channels-inner, time-inner and measured walks (specialised and generic Romein kernels)
    against the oracle; the measured order on this wide band is time inner."""
    print("case walk", mode, precision, path, walk, generic, flush=True)
    obs = round_to(wideband_obs(mode), precision)
    spec = make_spec(obs, mode, precision=precision)
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", precision)
    reserved = (ctypes.c_uint32 * 9)(1 if generic else 0)
    set_options(lib, compute_path=path, walk_order=walk, reserved=reserved)
    b = Binding(precision, obs, spec, with_psf=mode in MODES_PSF)
    norm = run_library(lib, b, mode)
    st = bullseye_stats()
    assert lib.bullseye_get_stats(ctypes.byref(st)) == 0
    assert st.time_inner == (0 if walk == 1 else 1)
    got = b.psf if mode in MODES_PSF else b.output
    rms, peak = rel_err(got, want)
    print("rms %.3e peak %.3e" % (rms, peak))
    tol = tolerance(precision, path)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        # FP64 sums of ~3e4 terms in another order: the bound is ~n * eps = 3e-12
        np.testing.assert_allclose(norm, want_norm, rtol=1e-11, atol=0)


def test_measured_walk_is_channels_inner_on_a_narrow_band():
    """1.40-1.42 GHz in 8 channels, 30 time steps over 8 h: channels inner."""
    obs = round_to(small_obs("grid_single_pol", seed=3), "mixed")
    spec = make_spec(obs, "grid_single_pol", precision="mixed")
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP32_KAHAN)
    b = Binding("mixed", obs, spec)
    run_library(lib, b, "grid_single_pol")
    st = bullseye_stats()
    assert lib.bullseye_get_stats(ctypes.byref(st)) == 0
    assert st.time_inner == 0
