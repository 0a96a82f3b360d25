"""bullseye_grid_noalloc: the allocation-free gridding entry point (no counterpart in the
reference) against the entry point of the same mode, through the C ABI, on two calls in a row --
the second with fewer rows -- in one library context.  Bars: FP64 path of the DOUBLE build
1e-12 of the peak (the grids differ only in the order of the FP64 flush atomics), FP32
compensated path of the MIXED build 1e-7 (the same, then rounded to complex64); normalisation
sums to 1e-12.  Also: inputs whose tables exceed the scratch, host buffers and w-projection are
refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from common import make_spec, rel_err, round_to, set_options, small_obs, wproj_spec
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, load_library

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
MODES = ["grid_single_pol", "facet_single_pol", "grid_duel_pol", "facet_duel_pol", "grid_4_cor",
         "facet_4_cor", "facet_4_cor_corrections", "grid_sampling_function",
         "facet_sampling_function"]


def outputs(b, psf, nnorm):
    torch.cuda.synchronize()
    g = b.host("sampling_function_buffer" if psf else "output_buffer").copy()
    norm = np.ctypeslib.as_array(ctypes.cast(b.params.normalization_terms,
                                             ctypes.POINTER(ctypes.c_double)), (nnorm,)).copy()
    return g, norm


@pytest.mark.parametrize("mode,precision,path", [
    ("grid_single_pol", "double", PATH_FP64),
    ("facet_4_cor_corrections", "mixed", PATH_FP32_KAHAN),
    ("facet_sampling_function", "mixed", PATH_FP32_KAHAN),
    ("facet_duel_pol", "double", PATH_FP64)])
def test_noalloc_matches_entry_point(mode, precision, path):
    print("case noalloc", mode, precision, flush=True)
    psf = mode.endswith("sampling_function")
    obs = round_to(small_obs(mode, seed=41, n_ant=12, n_time=30, n_chan=8), precision)
    spec = make_spec(obs, mode, precision=precision, npix=128, n_facets=3)
    lib = load_library("GPU", precision)
    set_options(lib, device_resident=1, compute_path=path)
    b = Binding(precision, obs, spec, device=DEV, with_psf=psf)
    nnorm = spec.n_facets * spec.cube * spec.n_pol_gridded
    bar = 1e-12 if precision == "double" else 1e-7
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        for rows in (obs.rows, obs.rows // 2 + 7):   # the second call grids fewer rows
            b.params.row_count = rows
            assert lib.bullseye_reset_grids() == 0
            assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
            lib.gridding_barrier()
            want, nwant = outputs(b, psf, nnorm)
            assert np.abs(want).max() > 0
            assert lib.bullseye_reset_grids() == 0
            rc = lib.bullseye_grid_noalloc(b.byref(), MODES.index(mode))
            assert rc == 0, lib.bullseye_last_error()
            lib.gridding_barrier()
            got, ngot = outputs(b, psf, nnorm)
            rms, peak = rel_err(got, want)
            print("rows %d: rms %.3e peak %.3e" % (rows, rms, peak), flush=True)
            assert rms <= bar and peak <= 10 * bar, (rows, rms, peak)
            if not psf:
                np.testing.assert_allclose(ngot, nwant, rtol=1e-12)
    finally:
        lib.releaseLibrary()
        set_options(lib)


def test_noalloc_refusals():
    """Refused before any device work: tables beyond the scratch (row_count raised past what
    64 MiB holds; no array is read), host buffers, w-projection."""
    print("case noalloc refusals", flush=True)
    mode = "grid_single_pol"
    obs = round_to(small_obs(mode, seed=42, n_ant=8, n_time=10, n_chan=4), "double")
    spec = make_spec(obs, mode, precision="double", npix=128)
    lib = load_library("GPU", "double")
    set_options(lib, device_resident=1)
    b = Binding("double", obs, spec, device=DEV)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        b.params.row_count = 2_000_000
        assert lib.bullseye_grid_noalloc(b.byref(), 0) == 1
        assert b"scratch" in lib.bullseye_last_error()
        b.params.row_count = obs.rows
        assert lib.bullseye_grid_noalloc(b.byref(), 0) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
    finally:
        lib.releaseLibrary()
        set_options(lib)
    # host buffers
    set_options(lib, device_resident=0)
    bh = Binding("double", obs, spec)
    try:
        assert lib.initLibrary(bh.byref()) == 0
        assert lib.bullseye_grid_noalloc(bh.byref(), 0) == 1
        assert b"device_resident" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
    # w-projection
    spec_w, halves = wproj_spec(obs, "grid_single_pol", "double", S=3, OS=4, planes=4)
    set_options(lib, device_resident=1)
    bw = Binding("double", obs, spec_w, device=DEV)
    try:
        assert lib.initLibrary(bw.byref()) == 0, lib.bullseye_last_error()
        assert lib.bullseye_grid_noalloc(bw.byref(), 0) == 1
        assert b"AA kernel" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
        set_options(lib)
