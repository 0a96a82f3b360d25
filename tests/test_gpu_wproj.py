"""w-projection gridding (2D complex kernels, conj for w < 0, plane choice, per-plane supports)
through the C ABI against the oracle (cpu_algorithm/convolution_policies.h:231-289).

Kernels are random complex stacks truncated per plane (kernels.truncate_wkernels), so no
symmetry hides a transposed or mirrored tap.  Bars as in test_gpu_parity."""
import ctypes

import numpy as np
import pytest

from common import (ALL_MODES, MODES_PSF, make_spec, rel_err, round_to, set_options,
                    small_obs, wproj_spec)
from bullseye_amd import kernels
from bullseye_amd.base_types import force_precision
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu


def wspec(obs, mode, precision, S=7, OS=4, planes=8, halves=None, seed=0, npix=128):
    return wproj_spec(obs, mode, precision, S=S, OS=OS, planes=planes, halves=halves,
                      seed=seed, npix=npix)


def run_case(precision, path, mode, halves_fn, seed=3, S=7, npix=128, walk=0):
    obs = round_to(small_obs(mode, seed=seed), precision)
    spec, halves = wspec(obs, mode, precision, S=S, halves=halves_fn(8, S), npix=npix)
    ref = Binding("double", obs, spec, with_psf=mode in MODES_PSF)
    want_norm = oracle.run(ref, mode)
    want = ref.psf if mode in MODES_PSF else ref.output
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path, walk_order=walk)
    arr = (ctypes.c_uint32 * len(halves))(*halves)
    assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    b = Binding(precision, obs, spec, with_psf=mode in MODES_PSF)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        norm = np.ctypeslib.as_array(
            ctypes.cast(b.params.normalization_terms, ctypes.POINTER(ctypes.c_double)),
            (len(want_norm),)).copy()
        if mode not in MODES_PSF:
            assert lib.normalize(b.byref()) == 0
    finally:
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
    got = b.psf if mode in MODES_PSF else b.output
    return rel_err(got, want), norm, want_norm, want


HALVES = {"uniform": lambda n, s: [s] * n,
          "growing": lambda n, s: [1, 2, 3, 4, 5, 6, 7, 7][:n]}
CASES = [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN), ("mixed", PATH_FP64)]


@pytest.mark.parametrize("halves", ["uniform", "growing"])
@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", ALL_MODES)
def test_wproj_parity(mode, precision, path, halves):
    print("case wproj", mode, precision, path, halves, flush=True)
    (rms, peak), norm, want_norm, want = run_case(precision, path, mode, HALVES[halves])
    print("rms %.3e peak %.3e" % (rms, peak))
    assert np.abs(want).max() > 0
    tol = (1e-11, 1e-9) if (precision == "double" and path == PATH_FP64) else (1e-6, 2e-5)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-9)


# Class widths 7 and 13 take the column-ownership kernel, 19..31 the residue-pair kernel with
# one pair per lane, 39 two pairs and 47 four pairs per lane.
WIDE = lambda n, s: [3, 6, 9, 12, 15, 19, 23, 23][:n]


@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", ["grid_single_pol", "grid_duel_pol", "facet_4_cor_corrections",
                                  "grid_sampling_function"])
def test_wproj_wide_classes(mode, precision, path):
    print("case wproj wide", mode, precision, path, flush=True)
    (rms, peak), norm, want_norm, want = run_case(precision, path, mode, WIDE, S=23, npix=256)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert np.abs(want).max() > 0
    tol = (1e-11, 1e-9) if (precision == "double" and path == PATH_FP64) else (1e-6, 2e-5)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-9)


@pytest.mark.parametrize("walk", [1, 2])
@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", ["grid_single_pol", "facet_4_cor_corrections"])
def test_wproj_walk_orders(mode, precision, path, walk):
    print("case wproj walk", mode, precision, path, walk, flush=True)
    (rms, peak), norm, want_norm, want = run_case(precision, path, mode, WIDE, S=23, npix=256,
                                                  walk=walk)
    print("rms %.3e peak %.3e" % (rms, peak))
    tol = (1e-11, 1e-9) if (precision == "double" and path == PATH_FP64) else (1e-6, 2e-5)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-9)


# Grids of at least 4096^2 cells take the residue-pair kernels' on-chip windows (the window of
# each stream is flushed to the grid once at its end; flushes outside it go straight to the
# grid): the same oracle bars on a 4096^2 grid, where the 20" cells make the footprints move
# across many cells (in and out of the windows).
@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", ["grid_single_pol", "facet_4_cor_corrections"])
def test_wproj_windows(mode, precision, path):
    print("case wproj windows", mode, precision, path, flush=True)
    (rms, peak), norm, want_norm, want = run_case(precision, path, mode, WIDE, S=23, npix=4096)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert np.abs(want).max() > 0
    tol = (1e-11, 1e-9) if (precision == "double" and path == PATH_FP64) else (1e-6, 2e-5)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)
    if mode not in MODES_PSF:
        np.testing.assert_allclose(norm, want_norm, rtol=1e-9)
