"""Direct Fourier transforms (bullseye_dft_image, bullseye_dft_predict) on the GPU.

* Against NumPy float64 brute force of the same sums (the image: weight * Re(V e^{2 pi i (u l +
  v m)}) over unflagged samples, l = (x - n/2) cell, m = -(y - n/2) cell; the predictor: the
  synthesiser's point-source model), to 1e-10 of the peak.
* Cross-checks of the gridder and degridder: a taper-corrected gridded image equals the DFT image
  over the central half of the image up to the gridding kernel's aliasing (<= 1e-2 of the peak
  with the 7x7 PSWF, 6.5e-3 on the oracle; a mirrored axis or a sign error would be O(1)), and
  degridding the FFT of a taper-divided point model reproduces the predicted visibilities to
  the same level.
"""
import ctypes

import numpy as np
import pytest
import torch

from common import make_spec, set_options, small_obs
from bullseye_amd import kernels, synth
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP64, load_library

pytestmark = pytest.mark.gpu
A2R = float(np.float32(np.pi / (180.0 * 3600.0)))


def os_squared(spec):
    return float(spec.oversample) ** 2


def lm_axes(n, cell):
    k = np.arange(n) - n // 2
    return k * cell * A2R, -k * cell * A2R


def test_dft_image_matches_numpy():
    obs = small_obs("grid_single_pol", seed=41, n_ant=6, n_time=6, n_chan=3)
    n, cell = 40, 30.0
    spec = make_spec(obs, "grid_single_pol", precision="double", npix=n, cell=cell)
    lib = load_library("GPU", "double")
    set_options(lib)
    b = Binding("double", obs, spec)
    img = np.zeros((n, n))
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert lib.bullseye_dft_image(b.byref(), img.ctypes.data) == 0, lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
    l, m = lm_axes(n, cell)
    lam = obs.wavelengths[0]
    u = obs.uvw[:, 0:1] / lam[None, :]
    v = obs.uvw[:, 1:2] / lam[None, :]
    vis = obs.vis[..., 0] * obs.weights[..., 0] * ~obs.flags[..., 0]
    want = np.zeros((n, n))
    for y in range(n):
        ph = np.exp(2j * np.pi * (u[..., None] * l[None, None, :] + (v * m[y])[..., None]))
        want[y] = (vis[..., None] * ph).real.sum(axis=(0, 1))
    np.testing.assert_allclose(img, want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_dft_predict_matches_synthesiser():
    obs = small_obs("grid_single_pol", seed=42, n_ant=6, n_time=5, n_chan=3, flag_fraction=0)
    spec = make_spec(obs, "grid_single_pol", precision="double", npix=32)
    comps = np.array([[1e-3, -2e-3, 1.0], [-4e-3, 5e-4, 0.5], [0.0, 0.0, 0.25]])
    lib = load_library("GPU", "double")
    set_options(lib)
    b = Binding("double", obs, spec)
    out = np.zeros(obs.vis.shape, np.complex128)
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert lib.bullseye_dft_predict(b.byref(), comps.ctypes.data, len(comps),
                                        out.ctypes.data) == 0, lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
    lam = obs.wavelengths[0]
    n_ = np.sqrt(1 - comps[:, 0] ** 2 - comps[:, 1] ** 2) - 1
    d = obs.uvw[:, 0:1] * comps[:, 0] + obs.uvw[:, 1:2] * comps[:, 1] + obs.uvw[:, 2:3] * n_
    want = (np.exp(2j * np.pi * d[:, None, :] / lam[None, :, None]) * comps[:, 2]).sum(-1)
    np.testing.assert_allclose(out[..., 0], want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_gridded_image_agrees_with_dft():
    """Grid (FP64, no normalisation), finalize with the taper correction: equals the DFT image
    (times 1/OS^2, the per-axis tap sum) up to aliasing of the 7x7 PSWF kernel."""
    obs = small_obs("grid_single_pol", seed=43, n_ant=8, n_time=10, n_chan=2, size_m=300.0)
    n, cell = 128, 20.0
    spec = make_spec(obs, "grid_single_pol", precision="double", npix=n, cell=cell)
    lib = load_library("GPU", "double")
    set_options(lib, compute_path=PATH_FP64, apply_taper_correction=1)
    b = Binding("double", obs, spec)
    dft = np.zeros((n, n))
    try:
        assert lib.initLibrary(b.byref()) == 0
        assert lib.grid_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.finalize(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.bullseye_dft_image(b.byref(), dft.ctypes.data) == 0
    finally:
        lib.releaseLibrary()
        set_options(lib)
    # finalize leaves a float32 image at the start of the facet's buffer (the reference's
    # in-place layout)
    img = b.output.reshape(-1).view(np.float32)[:n * n].reshape(n, n).astype(np.float64)
    # the AA array has unit sum over all its OS-times oversampled values, so a sample's 2S+1
    # taps per axis sum to ~1/OS: the unnormalised image is the DFT image over OS^2
    img *= spec.oversample ** 2
    # the central half: towards the edges the taper divisor falls to 2e-4 (T = 0.26 at n/4
    # from the centre) and amplifies the kernel's aliasing
    c = slice(n // 4, 3 * n // 4)
    err = np.abs(img[c, c] - dft[c, c]).max() / np.abs(dft).max()
    print("gridded vs DFT: max err %.3e of peak" % err)
    assert err <= 1e-2, err


def test_degridded_point_model_agrees_with_prediction():
    """Degrid the FFT of a point model divided by the kernel taper: equals the DFT prediction up
    to the kernel's interpolation error."""
    obs = small_obs("grid_single_pol", seed=44, n_ant=8, n_time=10, n_chan=2, size_m=300.0,
                    flag_fraction=0)
    obs.weights[:] = 1.0
    n, cell = 128, 20.0
    spec = make_spec(obs, "grid_single_pol", precision="double", npix=n, cell=cell)
    S, OS = spec.support, spec.oversample
    taper = kernels.grid_taper(spec.conv, S, OS, n)
    px, py = n // 2 + 9, n // 2 - 5
    model = np.zeros((n, n))
    model[py, px] = 1.0 / (taper[py] * taper[px])
    grid = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(model)))
    l = (px - n // 2) * cell * A2R
    m = -(py - n // 2) * cell * A2R
    # the image convention maps a sky direction (l0, m0) to pixel l = -l0 (phase sign of V)
    comps = np.array([[-l, -m, 1.0]])
    lib = load_library("GPU", "double")
    set_options(lib, compute_path=PATH_FP64)
    b = Binding("double", obs, spec)
    got = np.zeros(obs.vis.shape, np.complex128)
    want = np.zeros(obs.vis.shape, np.complex128)
    try:
        assert lib.initLibrary(b.byref()) == 0
        g = np.ascontiguousarray(grid.reshape(b.output.shape))
        assert lib.bullseye_degrid(b.byref(), 0, g.ctypes.data, got.ctypes.data) == 0, \
            lib.bullseye_last_error()
        assert lib.bullseye_dft_predict(b.byref(), comps.ctypes.data, 1, want.ctypes.data) == 0
    finally:
        lib.releaseLibrary()
    # the 2S+1 taps per axis sum to ~1/OS (unit-sum oversampled array); pixel (l, m) of the image
    # is the sky direction (-l, -m) of the predictor's convention; the w term of the prediction
    # (n - 1 ~ -5e-7 here) is below the bar
    sel = np.abs(got[..., 0]) > 0
    g0 = got[..., 0][sel] * os_squared(spec)
    err = np.abs(g0 - want[..., 0][sel]).max()
    print("degrid vs predict: max err %.3e (of 1), %d samples" % (err, sel.sum()))
    assert sel.sum() > 100 and err <= 2e-2, err
