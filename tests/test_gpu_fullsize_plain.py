"""The plain single-facet workload at full size against the FP64 oracle: 64 antennas in a 3 km
disc (2016 baselines) x 240 timesteps x 16 channels (1.40-1.42 GHz), 1 correlation, one 1024^2
grid at 2", a 7x7 PSWF oversampled x63, no w-projection -- the shape of the reference's own
benchmark (bullseye/mo/benchmarks/benchmark.cpp:97-289, one grid_single_pol facet).  The oracle
grids all 7.7 M samples (3.6e8 taps) on one core in seconds.

* FP64 path of the DOUBLE build: rms <= 1e-11 and peak <= 1e-9 of the oracle's peak.  The
  oracle runs the AA path under FE_DOWNWARD (cpu_algorithm/convolution_policies.h:73-75) and the
  device rounds its coordinate products to nearest; this bar shows that no sample of the
  workload changes its taps over that difference.
* The normalisation sum to 2e-9 relative: the oracle adds its 7.4 M per-sample terms in one
  running sum, every addition rounded towards -inf, which biases the total by up to one ulp per
  term (~4e-10 of it here); the device sums in another order, rounded to nearest.
* FP32-compensated path of the MIXED build: rms <= 1e-6 of the peak on the grid, and <= 1e-5
  RMS on the image after the same float64 shift / FFT on both sides.
"""
import ctypes

import numpy as np
import pytest

from common import rel_err, set_options
from bullseye_amd import kernels, synth
from bullseye_amd.binding import Binding, ImagingSpec
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plain():
    import torch
    print("case full-size plain: synthesising", flush=True)
    cfg = synth.CONFIGS["c0"]
    im = cfg["image"]
    obs = synth.simulate(**cfg["sim"], seed=5, device=torch.device("cuda:0"))
    # values both builds represent exactly: float32 wavelengths and kernel taps
    obs.wl = obs.wavelengths.astype(np.float32).astype(np.float64)
    conv = kernels.pswf_aa_kernel(im["support"], im["oversample"]).astype(np.float32)
    spec = ImagingSpec(npix=im["npix"], cell=im["cell"], conv=conv.astype(np.float64),
                       support=im["support"], oversample=im["oversample"])
    print("case full-size plain: oracle", flush=True)
    ref = Binding("double", obs, spec)
    norm = oracle.run(ref, "grid_single_pol", normalize=False)
    return obs, spec, ref.output.copy(), norm


def grid(precision, path, obs, spec):
    import torch
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path, device_resident=1)
    b = Binding(precision, obs, spec, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.grid_single_pol(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        out = b.host("output_buffer").astype(np.complex128)
        norm = np.ctypeslib.as_array(ctypes.cast(b.params.normalization_terms,
                                                 ctypes.POINTER(ctypes.c_double)), (1,)).copy()
    finally:
        lib.releaseLibrary()
        set_options(lib)
        del b
        torch.cuda.empty_cache()
    return out, norm


def image(g):
    n = g.shape[-1] * g.shape[-2]
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g, axes=(-2, -1))),
                           axes=(-2, -1)).real * n


def test_fp64_path_matches_oracle(plain):
    obs, spec, want, norm_want = plain
    print("case full-size plain FP64 vs oracle", flush=True)
    got, norm = grid("double", PATH_FP64, obs, spec)
    rms, peak = rel_err(got, want)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert rms <= 1e-11 and peak <= 1e-9, (rms, peak)
    np.testing.assert_allclose(norm, norm_want, rtol=2e-9)


def test_fp32_compensated_matches_oracle(plain):
    obs, spec, want, norm_want = plain
    print("case full-size plain FP32-comp vs oracle", flush=True)
    got, norm = grid("mixed", PATH_FP32_KAHAN, obs, spec)
    rms, peak = rel_err(got, want)
    irms, ipeak = rel_err(image(got[0, 0, 0] / norm[0]), image(want[0, 0, 0] / norm_want[0]))
    print("grid rms %.3e peak %.3e; image rms %.3e peak %.3e" % (rms, peak, irms, ipeak))
    assert rms <= 1e-6 and peak <= 2e-5, (rms, peak)
    assert irms <= 1e-5, irms
    np.testing.assert_allclose(norm, norm_want, rtol=2e-9)
