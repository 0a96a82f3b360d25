"""The full-size faceting workload through properties that need no oracle run (the oracle takes
minutes at this size; the parity tests cover it at small sizes).  The workload is 64 antennas
(2016 baselines) x 240 timesteps x 64 channels, 4 correlations with Jones, 25 facets of 512^2
and a 7x7 PSWF oversampled x63.

* The FP32-compensated path (MIXED build) agrees with the FP64 path (DOUBLE build) on the same
  input values: rms <= 1e-6 and peak <= 2e-5 of each facet's peak.  The normalisation sums agree
  to 1e-9: both take the tap choice in FP64, and their ~3e7-term FP64 sums run in different
  orders.
* Gridding is additive in the visibilities: on the FP64 path grid(V1 + V2) equals
  grid(V1) + grid(V2) to 1e-12 of the peak (the grid atomics' order is the only difference).
* Reversing the facet list reverses the grids and the normalisation sums (FP64, 1e-12).
"""
import ctypes
import dataclasses

import numpy as np
import pytest

from common import set_options
from bullseye_amd import kernels, synth
from bullseye_amd.binding import Binding, ImagingSpec, jones_reference_layout
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64, load_library

pytestmark = pytest.mark.gpu

MODE = "facet_4_cor_corrections"


@pytest.fixture(scope="module")
def workload():
    print("case full-size faceting: synthesising", flush=True)
    cfg = synth.CONFIGS["c1"]
    im = cfg["image"]
    obs = synth.simulate(**cfg["sim"], seed=7)
    rng = np.random.default_rng(107)
    facets = synth.facet_centres(im["facets"], im["facet_radius_deg"], obs.phase_centre, rng)
    conv = kernels.pswf_aa_kernel(im["support"], im["oversample"]).astype(np.float32)
    # Jones values representable in both builds
    jones = jones_reference_layout(obs, im["facets"], np.complex64).astype(np.complex128)
    spec = ImagingSpec(npix=im["npix"], cell=im["cell"], conv=conv.astype(np.float64),
                       support=im["support"], oversample=im["oversample"],
                       facet_centres=facets.astype(np.float32).astype(np.float64),
                       n_pol_gridded=4, jones=jones)
    obs.uvw = obs.uvw.astype(np.float32).astype(np.float64)
    obs.wl = obs.wavelengths.astype(np.float32).astype(np.float64)
    return obs, spec


def grid(precision, path, obs, spec):
    """One facet_4_cor_corrections pass with device-resident buffers: grids and norm sums."""
    import torch
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path, device_resident=1)
    b = Binding(precision, obs, spec, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, MODE)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        out = b.host("output_buffer").astype(np.complex128)
        n = spec.n_facets * 4
        norm = np.ctypeslib.as_array(ctypes.cast(b.params.normalization_terms,
                                                 ctypes.POINTER(ctypes.c_double)), (n,)).copy()
    finally:
        lib.releaseLibrary()
        set_options(lib)
        del b
        torch.cuda.empty_cache()
    return out, norm


def per_facet_err(got, want):
    F = want.shape[0]
    g, w = got.reshape(F, -1), want.reshape(F, -1)
    pk = np.abs(w).max(axis=1)
    assert np.all(pk > 0)
    rms = np.sqrt(np.mean(np.abs(g - w) ** 2, axis=1)) / pk
    peak = np.abs(g - w).max(axis=1) / pk
    return rms.max(), peak.max()


def test_fp32_compensated_matches_fp64(workload):
    obs, spec = workload
    print("case full-size FP32-comp vs FP64", flush=True)
    g32, n32 = grid("mixed", PATH_FP32_KAHAN, obs, spec)
    g64, n64 = grid("double", PATH_FP64, obs, spec)
    rms, peak = per_facet_err(g32, g64)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert rms <= 1e-6 and peak <= 2e-5, (rms, peak)
    np.testing.assert_allclose(n32, n64, rtol=1e-9)


def test_gridding_is_additive(workload):
    obs, spec = workload
    print("case full-size additivity", flush=True)
    rng = np.random.default_rng(11)
    scale = np.abs(obs.vis).mean()
    v1 = obs.vis.astype(np.complex128)
    v2 = scale * (rng.standard_normal(v1.shape) + 1j * rng.standard_normal(v1.shape))
    g1, n1 = grid("double", PATH_FP64, dataclasses.replace(obs, vis=v1), spec)
    g2, n2 = grid("double", PATH_FP64, dataclasses.replace(obs, vis=v2), spec)
    g12, n12 = grid("double", PATH_FP64, dataclasses.replace(obs, vis=v1 + v2), spec)
    rms, peak = per_facet_err(g12, g1 + g2)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert rms <= 1e-12 and peak <= 1e-12, (rms, peak)
    # the normalisation sums do not depend on the visibility values
    np.testing.assert_allclose(n12, n1, rtol=1e-12)
    np.testing.assert_allclose(n2, n1, rtol=1e-12)


def test_reversed_facets_reverse_the_grids(workload):
    obs, spec = workload
    print("case full-size facet reversal", flush=True)
    g, n = grid("double", PATH_FP64, obs, spec)
    F = spec.n_facets
    rev = dataclasses.replace(spec, facet_centres=spec.facet_centres[::-1].copy(),
                              jones=np.ascontiguousarray(spec.jones[:, :, ::-1]))
    gr, nr = grid("double", PATH_FP64, obs, rev)
    rms, peak = per_facet_err(gr[::-1], g)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert rms <= 1e-12 and peak <= 1e-12, (rms, peak)
    np.testing.assert_allclose(nr.reshape(F, 4)[::-1], n.reshape(F, 4), rtol=1e-12)
