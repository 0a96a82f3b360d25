"""The degridding adjoint (bullseye_degrid) through the C ABI.

* Parity with the oracle's adjoint (oracle_degrid, pinned by test_oracle_adjoint) on identical
  seeded inputs, for every visibility mode, the AA and w-projection kernels, each ABI and
  compute path.  Bars relative to the oracle's peak: FP64 into float64 rms <= 1e-11, peak
  <= 1e-9; FP64 into float32 visibilities rms <= 2e-7, peak <= 1e-6 (output rounding); FP32
  paths rms <= 1e-6, peak <= 1e-5.
* The dot-product test on the library's own pair: Re<G v, g> = Re<v, G^T g> to <= 1e-6
  relative on the FP32 path (<= 1e-11 on FP64 into float64), with g = G v / |G v| plus an
  independent random grid of the same norm, so <G v, g> is of order |G v| and no cancellation
  hides an error.
"""
import ctypes

import numpy as np
import pytest

from common import make_spec, rel_err, round_to, set_options, small_obs, wproj_spec
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import (PATH_FP32, PATH_FP32_KAHAN, PATH_FP64, load_library)
from oracle import oracle

pytestmark = pytest.mark.gpu

VIS_MODES = ["grid_single_pol", "facet_single_pol", "grid_duel_pol", "facet_duel_pol",
             "grid_4_cor", "facet_4_cor", "facet_4_cor_corrections"]
MODE_ID = {m: i for i, m in enumerate(VIS_MODES)}
CASES = [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN), ("mixed", PATH_FP64),
         ("mixed", PATH_FP32), ("single", PATH_FP32_KAHAN)]


def tolerance(precision, path):
    if path == PATH_FP64:
        return (1e-11, 1e-9) if precision == "double" else (2e-7, 1e-6)
    return (1e-6, 1e-5)


def random_obs(mode, seed):
    obs = small_obs(mode, seed=seed, n_ant=10, n_time=20, n_chan=6)
    rng = np.random.default_rng(seed + 100)
    obs.vis = (rng.normal(size=obs.vis.shape) + 1j * rng.normal(size=obs.vis.shape))
    obs.weights = rng.uniform(0.5, 2.0, obs.weights.shape).astype(np.float32)
    return obs


def spec_for(obs, mode, precision, kernel):
    if kernel == "aa":
        return make_spec(obs, mode, precision=precision, npix=96), None
    return wproj_spec(obs, mode, precision, S=5, OS=4, planes=6, halves=[1, 2, 3, 4, 5, 5],
                      npix=96)


def open_lib(precision, path, halves):
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    if halves is not None:
        arr = (ctypes.c_uint32 * len(halves))(*halves)
        assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    else:
        lib.bullseye_set_wplane_support(None, 0)
    return lib


def degrid(lib, b, mode, model):
    out = np.zeros(b.keep["visibilities"].shape, dtype=b.bt.visibility_type)
    m = np.ascontiguousarray(model.astype(b.bt.grid_type))
    rc = lib.bullseye_degrid(b.byref(), MODE_ID[mode], m.ctypes.data, out.ctypes.data)
    assert rc == 0, lib.bullseye_last_error()
    return out


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
@pytest.mark.parametrize("precision,path", CASES)
@pytest.mark.parametrize("mode", VIS_MODES)
def test_degrid_parity(mode, precision, path, kernel):
    print("case degrid", mode, precision, path, kernel, flush=True)
    obs = round_to(random_obs(mode, seed=21), precision)
    spec, halves = spec_for(obs, mode, precision, kernel)
    ref = Binding("double", obs, spec)
    rng = np.random.default_rng(5)
    model = rng.normal(size=ref.output.shape) + 1j * rng.normal(size=ref.output.shape)
    model = model.astype(ref.bt.grid_type) if precision == "double" else \
        model.astype(np.complex64).astype(np.complex128)
    want = oracle.degrid(ref, mode, model)
    lib = open_lib(precision, path, halves)
    b = Binding(precision, obs, spec)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        got = degrid(lib, b, mode, model)
    finally:
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
    rms, peak = rel_err(got, want)
    print("rms %.3e peak %.3e" % (rms, peak))
    assert np.abs(want).max() > 0
    tol = tolerance(precision, path)
    assert rms <= tol[0] and peak <= tol[1], (rms, peak)


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
@pytest.mark.parametrize("precision,path", [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN)])
@pytest.mark.parametrize("mode", ["facet_single_pol", "facet_4_cor_corrections", "grid_duel_pol"])
def test_library_dot_product(mode, precision, path, kernel):
    print("case dot", mode, precision, path, kernel, flush=True)
    obs = round_to(random_obs(mode, seed=22), precision)
    spec, halves = spec_for(obs, mode, precision, kernel)
    lib = open_lib(precision, path, halves)
    b = Binding(precision, obs, spec)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        gv = b.output.astype(np.complex128)
        rng = np.random.default_rng(9)
        noise = rng.normal(size=gv.shape) + 1j * rng.normal(size=gv.shape)
        g = gv / np.linalg.norm(gv) + noise / np.linalg.norm(noise)
        g = g.astype(b.bt.grid_type).astype(np.complex128)
        gtg = degrid(lib, b, mode, g).astype(np.complex128)
    finally:
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
    v = b.keep["visibilities"].astype(np.complex128)
    lhs = np.vdot(gv, g).real
    rhs = np.vdot(v, gtg).real
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("lhs %.9e rhs %.9e rel %.3e" % (lhs, rhs, rel))
    assert rel <= (1e-11 if precision == "double" else 1e-6), rel


def test_degrid_device_resident():
    import torch
    mode = "facet_4_cor_corrections"
    obs = round_to(random_obs(mode, seed=23), "mixed")
    spec, _ = spec_for(obs, mode, "mixed", "aa")
    ref = Binding("double", obs, spec)
    rng = np.random.default_rng(6)
    model = (rng.normal(size=ref.output.shape) + 1j * rng.normal(size=ref.output.shape))
    model = model.astype(np.complex64)
    want = oracle.degrid(ref, mode, model.astype(np.complex128))
    lib = open_lib("mixed", PATH_FP32_KAHAN, None)
    set_options(lib, compute_path=PATH_FP32_KAHAN, device_resident=1)
    dev = torch.device("cuda:0")
    b = Binding("mixed", obs, spec, device=dev)
    m = torch.from_numpy(model).to(dev)
    out = torch.zeros(b.keep["visibilities"].shape, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert lib.bullseye_degrid(b.byref(), MODE_ID[mode], m.data_ptr(), out.data_ptr()) == 0, \
            lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
        set_options(lib)
    rms, peak = rel_err(out.cpu().numpy(), want)
    assert rms <= 1e-6 and peak <= 1e-5, (rms, peak)


def test_degrid_rejects_sampling_function_modes():
    obs = round_to(random_obs("grid_single_pol", seed=24), "mixed")
    spec, _ = spec_for(obs, "grid_single_pol", "mixed", "aa")
    lib = open_lib("mixed", PATH_FP32_KAHAN, None)
    b = Binding("mixed", obs, spec)
    out = np.zeros(b.keep["visibilities"].shape, np.complex64)
    try:
        assert lib.initLibrary(b.byref()) == 0
        rc = lib.bullseye_degrid(b.byref(), 7, b.output.ctypes.data, out.ctypes.data)
        assert rc != 0 and b"visibility mode" in lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
