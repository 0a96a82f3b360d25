"""bullseye_form_stokes (Stokes formation + compaction on the device) against the oracle's
restatement of the pipeliner's host steps, bit for bit, with host and device output buffers."""
import numpy as np
import pytest
import torch

from common import make_spec, set_options, small_obs
from bullseye_amd import stokes
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import load_library
from oracle import oracle

pytestmark = pytest.mark.gpu

LIN, CIRC = [9, 10, 11, 12], [5, 6, 7, 8]
CASES = [  # (mode giving P, jones, data correlations, wanted)
    ("facet_duel_pol", False, LIN, "I"), ("facet_duel_pol", False, CIRC, "V"),
    ("facet_duel_pol", False, LIN, "U"), ("facet_4_cor_corrections", True, LIN, "I"),
    ("facet_single_pol", False, [1, 2, 3, 4], "I"), ("facet_duel_pol", False, LIN, "XY"),
]


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("mode,jones,data,wanted", CASES)
def test_form_stokes_matches_pipeliner(mode, jones, data, wanted, precision, device):
    print("case stokes", mode, jones, wanted, precision, device, flush=True)
    grid_c, feeds = stokes.correlations_to_grid(jones, wanted, data)
    first, second, sign = stokes.stokes_slots(jones, wanted, feeds, data, grid_c)
    obs = small_obs(mode, seed=1)
    cube = 3
    spec = make_spec(obs, mode, precision=precision, npix=32, n_facets=2, cube=cube)
    lib = load_library("GPU", precision)
    set_options(lib, device_resident=1 if device else 0)
    b = Binding(precision, obs, spec, device=torch.device("cuda:0") if device else None)
    P = b.params.number_of_polarization_terms_being_gridded
    rng = np.random.default_rng(7)
    shape = (2, cube, P, 32, 32)
    g0 = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(b.bt.grid_type)
    want = oracle.pipeliner_stokes(g0, jones, data, grid_c if not jones else data, feeds, wanted)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        if device:
            b.keep["output_buffer"].copy_(torch.from_numpy(g0))
            torch.cuda.synchronize()
        else:
            b.output[...] = g0
        stokes.form_stokes(lib, b, first, second, sign, wanted)
        got = b.host("output_buffer")
    finally:
        lib.releaseLibrary()
        set_options(lib)
    np.testing.assert_array_equal(got, want)


def test_contiguous_compaction():
    obs = small_obs("facet_4_cor", seed=2)
    spec = make_spec(obs, "facet_4_cor", precision="mixed", npix=16, n_facets=2, cube=3)
    lib = load_library("GPU", "mixed")
    set_options(lib)
    b = Binding("mixed", obs, spec)
    rng = np.random.default_rng(8)
    g0 = (rng.normal(size=b.output.shape) + 1j * rng.normal(size=b.output.shape)).astype(np.complex64)
    try:
        assert lib.initLibrary(b.byref()) == 0
        b.output[...] = g0
        stokes.form_stokes(lib, b, 0, 3, 1, "I", contiguous=True)
        got = b.output.copy()
    finally:
        lib.releaseLibrary()
    st = (g0[:, :, 0] + g0[:, :, 3]) / np.float32(2)
    flat = got.reshape(2, 12, 16, 16)
    for c in range(3):
        np.testing.assert_array_equal(flat[:, c], st[:, c])
