"""The AA gridder's sample preparation from the sorted row table and the channel table.

* Device buffers that are element-aligned but not 16-byte aligned take the per-correlation loads
  (the vector loads need 16-byte alignment); the grids must still match the oracle.
* Id checks: an out-of-range timestamp id on a usable row raises the range error when its sample
  reaches the Jones correction; the same id on a flagged row is never read; an out-of-range spw
  id raises the error whether the row is flagged or not (the spw is checked first).
Bars as in test_gpu_parity.
"""
import numpy as np
import pytest

from common import make_spec, rel_err, round_to, run_library, set_options, small_obs
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP32_KAHAN, load_library
from oracle import oracle

pytestmark = pytest.mark.gpu


def oracle_grid(obs, spec, mode, normalize=True):
    ref = Binding("double", obs, spec)
    norm = oracle.run(ref, mode, normalize=normalize)
    return ref.output, norm


def shift_device_array(b, name, arr, dtype, pad_bytes):
    """Copy `arr` to the device `pad_bytes` past the start of a fresh allocation and point the
    call's `name` there (still aligned to the element, no longer to 16 bytes)."""
    import torch
    raw = np.ascontiguousarray(np.asarray(arr).astype(dtype)).view(np.uint8).ravel()
    t = torch.zeros(raw.size + pad_bytes, dtype=torch.uint8, device="cuda:0")
    t[pad_bytes:] = torch.from_numpy(raw).to("cuda:0")
    b.keep[name] = t
    setattr(b.params, name, t.data_ptr() + pad_bytes)
    assert (t.data_ptr() + pad_bytes) % 16 != 0


@pytest.mark.parametrize("mode", ["facet_4_cor_corrections", "grid_4_cor"])
def test_unaligned_device_buffers(mode):
    import torch
    print("case unaligned device buffers", mode, flush=True)
    obs = round_to(small_obs(mode, seed=61), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    want, want_norm = oracle_grid(obs, spec, mode)
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP32_KAHAN, device_resident=1)
    b = Binding("mixed", obs, spec, device=torch.device("cuda:0"))
    bt = b.bt
    shift_device_array(b, "visibilities", obs.vis, bt.visibility_type, 8)
    shift_device_array(b, "visibility_weights", obs.weights, bt.weight_type, 4)
    shift_device_array(b, "flags", obs.flags, np.bool_, 1)
    torch.cuda.synchronize()
    try:
        norm = run_library(lib, b, mode)
    finally:
        set_options(lib)
    rms, peak = rel_err(b.host("output_buffer"), want)
    assert np.abs(want).max() > 0
    assert rms <= 1e-6 and peak <= 2e-5, (rms, peak)
    np.testing.assert_allclose(norm, want_norm, rtol=1e-12)


def grid_once(obs, spec, mode):
    lib = load_library("GPU", "mixed")
    set_options(lib, compute_path=PATH_FP32_KAHAN)
    b = Binding("mixed", obs, spec)
    assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
    try:
        rc = getattr(lib, mode)(b.byref())
        err = lib.bullseye_last_error()
    finally:
        lib.releaseLibrary()
    return rc, err, b


def test_bad_timestamp_on_usable_row_raises():
    print("case bad timestamp, usable row", flush=True)
    mode = "facet_4_cor_corrections"
    obs = round_to(small_obs(mode, seed=63, flag_fraction=0.0), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    k = obs.rows // 3
    obs.time_idx[k] = obs.n_time
    rc, err, _ = grid_once(obs, spec, mode)
    assert rc != 0 and b"timestamp id out of range" in err, err


def test_bad_timestamp_on_flagged_row_is_not_read():
    print("case bad timestamp, flagged row", flush=True)
    mode = "facet_4_cor_corrections"
    obs = round_to(small_obs(mode, seed=63, flag_fraction=0.0), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    k = obs.rows // 3
    obs.flagged_rows[k] = True
    want, _ = oracle_grid(obs, spec, mode, normalize=False)   # grid_once does not normalise
    obs.time_idx[k] = obs.n_time
    rc, err, b = grid_once(obs, spec, mode)
    assert rc == 0, err
    rms, peak = rel_err(b.output, want)
    assert rms <= 1e-6 and peak <= 2e-5, (rms, peak)


@pytest.mark.parametrize("flagged", [False, True])
def test_bad_spw_raises(flagged):
    print("case bad spw, flagged", flagged, flush=True)
    mode = "facet_4_cor_corrections"
    obs = round_to(small_obs(mode, seed=65, flag_fraction=0.0), "mixed")
    spec = make_spec(obs, mode, precision="mixed")
    k = obs.rows // 2
    obs.spw[k] = 3
    obs.flagged_rows[k] = flagged
    rc, err, _ = grid_once(obs, spec, mode)
    assert rc != 0 and b"spw id" in err, err
