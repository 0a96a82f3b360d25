"""The gridding entry points return once their work is queued, as the reference's do
(cpu_algorithm/wrapper.cpp:184 runs the gridder in std::async; gpu_algorithm/wrapper.cu launches
on a stream; bullseye_pipeliner.py:325 calls gridding_barrier before touching the buffers): with
device-resident buffers and a chunk that is not the final one, a call comes back in a fraction of
its kernel time and gridding_barrier then waits for the rest.  The grids after the barrier are
the same as those of a call followed at once by the barrier."""
import ctypes
import time

import numpy as np
import pytest
import torch

from common import make_spec, rel_err, round_to, set_options, small_obs, wproj_spec
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import PATH_FP32_KAHAN, bullseye_stats, load_library

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def timed_call(lib, b, mode):
    """(seconds in the call, seconds in the barrier after it, kernel ms of the call)."""
    assert lib.bullseye_reset_grids() == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = getattr(lib, mode)(b.byref())
    t1 = time.perf_counter()
    lib.gridding_barrier()
    t2 = time.perf_counter()
    assert rc == 0, lib.bullseye_last_error()
    st = bullseye_stats()
    lib.bullseye_get_stats(ctypes.byref(st))
    return t1 - t0, t2 - t1, st.kernel_ms


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
def test_grid_call_returns_before_its_work_is_done(kernel):
    print("case async", kernel, flush=True)
    mode = "facet_4_cor_corrections" if kernel == "aa" else "facet_single_pol"
    obs = round_to(small_obs(mode, seed=51, n_ant=64, n_time=40, n_chan=64, size_m=1000.0),
                   "mixed")
    if kernel == "aa":
        spec = make_spec(obs, mode, precision="mixed", npix=256, n_facets=25)
        halves = None
    else:
        spec, halves = wproj_spec(obs, mode, "mixed", S=15, OS=8, planes=16, npix=512)
    lib = load_library("GPU", "mixed")
    set_options(lib, device_resident=1, compute_path=PATH_FP32_KAHAN)
    if halves is not None:
        arr = (ctypes.c_uint32 * len(halves))(*halves)
        assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    b = Binding("mixed", obs, spec, device=DEV)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        b.params.is_final_data_chunk = True
        timed_call(lib, b, mode)                       # warm-up: code objects, buffers
        lib.gridding_barrier()
        want = b.host("output_buffer").copy()
        b.params.is_final_data_chunk = False
        t_call, t_wait, kms = timed_call(lib, b, mode)
        print("call %.3f ms, barrier %.3f ms, kernel %.3f ms" % (t_call * 1e3, t_wait * 1e3, kms),
              flush=True)
        assert kms > 2.0, kms                          # enough work to tell the two apart
        assert t_call < 0.5 * kms * 1e-3, (t_call, kms)
        assert t_wait > 0.5 * kms * 1e-3, (t_wait, kms)
        # the same chunk as the final one publishes the same grids
        b.params.is_final_data_chunk = True
        assert lib.bullseye_reset_grids() == 0
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        got = b.host("output_buffer")
        assert rel_err(got, want)[0] < 1e-6
    finally:
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
        set_options(lib)
