"""The batched uneven-facet workload at full size, through properties (the oracle would take
hours at this size; tests/test_gpu_batch.py pins every facet to the single-facet entry points
and the oracle at small sizes): the w-projection visibilities (128 antennas, 8128 baselines,
480 timesteps x 32 channels 1.0-1.5 GHz, 1 correlation) gridded into 200 facets of 128^2-2048^2
(log-uniform) at 1.5", seven w-kernel stacks of 16-128 planes (supports 7 -> 63 by |w|, OS x8),
50 sampling-function facets, in one persistent HIP-graph launch.

* FP32-compensated vs FP64 path, facet by facet: rms <= 1e-6 and peak <= 2e-5 of each facet's
  peak (the north-star bar is 1e-5 rms).
* The graph schedule equals one launch per (facet, class) of the single-facet kernels facet by
  facet (same compute path: 1e-6 rms; the flush order differs).
* Every schedule and path counts the same taps.
"""
import ctypes

import numpy as np
import pytest
import torch

from bullseye_amd import kernels, synth
from bullseye_amd.batch import FacetBatch
from bullseye_amd.binding import Binding, ImagingSpec
from bullseye_amd.library_loader import (BATCH_PER_FACET, BATCH_PERSISTENT_GRAPH,
                                         PATH_FP32_KAHAN, PATH_FP64, bullseye_options,
                                         bullseye_stats, load_library)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def per_facet_errors(got, want):
    out = []
    for g, w in zip(got, want):
        w = w.to(torch.complex128)
        d = (g.to(torch.complex128) - w).abs()
        pk = w.abs().max().item()
        assert pk > 0
        out.append((torch.sqrt((d ** 2).mean()).item() / pk, d.max().item() / pk))
    return np.array(out)


@pytest.fixture(scope="module")
def passes():
    print("case full-size batch: synthesising", flush=True)
    cfg = synth.CONFIGS["c2"]
    obs = synth.simulate(**cfg["sim"], seed=1, device=DEV)
    facets, ksets, _, wmax, _, _ = synth.batched_facets(obs, 200, 201, device=DEV)
    lib = load_library("GPU", "mixed")
    o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                         compute_path=PATH_FP32_KAHAN)
    assert lib.bullseye_set_options(ctypes.byref(o)) == 0
    spec0 = ImagingSpec(npix=256, cell=1.5, conv=kernels.pswf_aa_kernel(3, 63), support=3,
                        oversample=63)
    b = Binding("mixed", obs, spec0, device=DEV, with_psf=True)
    torch.cuda.synchronize()
    assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
    out = {}
    try:
        batch = FacetBatch(lib, b, "facet_single_pol")
        im = cfg["image"]
        for k, (conv, halves) in enumerate(ksets):
            batch.kernel_set(k, conv, im["support"], im["oversample"], halves, wmax)
        del ksets
        st = bullseye_stats()
        for name, path, schedule in (("graph32", PATH_FP32_KAHAN, BATCH_PERSISTENT_GRAPH),
                                     ("graph64", PATH_FP64, BATCH_PERSISTENT_GRAPH),
                                     ("facet32", PATH_FP32_KAHAN, BATCH_PER_FACET)):
            print("case full-size batch: %s" % name, flush=True)
            o.compute_path = path
            assert lib.bullseye_set_options(ctypes.byref(o)) == 0
            batch.plan(facets, schedule)
            batch.execute()
            lib.bullseye_get_stats(ctypes.byref(st))
            out[name] = ([t.clone() for t in batch.outputs], int(st.taps))
            print("  %.1f ms of kernels, %.3e taps" % (st.kernel_ms, st.taps), flush=True)
        batch.release()
    finally:
        lib.releaseLibrary()
        o.compute_path = 0
        o.device_resident = 0
        lib.bullseye_set_options(ctypes.byref(o))
    return out


def test_fp32_compensated_matches_fp64(passes):
    e = per_facet_errors(passes["graph32"][0], passes["graph64"][0])
    print("worst facet rms %.3e peak %.3e" % tuple(e.max(axis=0)))
    assert e[:, 0].max() <= 1e-6 and e[:, 1].max() <= 2e-5


def test_graph_schedule_matches_per_facet_launches(passes):
    e = per_facet_errors(passes["graph32"][0], passes["facet32"][0])
    print("worst facet rms %.3e peak %.3e" % tuple(e.max(axis=0)))
    assert e[:, 0].max() <= 1e-6 and e[:, 1].max() <= 2e-5


def test_tap_counts_agree(passes):
    taps = {k: v[1] for k, v in passes.items()}
    assert taps["graph32"] > 1e12
    assert taps["graph32"] == taps["graph64"] == taps["facet32"], taps
