"""Batched facets (bullseye_batch_*) on the GPU.

A batch grids facets of different size, cell, centre, w-kernel stack and kind from one
visibility set.  Each facet must equal what the single-facet entry points give for that facet
alone (facet_single_pol / facet_sampling_function with its own initLibrary geometry, kernels and
per-plane supports), which test_gpu_wproj pins to the oracle; two facets are also checked
against the oracle directly.  Schedules (persistent graph, persistent eager, one launch per
facet), tile choices (lane shape, chunk length, channels per stream, base rows) and graph
replays must agree with each other.  Bars: FP64 path of the double build 1e-11 rms / 1e-9 peak
of the reference's peak (summation order of the grid atomics differs), FP32 compensated path
1e-6 / 2e-5; normalisation sums to 1e-9 relative; tap counts exactly.
"""
import ctypes

import numpy as np
import pytest
import torch

from common import rel_err, round_to, set_options, small_obs, wproj_spec
from bullseye_amd import synth
from bullseye_amd.base_types import force_precision
from bullseye_amd.batch import GRID_MODES, BatchFacet, FacetBatch
from bullseye_amd.binding import Binding
from bullseye_amd.library_loader import (BATCH_PER_FACET, BATCH_PERSISTENT,
                                         BATCH_PERSISTENT_GRAPH, PATH_FP32_KAHAN, PATH_FP64,
                                         SHAPE_COLUMN, SHAPE_PAIRS1, SHAPE_PAIRS2, SHAPE_PAIRS4,
                                         bullseye_batch_facet, bullseye_stats, load_library)
from oracle import oracle

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

# kernel sets: (S, OS, planes, per-plane half supports, wmax factor)
SETS = {0: (7, 4, 8, [1, 2, 3, 4, 5, 6, 7, 7], 0.9),
        1: (12, 4, 6, [3, 5, 8, 10, 12, 12], 1.1)}
# facets: (npix, cell arcsec, kernel set, sampling function)
FACETS = [(96, 15.0, 0, False), (128, 18.0, 1, False), (160, 15.0, 0, True),
          (112, 16.0, 1, True), (144, 12.0, 0, False), (64, 30.0, 1, False)]


def observation(precision, seed=5):
    return round_to(small_obs("facet_single_pol", seed=seed, n_ant=12, n_time=30, n_chan=8),
                    precision)


def kernel_specs(obs, precision):
    out = {}
    for kid, (S, OS, planes, halves, wf) in SETS.items():
        spec, halves = wproj_spec(obs, "facet_single_pol", precision, S=S, OS=OS, planes=planes,
                                  halves=halves, seed=11 + kid)
        spec.wmax = float(np.float32(spec.wmax / 0.9 * wf))
        out[kid] = (spec, halves)
    return out


def facet_list(obs, precision, facets=FACETS, seed=2):
    bt = force_precision(precision)
    rng = np.random.default_rng(seed)
    cen = synth.facet_centres(len(facets), 0.5, obs.phase_centre, rng)
    cen = cen.astype(bt.uvw_type).astype(np.float64)
    return [BatchFacet(n, cell, (float(cen[i, 0]), float(cen[i, 1])), kid, psf)
            for i, (n, cell, kid, psf) in enumerate(facets)]


NC = {"facet_single_pol": 1, "facet_duel_pol": 2, "facet_4_cor": 4}


def single_facet_spec(kspec, f, mode="facet_single_pol", cube=1):
    spec = kspec.__class__(**kspec.__dict__)
    spec.npix, spec.cell = f.npix, f.cell
    spec.facet_centres = np.array([f.centre], np.float64)
    spec.n_pol_gridded = NC[mode]
    spec.pol_index, spec.second_pol_index = (0, 3) if mode == "facet_duel_pol" else (0, 1)
    spec.cube = cube
    if cube > 1:
        spec.channel_grid_index = CGI(cube)
        spec.sf_channel_grid_index = CGI(cube)
    if f.sampling_function:
        spec.sf_count = cube
    return spec


def CGI(cube, n_chan=8):
    return (np.arange(n_chan) % cube).astype(np.uint64)


def single_facet(lib, precision, obs, kspec, halves, f, mode="facet_single_pol", cube=1):
    """The facet through its single-facet entry point alone: (grid, norm, taps)."""
    spec = single_facet_spec(kspec, f, mode, cube)
    nnorm = cube * NC[mode]
    mode = "facet_sampling_function" if f.sampling_function else mode
    arr = (ctypes.c_uint32 * len(halves))(*halves)
    assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0
    b = Binding(precision, obs, spec, with_psf=f.sampling_function)
    try:
        assert lib.initLibrary(b.byref()) == 0, lib.bullseye_last_error()
        assert getattr(lib, mode)(b.byref()) == 0, lib.bullseye_last_error()
        lib.gridding_barrier()
        st = bullseye_stats()
        lib.bullseye_get_stats(ctypes.byref(st))
        single_facet.time_inner = int(st.time_inner)
        norm = np.ctypeslib.as_array(
            ctypes.cast(b.params.normalization_terms, ctypes.POINTER(ctypes.c_double)),
            (nnorm,)).copy()
    finally:
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
    grid = (b.psf if f.sampling_function else b.output)[0]   # [cube][P][ny][nx]
    return grid, (np.zeros(1) if f.sampling_function else norm), int(st.taps)


class Session:
    """initLibrary on the visibility set (any geometry), kernel sets defined, batch ready."""

    def __init__(self, lib, precision, path, obs, kspecs, device=DEV, mode="facet_single_pol",
                 cube=1):
        self.lib, self.path = lib, path
        spec0 = single_facet_spec(kspecs[0][0], BatchFacet(128, 15.0, (0.0, 0.0), 0, True),
                                  mode, cube)
        self.binding = Binding(precision, obs, spec0, device=device, with_psf=True)
        set_options(lib, device_resident=1 if device is not None else 0,
                    compute_path=self.path)
        halves0 = kspecs[0][1]
        arr = (ctypes.c_uint32 * len(halves0))(*halves0)
        assert lib.bullseye_set_wplane_support(arr, len(halves0)) == 0
        assert lib.initLibrary(self.binding.byref()) == 0, lib.bullseye_last_error()
        lib.bullseye_set_wplane_support(None, 0)
        self.batch = FacetBatch(lib, self.binding, mode)
        for kid, (spec, halves) in kspecs.items():
            self.batch.kernel_set(kid, spec.conv, spec.support, spec.oversample, halves,
                                  spec.wmax)

    def run(self, facets, schedule=BATCH_PERSISTENT_GRAPH, repeats=1):
        self.batch.plan(facets, schedule)
        outs = []
        for _ in range(repeats):
            grids, norms = self.batch.execute()
            st = bullseye_stats()
            self.lib.bullseye_get_stats(ctypes.byref(st))
            outs.append(([to_np(g) for g in grids], [n.copy() for n in norms], int(st.taps)))
        return outs if repeats > 1 else outs[0]

    def close(self):
        self.lib.releaseLibrary()
        set_options(self.lib)


def to_np(g):
    return g.cpu().numpy() if isinstance(g, torch.Tensor) else np.array(g)


TOL = {PATH_FP64: (1e-11, 1e-9), PATH_FP32_KAHAN: (1e-6, 2e-5)}


def check_close(got, want, path):
    rms, peak = rel_err(got, want)
    assert np.abs(want).max() > 0
    assert rms <= TOL[path][0] and peak <= TOL[path][1], (rms, peak)
    return rms, peak


def open_session(precision, path, obs, kspecs, device=DEV, mode="facet_single_pol", cube=1):
    lib = load_library("GPU", precision)
    return lib, Session(lib, precision, path, obs, kspecs, device=device, mode=mode, cube=cube)


@pytest.mark.parametrize("precision,path", [("double", PATH_FP64), ("mixed", PATH_FP32_KAHAN)])
def test_batch_matches_single_facet_calls(precision, path):
    print("case batch vs single", precision, path, flush=True)
    obs = observation(precision)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision)
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    want = [single_facet(lib, precision, obs, *kspecs[f.kernel_set], f) for f in facets]
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        grids, norms, taps = s.run(facets)
    finally:
        s.close()
    for i, f in enumerate(facets):
        rms, peak = check_close(grids[i], want[i][0], path)
        print("facet %d (%d px, set %d, psf %d): rms %.2e peak %.2e" %
              (i, f.npix, f.kernel_set, f.sampling_function, rms, peak))
        if not f.sampling_function:
            np.testing.assert_allclose(norms[i].reshape(-1), want[i][1], rtol=1e-9)
    assert taps == sum(w[2] for w in want), (taps, [w[2] for w in want])


def test_batch_matches_oracle():
    print("case batch vs oracle", flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=7)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, facets=[(128, 18.0, 1, False), (96, 15.0, 0, True)],
                        seed=4)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        grids, norms, _ = s.run(facets)
    finally:
        s.close()
    for i, f in enumerate(facets):
        kspec, halves = kspecs[f.kernel_set]
        spec = single_facet_spec(kspec, f)
        mode = "facet_sampling_function" if f.sampling_function else "facet_single_pol"
        ref = Binding("double", obs, spec, with_psf=f.sampling_function)
        want_norm = oracle.run(ref, mode, normalize=False)
        want = (ref.psf if f.sampling_function else ref.output)[0]
        check_close(grids[i], want, path)
        if not f.sampling_function:
            np.testing.assert_allclose(norms[i].reshape(-1), want_norm, rtol=1e-9)


def test_batch_schedules_and_graph_replays_agree():
    print("case batch schedules", flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=9)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, seed=6)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        eager, _, taps_e = s.run(facets, BATCH_PERSISTENT)
        per_facet, _, taps_p = s.run(facets, BATCH_PER_FACET)
        replays = s.run(facets, BATCH_PERSISTENT_GRAPH, repeats=3)
        # eager work and synchronisation between replays, then replay again
        s.run(facets, BATCH_PERSISTENT)
        s.batch.plan(facets, BATCH_PERSISTENT_GRAPH)
        for _ in range(2):
            x = torch.randn(1 << 20, device=DEV)
            (x * 2).sum().item()
            torch.cuda.synchronize()
            grids, _ = s.batch.execute()
            replays.append(([to_np(g) for g in grids], None, None))
    finally:
        s.close()
    assert taps_e == taps_p
    for i in range(len(facets)):
        check_close(per_facet[i], eager[i], path)
        for r in replays:
            check_close(r[0][i], eager[i], path)


TILES = [dict(chunk=2), dict(chunk=4), dict(shape=SHAPE_PAIRS1), dict(shape=SHAPE_PAIRS2),
         dict(shape=SHAPE_PAIRS4, chunk=2), dict(shape=SHAPE_COLUMN),
         dict(channels_per_stream=2, base_rows=16), dict(channels_per_stream=1, base_rows=8),
         dict(base_rows=64, chunk=4), dict(facet_major=1), dict(facet_major=1, chunk=2)]


@pytest.mark.parametrize("tile", range(len(TILES)))
def test_batch_tiles_agree(tile):
    kw = TILES[tile]
    print("case batch tile", kw, flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=13)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, seed=8)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        want, wnorm, wtaps = s.run(facets)
        s.batch.set_tuning(**kw)
        got, gnorm, gtaps = s.run(facets)
        s.batch.set_tuning()
    finally:
        s.close()
    assert gtaps == wtaps
    for i in range(len(facets)):
        check_close(got[i], want[i], path)
        np.testing.assert_allclose(gnorm[i], wnorm[i], rtol=1e-9)


def test_batch_autotune_then_plan():
    print("case batch autotune", flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    obs = observation(precision, seed=15)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, seed=10)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        want, _, wtaps = s.run(facets)
        t = s.batch.autotune(facets)
        print("tuning", {k: v for k, v in t.items() if k not in ("shape", "chunk")})
        assert t["tuned"] == 1 and t["tuned_ms"] > 0
        assert t["tuned_ms"] <= t["default_ms"] * 1.5 + 0.05
        assert t["channels_per_stream"] in (4, 8, 16) and t["base_rows"] in (16, 32, 64)
        assert s.batch.tuning()["tuned"] == 1
        got, _, gtaps = s.run(facets)
        s.batch.set_tuning()
    finally:
        s.close()
    assert gtaps == wtaps
    for i in range(len(facets)):
        check_close(got[i], want[i], path)


def test_batch_host_arrays_match_device_resident():
    print("case batch host mode", flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    obs = observation(precision, seed=17)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, seed=12)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        want, wnorm, wtaps = s.run(facets)
    finally:
        s.close()
    lib, s = open_session(precision, path, obs, kspecs, device=None)
    try:
        got, gnorm, gtaps = s.run(facets)
    finally:
        s.close()
    assert gtaps == wtaps
    for i in range(len(facets)):
        check_close(got[i], want[i], path)
        np.testing.assert_allclose(gnorm[i], wnorm[i], rtol=1e-9)


def test_batch_reports_data_changed_after_plan():
    """w scaled up after planning: samples wider than their stream's class are dropped with an
    error instead of being gridded outside the class footprint."""
    print("case batch stale plan", flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=19)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, facets=[(128, 15.0, 0, False)], seed=14)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        s.batch.plan(facets)
        s.batch.execute()
        uvw = s.binding.keep["uvw_coords"]
        uvw[:, 2] *= 1.6
        torch.cuda.synchronize()
        with pytest.raises(RuntimeError, match="support class"):
            s.batch.execute()
    finally:
        s.close()


def test_batch_rejects_bad_tables():
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=21)
    kspecs = kernel_specs(obs, precision)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        bad = [BatchFacet(128, 15.0, (0.0, 0.0), 7)]
        with pytest.raises(RuntimeError, match="kernel set 7"):
            s.batch.plan(bad)
        with pytest.raises(RuntimeError, match="does not fit"):
            s.batch.plan([BatchFacet(20, 15.0, (0.0, 0.0), 1)])
        with pytest.raises(RuntimeError, match="no batch plan"):
            s.batch.execute()
    finally:
        s.close()


@pytest.mark.parametrize("mode,cube", [("facet_duel_pol", 1), ("facet_4_cor", 1),
                                       ("facet_single_pol", 2), ("facet_4_cor", 3)])
def test_batch_modes_and_cubes_match_single_facet_calls(mode, cube):
    """2- and 4-correlation facets, and channels spread over several grid slices (every facet's
    output and sums [cube][ncorr]; sampling-function facets over the same slices)."""
    print("case batch", mode, "cube", cube, flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    obs = round_to(small_obs("facet_4_cor", seed=23, n_ant=12, n_time=30, n_chan=8), precision)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, facets=FACETS[:4], seed=16)
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    want = [single_facet(lib, precision, obs, *kspecs[f.kernel_set], f, mode, cube)
            for f in facets]
    lib, s = open_session(precision, path, obs, kspecs, mode=mode, cube=cube)
    try:
        grids, norms, taps = s.run(facets)
    finally:
        s.close()
    for i, f in enumerate(facets):
        assert grids[i].shape == want[i][0].shape, (grids[i].shape, want[i][0].shape)
        rms, peak = check_close(grids[i], want[i][0], path)
        print("facet %d: rms %.2e peak %.2e" % (i, rms, peak))
        if not f.sampling_function:
            np.testing.assert_allclose(norms[i].reshape(-1), want[i][1], rtol=1e-9)
    assert taps == sum(w[2] for w in want)


def test_batch_non_square_facet_matches_oracle():
    """nx != ny and cell_x != cell_y: u scales with nx * cell_x, v with ny * cell_y."""
    print("case batch non-square", flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=25)
    kspecs = kernel_specs(obs, precision)
    f = facet_list(obs, precision, facets=[(128, 15.0, 0, False)], seed=18)[0]
    f.ny, f.cell_y = 96, 19.0
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        grids, norms, _ = s.run([f])
    finally:
        s.close()
    assert grids[0].shape == (1, 1, 96, 128)
    spec = single_facet_spec(kspecs[0][0], f)
    ref = Binding("double", obs, spec)
    ref.params.ny = 96
    ref.params.cell_size_y = 19.0
    out = ref.set_array("output_buffer", np.zeros((1, 1, 1, 96, 128), np.complex128),
                        np.complex128)
    want_norm = oracle.run(ref, "facet_single_pol", normalize=False)
    check_close(grids[0], out[0], path)
    np.testing.assert_allclose(norms[0].reshape(-1), want_norm, rtol=1e-9)


def test_batch_holds_300_facets():
    """A batch wider than an 8-bit facet index: 300 small facets of both kinds; every facet's
    taps add up to those of its single-facet call (one in 16 checked against the full grid), and
    a plan of more than 65536 facets is refused before it allocates anything."""
    print("case batch 300 facets", flush=True)
    precision, path = "double", PATH_FP64
    obs = observation(precision, seed=27)
    kspecs = kernel_specs(obs, precision)
    rng = np.random.default_rng(20)
    table = [(int(rng.choice([32, 48, 64])), float(rng.choice([20.0, 30.0, 40.0])),
              int(rng.integers(0, 2)), bool(rng.random() < 0.25)) for _ in range(300)]
    table = [(n, c, 0 if n < 40 else k, p) for n, c, k, p in table]  # set 1 needs > 2*12+4 px
    facets = facet_list(obs, precision, facets=table, seed=22)
    lib, s = open_session(precision, path, obs, kspecs)
    try:
        grids, norms, taps = s.run(facets)
        too_many = (bullseye_batch_facet * 65537)()
        assert lib.bullseye_batch_plan(s.batch.binding.byref(), GRID_MODES["facet_single_pol"],
                                       too_many, 65537, 0) != 0
        assert b"at most 65536" in lib.bullseye_last_error()
    finally:
        s.close()
    lib = load_library("GPU", precision)
    set_options(lib, compute_path=path)
    total = 0
    for i, f in enumerate(facets):
        want, wnorm, wtaps = single_facet(lib, precision, obs, *kspecs[f.kernel_set], f)
        total += wtaps
        if i % 16 == 0 and np.abs(want).max() > 0:
            check_close(grids[i], want, path)
            if not f.sampling_function:
                np.testing.assert_allclose(norms[i].reshape(-1), wnorm, rtol=1e-9)
    assert taps == total


def test_batch_walk_order_matches_single_facet():
    """The batch measures the walk order (time or channels inner) from the data as the
    single-facet path does: the same choice on a narrow and on a wide band, and the two bands
    choose differently."""
    print("case batch walk order", flush=True)
    precision, path = "double", PATH_FP64
    orders = []
    for f_lo, f_hi in ((1.40e9, 1.42e9), (1.0e9, 1.5e9)):
        obs = round_to(small_obs("facet_single_pol", seed=29, n_ant=12, n_time=30, n_chan=8,
                                 f_lo=f_lo, f_hi=f_hi), precision)
        kspecs = kernel_specs(obs, precision)
        f = facet_list(obs, precision, facets=[(128, 15.0, 0, False)], seed=24)[0]
        lib = load_library("GPU", precision)
        set_options(lib, compute_path=path)
        single_facet(lib, precision, obs, *kspecs[0], f)
        want = single_facet.time_inner
        st = bullseye_stats()
        lib, s = open_session(precision, path, obs, kspecs)
        try:
            s.run([f])
            lib.bullseye_get_stats(ctypes.byref(st))
            got = int(st.time_inner)
        finally:
            s.close()
        print("band %.2f-%.2f GHz: single %d batch %d" % (f_lo / 1e9, f_hi / 1e9, want, got))
        assert got == want
        orders.append(got)
    assert orders[0] != orders[1], orders


@pytest.mark.parametrize("device", [DEV, None])
def test_batch_of_an_empty_visibility_set(device):
    """No rows: the plan has no work, every facet's grid and sums are zero."""
    print("case batch empty", device, flush=True)
    precision, path = "mixed", PATH_FP32_KAHAN
    obs = observation(precision, seed=31)
    kspecs = kernel_specs(obs, precision)
    facets = facet_list(obs, precision, facets=FACETS[:3], seed=26)
    lib = load_library("GPU", precision)
    s = Session(lib, precision, path, obs, kspecs, device=device)
    try:
        empty = Binding(precision, obs, s.binding.spec, rows=(0, 0), device=device, with_psf=True)
        batch = FacetBatch(lib, empty, "facet_single_pol")
        batch.plan(facets)
        grids, norms = batch.execute()
        st = bullseye_stats()
        lib.bullseye_get_stats(ctypes.byref(st))
        assert st.taps == 0 and st.samples_gridded == 0
        for g, n in zip(grids, norms):
            assert np.abs(to_np(g)).max() == 0 and np.all(n == 0)
    finally:
        s.close()
