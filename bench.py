"""Benchmark: Gtaps/s of gridding + phase rotation, FP32-compensated vs FP64, on one MI355X.

Workload (one GPU): targeted faceting with 4 correlations -- 64 antennas uniform in a 3 km disc
(2016 baselines), 240 timesteps over 8 h at dec -30 deg, 64 channels 1.40-1.42 GHz, 4
correlations complex64 corrupted by per-antenna per-timestep Jones I + N(0, 0.05^2); 25 facets
of 512^2 at 2" within 2 deg of the phase centre; 7x7 PSWF AA kernel oversampled x63; gridded
with facet_4_cor_corrections (phase rotation, uvw rotation, Jones correction).  Inputs are
synthetic (seeded) and resident in HBM before timing.

A step is one imaging pass: zero the accumulation grids, grid every row (facet_4_cor_corrections,
device-resident buffers) and publish the grids to the caller's buffer.  One tap is one
(facet, row, channel, kernel point) accumulate of an unflagged in-grid sample (4 correlations
each).  The FP64 path runs the same step on the same inputs; the error of the FP32-compensated
grids is stated against it (full size) and against the FP64 CPU oracle (a row sample).

Prints one JSON line.  --gpus N with N > 1 exits: the measurement is of one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gtaps/s gridding+phase-rot, FP32-comp vs FP64 on 1 GPU at stated RMS error"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOPS_PER_TAP_PER_CORR = 5   # conv_u*conv_v, complex x real, complex add (real AA kernel)
HOST_CHUNKS = 8               # pipeliner-style chunks of the host-buffer measurement


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-rows", type=int, default=0,
                    help="rows of the oracle's CPU sample (0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--channels-per-stream", type=int, default=0)
    ap.add_argument("--rows-per-stream", type=int, default=0)
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="before timing, copy this many MiB device to device once (a plain copy "
                         "of known size that scales HBM byte counters)")
    ap.add_argument("--workload", choices=["faceting", "plain", "wproj", "adjoint", "batch"],
                    default="faceting",
                    help="faceting: the 25-facet Jones workload (default); plain: one 1024^2 "
                         "grid of 64 ant x 240 t x 16 ch, 1 correlation (the shape of the "
                         "reference's own benchmark.cpp); wproj: the 4096^2 "
                         "w-projection workload; adjoint: chi-squared value and gradient of a "
                         "2048^2 facet model through gridding and its degridding adjoint; "
                         "batch: the w-projection visibilities into 200 uneven facets in one "
                         "persistent graph launch")
    ap.add_argument("--no-host", action="store_true",
                    help="faceting workload: skip the host-buffer passes (profiling runs)")
    ap.add_argument("--no-batch", action="store_true",
                    help="faceting: leave out the batched-facets figures of the default line")
    ap.add_argument("--batch-facets", type=int, default=200)
    ap.add_argument("--batch-timesteps", type=int, default=480,
                    help="timesteps of the w-projection visibilities the batch grids")
    ap.add_argument("--tune-timesteps", type=int, default=30,
                    help="timesteps of the sample the batch autotuner times (0: no autotune)")
    ap.add_argument("--tuning", default="",
                    help="batch tiles as the JSON 'tuning' object a batch run prints (replaces "
                         "the autotuner)")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch: skip the FP64, default-tile and per-facet passes")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def calibration_copy(mib):
    """One device-to-device copy of `mib` MiB: reads and writes exactly mib * 1024 KiB."""
    if mib <= 0:
        return
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("calibration copy: %d MiB" % mib)
    del a, b


def profiled_value(name, key):
    """`key = value` from the committed profile summary `profiles/<name>`; None when absent."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith(key + " ="):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        return None
    return None


def profiled_traffic(name):
    """HBM bytes per launch of the workload's gridding kernel(s): rocprofv3 FETCH_SIZE plus
    WRITE_SIZE, each scaled by what the counter reads for a 1 GiB device copy, as summarised in
    the committed profile `profiles/<name>`; None when the summary is absent."""
    return profiled_value(name, "traffic_bytes_per_launch")


def flush_atomics(lib, profile, kernel_ms):
    """The grid flushes against the device's FP64 atomic-add rate: atomics per launch
    (rocprofv3 TCC_ATOMIC in the committed profile), their rate over the measured kernel time,
    and the rate bullseye_fp64_atomic_rate measures now (every flush is such an add, executed
    on the memory side)."""
    rate = ctypes.c_double(0.0)
    ceiling = None
    if lib.bullseye_fp64_atomic_rate(ctypes.byref(rate)) == 0:
        ceiling = rate.value / 1e9
    else:
        log("atomic rate: %s" % lib.bullseye_last_error().decode())
    per = profiled_value(profile, "atomics_per_launch")
    out = {"per_launch": per, "gops": None, "ceiling_gops": None if ceiling is None
           else round(ceiling, 3), "frac": None}
    if per is not None and kernel_ms > 0:
        out["gops"] = round(per / (kernel_ms * 1e-3) / 1e9, 3)
        if ceiling:
            out["frac"] = round(out["gops"] / ceiling, 4)
    return out


def rel_errors(got, want):
    import numpy as np
    out = []
    for f in range(want.shape[0]):
        w = want[f].astype(np.complex128)
        d = np.abs(got[f].astype(np.complex128) - w)
        peak = np.abs(w).max()
        out.append((float(np.sqrt(np.mean(d ** 2)) / peak), float(d.max() / peak)))
    return max(e[0] for e in out), max(e[1] for e in out)


def facet_images(grids, taper=None):
    """Float64 images of grids [F, P, ny, nx] by the finalize chain of
    fft_and_repacking_routines.cpp:83-118 (ifftshift, unnormalised backward FFT, fftshift, real
    part), optionally divided by the separable grid taper T(l) T(m)."""
    import numpy as np
    g = np.asarray(grids, dtype=np.complex128)
    ny, nx = g.shape[-2:]
    img = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g, axes=(-2, -1)), axes=(-2, -1)),
                          axes=(-2, -1)).real * (nx * ny)
    if taper is not None:
        img = img / np.outer(taper, taper)
    return img


def facet_error_table(got, want, norm_got, norm_want, taper):
    """Per-facet error of `got` against `want` (grids [F, P, ny, nx], normalisation sums
    [F * P]), relative to each facet's peak: (i) the normalised uv-grids, (ii) the images after
    the same float64 shift / FFT / taper on both sides, over the central half of each axis (the
    taper falls towards 1e-4 at the edges, where dividing by it only amplifies rounding)."""
    import numpy as np
    F, P, ny, nx = want.shape
    ng = np.asarray(norm_got, np.float64).reshape(F, P, 1, 1)
    nw = np.asarray(norm_want, np.float64).reshape(F, P, 1, 1)
    gg = np.asarray(got, np.complex128) / ng
    gw = np.asarray(want, np.complex128) / nw
    cy, cx = slice(ny // 4, ny - ny // 4), slice(nx // 4, nx - nx // 4)
    rows = {"grid_rms": [], "grid_peak": [], "image_rms": [], "image_peak": []}
    for f in range(F):
        d = np.abs(gg[f] - gw[f])
        pk = np.abs(gw[f]).max()
        rows["grid_rms"].append(float(np.sqrt(np.mean(d ** 2)) / pk))
        rows["grid_peak"].append(float(d.max() / pk))
        ig = facet_images(gg[f], taper)[:, cy, cx]
        iw = facet_images(gw[f], taper)[:, cy, cx]
        di = np.abs(ig - iw)
        pi = np.abs(iw).max()
        rows["image_rms"].append(float(np.sqrt(np.mean(di ** 2)) / pi))
        rows["image_peak"].append(float(di.max() / pi))
    out = {k: [float("%.3g" % x) for x in v] for k, v in rows.items()}
    out["worst"] = {k: max(v) for k, v in out.items()}
    return out


def host_chunk_bindings(obs, spec, k):
    """The pipeliner's chunking (bullseye_pipeliner.py:312-386) over host arrays: k chunks of
    whole timesteps (rows are time-ordered), each with the Jones terms of its own timesteps;
    the last chunk is the final one."""
    import numpy as np
    from bullseye_amd.binding import Binding
    t = obs.time_idx.astype(np.int64)
    bounds = np.searchsorted(t, np.linspace(0, obs.n_time, k + 1).round().astype(np.int64))
    out = []
    for i in range(k):
        r0, r1 = int(bounds[i]), int(bounds[i + 1])
        b = Binding("mixed", obs, spec, rows=(r0, r1))
        t0, t1 = int(t[r0]), int(t[r1 - 1]) + 1
        b._set("timestamp_ids", (t[r0:r1] - t0).astype(np.uint64), np.uint64)
        b._set("jones_terms", spec.jones[t0:t1], b.bt.visibility_type)
        b.params.no_timestamps_read = t1 - t0
        b.params.is_final_data_chunk = i == k - 1
        out.append(b)
    return out


def norm_terms(binding, n):
    """The normalisation sums the library published into the caller's struct."""
    import numpy as np
    return np.ctypeslib.as_array(ctypes.cast(binding.params.normalization_terms,
                                             ctypes.POINTER(ctypes.c_double)), (n,)).copy()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is supported; the facet-sharded multi-GPU layout is a "
              "design, not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "plain":
        return main_plain(args)
    if args.workload == "wproj":
        return main_wproj(args)
    if args.workload == "adjoint":
        return main_adjoint(args)
    if args.workload == "batch":
        return main_batch(args)
    import numpy as np
    import torch
    from bullseye_amd import kernels, synth
    from bullseye_amd.binding import Binding, ImagingSpec, jones_reference_layout
    from bullseye_amd.library_loader import (PATH_FP32_KAHAN, PATH_FP64, bullseye_options,
                                             bullseye_stats, load_library)

    cfg = synth.CONFIGS["c1"]
    im = cfg["image"]
    t0 = time.time()
    obs = synth.simulate(**cfg["sim"], seed=args.seed)
    rng = np.random.default_rng(args.seed + 100)
    facets = synth.facet_centres(im["facets"], im["facet_radius_deg"], obs.phase_centre, rng)
    conv = kernels.pswf_aa_kernel(im["support"], im["oversample"])
    spec = ImagingSpec(npix=im["npix"], cell=im["cell"], conv=conv, support=im["support"],
                       oversample=im["oversample"], facet_centres=facets, n_pol_gridded=4,
                       jones=jones_reference_layout(obs, im["facets"], np.complex64))
    log("synthesised %d rows x %d ch x %d corr in %.1fs" % (obs.rows, obs.n_chan, obs.n_corr,
                                                           time.time() - t0))
    dev = torch.device("cuda:0")
    lib = load_library("GPU", "mixed")
    b = Binding("mixed", obs, spec, device=dev)
    torch.cuda.synchronize()

    def set_opts(path):
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                             compute_path=path, channels_per_stream=args.channels_per_stream,
                             rows_per_stream=args.rows_per_stream)
        assert lib.bullseye_set_options(ctypes.byref(o)) == 0, lib.bullseye_last_error()

    def check(rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib.bullseye_last_error().decode()))

    set_opts(PATH_FP32_KAHAN)
    check(lib.initLibrary(b.byref()), "initLibrary")
    stats = bullseye_stats()
    calibration_copy(args.calibration_copy_mib)

    def run(path, steps, warmup):
        set_opts(path)
        for _ in range(warmup):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.facet_4_cor_corrections(b.byref()), "grid")
            lib.gridding_barrier()
        torch.cuda.synchronize()
        kms, taps = [], 0
        t = time.perf_counter()
        for _ in range(steps):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.facet_4_cor_corrections(b.byref()), "grid")
            lib.gridding_barrier()
            lib.bullseye_get_stats(ctypes.byref(stats))
            kms.append(stats.kernel_ms)
            taps = stats.taps
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        return wall, taps, float(np.mean(kms)), stats.samples_gridded

    F, npx = im["facets"], im["npix"]
    wall32, taps, k32, nsamp = run(PATH_FP32_KAHAN, args.steps, args.warmup)
    grid32 = b.host("output_buffer").copy()
    norm32 = norm_terms(b, F * 4)
    log("FP32-comp: %.2f ms/step, kernel %.2f ms, %.3e taps" % (wall32 * 1e3, k32, taps))
    wall64, taps64, k64, _ = run(PATH_FP64, max(1, args.steps // 2), 1)
    grid64 = b.host("output_buffer").copy()
    norm64 = norm_terms(b, F * 4)
    log("FP64: %.2f ms/step, kernel %.2f ms" % (wall64 * 1e3, k64))
    lib.releaseLibrary()
    rms_vs_gpu64, peak_vs_gpu64 = rel_errors(grid32.reshape(F, -1), grid64.reshape(F, -1))
    taper = kernels.grid_taper(conv, im["support"], im["oversample"], npx)
    err_gpu = facet_error_table(grid32.reshape(F, 4, npx, npx), grid64.reshape(F, 4, npx, npx),
                                norm32, norm64, taper)
    del grid64
    wall_host = wall_chunks = None
    chunk_err = (None, None)
    if not args.no_host:
        # the reference's model: host arrays handed over every call (inputs copied to the device,
        # grids copied back on the final chunk); reported beside `value`, never as it
        bh = Binding("mixed", obs, spec)
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=0,
                             compute_path=PATH_FP32_KAHAN)
        check(lib.bullseye_set_options(ctypes.byref(o)), "options")
        check(lib.initLibrary(bh.byref()), "initLibrary(host)")
        host_ms = []
        for i in range(3):
            t = time.perf_counter()
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.facet_4_cor_corrections(bh.byref()), "grid(host)")
            lib.gridding_barrier()
            host_ms.append((time.perf_counter() - t) * 1e3)
        lib.releaseLibrary()
        del bh
        wall_host = min(host_ms[1:]) * 1e-3
        log("FP32-comp with host buffers (PCIe both ways): %.2f ms/step" % (wall_host * 1e3))
        # the same pass as the pipeliner runs it: HOST_CHUNKS calls of whole timesteps, each
        # call's staging copies overlapping the previous call's kernels (two staging sets)
        chunks = host_chunk_bindings(obs, spec, HOST_CHUNKS)
        check(lib.bullseye_set_options(ctypes.byref(o)), "options")
        check(lib.initLibrary(chunks[-1].byref()), "initLibrary(host chunks)")
        chunk_ms = []
        for i in range(3):
            t = time.perf_counter()
            check(lib.bullseye_reset_grids(), "reset")
            for bc in chunks:
                check(lib.facet_4_cor_corrections(bc.byref()), "grid(host chunk)")
            lib.gridding_barrier()
            chunk_ms.append((time.perf_counter() - t) * 1e3)
        lib.releaseLibrary()
        grid_chunks = chunks[-1].host("output_buffer").astype(np.complex128)
        del chunks
        wall_chunks = min(chunk_ms[1:]) * 1e-3
        chunk_err = rel_errors(grid_chunks.reshape(F, -1), grid32.reshape(F, -1))
        del grid_chunks
        log("FP32-comp with host buffers in %d chunks: %.2f ms/step (grid vs one call: rms %.1e)"
            % (HOST_CHUNKS, wall_chunks * 1e3, chunk_err[0]))

    gtaps32 = taps / wall32 / 1e9
    gtaps64 = taps64 / wall64 / 1e9
    flops = taps * 4 * FLOPS_PER_TAP_PER_CORR
    achieved_tflops = flops / (k32 * 1e-3) / 1e12

    result = {
        "metric": METRIC, "value": round(gtaps32, 3), "unit": "Gtaps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall32 * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded; inputs resident in HBM)",
        "config": {"workload": "targeted faceting, 4 correlations with Jones: 64 ant (2016 "
                               "baselines) x 240 t x 64 ch, 25 facets of 512^2 at 2\", 7x7 "
                               "PSWF x63 (facet_4_cor_corrections)",
                   "rows": obs.rows, "channels": obs.n_chan, "correlations": 4,
                   "facets": im["facets"], "npix": im["npix"], "taps_per_step": taps,
                   "samples_per_step": nsamp},
        "fp64_gtaps_per_s": round(gtaps64, 3), "fp32comp_over_fp64": round(gtaps32 / gtaps64, 3),
        "host_buffers_ms_per_step": None if wall_host is None else round(wall_host * 1e3, 3),
        "host_buffers_gtaps_per_s": None if wall_host is None else round(taps / wall_host / 1e9, 3),
        "host_buffers_chunked": None if wall_chunks is None else {
            "chunks": HOST_CHUNKS, "ms_per_step": round(wall_chunks * 1e3, 3),
            "gtaps_per_s": round(taps / wall_chunks / 1e9, 3), "rms_vs_one_call": chunk_err[0]},
        "rms_err_vs_fp64_gpu": rms_vs_gpu64, "peak_err_vs_fp64_gpu": peak_vs_gpu64,
        "error_per_facet_vs_fp64_gpu": err_gpu,
        "roofline": {"bound": "mfma", "achieved": round(achieved_tflops, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic("romein_w_facet_4_cor_corrections_fp32comp_pmc.md"),
                     "kernel_ms": round(k32, 3),
                     "flops_per_launch": flops},
        "flush_atomics": flush_atomics(
            lib, "romein_w_facet_4_cor_corrections_fp32comp_pmc.md", k32),
    }

    if not args.no_cpu:
        from oracle import oracle
        n = args.cpu_sample_rows or (obs.rows * 3 // 4)
        sub = (0, n)
        spec64 = ImagingSpec(**{**spec.__dict__, "jones": jones_reference_layout(
            obs, im["facets"], np.complex128)})
        ref = Binding("double", obs, spec64, rows=sub)
        t = time.perf_counter()
        norm_or = oracle.run(ref, "facet_4_cor_corrections", normalize=False)
        t_or64 = time.perf_counter() - t
        # GPU FP32-comp on the same rows
        set_opts(PATH_FP32_KAHAN)
        bs = Binding("mixed", obs, spec, rows=sub, device=dev)
        torch.cuda.synchronize()
        check(lib.initLibrary(bs.byref()), "initLibrary(sample)")
        check(lib.facet_4_cor_corrections(bs.byref()), "grid(sample)")
        lib.gridding_barrier()
        lib.bullseye_get_stats(ctypes.byref(stats))
        sample_taps = stats.taps
        norm_s = norm_terms(bs, F * 4)
        lib.releaseLibrary()
        gs = bs.host("output_buffer")
        rms_o, peak_o = rel_errors(gs.reshape(F, -1), ref.output.reshape(F, -1))
        result["error_per_facet_vs_fp64_oracle"] = facet_error_table(
            gs.reshape(F, 4, npx, npx), ref.output.reshape(F, 4, npx, npx), norm_s, norm_or,
            taper)
        del gs
        # CPU baseline: the single-precision restatement on the same sample, all host cores
        ref32 = Binding("single", obs, spec, rows=sub)
        t = time.perf_counter()
        oracle.run(ref32, "facet_4_cor_corrections", precision="single", normalize=False)
        t_or32 = time.perf_counter() - t
        threads = oracle.load("single").oracle_max_threads()
        result["rms_err_vs_fp64_oracle"] = rms_o
        result["peak_err_vs_fp64_oracle"] = peak_o
        result["cpu_baseline"] = {
            "value": round(sample_taps / t_or32 / 1e9, 4), "unit": "Gtaps/s",
            "cores": min(threads, im["facets"]), "kind": "port",
            "sample": "first %d of %d rows (%d timesteps) of the same workload, %.2e taps, "
                      "%.1f s; OpenMP over facets (reference scheme)" % (
                          n, obs.rows, n // 2016, sample_taps, t_or32)}
        result["fp64_oracle_cpu_gtaps_per_s"] = round(sample_taps / t_or64 / 1e9, 4)
        del ref, ref32, bs
    if not args.no_batch:
        # the batched uneven-facet workload (pattern 3's target configuration), one measured
        # replay of each schedule after one warm-up replay
        del b, obs
        torch.cuda.empty_cache()
        r = batch_workload(args, 1, 1, True, False, False)
        result["batched_facets"] = {
            "workload": r["config"]["workload"], "facets": r["config"]["facets"],
            "taps_per_step": r["config"]["taps_per_step"], "gtaps_per_s": r["value"],
            "ms_per_step": r["ms_per_step"], "fp64_gtaps_per_s": r["fp64_gtaps_per_s"],
            "fp32comp_over_fp64": r["fp32comp_over_fp64"],
            "per_facet_launch_gtaps_per_s": r["per_facet_launch_gtaps_per_s"],
            "batched_over_per_facet": r["batched_over_per_facet"],
            "rms_err_vs_fp64_gpu_worst_facet": r["rms_err_vs_fp64_gpu_worst_facet"],
            "peak_err_vs_fp64_gpu_worst_facet": r["peak_err_vs_fp64_gpu_worst_facet"],
            "roofline_frac": r["roofline"]["frac"], "flush_atomics": r["flush_atomics"]}
    print(json.dumps(result), flush=True)


def main_plain(args):
    """Plain single-facet gridding, the shape of the reference's benchmark
    (bullseye/mo/benchmarks/benchmark.cpp:97-289, one grid_single_pol facet): 64 antennas
    uniform in a 3 km disc (2016 baselines), 240 timesteps over 8 h at dec -30 deg, 16 channels
    1.40-1.42 GHz, 1 correlation complex64 (20 point sources + noise, 5% flags, natural
    weights); one 1024^2 grid at 2", 7x7 PSWF oversampled x63, no w-projection.  The whole
    workload is also gridded by the FP64 CPU oracle (one thread: one grid) for the error, and by
    the single-precision restatement for the CPU baseline."""
    import numpy as np
    import torch
    from bullseye_amd import kernels, synth
    from bullseye_amd.binding import Binding, ImagingSpec
    from bullseye_amd.library_loader import (PATH_FP32_KAHAN, PATH_FP64, bullseye_options,
                                             bullseye_stats, load_library)
    cfg = synth.CONFIGS["c0"]
    im = cfg["image"]
    dev = torch.device("cuda:0")
    t0 = time.time()
    obs = synth.simulate(**cfg["sim"], seed=args.seed, device=dev)
    conv = kernels.pswf_aa_kernel(im["support"], im["oversample"])
    spec = ImagingSpec(npix=im["npix"], cell=im["cell"], conv=conv, support=im["support"],
                       oversample=im["oversample"])
    log("synthesised %d rows x %d ch in %.1fs" % (obs.rows, obs.n_chan, time.time() - t0))
    lib = load_library("GPU", "mixed")
    b = Binding("mixed", obs, spec, device=dev)
    torch.cuda.synchronize()

    def check(rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib.bullseye_last_error().decode()))

    def set_opts(path):
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                             compute_path=path, channels_per_stream=args.channels_per_stream,
                             rows_per_stream=args.rows_per_stream)
        check(lib.bullseye_set_options(ctypes.byref(o)), "options")

    set_opts(PATH_FP32_KAHAN)
    check(lib.initLibrary(b.byref()), "initLibrary")
    stats = bullseye_stats()
    calibration_copy(args.calibration_copy_mib)

    def run(path, steps, warmup):
        set_opts(path)
        for _ in range(warmup):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.grid_single_pol(b.byref()), "grid")
            lib.gridding_barrier()
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.grid_single_pol(b.byref()), "grid")
            lib.gridding_barrier()
            lib.bullseye_get_stats(ctypes.byref(stats))
            kms.append(stats.kernel_ms)
        torch.cuda.synchronize()
        return ((time.perf_counter() - t) / steps, stats.taps, float(np.mean(kms)),
                stats.samples_gridded)

    steps = max(args.steps, 20)
    wall32, taps, k32, nsamp = run(PATH_FP32_KAHAN, steps, max(args.warmup, 3))
    grid32 = b.host("output_buffer").copy()
    norm32 = norm_terms(b, 1)
    log("FP32-comp: %.3f ms/step, kernel %.3f ms, %.3e taps" % (wall32 * 1e3, k32, taps))
    wall64, taps64, k64, _ = run(PATH_FP64, steps, 3)
    grid64 = b.host("output_buffer").copy()
    norm64 = norm_terms(b, 1)
    log("FP64: %.3f ms/step, kernel %.3f ms" % (wall64 * 1e3, k64))
    lib.releaseLibrary()
    npx = im["npix"]
    taper = kernels.grid_taper(conv, im["support"], im["oversample"], npx)
    shape = (1, 1, npx, npx)
    # algorithmic bytes of one pass: 13 B per sample (c64 visibility, f32 weight, bool flag),
    # 33 B per row (f64 uvw, spw, field, row flag) and one complex64 grid write
    alg_bytes = obs.rows * obs.n_chan * 13 + obs.rows * 33 + npx * npx * 8
    result = {
        "metric": METRIC, "value": round(taps / wall32 / 1e9, 3), "unit": "Gtaps/s",
        "n_gpus": 1, "steps": steps, "warmup": max(args.warmup, 3),
        "ms_per_step": round(wall32 * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded; inputs resident in HBM)",
        "config": {"workload": "plain single-facet gridding: 64 ant (2016 baselines) x 240 t x "
                               "16 ch (1.40-1.42 GHz), 1 corr, one 1024^2 grid at 2\", 7x7 PSWF "
                               "x63, no w-projection (grid_single_pol)",
                   "rows": obs.rows, "channels": obs.n_chan, "taps_per_step": taps,
                   "samples_per_step": nsamp},
        "fp64_gtaps_per_s": round(taps64 / wall64 / 1e9, 3),
        "fp32comp_over_fp64": round(wall64 / wall32, 3),
        "error_vs_fp64_gpu": facet_error_table(grid32.reshape(shape), grid64.reshape(shape),
                                               norm32, norm64, taper)["worst"],
        "roofline": {"bound": "hbm", "achieved": round(alg_bytes / (k32 * 1e-3) / 1e9, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg_bytes / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                     "traffic": profiled_traffic("romein_w_grid_single_pol_plain_pmc.md"),
                     "kernel_ms": round(k32, 4), "bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        from oracle import oracle
        ref = Binding("double", obs, spec)
        t = time.perf_counter()
        norm_or = oracle.run(ref, "grid_single_pol", normalize=False)
        t_or64 = time.perf_counter() - t
        result["error_vs_fp64_oracle"] = facet_error_table(
            grid32.reshape(shape), ref.output.reshape(shape), norm32, norm_or, taper)["worst"]
        result["fp64_path_error_vs_fp64_oracle"] = facet_error_table(
            grid64.reshape(shape), ref.output.reshape(shape), norm64, norm_or, taper)["worst"]
        del ref
        ref32 = Binding("single", obs, spec)
        t = time.perf_counter()
        oracle.run(ref32, "grid_single_pol", precision="single", normalize=False)
        t_or32 = time.perf_counter() - t
        result["cpu_baseline"] = {
            "value": round(taps / t_or32 / 1e9, 4), "unit": "Gtaps/s", "cores": 1,
            "kind": "port",
            "sample": "the whole workload, %.2e taps, %.1f s; one thread (the reference "
                      "parallelises over facets and this workload has one)" % (taps, t_or32)}
        result["fp64_oracle_cpu_gtaps_per_s"] = round(taps / t_or64 / 1e9, 4)
    print(json.dumps(result), flush=True)


def main_wproj(args):
    """W-projection workload: 128 antennas in a 10 km log-radial layout (8128 baselines), 480
    timesteps, 32 channels 1.0-1.5 GHz, 1 correlation; one 4096^2 grid at 1.5", 128 w-planes
    whose kernels grow from 7 to 63 taps with |w|, oversampling x8, complex64 kernels."""
    import numpy as np
    import torch
    from bullseye_amd import kernels, synth
    from bullseye_amd.binding import Binding, ImagingSpec
    from bullseye_amd.library_loader import (PATH_FP32_KAHAN, PATH_FP64, bullseye_options,
                                             bullseye_stats, load_library)
    cfg = synth.CONFIGS["c2"]
    im = cfg["image"]
    dev = torch.device("cuda:0")
    t0 = time.time()
    obs = synth.simulate(**cfg["sim"], seed=args.seed, device=dev)
    wl = np.abs(obs.uvw[:, 2:3] / obs.wavelengths[0][None, :])
    wmax = float(np.float32(wl.max()))
    halves = kernels.linear_plane_supports(im["wplanes"], 3, im["support"])
    conv = kernels.wproj_kernels(im["support"], im["oversample"], im["wplanes"], wmax,
                                 im["npix"], im["cell"], half_supports=halves,
                                 dtype=np.complex64, device=dev)
    spec = ImagingSpec(npix=im["npix"], cell=im["cell"], conv=conv, support=im["support"],
                       oversample=im["oversample"], wplanes=im["wplanes"], wmax=wmax)
    log("synthesised %d rows x %d ch and %d w-kernels in %.1fs" % (obs.rows, obs.n_chan,
                                                                  im["wplanes"], time.time() - t0))
    lib = load_library("GPU", "mixed")
    b = Binding("mixed", obs, spec, device=dev)
    torch.cuda.synchronize()
    arr = (ctypes.c_uint32 * len(halves))(*halves)
    assert lib.bullseye_set_wplane_support(arr, len(halves)) == 0

    def check(rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib.bullseye_last_error().decode()))

    def set_opts(path):
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                             compute_path=path, channels_per_stream=args.channels_per_stream)
        check(lib.bullseye_set_options(ctypes.byref(o)), "options")

    set_opts(PATH_FP32_KAHAN)
    t = time.time()
    check(lib.initLibrary(b.byref()), "initLibrary")
    log("initLibrary (tap re-layout) %.1fs" % (time.time() - t))
    stats = bullseye_stats()
    calibration_copy(args.calibration_copy_mib)

    def run(path, steps, warmup):
        set_opts(path)
        for _ in range(warmup):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.grid_single_pol(b.byref()), "grid")
            lib.gridding_barrier()
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            check(lib.bullseye_reset_grids(), "reset")
            check(lib.grid_single_pol(b.byref()), "grid")
            lib.gridding_barrier()
            lib.bullseye_get_stats(ctypes.byref(stats))
            kms.append(stats.kernel_ms)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, stats.taps, float(np.mean(kms)), stats.samples_gridded

    wall32, taps, k32, nsamp = run(PATH_FP32_KAHAN, args.steps, args.warmup)
    grid32 = b.host("output_buffer").copy()
    log("FP32-comp: %.2f ms/step, kernels %.2f ms, %.3e taps" % (wall32 * 1e3, k32, taps))
    wall64, taps64, k64, _ = run(PATH_FP64, max(1, args.steps // 2), 1)
    grid64 = b.host("output_buffer").copy()
    log("FP64: %.2f ms/step, kernels %.2f ms" % (wall64 * 1e3, k64))
    lib.releaseLibrary()
    lib.bullseye_set_wplane_support(None, 0)
    rms, peak = rel_errors(grid32.reshape(1, -1), grid64.reshape(1, -1))
    flops = taps * 8
    result = {
        "metric": METRIC, "value": round(taps / wall32 / 1e9, 3), "unit": "Gtaps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall32 * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded; inputs resident in HBM)",
        "config": {"workload": "w-projection: 128 ant (8128 baselines, 10 km log-radial) x 480 t "
                               "x 32 ch (1.0-1.5 GHz), 1 corr, 4096^2 at 1.5 arcsec, 128 w-planes, "
                               "support 7->63 by |w|, OS x8, complex64 kernels (grid_single_pol)",
                   "rows": obs.rows, "channels": obs.n_chan, "taps_per_step": taps,
                   "samples_per_step": nsamp, "wmax_lambda": wmax},
        "fp64_gtaps_per_s": round(taps64 / wall64 / 1e9, 3),
        "fp32comp_over_fp64": round((taps / wall32) / (taps64 / wall64), 3),
        "rms_err_vs_fp64_gpu": rms, "peak_err_vs_fp64_gpu": peak,
        "roofline": {"bound": "mfma", "achieved": round(flops / (k32 * 1e-3) / 1e12, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (k32 * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic("wproj_grid_single_pol_fp32comp_pmc.md"),
                     "kernel_ms": round(k32, 3), "flops_per_launch": flops},
        "flush_atomics": flush_atomics(lib, "wproj_grid_single_pol_fp32comp_pmc.md", k32),
    }
    if not args.no_cpu:
        from oracle import oracle
        n = args.cpu_sample_rows or 4 * 8128
        sub = (0, n)
        # GPU FP32-comp on the sample: its algorithmic tap count and its error vs the oracle
        lib.bullseye_set_wplane_support(arr, len(halves))
        set_opts(PATH_FP32_KAHAN)
        bs = Binding("mixed", obs, spec, rows=sub, device=dev)
        torch.cuda.synchronize()
        check(lib.initLibrary(bs.byref()), "initLibrary(sample)")
        check(lib.grid_single_pol(bs.byref()), "grid(sample)")
        lib.gridding_barrier()
        lib.bullseye_get_stats(ctypes.byref(stats))
        sample_taps = stats.taps
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
        gpu_sample = bs.host("output_buffer").copy()
        del bs
        ref = Binding("double", obs, spec, rows=sub)
        t = time.perf_counter()
        oracle.run(ref, "grid_single_pol", normalize=False)
        t_or64 = time.perf_counter() - t
        rms_o, peak_o = rel_errors(gpu_sample.reshape(1, -1), ref.output.reshape(1, -1))
        del ref
        ref32 = Binding("single", obs, spec, rows=sub)
        t = time.perf_counter()
        oracle.run(ref32, "grid_single_pol", precision="single", normalize=False)
        t_or32 = time.perf_counter() - t
        result["rms_err_vs_fp64_oracle"] = rms_o
        result["peak_err_vs_fp64_oracle"] = peak_o
        result["cpu_baseline"] = {
            "value": round(sample_taps / t_or32 / 1e9, 4), "unit": "Gtaps/s", "cores": 1,
            "kind": "port",
            "sample": "first %d of %d rows (%d timesteps) of the same workload, %.2e in-support "
                      "taps, %.1f s; one thread (the reference parallelises over facets; this "
                      "workload has one grid); the oracle visits the full 63x63 support"
                      % (n, obs.rows, n // 8128, sample_taps, t_or32)}
        result["fp64_oracle_cpu_gtaps_per_s"] = round(sample_taps / t_or64 / 1e9, 4)
    print(json.dumps(result), flush=True)


def main_adjoint(args):
    """Differentiable facet imaging: 1M visibilities (a seeded row subset of the w-projection
    observation, all 32 channels), one 2048^2 facet at 1.5 arcsec centred 0.2 deg from the
    phase centre, 32 w-planes (support 7->63 by |w|, OS x8, complex64 kernels), FP32
    compensated path.  A step is one chi-squared evaluation between the degridded model
    visibilities of a sparse model image (500 pixels ~ U(0, 1)) and the data, and its
    gradient with respect to the image: FFT, degridding, loss, gridding, FFT, all on the
    device."""
    import numpy as np
    import torch
    from bullseye_amd import synth
    from bullseye_amd.autograd import FacetImager, chi_squared
    from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64
    cfg = synth.CONFIGS["c2"]
    dev = torch.device("cuda:0")
    t0 = time.time()
    obs = synth.simulate(**cfg["sim"], seed=args.seed, device=dev)
    obs, spec, halves, img = synth.differentiable_facet(obs, args.seed + 7, device=dev)
    log("synthesised %d rows x %d ch (%d samples), kernels in %.1fs" % (
        obs.rows, obs.n_chan, obs.rows * obs.n_chan, time.time() - t0))
    calibration_copy(args.calibration_copy_mib)

    def step(im, image_t, data):
        image_t.grad = None
        loss = chi_squared(image_t, data, im)
        loss.backward()
        return loss

    def run(path, dtype, steps, warmup):
        with FacetImager(obs, spec, mode="facet_single_pol", precision="mixed", path=path,
                         device=dev, wplane_support=halves) as im:
            data = im.b.keep["visibilities"]
            image_t = torch.tensor(img, dtype=dtype, device=dev, requires_grad=True)
            for _ in range(warmup):
                step(im, image_t, data)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                loss = step(im, image_t, data)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t) / steps
            taps = im.stats().taps
            grad = image_t.grad.double().cpu().numpy()
            # dot-product test of the pair on this data
            g = im.grid(data)
            noise = torch.randn(g.shape, dtype=g.dtype, device=dev,
                                generator=torch.Generator(device=dev).manual_seed(3))
            gg = (g / g.norm() + noise / noise.norm()).to(g.dtype)
            lhs = torch.vdot(g.reshape(-1).to(torch.complex128),
                             gg.reshape(-1).to(torch.complex128)).real.item()
            rhs = torch.vdot(data.reshape(-1).to(torch.complex128),
                             im.degrid(gg).reshape(-1).to(torch.complex128)).real.item()
            dot = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        return wall, taps, grad, loss.item(), dot

    wall32, taps, g32, loss32, dot32 = run(PATH_FP32_KAHAN, torch.float32, args.steps,
                                            args.warmup)
    log("FP32-comp: %.2f ms/step (value+gradient), %.3e taps per pass, dot %.2e" % (
        wall32 * 1e3, taps, dot32))
    wall64, _, g64, loss64, dot64 = run(PATH_FP64, torch.float64, max(1, args.steps // 2), 1)
    log("FP64: %.2f ms/step, dot %.2e" % (wall64 * 1e3, dot64))
    peak = np.abs(g64).max()
    rms = float(np.sqrt(np.mean((g32 - g64) ** 2)) / peak)
    pk = float(np.abs(g32 - g64).max() / peak)
    result = {
        "metric": METRIC, "value": round(2 * taps / wall32 / 1e9, 3), "unit": "Gtaps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall32 * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded; inputs resident in HBM)",
        "config": {"workload": "differentiable facet imaging: 1M visibilities of the "
                               "w-projection observation, one 2048^2 facet at 1.5 arcsec, 32 "
                               "w-planes (support 7->63, OS x8), chi-squared value + gradient "
                               "of a 500-pixel model (degrid + grid per step)",
                   "rows": obs.rows, "channels": obs.n_chan, "taps_per_pass": taps},
        "fp64_ms_per_step": round(wall64 * 1e3, 3),
        "dot_product_rel_fp32": dot32, "dot_product_rel_fp64": dot64,
        "grad_rms_err_vs_fp64_gpu": rms, "grad_peak_err_vs_fp64_gpu": pk,
        "loss_fp32": loss32, "loss_fp64": loss64,
    }
    print(json.dumps(result), flush=True)

def main_batch(args):
    print(json.dumps(batch_workload(args, args.steps, args.warmup, not args.no_compare, True,
                                    not args.no_cpu)), flush=True)


def batch_workload(args, steps, warmup, compare_fp64, compare_tiles, cpu):
    """Batched uneven facets: the w-projection workload's visibilities (128 antennas, 10 km
    log-radial, 32 channels 1.0-1.5 GHz, 1 correlation; --batch-timesteps of its 480 timesteps)
    gridded into --batch-facets facets with sizes log-uniform over 128^2-2048^2 at 1.5", each
    with one of seven w-kernel stacks (16, 24, 32, 48, 64, 96 or 128 planes, supports 7->63 by
    |w|, OS x8, complex64) and a quarter of them sampling-function facets.  A step is one
    bullseye_batch_execute: replay of the plan's HIP graph (clear, one persistent launch per
    (kind, support class) over all facets) and the copy of every facet's grid to its output."""
    import numpy as np
    import torch
    from bullseye_amd import kernels, synth
    from bullseye_amd.batch import FacetBatch
    from bullseye_amd.binding import Binding, ImagingSpec
    from bullseye_amd.library_loader import (BATCH_PER_FACET, BATCH_PERSISTENT_GRAPH,
                                             PATH_FP32_KAHAN, PATH_FP64, bullseye_options,
                                             bullseye_stats, load_library)
    cfg = synth.CONFIGS["c2"]
    im = cfg["image"]
    dev = torch.device("cuda:0")
    t0 = time.time()
    sim = dict(cfg["sim"])
    sim["n_time"] = args.batch_timesteps
    obs = synth.simulate(**sim, seed=args.seed, device=dev)
    nf = args.batch_facets
    facets, ksets, plane_counts, wmax, sizes, psf = synth.batched_facets(
        obs, nf, args.seed + 200, device=dev)
    log("synthesised %d rows x %d ch, %d kernel stacks, %d facets (%.0f Mpix) in %.1fs" % (
        obs.rows, obs.n_chan, len(ksets), nf, float((sizes.astype(float) ** 2).sum()) / 1e6,
        time.time() - t0))

    lib = load_library("GPU", "mixed")

    def check(rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib.bullseye_last_error().decode()))

    def set_opts(path):
        o = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                             compute_path=path)
        check(lib.bullseye_set_options(ctypes.byref(o)), "options")

    spec0 = ImagingSpec(npix=256, cell=im["cell"], conv=kernels.pswf_aa_kernel(3, 63),
                        support=3, oversample=63)
    b = Binding("mixed", obs, spec0, device=dev, with_psf=True)
    torch.cuda.synchronize()
    set_opts(PATH_FP32_KAHAN)
    check(lib.initLibrary(b.byref()), "initLibrary")
    batch = FacetBatch(lib, b, "facet_single_pol")
    for k, (conv, halves) in enumerate(ksets):
        batch.kernel_set(k, conv, im["support"], im["oversample"], halves, wmax)
    # the CPU baseline's facet: the median size among the visibility facets
    vis_idx = [i for i in range(nf) if not psf[i]]
    cpu_f = facets[sorted(vis_idx, key=lambda i: sizes[i])[len(vis_idx) // 2]]
    cpu_conv, cpu_halves = ksets[cpu_f.kernel_set]
    del ksets
    stats = bullseye_stats()
    calibration_copy(args.calibration_copy_mib)

    tune = None
    if args.tuning:
        tj = json.loads(args.tuning)
        batch.set_tuning(channels_per_stream=tj.get("channels_per_stream", 0),
                         base_rows=tj.get("base_rows", 0),
                         shape={int(m): v for m, v in tj.get("shape", {}).items()},
                         chunk={int(m): v for m, v in tj.get("chunk", {}).items()},
                         facet_major=tj.get("facet_major", 0))
    elif args.tune_timesteps > 0:
        t = time.time()
        bt_ = Binding("mixed", obs, spec0, rows=(0, args.tune_timesteps * obs.baseline_count),
                      device=dev, with_psf=True)
        tuner = FacetBatch(lib, bt_, "facet_single_pol")
        tuner.set_tuning(channels_per_stream=8)
        tune = tuner.autotune(facets)
        del bt_, tuner
        log("autotune on %d timesteps: %.1fs, tuned %.1f ms vs default tiles %.1f ms; cps %d, "
            "base rows %d" % (args.tune_timesteps, time.time() - t, tune["tuned_ms"],
                              tune["default_ms"], tune["channels_per_stream"], tune["base_rows"]))

    def run(path, schedule, steps, warmup, what):
        set_opts(path)
        t = time.time()
        batch.plan(facets, schedule)
        tplan = time.time() - t
        for _ in range(warmup):
            batch.execute()
            log("%s: warmup step done" % what)
        torch.cuda.synchronize()
        kms, taps = [], 0
        t = time.perf_counter()
        for i in range(steps):
            batch.execute()
            lib.bullseye_get_stats(ctypes.byref(stats))
            kms.append(stats.kernel_ms)
            taps = stats.taps
            log("%s: step %d, kernels %.1f ms" % (what, i, stats.kernel_ms))
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        log("%s: %.1f ms/step (plan %.1fs), %.3e taps, %.1f Gtaps/s" % (
            what, wall * 1e3, tplan, taps, taps / wall / 1e9))
        return wall, taps, float(np.mean(kms)), stats.samples_gridded, tplan

    wall32, taps, k32, nsamp, tplan = run(PATH_FP32_KAHAN, BATCH_PERSISTENT_GRAPH, steps,
                                          warmup, "persistent graph, tuned")
    tuned = batch.tuning()
    compare = {}
    if compare_fp64:
        out32 = [o.clone() for o in batch.outputs]
        wall64, taps64, _, _, _ = run(PATH_FP64, BATCH_PERSISTENT_GRAPH, 1, 0, "FP64 persistent")
        errs = []
        for o32, o64 in zip(out32, batch.outputs):
            w = o64.to(torch.complex128)
            d = (o32.to(torch.complex128) - w).abs()
            peak = w.abs().max().item()
            if peak > 0:
                errs.append((torch.sqrt((d ** 2).mean()).item() / peak, d.max().item() / peak))
        del out32
        # one launch per (facet, class) of the single-facet kernels, with the tuned tiles
        wall_pf, taps_pf, _, _, _ = run(PATH_FP32_KAHAN, BATCH_PER_FACET, 1, 0,
                                        "one launch per facet and class")
        assert taps_pf == taps and taps64 == taps
        compare = {
            "fp64_gtaps_per_s": round(taps64 / wall64 / 1e9, 3),
            "fp32comp_over_fp64": round(wall64 / wall32, 3),
            "per_facet_launch_gtaps_per_s": round(taps_pf / wall_pf / 1e9, 3),
            "batched_over_per_facet": round(wall_pf / wall32, 3),
            "rms_err_vs_fp64_gpu_worst_facet": max(e[0] for e in errs),
            "peak_err_vs_fp64_gpu_worst_facet": max(e[1] for e in errs)}
        if compare_tiles:
            batch.set_tuning()
            wall_d, taps_d, _, _, _ = run(PATH_FP32_KAHAN, BATCH_PERSISTENT_GRAPH, 1, 0,
                                          "persistent graph, default tiles")
            assert taps_d == taps
            compare["default_tiles_gtaps_per_s"] = round(taps_d / wall_d / 1e9, 3)
    batch.release()
    lib.releaseLibrary()
    flops = taps * 8
    n_psf = int(psf.sum())
    result = {
        "metric": METRIC, "value": round(taps / wall32 / 1e9, 3), "unit": "Gtaps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall32 * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded; inputs resident in HBM)",
        "config": {"workload": "batched uneven facets: w-projection visibilities (128 ant, 8128 "
                               "baselines, %d t x 32 ch 1.0-1.5 GHz, 1 corr) into %d facets, "
                               "sizes log-uniform 128^2-2048^2 at 1.5 arcsec, 7 w-kernel stacks "
                               "of 16-128 planes (supports 7->63 by |w|, OS x8), %d sampling-"
                               "function facets; persistent class launches in one HIP graph"
                               % (args.batch_timesteps, nf, n_psf),
                   "rows": obs.rows, "channels": obs.n_chan, "facets": nf,
                   "facet_mpix": round(float((sizes.astype(float) ** 2).sum()) / 1e6, 2),
                   "taps_per_step": taps, "samples_per_step": nsamp,
                   "plan_s": round(tplan, 2),
                   "tuning": {"channels_per_stream": tuned["channels_per_stream"],
                              "base_rows": tuned["base_rows"],
                              "facet_major": tuned["facet_major"],
                              "shape": {m: v for m, v in tuned["shape"].items() if v},
                              "chunk": {m: v for m, v in tuned["chunk"].items() if v > 1}}},
        **compare,
        "roofline": {"bound": "mfma", "achieved": round(flops / (k32 * 1e-3) / 1e12, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (k32 * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic("batch_facets_fp32comp_pmc.md"),
                     "kernel_ms": round(k32, 3), "flops_per_launch": flops},
        "flush_atomics": flush_atomics(lib, "batch_facets_fp32comp_pmc.md", k32),
    }
    if tune is not None:
        result["autotune"] = {"timesteps": args.tune_timesteps,
                              "tuned_ms": round(tune["tuned_ms"], 3),
                              "default_ms": round(tune["default_ms"], 3)}
    if cpu:
        # one facet of the batch on the first two timesteps: its GPU tap count through the
        # single-facet entry point, then the oracle's single-precision build on the same rows
        from oracle import oracle
        P = plane_counts[cpu_f.kernel_set]
        spec_f = ImagingSpec(npix=cpu_f.npix, cell=cpu_f.cell, conv=cpu_conv,
                             support=im["support"], oversample=im["oversample"], wplanes=P,
                             wmax=wmax, facet_centres=np.array([cpu_f.centre]))
        sub = (0, 4 * obs.baseline_count)
        arr = (ctypes.c_uint32 * len(cpu_halves))(*cpu_halves)
        check(lib.bullseye_set_wplane_support(arr, len(cpu_halves)), "supports")
        set_opts(PATH_FP32_KAHAN)
        bs = Binding("mixed", obs, spec_f, rows=sub, device=dev)
        torch.cuda.synchronize()
        check(lib.initLibrary(bs.byref()), "initLibrary(sample)")
        check(lib.facet_single_pol(bs.byref()), "grid(sample)")
        lib.gridding_barrier()
        lib.bullseye_get_stats(ctypes.byref(stats))
        sample_taps = stats.taps
        lib.releaseLibrary()
        lib.bullseye_set_wplane_support(None, 0)
        del bs
        ref32 = Binding("single", obs, spec_f, rows=sub)
        t = time.perf_counter()
        oracle.run(ref32, "facet_single_pol", precision="single", normalize=False)
        t_or32 = time.perf_counter() - t
        result["cpu_baseline"] = {
            "value": round(sample_taps / t_or32 / 1e9, 4), "unit": "Gtaps/s", "cores": 1,
            "kind": "port",
            "sample": "one %d^2 facet of the batch (%d w-planes) on the first 4 of %d timesteps, "
                      "%.2e in-support taps, %.1f s; one thread (the reference parallelises over "
                      "the facets of one geometry); the oracle visits the full 63x63 support"
                      % (cpu_f.npix, P, args.batch_timesteps, sample_taps, t_or32)}
    return result


if __name__ == "__main__":
    main()
