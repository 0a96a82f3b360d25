// wrapper.hip -- the C ABI of the gfx950 imaging library (include/bullseye_amd.h).
//
// One library context per process, like the reference (cpu_algorithm/wrapper.cpp:56-64,
// gpu_algorithm/wrapper.cu:58-71).  Gridding accumulates on the device across chunks into an
// FP64 grid that initLibrary seeds from the caller's output buffer; the final chunk converts it
// back into the caller's buffer (gpu_algorithm/wrapper.cu:293-314).  Calls return an int status
// and never throw across the ABI.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "host_common.h"

namespace bam {



// Two timing events released on every return path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create() {
        hipError_t e = hipEventCreate(&e0);
        return e != hipSuccess ? e : hipEventCreate(&e1);
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

struct TimedCall {
    // start/stop on the stream; cstart/copied around the staging copies (the copy stream)
    hipEvent_t start, cstart, copied, k0, k1, stop;
    bool has_kernel;
    bool recorded;   // every event recorded (a call that failed midway is only released)
};

struct State {
    bool initialized = false;
    bullseye_options opts{};
    std::string err;
    hipStream_t stream = nullptr;
    // w-projection class launches run on the stream and WSIDE side streams (default priority),
    // forked after the class lists are built and joined before the call's end
    static constexpr int WSIDE = 3;
    hipStream_t wside[WSIDE] = {};
    hipEvent_t wfork = nullptr, wjoin[WSIDE] = {};
    // geometry fixed at initLibrary
    size_t F = 0, cube = 0, P = 0, pstride = 0, nx = 0, ny = 0, S = 0, OS = 0, csz = 0;
    size_t sfC = 0, wplanes = 1;
    // normalisation sums and counters: nslots copies of norm_stride = F*cube*P (and 2) elements
    size_t norm_stride = 0;
    uint32_t nslots = 1;
    bool psf = false;
    int path = PATH_FP64;
    // device buffers
    DevBuf grid, psfgrid, norm, counters, errflag, conv_f, conv_d, tapsum, geom, taper;
    DevBuf perm, blstart, keys, vals, keys_out, counts, ws, rowrec, chanrec;
    // Host-buffer mode: two sets of staging buffers used by alternate calls.  A call's copies
    // run on the copy stream once the call before the previous one (the last reader of the
    // set) has finished on the stream, so they overlap the previous call's kernels; the
    // stream waits for the copies before its first reader.
    struct InputSet {
        DevBuf vis, uvw, w, flags, rowflags, field, spw, a1, a2, time, jones, wl, enabled, cgi,
            fc;
        hipEvent_t staged = nullptr, done = nullptr;
        bool used = false;
    };
    InputSet ins[2];
    int ins_cur = 0;
    hipStream_t cstream = nullptr;
    DevBuf out_tmp, fft_scratch, fft_real, jinv, walk, dg_model, dg_out, stokes_tmp, dft_smp,
        dft_comp;
    // w-projection: taps re-laid per (plane, frac_v, frac_u), per-plane supports and sums
    DevBuf wtaps_f, wtaps_d, wplane_off, wplane_half, wplane_norm, wcls, wlist;
    DevBuf wctaps_f, wctaps_d, wcplane_off;   // column tables of the narrow classes
    std::vector<uint32_t> wplane_half_req;   // from bullseye_set_wplane_support
    std::vector<uint32_t> wplane_half_host;
    uint32_t w_nf = 0, w_fbase = 0;
    std::vector<normalization_base_type> host_norm;
    std::vector<double> taper_x, taper_y;
    // timing
    std::vector<TimedCall> pending;
    double grid_seconds = 0, inv_seconds = 0;
    bullseye_stats stats{};
    unsigned long long last_counters[2] = {0, 0};
    int device = 0;                 // the HIP device of the context (initLibrary's current one)
    // inside bullseye_grid_noalloc: the call's tables are carved from this arena
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};
static State g;

// Scratch of the allocation-free gridding entry point (bullseye_grid_noalloc): the per-call
// tables of one gridding call -- inverted Jones terms, facet geometry, the baseline sort's keys,
// permutation, counts and workspace, the walk-order measurement, the row and channel tables --
// carved out of one static device array.  Every GPU the library runs on holds its own copy
// until the process ends.  Host code never uses the array's name as an address: initLibrary
// resolves each device's copy with hipGetSymbolAddress and keeps the address per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g.err = buf;
    return 1;
}


static bool dev_mode() { return g.opts.device_resident != 0; }

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Device memory for one of a gridding call's tables: carved from the scratch arena inside
// bullseye_grid_noalloc, otherwise the library buffer b, grown as needed.
static hipError_t table(DevBuf &b, size_t bytes, void **out) {
    if (g.arena) {
        const size_t off = align256(g.arena_used);
        if (off > g.arena_cap || bytes > g.arena_cap - off) return hipErrorOutOfMemory;
        *out = g.arena + off;
        g.arena_used = off + bytes;
        return hipSuccess;
    }
    hipError_t e = b.ensure(bytes);
    *out = b.p;
    return e;
}

// Copy `bytes` from the caller's array to a device staging buffer (host mode) or use it as
// is (device-resident mode).
hipError_t stage(DevBuf &b, const void *src, size_t bytes, const void **out, hipStream_t st) {
    if (!src) {
        *out = nullptr;
        return hipSuccess;
    }
    if (dev_mode()) {
        *out = src;
        return hipSuccess;
    }
    hipError_t e = b.ensure(bytes);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st);
    *out = b.p;
    return e;
}

hipError_t stage(DevBuf &b, const void *src, size_t bytes, const void **out) {
    return stage(b, src, bytes, out, g.stream);
}

// Image-plane response of the AA kernel along one axis, normalised to 1 at the centre
// (the taper the finalize step divides out).
static std::vector<double> taper_vector(const std::vector<double> &conv, size_t S, size_t OS,
                                        size_t n) {
    std::vector<double> t(n);
    const double centre = (double)((S + 1) * OS);
    double sum = 0;
    for (double c : conv) sum += c;
    for (size_t k = 0; k < n; ++k) {
        const double l = (double)k - (double)(n / 2);
        double acc = 0;
        for (size_t i = 0; i < conv.size(); ++i)
            acc += conv[i] * cos(2.0 * M_PI * ((double)i - centre) / (double)OS * l / (double)n);
        t[k] = acc / sum;
    }
    return t;
}

static size_t grid_elems() { return g.F * g.cube * g.ny * g.nx * g.pstride * 2; }
static size_t psf_elems() { return g.F * g.sfC * g.ny * g.nx * 2; }
static size_t abi_grid_bytes() { return g.F * g.cube * g.P * g.ny * g.nx * 2 * sizeof(grid_base_type); }
static size_t abi_psf_bytes() { return g.F * g.sfC * g.ny * g.nx * 2 * sizeof(grid_base_type); }

int check_device_errors() {
    int flag = 0;
    if (g.errflag.p) {
        hipError_t e = hipMemcpy(&flag, g.errflag.p, sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("reading the device error flag: %s", hipGetErrorString(e));
        if (flag) {
            int zero = 0;
            (void)hipMemcpy(g.errflag.p, &zero, sizeof(int), hipMemcpyHostToDevice);
        }
    }
    if (flag & 1) return fail("JONES MATRIX SET CONTAINS SINGULAR MATRICIES");
    if (flag & 2) return fail("antenna id out of range (antenna_count or baseline_count too small)");
    if (flag & 4) return fail("channel grid index, spw id or timestamp id out of range");
    if (flag & 8)
        return fail("a w-projection sample is wider than its stream's support class (the data "
                    "changed after the batch was planned)");
    return 0;
}

static int barrier_impl() {
    if (!g.stream) return 0;
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return fail("gridding stream failed: %s", hipGetErrorString(e));
    for (auto &c : g.pending) {
        float ms = 0, kms = 0;
        if (!c.recorded) {
            for (hipEvent_t ev : {c.start, c.cstart, c.copied, c.k0, c.k1, c.stop})
                if (ev) (void)hipEventDestroy(ev);
            continue;
        }
        (void)hipEventElapsedTime(&ms, c.start, c.stop);
        g.grid_seconds += ms * 1e-3;
        float cms = 0;
        (void)hipEventElapsedTime(&cms, c.cstart, c.copied);
        g.stats.copy_ms = dev_mode() ? 0.0 : cms;
        if (c.has_kernel) {
            (void)hipEventElapsedTime(&kms, c.k0, c.k1);
            g.stats.kernel_ms = kms;
            float pms = 0;
            (void)hipEventElapsedTime(&pms, c.start, c.k0);
            g.stats.prep_ms = pms;
        }
        (void)hipEventDestroy(c.start);
        (void)hipEventDestroy(c.cstart);
        (void)hipEventDestroy(c.copied);
        (void)hipEventDestroy(c.k0);
        (void)hipEventDestroy(c.k1);
        (void)hipEventDestroy(c.stop);
    }
    const bool had_pending = !g.pending.empty();
    g.pending.clear();
    if (had_pending && g.counters.p) {
        std::vector<unsigned long long> cs(2 * (size_t)g.nslots);
        e = hipMemcpy(cs.data(), g.counters.p, cs.size() * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail("reading counters: %s", hipGetErrorString(e));
        unsigned long long c[2] = {0, 0};
        for (size_t i = 0; i < g.nslots; ++i) {
            c[0] += cs[2 * i];
            c[1] += cs[2 * i + 1];
        }
        g.stats.samples_gridded = c[0] - g.last_counters[0];
        g.stats.taps = c[1] - g.last_counters[1];
        g.last_counters[0] = c[0];
        g.last_counters[1] = c[1];
        if (g.walk.p) {
            int ti = 0;
            e = hipMemcpy(&ti, g.walk.p, sizeof ti, hipMemcpyDeviceToHost);
            if (e != hipSuccess) return fail("reading the walk order: %s", hipGetErrorString(e));
            g.stats.time_inner = (uint32_t)ti;
        }
    }
    return check_device_errors();
}

// The device results of all chunks so far, copied into the caller's buffers (final chunk).
static int publish_results(gridding_parameters *p, bool psf) {
    if (!psf) {
        void *dst = dev_mode() ? p->output_buffer : nullptr;
        if (!dev_mode()) {
            HIP_TRY(g.out_tmp.ensure(abi_grid_bytes()));
            dst = g.out_tmp.p;
        }
        if (dst)
            HIP_TRY(grid_to_abi(g.grid.as<double>(), dst, g.F * g.cube, (uint32_t)g.P,
                                (uint32_t)g.pstride, g.nx * g.ny, g.stream));
        HIP_TRY(fold_slots(g.norm.as<double>(), g.norm_stride, g.nslots, g.stream));
        HIP_TRY(hipMemcpyAsync(g.host_norm.data(), g.norm.p,
                               g.host_norm.size() * sizeof(normalization_base_type),
                               hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        if (!dev_mode() && p->output_buffer)
            HIP_TRY(hipMemcpy(p->output_buffer, g.out_tmp.p, abi_grid_bytes(),
                              hipMemcpyDeviceToHost));
    } else {
        void *dst = dev_mode() ? p->sampling_function_buffer : nullptr;
        if (!dev_mode()) {
            HIP_TRY(g.out_tmp.ensure(std::max(abi_psf_bytes(), g.out_tmp.bytes)));
            dst = g.out_tmp.p;
        }
        if (dst)
            HIP_TRY(grid_to_abi(g.psfgrid.as<double>(), dst, g.F * g.sfC, 1, 1, g.nx * g.ny,
                                g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        if (!dev_mode() && p->sampling_function_buffer)
            HIP_TRY(hipMemcpy(p->sampling_function_buffer, g.out_tmp.p, abi_psf_bytes(),
                              hipMemcpyDeviceToHost));
    }
    return 0;
}

// Streams: channels walked together and time segments per baseline.
static void choose_streams(const gridding_parameters *p, size_t F, size_t B, size_t rows, int W,
                           int mode, bool same_slice, uint32_t &cps, uint32_t &nseg) {
    const size_t C = p->channel_count;
    cps = g.opts.channels_per_stream ? g.opts.channels_per_stream
                                     : (uint32_t)(same_slice ? std::min<size_t>(C, 32) : 1);
    cps = (uint32_t)std::max<size_t>(1, std::min<size_t>(cps, C));
    const size_t nblocks = (C + cps - 1) / cps;
    if (g.opts.rows_per_stream) {
        const size_t per_bl = std::max<size_t>(1, rows / std::max<size_t>(1, B));
        nseg = (uint32_t)std::max<size_t>(1, per_bl / g.opts.rows_per_stream);
    } else {
        // enough lane groups to fill 256 CUs x 4 SIMDs several waves deep, segments of at least
        // 8 rows (one facet of 2016 baselines x 240 rows x 16 channels: 15-row segments grid in
        // 0.44 ms, 30-row ones in 0.59 ms, 8-row ones in 0.51 ms)
        int gpw = romein_w_groups_per_wave(W, mode, g.path);
        if (gpw == 0) gpw = 64 / std::max(1, W * ((W + 7) / 8));
        const size_t want_groups = (size_t)256 * 4 * 4 * gpw;
        const size_t have = F * B * nblocks;
        nseg = 1;
        const size_t per_bl = std::max<size_t>(1, rows / std::max<size_t>(1, B));
        while (have * nseg < want_groups && per_bl / (nseg * 2) >= 8) nseg *= 2;
    }
}

// Inputs of a gridding or degridding call on the device (staged copies, or the caller's own
// device pointers), the inverted Jones terms, the per-facet geometry and the scalar fields of
// GridArgs that do not depend on the stream decomposition.
// `cstart` and `copied` (optional) are recorded around the staging copies.
static int setup_inputs(gridding_parameters *p, int mode, bool facet, bool phase, bool psf,
                        bool need_vis, GridArgs &a, hipEvent_t cstart = nullptr,
                        hipEvent_t copied = nullptr) {
    const size_t rows = p->row_count, C = p->channel_count, npol = p->number_of_polarization_terms;
    const size_t nspw = std::max<size_t>(1, p->spw_count);
    // staging set of this call (host mode): everything queued on the stream so far is the
    // other set's last reader; this set's last reader was the call before that
    hipStream_t cs = g.stream;
    State::InputSet &in = g.ins[g.ins_cur ^ 1];
    if (!dev_mode()) {
        HIP_TRY(hipEventRecord(g.ins[g.ins_cur].done, g.stream));
        g.ins[g.ins_cur].used = true;
        g.ins_cur ^= 1;
        cs = g.cstream;
        if (in.used) HIP_TRY(hipStreamWaitEvent(cs, in.done, 0));
    }
    if (cstart) HIP_TRY(hipEventRecord(cstart, cs));
    const void *d;
    HIP_TRY(stage(in.vis, (psf || !need_vis) ? nullptr : p->visibilities,
                  rows * C * npol * 2 * sizeof(visibility_base_type), &d, cs));
    a.vis = (const visibility_base_type *)d;
    HIP_TRY(stage(in.uvw, p->uvw_coords, rows * 3 * sizeof(uvw_base_type), &d, cs));
    a.uvw = (const uvw_base_type *)d;
    HIP_TRY(stage(in.wl, p->reference_wavelengths,
                  nspw * C * sizeof(reference_wavelengths_base_type), &d, cs));
    a.wavelengths = (const reference_wavelengths_base_type *)d;
    HIP_TRY(stage(in.w, psf ? nullptr : p->visibility_weights,
                  rows * C * npol * sizeof(visibility_weights_base_type), &d, cs));
    a.weights = (const visibility_weights_base_type *)d;
    HIP_TRY(stage(in.flags, psf ? nullptr : p->flags, rows * C * npol, &d, cs));
    a.flags = (const uint8_t *)d;
    HIP_TRY(stage(in.rowflags, p->flagged_rows, rows, &d, cs));
    a.row_flags = (const uint8_t *)d;
    HIP_TRY(stage(in.field, p->field_array, rows * 4, &d, cs));
    a.field = (const uint32_t *)d;
    HIP_TRY(stage(in.spw, p->spw_index_array, rows * 4, &d, cs));
    a.spw = (const uint32_t *)d;
    HIP_TRY(stage(in.a1, p->antenna_1_ids, rows * 4, &d, cs));
    a.ant1 = (const uint32_t *)d;
    HIP_TRY(stage(in.a2, p->antenna_2_ids, rows * 4, &d, cs));
    a.ant2 = (const uint32_t *)d;
    HIP_TRY(stage(in.enabled, p->enabled_channels, nspw * C, &d, cs));
    a.enabled = (const uint8_t *)d;
    HIP_TRY(stage(in.cgi, psf ? p->sampling_function_channel_grid_indicies
                              : p->channel_grid_indicies,
                  nspw * C * sizeof(size_t), &d, cs));
    a.cgi = (const uint64_t *)d;
    // per-row arrays may be NULL in an empty chunk (the reference's loops then never run)
    const bool any_rows = rows > 0;
    const size_t nj = p->no_timestamps_read * p->antenna_count * g.F * nspw * C;
    if (mode == MODE_4J && any_rows) {
        HIP_TRY(stage(in.time, p->timestamp_ids, rows * sizeof(size_t), &d, cs));
        a.time_ids = (const uint64_t *)d;
        HIP_TRY(stage(in.jones, p->jones_terms, nj * 8 * sizeof(visibility_base_type), &d, cs));
        a.jones = (const visibility_base_type *)d;
    }
    // facet centres (facet_* modes) with the other inputs
    const void *fc = nullptr;
    if (facet) {
        if (!p->facet_centres) return fail("facet_centres is NULL");
        HIP_TRY(stage(in.fc, p->facet_centres, g.F * 2 * sizeof(uvw_base_type), &fc, cs));
    }
    if (copied) HIP_TRY(hipEventRecord(copied, cs));
    if (!dev_mode()) {
        HIP_TRY(hipEventRecord(in.staged, cs));
        HIP_TRY(hipStreamWaitEvent(g.stream, in.staged, 0));
    }
    if (!a.wavelengths || !a.enabled || !a.cgi ||
        (any_rows && (!a.uvw || !a.row_flags || !a.field || !a.spw ||
                      (!psf && ((need_vis && !a.vis) || !a.weights || !a.flags)))))
        return fail("a required array pointer is NULL");
    if (mode == MODE_4J && any_rows) {
        if (!a.time_ids) return fail("timestamp_ids is NULL");
        // jones_2x2.h:113-126 invert_all, into a library buffer in the compute precision
        const bool fp64 = g.path == PATH_FP64;
        void *jinv = nullptr;
        HIP_TRY(table(g.jinv, nj * 8 * (fp64 ? 8 : 4), &jinv));
        HIP_TRY(invert_jones(a.jones, jinv, nj, fp64, g.errflag.as<int>(), g.stream));
        a.jones_inv = jinv;
    }

    // per-facet geometry in FP64 (facet_* modes; grid_* modes use the phase centre), computed
    // on the device from the staged facet centres: no host buffer outlives the call
    void *geom = nullptr;
    HIP_TRY(table(g.geom, g.F * sizeof(FacetGeom), &geom));
    HIP_TRY(launch_facet_geom((const uvw_base_type *)fc, (double)p->phase_centre_ra,
                              (double)p->phase_centre_dec, (uint32_t)g.F, facet, phase,
                              (FacetGeom *)geom, g.stream));
    a.geom = reinterpret_cast<const FacetGeom *>(geom);
    a.conv = g.path == PATH_FP64 ? g.conv_d.p : g.conv_f.p;
    a.tapsum = g.tapsum.as<double>();
    a.error_flag = g.errflag.as<int>();
    a.nx = (uint32_t)g.nx;
    a.ny = (uint32_t)g.ny;
    a.S = (uint32_t)g.S;
    a.W = (uint32_t)(2 * p->conv_support + 1);
    a.OS = (uint32_t)g.OS;
    a.csz = (uint32_t)g.csz;
    a.chan = (uint32_t)C;
    a.npol = (uint32_t)npol;
    a.spw_count = (uint32_t)nspw;
    a.antenna_count = (uint32_t)p->antenna_count;
    a.ntime = (uint32_t)p->no_timestamps_read;
    a.facets = (uint32_t)g.F;
    a.cube = (uint32_t)(psf ? g.sfC : g.cube);
    a.pstride = (uint32_t)(psf ? 1 : g.pstride);
    for (int k = 0; k < 4; ++k) a.pol_idx[k] = (uint32_t)k;
    if (mode == MODE_1) a.pol_idx[0] = (uint32_t)p->polarization_index;
    if (mode == MODE_2) {
        a.pol_idx[0] = (uint32_t)p->polarization_index;
        a.pol_idx[1] = (uint32_t)p->second_polarization_index;
    }
    a.imaging_field = p->imaging_field;
    // templated_gridder.h:23-27 (u_scale uses the float ARCSEC_TO_RAD constant)
    a.u_scale = (double)((double)p->nx * (double)p->cell_size_x * (double)ARCSEC_TO_RAD);
    a.v_scale = -(double)((double)p->ny * (double)p->cell_size_y * (double)ARCSEC_TO_RAD);
    a.off_x = (double)(p->nx / 2 - p->conv_support);
    a.off_y = (double)(p->ny / 2 - p->conv_support);
    a.phase = phase ? 1 : 0;
    a.transform = facet ? 1 : 0;
    return 0;
}

// Checks shared by the gridding and degridding entry points.
static int validate_call(const gridding_parameters *p, int mode) {
    const bool psf = mode == MODE_PSF;
    const bool wproj = g.wplanes > 1;
    if ((p->wplanes > 1) != wproj) return fail("wplanes differs from initLibrary");
    const int W = (int)(2 * p->conv_support + 1);
    if (!wproj && W > romein_max_width())
        return fail("conv_support %zu too large for the AA gridder (max full support %d)",
                    (size_t)p->conv_support, romein_max_width());
    if (p->conv_support != g.S || p->conv_oversample != g.OS)
        return fail("conv_support/conv_oversample differ from initLibrary");
    const int ncorr = (mode == MODE_2) ? 2 : (mode == MODE_4 || mode == MODE_4J) ? 4 : 1;
    if (!psf && (size_t)ncorr > g.P)
        return fail("mode grids %d correlations but number_of_polarization_terms_being_gridded is %zu",
                    ncorr, g.P);
    if (psf && !g.psf) return fail("sampling function buffer was not set up at initLibrary");
    if (p->row_count && (!p->antenna_1_ids || !p->antenna_2_ids))
        return fail("antenna_1_ids/antenna_2_ids are required (Romein gridding walks baselines)");
    if (mode == MODE_4J && p->row_count && !p->jones_terms) return fail("jones_terms is NULL");
    if (p->row_count > 0xffffffffull) return fail("row_count too large");
    const size_t npol = p->number_of_polarization_terms;
    if (mode == MODE_1) {
        if (p->polarization_index >= npol) return fail("polarization_index out of range");
    } else if (mode == MODE_2) {
        if (p->polarization_index >= npol || p->second_polarization_index >= npol)
            return fail("polarization indexes out of range");
    } else if ((mode == MODE_4 || mode == MODE_4J) && npol < 4) {
        return fail("4-correlation gridding needs 4 correlations per sample");
    }
    return 0;
}

static size_t call_baselines(const gridding_parameters *p) {
    return p->baseline_count ? p->baseline_count : p->antenna_count * (p->antenna_count + 1) / 2;
}

// Bytes of the tables an AA gridding call carves (table()), each rounded to 256 bytes: an
// upper bound of what bullseye_grid_noalloc needs from its scratch.
static size_t aa_call_table_bytes(const gridding_parameters *p, int mode) {
    const size_t rows = p->row_count, B = call_baselines(p);
    const size_t nfc = std::max<size_t>(1, p->spw_count) * p->channel_count;
    size_t n = align256(g.F * sizeof(FacetGeom));
    if (mode == MODE_4J && rows > 0)
        n += align256(p->no_timestamps_read * p->antenna_count * g.F *
                      std::max<size_t>(1, p->spw_count) * p->channel_count * 8 *
                      (g.path == PATH_FP64 ? 8 : 4));
    n += 4 * align256(rows * 4) + 2 * align256((B + 1) * 4);
    n += align256(ingest_workspace_bytes((uint32_t)rows, (uint32_t)B));
    n += align256(walk_order_bytes());
    n += align256(std::max<size_t>(1, rows) * sizeof(RowRec));
    n += align256(std::max<size_t>(1, nfc) * sizeof(ChanRec));
    return n;
}

// One gridding call.  noalloc (bullseye_grid_noalloc): the tables come from the scratch arena
// and no timing events are created, so the call allocates nothing.
static int grid_call(gridding_parameters *p, int mode, bool facet, bool phase,
                     bool noalloc = false) {
    if (!g.initialized) return fail("initLibrary has not been called");
    const bool psf = mode == MODE_PSF;
    const bool wproj = g.wplanes > 1;
    if (validate_call(p, mode)) return 1;
    const int W = (int)(2 * p->conv_support + 1);
    const size_t rows = p->row_count, C = p->channel_count;
    const size_t nspw = std::max<size_t>(1, p->spw_count);
    const size_t B = call_baselines(p);

    // registered before anything can fail, so the barrier releases the events of a failed call
    TimedCall none{};
    TimedCall *tcp = &none;
    if (!noalloc) {
        g.pending.push_back(TimedCall{});
        tcp = &g.pending.back();
        HIP_TRY(hipEventCreate(&tcp->start));
        HIP_TRY(hipEventCreate(&tcp->cstart));
        HIP_TRY(hipEventCreate(&tcp->copied));
        HIP_TRY(hipEventCreate(&tcp->k0));
        HIP_TRY(hipEventCreate(&tcp->k1));
        HIP_TRY(hipEventCreate(&tcp->stop));
        HIP_TRY(hipEventRecord(tcp->start, g.stream));
    }
    TimedCall &tc = *tcp;

    GridArgs a{};
    if (setup_inputs(p, mode, facet, phase, psf, true, a, tc.cstart, tc.copied)) return 1;

    // ingest: stable sort of rows by baseline
    void *perm, *keys, *vals, *keys_out, *counts, *blstart, *ws;
    HIP_TRY(table(g.perm, rows * 4, &perm));
    HIP_TRY(table(g.keys, rows * 4, &keys));
    HIP_TRY(table(g.vals, rows * 4, &vals));
    HIP_TRY(table(g.keys_out, rows * 4, &keys_out));
    HIP_TRY(table(g.counts, (B + 1) * 4, &counts));
    HIP_TRY(table(g.blstart, (B + 1) * 4, &blstart));
    const size_t wsb = ingest_workspace_bytes((uint32_t)rows, (uint32_t)B);
    HIP_TRY(table(g.ws, wsb, &ws));
    HIP_TRY(ingest_sort(a.ant1, a.ant2, (uint32_t)rows, (uint32_t)p->antenna_count, (uint32_t)B,
                        (uint32_t *)keys, (uint32_t *)vals, (uint32_t *)keys_out,
                        (uint32_t *)counts, (uint32_t *)perm, (uint32_t *)blstart, ws, wsb,
                        g.errflag.as<int>(), g.stream));

    // do all enabled channels map to one grid slice? (then channels can share a stream)
    bool same_slice = true;
    if (!dev_mode()) {
        const size_t *cg = psf ? p->sampling_function_channel_grid_indicies : p->channel_grid_indicies;
        long first = -1;
        for (size_t i = 0; i < nspw * C; ++i) {
            if (!p->enabled_channels[i]) continue;
            if (first < 0) first = (long)cg[i];
            else if ((long)cg[i] != first) same_slice = false;
        }
    } else {
        same_slice = g.cube == 1 && (!psf || g.sfC == 1);
    }
    uint32_t cps, nseg;
    choose_streams(p, g.F, B, rows, W, mode, same_slice, cps, nseg);
    if (wproj)
        cps = g.opts.channels_per_stream
                  ? (uint32_t)std::max<size_t>(1, std::min<size_t>(g.opts.channels_per_stream, C))
                  : same_slice ? (uint32_t)std::min<size_t>(C, 8) : 1;

    a.perm = (const uint32_t *)perm;
    a.bl_start = (const uint32_t *)blstart;
    a.grid = psf ? g.psfgrid.as<double>() : g.grid.as<double>();
    a.norm = g.norm.as<double>();
    a.counters = g.counters.as<unsigned long long>();
    a.norm_stride = g.norm_stride;
    a.nslots = g.nslots;
    a.baselines = (uint32_t)B;
    a.chan_per_stream = cps;
    a.n_chan_blocks = (uint32_t)((C + cps - 1) / cps);
    a.n_segments = nseg;
    a.n_streams = (uint64_t)B * a.n_chan_blocks * nseg * g.F;
    void *walk;
    HIP_TRY(table(g.walk, walk_order_bytes(), &walk));
    a.time_inner = (const int *)walk;
    HIP_TRY(launch_walk_order(a, (int)g.opts.walk_order, (int *)walk, g.stream));

    // sorted row table (every gridder's sample preparation) and channel table (the AA gridder)
    {
        const size_t nfc = nspw * C;
        void *rowrec, *chanrec;
        HIP_TRY(table(g.rowrec, std::max<size_t>(1, rows) * sizeof(RowRec), &rowrec));
        HIP_TRY(table(g.chanrec, std::max<size_t>(1, nfc) * sizeof(ChanRec), &chanrec));
        a.rowrec = (const RowRec *)rowrec;
        a.chanrec = (const ChanRec *)chanrec;
        a.chan_lds = nfc <= CHAN_LDS_MAX ? 1u : 0u;
        HIP_TRY(build_row_tables(a, (uint32_t)rows, mode == MODE_4J && rows > 0,
                                 (RowRec *)rowrec, (ChanRec *)chanrec, g.stream));
    }

    if (wproj) {
        // w-projection: streams are segments of 32 rows of a baseline; the device classifies
        // every (segment, channel block) by its support class and sorts them by class, and
        // one persistent launch per class width takes its list from there (no host round trip)
        const float wmax_f = (float)p->wmax_est;
        if (!(wmax_f > 0)) return fail("wmax_est must be > 0 for w-projection");
        const WListPlan lp = wproj_list_plan((uint32_t)rows, (uint32_t)B, a.n_chan_blocks);
        if (lp.entries >= 0x7fffffffull || lp.entries * g.F > 0xffffffffull)
            return fail("too many w-projection streams");
        a.n_segments = 0;   // streams come from the device lists
        a.n_streams = lp.entries * g.F;
        void *lws = nullptr;
        HIP_TRY(table(g.wlist, lp.total, &lws));
        HIP_TRY(wproj_build_lists(a, lp, lws, g.wplane_half.as<uint32_t>(), (uint32_t)g.wplanes,
                                  wmax_f, g.stream));
        const void *taps = g.path == PATH_FP64 ? g.wtaps_d.p : g.wtaps_f.p;
        const void *ctaps = g.path == PATH_FP64 ? g.wctaps_d.p : g.wctaps_f.p;
        uint32_t max_m = 1;
        for (uint32_t h : g.wplane_half_host) max_m = std::max(max_m, 2 * h + 1);
        HIP_TRY(hipEventRecord(tc.k0, g.stream));
        // one persistent launch per class width, dealt over the stream and the side streams so
        // that a launch's tail and the small classes overlap other classes' work
        HIP_TRY(hipEventRecord(g.wfork, g.stream));
        for (int i = 0; i < State::WSIDE; ++i) HIP_TRY(hipStreamWaitEvent(g.wside[i], g.wfork, 0));
        // Classes wider than 31 hold few samples each: one launch per lane shape (two or four
        // residue pairs per lane) walks all of them, each stream at its own width, instead of
        // one latency-bound launch per class.  They go first, so that their long streams overlap
        // the wide-parallel launches of the narrower classes, which follow widest first.
        std::vector<std::pair<uint32_t, uint32_t>> order;
        if (max_m > 45) order.push_back({47, max_m});
        if (max_m > 31) order.push_back({33, std::min<uint32_t>(max_m, 45)});
        for (int M = (int)std::min<uint32_t>(max_m, 31); M >= 1; M -= 2)
            order.push_back({(uint32_t)M, (uint32_t)M});
        hipError_t le = hipSuccess;
        for (size_t i = 0; i < order.size() && le == hipSuccess; ++i) {
            const int turn = (int)(i % (State::WSIDE + 1));
            le = launch_wproj_device_list(a, lp, lws, taps, g.wplane_off.as<uint64_t>(), ctaps,
                                          g.wcplane_off.as<uint64_t>(),
                                          g.wplane_half.as<uint32_t>(),
                                          g.wplane_norm.as<double>(), (uint32_t)g.wplanes,
                                          g.w_nf, g.w_fbase, order[i].first, order[i].second,
                                          wmax_f, mode, g.path,
                                          turn ? g.wside[turn - 1] : g.stream);
        }
        // joined also after a failed launch: later work on the stream (the next call's class
        // lists) must not overtake what was launched
        for (int i = 0; i < State::WSIDE; ++i) {
            HIP_TRY(hipEventRecord(g.wjoin[i], g.wside[i]));
            HIP_TRY(hipStreamWaitEvent(g.stream, g.wjoin[i], 0));
        }
        HIP_TRY(le);
    } else {
        const bool specialised = romein_w_groups_per_wave(W, mode, g.path) > 0 &&
                                 !g.opts.reserved[0];
        if (specialised) {
            const bool four = mode == MODE_4 || mode == MODE_4J;
            a.pol_contig = (four && !psf && p->number_of_polarization_terms == 4 &&
                            ((uintptr_t)a.vis % 16) == 0 && ((uintptr_t)a.weights % 16) == 0 &&
                            ((uintptr_t)a.flags % 4) == 0) ? 1u : 0u;
        }
        if (!noalloc) HIP_TRY(hipEventRecord(tc.k0, g.stream));
        if (specialised)
            HIP_TRY(launch_romein_w(a, mode, g.path, g.stream));
        else
            HIP_TRY(launch_romein_aa(a, mode, g.path, g.stream));
    }
    if (!noalloc) {
        HIP_TRY(hipEventRecord(tc.k1, g.stream));
        tc.has_kernel = true;
        HIP_TRY(hipEventRecord(tc.stop, g.stream));
        tc.recorded = true;
    }
    if (p->is_final_data_chunk) {
        if (barrier_impl()) return 1;
        if (publish_results(p, psf)) return 1;
    }
    return 0;
}

// W-projection taps: conv[plane][csz][csz] complex -> contiguous W_p x W_p blocks per
// (plane, frac_v, frac_u), FP32 and FP64, fraction indexes fbase .. fbase + nf - 1 (nearest
// rounding gives fractions in [OS/2, 3 OS/2]), plus the Re-sum of every block.
int build_wtaps(const void *conv, bool conv_on_device, size_t P, size_t S, size_t OS,
                const std::vector<uint32_t> &halves, DevBuf &taps_f, DevBuf &taps_d,
                DevBuf &plane_off, DevBuf &plane_half, DevBuf &plane_norm, DevBuf &ctaps_f,
                DevBuf &ctaps_d, DevBuf &cplane_off, uint32_t &w_nf, uint32_t &w_fbase) {
    const size_t csz = (2 * S + 3) + (2 * S + 2) * (OS - 1);
    w_fbase = (uint32_t)(OS / 2);
    w_nf = (uint32_t)(OS + 2);
    const size_t nf = w_nf;
    std::vector<uint64_t> off(P + 1, 0);
    for (size_t q = 0; q < P; ++q) {
        const size_t wq = 2 * halves[q] + 1;
        off[q + 1] = off[q] + nf * nf * wq * wq;
    }
    if (off[P] > 0xffffffffull) return fail("w-projection tap table exceeds 2^32 taps");
    const size_t nconv = P * csz * csz;
    std::vector<convolution_base_type> ch(2 * nconv);
    if (conv_on_device)
        HIP_TRY(hipMemcpy(ch.data(), conv, 2 * nconv * sizeof(convolution_base_type),
                          hipMemcpyDeviceToHost));
    else
        memcpy(ch.data(), conv, 2 * nconv * sizeof(convolution_base_type));
    std::vector<double> td(2 * off[P]);
    std::vector<float> tf(2 * off[P]);
    std::vector<double> nrm(P * nf * nf, 0.0);
    for (size_t q = 0; q < P; ++q) {
        const size_t sp = halves[q], wq = 2 * sp + 1;
        for (size_t fv = 0; fv < nf; ++fv)
            for (size_t fu = 0; fu < nf; ++fu) {
                double sum = 0;
                for (size_t tv = 0; tv < wq; ++tv)
                    for (size_t tu = 0; tu < wq; ++tu) {
                        const size_t iv = w_fbase + fv + (S - sp + tv) * OS;
                        const size_t iu = w_fbase + fu + (S - sp + tu) * OS;
                        double re = 0, im = 0;
                        if (iv < csz && iu < csz) {
                            const size_t ci = (q * csz + iv) * csz + iu;
                            re = ch[2 * ci];
                            im = ch[2 * ci + 1];
                        }
                        const size_t ti = off[q] + ((fv * nf + fu) * wq + tv) * wq + tu;
                        td[2 * ti] = re;
                        td[2 * ti + 1] = im;
                        tf[2 * ti] = (float)re;
                        tf[2 * ti + 1] = (float)im;
                        sum += re;
                    }
                nrm[(q * nf + fv) * nf + fu] = sum;
            }
    }
    // column tables of the classes M <= WCOL_MAX (wp::WArgs::ctaps): per plane no wider than
    // M and block, 2M rows of M taps, entry (j, t) window row j mod M of column t (zero
    // beyond W_p)
    constexpr int NCL = wp::WCOL_CLASSES;
    std::vector<uint64_t> coff((size_t)NCL * P, ~0ull);
    uint64_t cn = 0;
    for (size_t ci = 0; ci < (size_t)NCL; ++ci)
        for (size_t q = 0; q < P; ++q) {
            const size_t wq = 2 * halves[q] + 1, M = 2 * ci + 1;
            if (wq > M) continue;
            coff[ci * P + q] = cn;
            cn += nf * nf * 2 * M * M;
        }
    if (cn > 0xffffffffull) return fail("w-projection column tables exceed 2^32 taps");
    std::vector<double> cd(2 * cn, 0.0);
    std::vector<float> cf(2 * cn, 0.0f);
    for (size_t ci = 0; ci < (size_t)NCL; ++ci)
        for (size_t q = 0; q < P; ++q) {
            const size_t wq = 2 * halves[q] + 1, M = 2 * ci + 1;
            if (wq > M) continue;
            for (size_t blk = 0; blk < nf * nf; ++blk)
                for (size_t t = 0; t < wq; ++t)
                    for (size_t j = 0; j < 2 * M; ++j) {
                        const size_t row = j % M;
                        if (row >= wq) continue;
                        const size_t src = off[q] + (blk * wq + row) * wq + t;
                        const size_t dst = coff[ci * P + q] + (blk * 2 * M + j) * M + t;
                        cd[2 * dst] = td[2 * src];
                        cd[2 * dst + 1] = td[2 * src + 1];
                        cf[2 * dst] = tf[2 * src];
                        cf[2 * dst + 1] = tf[2 * src + 1];
                    }
        }
    HIP_TRY(ctaps_d.ensure(std::max<size_t>(cd.size(), 2) * 8));
    HIP_TRY(ctaps_f.ensure(std::max<size_t>(cf.size(), 2) * 4));
    HIP_TRY(cplane_off.ensure(coff.size() * 8));
    if (cn) {
        HIP_TRY(hipMemcpy(ctaps_d.p, cd.data(), cd.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctaps_f.p, cf.data(), cf.size() * 4, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(cplane_off.p, coff.data(), coff.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(taps_d.ensure(td.size() * 8));
    HIP_TRY(taps_f.ensure(tf.size() * 4));
    HIP_TRY(plane_off.ensure(P * 8));
    HIP_TRY(plane_half.ensure(P * 4));
    HIP_TRY(plane_norm.ensure(nrm.size() * 8));
    HIP_TRY(hipMemcpy(taps_d.p, td.data(), td.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(taps_f.p, tf.data(), tf.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(plane_off.p, off.data(), P * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(plane_half.p, halves.data(), P * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(plane_norm.p, nrm.data(), nrm.size() * 8, hipMemcpyHostToDevice));
    return 0;
}

LibView lib_view() {
    LibView v;
    v.initialized = g.initialized;
    v.stream = g.stream;
    v.device_resident = g.opts.device_resident != 0;
    v.path = g.path;
    v.walk_order = (int)g.opts.walk_order;
    v.errflag = g.errflag.as<int>();
    v.counters = g.counters.as<unsigned long long>();
    v.nslots = g.nslots;
    v.last_counters = g.last_counters;
    v.stats = &g.stats;
    v.grid_seconds = &g.grid_seconds;
    return v;
}

}  // namespace bam

using namespace bam;

extern "C" {

const char *bullseye_last_error(void) { return g.err.c_str(); }

int bullseye_set_options(const bullseye_options *o) {
    if (!o) return fail("NULL options");
    if (o->struct_size != sizeof(bullseye_options))
        return fail("bullseye_options.struct_size %u != %zu", o->struct_size, sizeof(bullseye_options));
    if (o->compute_path > BULLSEYE_PATH_FP32) return fail("bad compute_path %u", o->compute_path);
    if (o->walk_order > 2) return fail("bad walk_order %u", o->walk_order);
    if (g.initialized) {
        // initLibrary chose where the grids are seeded from and published to: the buffer
        // model cannot change under queued or future calls of the same context
        if ((o->device_resident != 0) != (g.opts.device_resident != 0))
            return fail("device_resident cannot change between initLibrary and releaseLibrary");
        // queued calls finish on the path they were launched with
        if (barrier_impl()) return 1;
    }
    g.opts = *o;
    if (g.initialized)
        g.path = o->compute_path ? (int)o->compute_path
#if defined(BULLSEYE_DOUBLE)
                                 : PATH_FP64;
#else
                                 : PATH_F32K;
#endif
    return 0;
}

int bullseye_get_options(bullseye_options *o) {
    if (!o) return fail("NULL options");
    *o = g.opts;
    o->struct_size = sizeof(bullseye_options);
    return 0;
}

int bullseye_get_stats(bullseye_stats *s) {
    if (!s) return fail("NULL stats");
    *s = g.stats;
    return 0;
}

int bullseye_set_wplane_support(const uint32_t *half_support, uint32_t nplanes) {
    if (g.initialized) return fail("bullseye_set_wplane_support must precede initLibrary");
    if (!half_support || !nplanes) {
        g.wplane_half_req.clear();
        return 0;
    }
    g.wplane_half_req.assign(half_support, half_support + nplanes);
    return 0;
}

int bullseye_reset_grids(void) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (barrier_impl()) return 1;
    HIP_TRY(fill_f64(g.grid.as<double>(), grid_elems(), 0.0, g.stream));
    if (g.psf) HIP_TRY(fill_f64(g.psfgrid.as<double>(), psf_elems(), 0.0, g.stream));
    HIP_TRY(fill_f64(g.norm.as<double>(), g.norm_stride * g.nslots, 0.0, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    std::fill(g.host_norm.begin(), g.host_norm.end(), 0.0);
    return 0;
}

double get_gridding_walltime(void) { return g.grid_seconds; }
double get_inversion_walltime(void) { return g.inv_seconds; }

void gridding_barrier(void) { (void)barrier_impl(); }

// Device buffers and the stream of the context (releaseLibrary, or a failed initLibrary).
static void release_resources() {
    DevBuf *bufs[] = {&g.grid, &g.psfgrid, &g.norm, &g.counters, &g.errflag, &g.conv_f,
                      &g.conv_d, &g.tapsum,
                      &g.geom, &g.taper, &g.perm, &g.blstart, &g.keys, &g.vals, &g.keys_out,
                      &g.rowrec, &g.chanrec,
                      &g.counts, &g.ws, &g.out_tmp,
                      &g.fft_scratch, &g.fft_real, &g.jinv, &g.wtaps_f, &g.wtaps_d,
                      &g.wplane_off, &g.wplane_half, &g.wplane_norm, &g.wcls, &g.wlist, &g.walk,
                      &g.wctaps_f, &g.wctaps_d, &g.wcplane_off,
                      &g.dg_model, &g.dg_out, &g.stokes_tmp, &g.dft_smp, &g.dft_comp};
    for (DevBuf *b : bufs) b->release();
    for (State::InputSet &s : g.ins) {
        DevBuf *ib[] = {&s.vis, &s.uvw, &s.w, &s.flags, &s.rowflags, &s.field, &s.spw, &s.a1,
                        &s.a2, &s.time, &s.jones, &s.wl, &s.enabled, &s.cgi, &s.fc};
        for (DevBuf *b : ib) b->release();
        if (s.staged) (void)hipEventDestroy(s.staged);
        if (s.done) (void)hipEventDestroy(s.done);
        s.staged = s.done = nullptr;
        s.used = false;
    }
    if (g.cstream) (void)hipStreamDestroy(g.cstream);
    g.cstream = nullptr;
    batch_release_all();
    for (int i = 0; i < State::WSIDE; ++i) {
        if (g.wside[i]) (void)hipStreamDestroy(g.wside[i]);
        if (g.wjoin[i]) (void)hipEventDestroy(g.wjoin[i]);
        g.wside[i] = nullptr;
        g.wjoin[i] = nullptr;
    }
    if (g.wfork) (void)hipEventDestroy(g.wfork);
    g.wfork = nullptr;
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g.stream = nullptr;
}

static int init_impl(gridding_parameters *p) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail("no HIP device");
    if (p->nx == 0 || p->ny == 0 || p->conv_oversample == 0)
        return fail("nx, ny and conv_oversample must be > 0");
    // kernels form grid row strides (nx * 4 correlations * 2) in 32 bits
    if (p->nx > (1u << 28) || p->ny > (1u << 28)) return fail("nx and ny must be <= 2^28");
    if (2 * p->conv_support + 3 >= p->nx || 2 * p->conv_support + 3 >= p->ny)
        return fail("conv_support too large for the grid");
    if (!p->conv) return fail("conv is NULL");
    HIP_TRY(hipGetDevice(&g.device));
    if (g.device < 0 || g.device >= NOALLOC_MAX_DEVICES)
        return fail("HIP device %d is beyond the library's %d", g.device, NOALLOC_MAX_DEVICES);
    if (!g_scratch_ptr[g.device]) {
        // this device's copy of the allocation-free call's scratch array, resolved once
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail("resolving the scratch array on device %d: %s", g.device,
                        hipGetErrorString(e));
        g_scratch_ptr[g.device] = static_cast<unsigned char *>(ptr);
    }
    HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&g.wfork, hipEventDisableTiming));
    HIP_TRY(hipStreamCreateWithFlags(&g.cstream, hipStreamNonBlocking));
    for (State::InputSet &s : g.ins) {
        HIP_TRY(hipEventCreateWithFlags(&s.staged, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        s.used = false;
    }
    g.ins_cur = 0;
    for (int i = 0; i < State::WSIDE; ++i) {
        HIP_TRY(hipStreamCreateWithFlags(&g.wside[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&g.wjoin[i], hipEventDisableTiming));
    }
    g.F = std::max<size_t>(1, p->num_facet_centres);
    g.cube = std::max<size_t>(1, p->cube_channel_dim_size);
    g.P = std::max<size_t>(1, p->number_of_polarization_terms_being_gridded);
    g.pstride = g.P;
    g.nx = p->nx;
    g.ny = p->ny;
    g.S = p->conv_support;
    g.OS = p->conv_oversample;
    g.wplanes = std::max<size_t>(1, p->wplanes);
    g.psf = p->should_grid_sampling_function;
    g.sfC = g.psf ? std::max<size_t>(1, p->sampling_function_channel_count) : 0;
    g.csz = (2 * g.S + 3) + (2 * g.S + 2) * (g.OS - 1);
    g.path = g.opts.compute_path ? (int)g.opts.compute_path
#if defined(BULLSEYE_DOUBLE)
                                 : PATH_FP64;
#else
                                 : PATH_F32K;
#endif
    HIP_TRY(g.grid.ensure(grid_elems() * sizeof(double)));
    g.norm_stride = g.F * g.cube * g.P;
    g.nslots = (uint32_t)std::max<size_t>(1, std::min<size_t>(64, ((size_t)1 << 22) / g.norm_stride));
    if (g.norm_stride > 0x7fffffffull) return fail("too many grid slices");
    HIP_TRY(g.norm.ensure(g.norm_stride * g.nslots * sizeof(double)));
    HIP_TRY(g.counters.ensure(2 * (size_t)g.nslots * sizeof(unsigned long long)));
    HIP_TRY(g.errflag.ensure(sizeof(int)));
    HIP_TRY(hipMemset(g.counters.p, 0, 2 * (size_t)g.nslots * sizeof(unsigned long long)));
    HIP_TRY(hipMemset(g.errflag.p, 0, sizeof(int)));
    g.last_counters[0] = g.last_counters[1] = 0;
    g.host_norm.assign(g.F * g.cube * g.P, 0.0);
    p->normalization_terms = g.host_norm.data();
    HIP_TRY(fill_f64(g.norm.as<double>(), g.norm_stride * g.nslots, 0.0, g.stream));
    // seed the accumulation grid from the caller's buffer (gpu_algorithm/wrapper.cu:165-168)
    const void *src = nullptr;
    if (p->output_buffer) {
        if (dev_mode()) {
            src = p->output_buffer;
        } else {
            HIP_TRY(g.out_tmp.ensure(abi_grid_bytes()));
            HIP_TRY(hipMemcpy(g.out_tmp.p, p->output_buffer, abi_grid_bytes(), hipMemcpyHostToDevice));
            src = g.out_tmp.p;
        }
    }
    HIP_TRY(abi_to_grid(src, g.grid.as<double>(), g.F * g.cube, (uint32_t)g.P, (uint32_t)g.pstride,
                        g.nx * g.ny, g.stream));
    if (g.psf) {
        HIP_TRY(g.psfgrid.ensure(psf_elems() * sizeof(double)));
        const void *ps = nullptr;
        if (p->sampling_function_buffer) {
            if (dev_mode()) {
                ps = p->sampling_function_buffer;
            } else {
                HIP_TRY(hipStreamSynchronize(g.stream));
                HIP_TRY(g.out_tmp.ensure(std::max(abi_psf_bytes(), g.out_tmp.bytes)));
                HIP_TRY(hipMemcpy(g.out_tmp.p, p->sampling_function_buffer, abi_psf_bytes(),
                                  hipMemcpyHostToDevice));
                ps = g.out_tmp.p;
            }
        }
        HIP_TRY(abi_to_grid(ps, g.psfgrid.as<double>(), g.F * g.sfC, 1, 1, g.nx * g.ny, g.stream));
    }
    // kernel taps: FP64 and FP32 copies (the compute path picks one), tap sums per fraction
    if (g.wplanes <= 1) {
        std::vector<convolution_base_type> ch(g.csz);
        if (dev_mode())
            HIP_TRY(hipMemcpy(ch.data(), p->conv, g.csz * sizeof(convolution_base_type),
                              hipMemcpyDeviceToHost));
        else
            memcpy(ch.data(), p->conv, g.csz * sizeof(convolution_base_type));
        std::vector<double> cd(ch.begin(), ch.end());
        std::vector<double> ts(g.OS + 1, 0.0);
        for (size_t f = 0; f <= g.OS; ++f)
            for (size_t k = 0; k < 2 * g.S + 1; ++k) ts[f] += cd[f + k * g.OS];
        // both precisions, so the compute path can change after initLibrary
        std::vector<float> cf(cd.begin(), cd.end());
        HIP_TRY(g.conv_d.ensure(g.csz * 8));
        HIP_TRY(g.conv_f.ensure(g.csz * 4));
        HIP_TRY(hipMemcpy(g.conv_d.p, cd.data(), g.csz * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g.conv_f.p, cf.data(), g.csz * 4, hipMemcpyHostToDevice));
        HIP_TRY(g.tapsum.ensure((g.OS + 1) * 8));
        HIP_TRY(hipMemcpy(g.tapsum.p, ts.data(), (g.OS + 1) * 8, hipMemcpyHostToDevice));
        g.taper_x = taper_vector(cd, g.S, g.OS, g.nx);
        g.taper_y = taper_vector(cd, g.S, g.OS, g.ny);
        HIP_TRY(g.taper.ensure((g.nx + g.ny) * 8));
        HIP_TRY(hipMemcpy(g.taper.p, g.taper_x.data(), g.nx * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g.taper.as<double>() + g.nx, g.taper_y.data(), g.ny * 8,
                          hipMemcpyHostToDevice));
    }
    if (g.wplanes > 1) {
        // w-projection taps: conv[plane][csz][csz] complex -> contiguous W_p x W_p blocks per
        // (plane, frac_v, frac_u), fraction indexes fbase .. fbase + nf - 1 (nearest rounding
        // gives fractions in [OS/2, 3 OS/2])
        const size_t P = g.wplanes, S = g.S, OS = g.OS;
        g.wplane_half_host.assign(P, (uint32_t)S);
        if (!g.wplane_half_req.empty()) {
            if (g.wplane_half_req.size() != P)
                return fail("bullseye_set_wplane_support gave %zu planes, wplanes is %zu",
                            g.wplane_half_req.size(), P);
            for (size_t q = 0; q < P; ++q) {
                if (g.wplane_half_req[q] > S) return fail("plane %zu support exceeds conv_support", q);
                g.wplane_half_host[q] = g.wplane_half_req[q];
            }
        }
        for (size_t q = 0; q < P; ++q)
            if (g.wplane_half_host[q] > 31)
                return fail("w-projection plane %zu: half support %u exceeds 31", q,
                            g.wplane_half_host[q]);
        if (build_wtaps(p->conv, dev_mode(), P, S, OS, g.wplane_half_host, g.wtaps_f, g.wtaps_d,
                        g.wplane_off, g.wplane_half, g.wplane_norm, g.wctaps_f, g.wctaps_d,
                        g.wcplane_off, g.w_nf, g.w_fbase))
            return 1;
    }
    HIP_TRY(hipStreamSynchronize(g.stream));
    g.grid_seconds = g.inv_seconds = 0;
    g.stats = bullseye_stats{};
    return 0;
}

int initLibrary(gridding_parameters *p) {
    if (g.initialized) return 0;
    if (!p) return fail("NULL params");
    g.err.clear();
    if (init_impl(p)) {
        // a failed set-up keeps nothing: the next initLibrary starts from scratch
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        release_resources();
        return 1;
    }
    g.initialized = true;
    return 0;
}

int releaseLibrary(void) {
    if (!g.initialized) return 0;
    (void)barrier_impl();
    release_resources();
    g.initialized = false;
    return 0;
}

// cpu_algorithm/wrapper.cpp:116-135
int weight_uniformly(gridding_parameters *p) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (barrier_impl()) return 1;
    if (!p->sampling_function_buffer || !p->output_buffer) return fail("buffers are NULL");
    const size_t F = g.F, C = g.cube, P = g.P, sfC = p->sampling_function_channel_count;
    const size_t npix = p->nx * p->ny;
    if (dev_mode()) {
        HIP_TRY(weight_uniformly_device(p->output_buffer, p->sampling_function_buffer,
                                        (const uint64_t *)p->channel_grid_indicies, (uint32_t)F,
                                        (uint32_t)C, (uint32_t)P, (uint32_t)sfC, npix, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        return 0;
    }
    grid_base_type *out = (grid_base_type *)p->output_buffer;
    const grid_base_type *sf = (const grid_base_type *)p->sampling_function_buffer;
    for (size_t f = 0; f < F; ++f)
        for (size_t gi = 0; gi < C; ++gi)
            for (size_t pix = 0; pix < npix; ++pix) {
                grid_base_type count = 0.0000001f;
                for (size_t c = 0; c < sfC; ++c)
                    count += (int)(p->channel_grid_indicies[c] == gi) * sf[((f * sfC + c) * npix + pix) * 2];
                count = 1 / count;
                for (size_t corr = 0; corr < P; ++corr) {
                    grid_base_type *cell = out + (((f * C + gi) * P + corr) * npix + pix) * 2;
                    cell[0] *= count;
                    cell[1] *= count;
                }
            }
    return 0;
}

// cpu_algorithm/wrapper.cpp:136-163
int normalize(gridding_parameters *p) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (barrier_impl()) return 1;
    const size_t slices = g.F * g.cube * g.P, npix = p->nx * p->ny;
    if (dev_mode()) {
        HIP_TRY(fold_slots(g.norm.as<double>(), g.norm_stride, g.nslots, g.stream));
        HIP_TRY(normalize_device(p->output_buffer, g.norm.as<double>(), slices, npix, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        return 0;
    }
    grid_base_type *out = (grid_base_type *)p->output_buffer;
    const normalization_base_type *nt = p->normalization_terms ? p->normalization_terms : g.host_norm.data();
    for (size_t s = 0; s < slices; ++s) {
        const grid_base_type d = (grid_base_type)nt[s];
        grid_base_type *sl = out + s * npix * 2;
        for (size_t i = 0; i < npix; ++i) {
            sl[2 * i] /= d;
            sl[2 * i + 1] /= d;
        }
    }
    return 0;
}

static int finalize_impl(gridding_parameters *p, bool psf) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (barrier_impl()) return 1;
    void *buf = psf ? p->sampling_function_buffer : p->output_buffer;
    if (!buf) return fail("buffer is NULL");
    const uint32_t F = (uint32_t)g.F;
    const uint32_t n = (uint32_t)(psf ? p->sampling_function_channel_count : p->cube_channel_dim_size);
    const uint32_t fstride = (uint32_t)(psf ? p->sampling_function_channel_count
                                            : p->cube_channel_dim_size * g.P);
    const size_t npix = p->nx * p->ny;
    const size_t total_bytes = (size_t)F * fstride * npix * 2 * sizeof(grid_base_type);
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e0, g.stream));
    void *dbuf = buf;
    if (!dev_mode()) {
        HIP_TRY(g.out_tmp.ensure(std::max(total_bytes, g.out_tmp.bytes)));
        HIP_TRY(hipMemcpyAsync(g.out_tmp.p, buf, total_bytes, hipMemcpyHostToDevice, g.stream));
        dbuf = g.out_tmp.p;
    }
    HIP_TRY(g.fft_scratch.ensure((size_t)F * n * npix * 2 * sizeof(grid_base_type)));
    HIP_TRY(g.fft_real.ensure((size_t)F * n * npix * sizeof(float)));
    const bool taper = g.opts.apply_taper_correction && !psf && g.taper.p;
    const double *tx = taper ? g.taper.as<double>() : nullptr;
    const double *ty = taper ? g.taper.as<double>() + g.nx : nullptr;
    HIP_TRY(finalize_device(dbuf, F, n, fstride, (uint32_t)p->nx, (uint32_t)p->ny, tx, ty,
                            g.fft_scratch.p, g.fft_real.as<float>(), dev_mode(), g.stream));
    HIP_TRY(hipEventRecord(ev.e1, g.stream));
    HIP_TRY(hipEventSynchronize(ev.e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
    g.inv_seconds += ms * 1e-3;
    if (!dev_mode()) {
        HIP_TRY(hipMemcpy(buf, g.out_tmp.p, total_bytes, hipMemcpyDeviceToHost));
        // fft_and_repacking_routines.cpp:106-116: real parts as float32, in place
        const grid_base_type *gt = (const grid_base_type *)buf;
        float *fl = (float *)buf;
        const size_t offset = (size_t)fstride * npix;
        for (size_t f = 0; f < F; ++f)
            for (size_t i = offset * f; i < offset * f + (size_t)n * npix; ++i) fl[i] = (float)gt[i * 2];
    }
    return 0;
}

int finalize(gridding_parameters *p) { return finalize_impl(p, false); }
int finalize_psf(gridding_parameters *p) { return finalize_impl(p, true); }

// gpu_algorithm/wrapper.cu:358-454: the reference's host-side reorder of the row arrays by
// baseline.  The gridder does not need it (it sorts on the device); it is kept so callers that
// invoke it see the same side effects.
int repack_input_data(gridding_parameters *p) {
    if (dev_mode()) return fail("repack_input_data works on host arrays; device-resident rows "
                                "need no repacking");
    const size_t rows = p->row_count, A = p->antenna_count, B = p->baseline_count;
    const size_t per = p->channel_count * p->number_of_polarization_terms;
    if (!p->baseline_starting_indexes) return fail("baseline_starting_indexes is NULL");
    auto bidx = [A](long a1, long a2) {
        long s = std::min(a1, a2);
        return (size_t)((s * (-s + (2 * (long)A + 1))) / 2 + std::labs(a1 - a2));
    };
    size_t *st = p->baseline_starting_indexes;
    memset(st, 0, sizeof(size_t) * (B + 1));
    for (size_t r = 0; r < rows; ++r) {
        size_t b = bidx(p->antenna_1_ids[r], p->antenna_2_ids[r]);
        if (b >= B) return fail("baseline index out of range");
        ++st[b + 1];
    }
    for (size_t b = 0; b < B; ++b) st[b + 1] += st[b];
    std::vector<size_t> cur(B, 0), dst(rows);
    for (size_t r = 0; r < rows; ++r) {
        size_t b = bidx(p->antenna_1_ids[r], p->antenna_2_ids[r]);
        dst[r] = st[b] + cur[b]++;
    }
    auto permute = [&](void *base, size_t elem) {
        if (!base) return;
        std::vector<unsigned char> tmp(rows * elem);
        for (size_t r = 0; r < rows; ++r)
            memcpy(&tmp[dst[r] * elem], (unsigned char *)base + r * elem, elem);
        memcpy(base, tmp.data(), rows * elem);
    };
    permute(p->uvw_coords, 3 * sizeof(uvw_base_type));
    permute(p->visibilities, per * 2 * sizeof(visibility_base_type));
    permute(p->visibility_weights, per * sizeof(visibility_weights_base_type));
    permute(p->flags, per);
    permute(p->flagged_rows, 1);
    permute(p->spw_index_array, 4);
    permute(p->antenna_1_ids, 4);
    permute(p->antenna_2_ids, 4);
    permute(p->field_array, 4);
    if (p->should_invert_jones_terms) permute(p->timestamp_ids, sizeof(size_t));
    return 0;
}

// No counterpart in the reference: grid_* / facet_* for callers that cannot allocate device
// memory (HIP graph capture, serving loops).  AA gridding with device-resident buffers only;
// the call's tables come from the library's static scratch array (NOALLOC_SCRATCH_BYTES per
// device), the grids and normalisation sums from initLibrary.  Calls that overlap on one device
// race on the scratch.  bullseye_stats is not updated by these calls.
int bullseye_grid_noalloc(gridding_parameters *p, int mode) {
    static const struct { int kmode; bool facet, phase; } modes[] = {
        {MODE_1, false, false},   {MODE_1, true, true},    {MODE_2, false, false},
        {MODE_2, true, true},     {MODE_4, false, false},  {MODE_4, true, true},
        {MODE_4J, true, true},    {MODE_PSF, false, false}, {MODE_PSF, true, false}};
    if (!g.initialized) return fail("initLibrary has not been called");
    if (!p) return fail("bullseye_grid_noalloc: NULL params");
    if (mode < 0 || mode > 8) return fail("bullseye_grid_noalloc: bad mode %d", mode);
    if (!dev_mode())
        return fail("bullseye_grid_noalloc needs device-resident buffers "
                    "(bullseye_options.device_resident = 1)");
    if (g.wplanes > 1)
        return fail("bullseye_grid_noalloc grids with the AA kernel only (wplanes <= 1)");
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != g.device)
        return fail("bullseye_grid_noalloc: device %d is current, the library runs on %d", cur,
                    g.device);
    const int km = modes[mode].kmode;
    if (validate_call(p, km)) return 1;
    const size_t need = aa_call_table_bytes(p, km);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail("bullseye_grid_noalloc: the call's tables need %zu bytes, the scratch holds "
                    "%zu (grid fewer rows per call)", need, NOALLOC_SCRATCH_BYTES);
    g.arena = g_scratch_ptr[g.device];
    g.arena_cap = NOALLOC_SCRATCH_BYTES;
    g.arena_used = 0;
    const int rc = grid_call(p, km, modes[mode].facet, modes[mode].phase, true);
    g.arena = nullptr;
    g.arena_cap = g.arena_used = 0;
    return rc;
}

int grid_single_pol(gridding_parameters *p) { return grid_call(p, MODE_1, false, false); }
int facet_single_pol(gridding_parameters *p) { return grid_call(p, MODE_1, true, true); }
int grid_duel_pol(gridding_parameters *p) { return grid_call(p, MODE_2, false, false); }
int facet_duel_pol(gridding_parameters *p) { return grid_call(p, MODE_2, true, true); }
int grid_4_cor(gridding_parameters *p) { return grid_call(p, MODE_4, false, false); }
int facet_4_cor(gridding_parameters *p) { return grid_call(p, MODE_4, true, true); }
int facet_4_cor_corrections(gridding_parameters *p) { return grid_call(p, MODE_4J, true, true); }
int grid_sampling_function(gridding_parameters *p) { return grid_call(p, MODE_PSF, false, false); }
// wrapper.cpp:402-425: faceting uvw transform but no phase shift for the sampling function
int facet_sampling_function(gridding_parameters *p) { return grid_call(p, MODE_PSF, true, false); }

// Stokes formation and compaction of the normalised grids in output_buffer (the pipeliner's
// stokes.py:90-172 and bullseye_pipeliner.py:406-416 steps), on the device.  Synchronous.
int bullseye_form_stokes(gridding_parameters *p, int first, int second, int sign, int compact,
                         int contiguous) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (!p || !p->output_buffer) return fail("bullseye_form_stokes: output_buffer is NULL");
    const uint32_t P = (uint32_t)g.P;
    if (first >= 0 && (first >= (int)P || second < 0 || second >= (int)P))
        return fail("bullseye_form_stokes: correlation slots %d, %d out of range (P = %u)",
                    first, second, P);
    if (first >= 0 && sign != 1 && sign != -1) return fail("bullseye_form_stokes: sign must be +-1");
    if (barrier_impl()) return 1;
    const uint64_t npix = (uint64_t)g.nx * g.ny;
    const size_t bytes = g.F * g.cube * P * npix * 2 * sizeof(grid_base_type);
    HIP_TRY(g.stokes_tmp.ensure(std::max<size_t>(16, g.F * g.cube * npix * 2 * sizeof(grid_base_type))));
    void *dg = p->output_buffer;
    if (!dev_mode()) {
        HIP_TRY(g.out_tmp.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(g.out_tmp.p, p->output_buffer, bytes, hipMemcpyHostToDevice,
                               g.stream));
        dg = g.out_tmp.p;
    }
    HIP_TRY(stokes_device(dg, g.stokes_tmp.p, (uint32_t)g.F, (uint32_t)g.cube, P, npix, first,
                          second, sign, compact, contiguous, g.stream));
    if (!dev_mode())
        HIP_TRY(hipMemcpyAsync(p->output_buffer, dg, bytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

// Direct Fourier image of the single-correlation samples (polarization_index), FP64.
int bullseye_dft_image(gridding_parameters *p, double *image) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (!p || !image) return fail("bullseye_dft_image: NULL argument");
    if (validate_call(p, MODE_1)) return 1;
    if (barrier_impl()) return 1;
    GridArgs a{};
    if (setup_inputs(p, MODE_1, false, false, false, true, a)) return 1;
    const uint64_t rows = p->row_count, npix = (uint64_t)p->nx * p->ny;
    HIP_TRY(g.dft_smp.ensure(std::max<size_t>(32, dft_sample_bytes(rows, (uint32_t)p->channel_count))));
    double *dimg = image;
    if (!dev_mode()) {
        HIP_TRY(g.dg_out.ensure(npix * sizeof(double)));
        dimg = g.dg_out.as<double>();
    }
    const double cx = (double)p->cell_size_x * (double)ARCSEC_TO_RAD;
    const double cy = (double)p->cell_size_y * (double)ARCSEC_TO_RAD;
    HIP_TRY(launch_dft_image(a, rows, g.dft_smp.p, (uint32_t)p->nx, (uint32_t)p->ny, cx, cy, dimg,
                             g.stream));
    if (!dev_mode())
        HIP_TRY(hipMemcpyAsync(image, dimg, npix * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return check_device_errors();
}

// Visibilities of point components (l, m, flux), FP64, into correlation polarization_index.
int bullseye_dft_predict(gridding_parameters *p, const double *components, uint32_t n,
                         void *vis_out) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (!p || !vis_out || (n && !components)) return fail("bullseye_dft_predict: NULL argument");
    if (validate_call(p, MODE_1)) return 1;
    if (barrier_impl()) return 1;
    GridArgs a{};
    if (setup_inputs(p, MODE_1, false, false, false, false, a)) return 1;
    const uint64_t rows = p->row_count;
    const size_t vbytes = rows * p->channel_count * p->number_of_polarization_terms * 2 *
                          sizeof(visibility_base_type);
    const double *dcomp = components;
    void *dvis = vis_out;
    if (!dev_mode()) {
        HIP_TRY(g.dft_comp.ensure(std::max<size_t>(24, (size_t)n * 3 * sizeof(double))));
        if (n)
            HIP_TRY(hipMemcpyAsync(g.dft_comp.p, components, (size_t)n * 3 * sizeof(double),
                                   hipMemcpyHostToDevice, g.stream));
        HIP_TRY(g.dg_out.ensure(std::max<size_t>(16, vbytes)));
        HIP_TRY(hipMemcpyAsync(g.dg_out.p, vis_out, vbytes, hipMemcpyHostToDevice, g.stream));
        dcomp = g.dft_comp.as<double>();
        dvis = g.dg_out.p;
    }
    HIP_TRY(launch_dft_predict(a, rows, dcomp, n, dvis, g.stream));
    if (!dev_mode())
        HIP_TRY(hipMemcpyAsync(vis_out, dvis, vbytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return check_device_errors();
}

// The degridding adjoint of a gridding entry point (visibility modes).  Synchronous.
int bullseye_degrid(gridding_parameters *p, int mode, const void *model, void *vis_out) {
    if (!g.initialized) return fail("initLibrary has not been called");
    if (!p || !model || !vis_out) return fail("bullseye_degrid: NULL argument");
    static const struct { int kmode; bool facet, phase; } modes[] = {
        {MODE_1, false, false}, {MODE_1, true, true}, {MODE_2, false, false},
        {MODE_2, true, true},   {MODE_4, false, false}, {MODE_4, true, true},
        {MODE_4J, true, true}};
    if (mode < 0 || mode > 6) return fail("bullseye_degrid: mode %d is not a visibility mode", mode);
    const int km = modes[mode].kmode;
    if (validate_call(p, km)) return 1;
    if (barrier_impl()) return 1;
    const size_t rows = p->row_count, C = p->channel_count, npol = p->number_of_polarization_terms;
    const size_t model_bytes = g.F * g.cube * g.P * g.ny * g.nx * 2 * sizeof(grid_base_type);
    const size_t out_bytes = rows * C * npol * 2 * sizeof(visibility_base_type);
    EventPair ev;
    HIP_TRY(ev.create());
    GridArgs a{};
    if (setup_inputs(p, km, modes[mode].facet, modes[mode].phase, false, false, a)) return 1;
    const void *dm = model;
    void *dout = vis_out;
    if (!dev_mode()) {
        HIP_TRY(g.dg_model.ensure(model_bytes));
        HIP_TRY(g.dg_out.ensure(std::max<size_t>(out_bytes, 16)));
        HIP_TRY(hipMemcpyAsync(g.dg_model.p, model, model_bytes, hipMemcpyHostToDevice, g.stream));
        dm = g.dg_model.p;
        dout = g.dg_out.p;
    }
    const bool wproj = g.wplanes > 1;
    const void *taps = wproj ? (g.path == PATH_FP64 ? g.wtaps_d.p : g.wtaps_f.p) : nullptr;
    HIP_TRY(hipEventRecord(ev.e0, g.stream));
    HIP_TRY(launch_degrid(a, dm, dout, rows, (uint32_t)g.pstride, taps,
                          g.wplane_off.as<uint64_t>(), g.wplane_half.as<uint32_t>(),
                          (uint32_t)g.wplanes, g.w_nf, g.w_fbase, (float)p->wmax_est, km, g.path,
                          g.stream));
    HIP_TRY(hipEventRecord(ev.e1, g.stream));
    if (!dev_mode())
        HIP_TRY(hipMemcpyAsync(vis_out, dout, out_bytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
    g.stats.kernel_ms = ms;
    return check_device_errors();
}

}  // extern "C"
