// batch.hip -- batched facets of uneven size, cell, w-kernel stack and kind in one plan
// (include/bullseye_amd.h, "batched facets").
//
// The reference images one facet geometry per initLibrary and its pipeliner loops over calls
// (bullseye/bin/bullseye_pipeliner.py:350-404); every facet is a separate pass over the
// visibilities.  Here a plan holds up to 65536 facets over one staged, baseline-sorted
// visibility set:
//   * every kernel stack ("kernel set") classifies the streams -- (baseline, channel block,
//     segment of base_rows rows) -- by the widest w-plane any of their samples uses
//     (k_stream_class, wproj.hip); facets sharing a kernel set share its classes;
//   * per (kind, class width M) one work list holds the entries (facet << 32 | stream) of every
//     facet, streams in order and the facets of a stream together, so the visibilities a
//     stream reads are reused from L2 by all facets in turn;
//   * one persistent launch per list drains it through an atomic work counter (the class
//     kernels of wproj.hip read each entry's facet arguments from a device table);
//   * the launches spread over four graph branches (longest first onto the least loaded
//     branch), after fill kernels that clear the FP64 accumulation grids, normalisation slots
//     and counters; the whole sequence is captured once as a HIP graph and replayed.
// A launch's visibility chunk can span `chunk` classification segments: it walks their rows
// and skips those of segments of another class (WArgs::filter), so each sample is gridded
// exactly once whatever the chunk.  The autotuner times each class launch in each lane shape
// and chunk length and keeps the fastest per class.
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <chrono>

#include "host_common.h"

namespace bam {

namespace {

constexpr uint32_t MAX_SETS = 64;
constexpr uint32_t MAX_FACETS = 1u << 16;
constexpr int BRANCHES = 4;
constexpr uint32_t NCLS = 32;   // class widths 1, 3, ..., 63
constexpr uint32_t LIST_PAD = 0xffffffffu;   // stream id of a no-op work-list entry

// Streams per column-kernel item of class width M (1 where the class has no column kernel).
uint32_t column_run(int M, int kind, int path) {
    return M <= wp::WCOL_MAX ? wproj_streams_per_item(M, kind, path, WSHAPE_COLUMN) : 1u;
}

struct KernelSet {
    bool defined = false;
    uint32_t wplanes = 0, S = 0, OS = 0, csz = 0, nf = 0, fbase = 0;
    float wmax = 0;
    std::vector<uint32_t> half;
    DevBuf taps_f, taps_d, off, halfd, norm, ctaps_f, ctaps_d, coff;
    void release() {
        taps_f.release();
        taps_d.release();
        ctaps_f.release();
        ctaps_d.release();
        coff.release();
        off.release();
        halfd.release();
        norm.release();
        defined = false;
    }
};

struct Launch {
    wp::WArgs w;         // launch-level arguments (list, queue, decomposition)
    int mode;            // MODE_1 / MODE_2 / MODE_4 / MODE_PSF
    int shape;
    uint32_t M, chunk;
    int branch;
    double work;         // entries x M^2: the branch balancing estimate
    bool single;         // one facet's launch of the single-facet kernels
};

struct Plan {
    bool ready = false;
    int mode = MODE_1, schedule = 0, path = PATH_F32K;
    uint32_t n = 0, NC = 1, cube = 1, sfC = 1, cubeB = 1;
    uint32_t B = 0, C = 0, cps = 1, ncb = 1, base_rows = 32, nseg_base = 1, rows = 0;
    uint64_t nbase = 0;
    std::vector<bullseye_batch_facet> facets;
    std::vector<int> fkind;                  // MODE_x of each facet
    std::vector<uint32_t> fcube, fnc, fset;  // per facet: slices, correlations, kernel set
    std::vector<size_t> acc_off;             // doubles
    size_t acc_elems = 0, norm_stride = 0;
    std::vector<uint32_t> set_slot;          // kernel set id -> row of cls_all (~0: unused)
    DevBuf acc, norm, fargs, geom, lists, lists64, queues, cls_all, out_tmp, walk;
    DevBuf in_vis, in_uvw, in_w, in_flags, in_rowflags, in_field, in_spw, in_a1, in_a2, in_wl,
        in_enabled, in_cgi, in_sfcgi;
    DevBuf perm, keys, vals, keys_out, counts, blstart, ws, rowrec;
    std::vector<Launch> launches;
    std::vector<wp::WArgs> fargs_host;
    std::vector<FacetGeom> geom_host;
    std::vector<uint32_t> lists_host;        // per-facet launches: plain stream ids
    std::vector<uint64_t> lists64_host;      // merged launches: facet << 32 | stream
    std::vector<uint8_t> cls_host;
    std::vector<double> norm_host;
    GridArgs base{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t side[BRANCHES - 1] = {};
    hipEvent_t fork = nullptr, join[BRANCHES - 1] = {}, t0 = nullptr, t1 = nullptr;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void release() {
        drop_graph();
        DevBuf *bufs[] = {&acc, &norm, &fargs, &geom, &lists, &lists64, &queues, &cls_all,
                          &out_tmp, &walk,
                          &in_vis, &in_uvw, &in_w, &in_flags, &in_rowflags, &in_field, &in_spw,
                          &in_a1, &in_a2, &in_wl, &in_enabled, &in_cgi, &in_sfcgi, &perm, &rowrec, &keys,
                          &vals, &keys_out, &counts, &blstart, &ws};
        for (DevBuf *b : bufs) b->release();
        for (int b = 0; b < BRANCHES - 1; ++b) {
            if (side[b]) (void)hipStreamDestroy(side[b]);
            if (join[b]) (void)hipEventDestroy(join[b]);
            side[b] = nullptr;
            join[b] = nullptr;
        }
        if (fork) (void)hipEventDestroy(fork);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        fork = t0 = t1 = nullptr;
        launches.clear();
        lists_host.clear();
        lists_host.shrink_to_fit();
        lists64_host.clear();
        lists64_host.shrink_to_fit();
        ready = false;
    }
};

KernelSet sets[MAX_SETS];
Plan plan;
bullseye_batch_tuning tuning = {sizeof(bullseye_batch_tuning), 0, 0, 0, {0}, {0}, 0, 0, 0, 0};

int mode_of(int grid_mode) {
    switch (grid_mode) {
        case BULLSEYE_FACET_SINGLE_POL: return MODE_1;
        case BULLSEYE_FACET_DUEL_POL: return MODE_2;
        case BULLSEYE_FACET_4_COR: return MODE_4;
    }
    return 0;
}

uint32_t chunk_of(const bullseye_batch_tuning &t, uint32_t M) {
    const uint32_t c = t.chunk[M / 2];
    return c == 2 || c == 4 ? c : 1;
}

// baseline_transform_policies.h:120-143 and phase_transform_policies.h:123-133 in FP64, as the
// single-facet entry points compute them
FacetGeom facet_geometry(double ra0, double dec0, double ra, double dec, bool phase) {
    const double ARC = (double)ARCSEC_TO_RAD;
    FacetGeom fg;
    memset(&fg, 0, sizeof fg);
    const double d_ra = (ra - ra0) * ARC;
    const double n_dec = dec * ARC, o_dec = dec0 * ARC;
    const double li0 = cos(n_dec) * sin(d_ra);
    const double mi0 = sin(n_dec) * cos(o_dec) - cos(n_dec) * sin(o_dec) * cos(d_ra);
    const double ni0 = sin(n_dec) * sin(o_dec) + cos(n_dec) * cos(o_dec) * cos(d_ra);
    fg.u_term = li0 / ni0;
    fg.v_term = mi0 / ni0;
    if (phase) {
        fg.dl = -cos(n_dec) * sin(d_ra);
        fg.dm = -(sin(n_dec) * cos(o_dec) - cos(n_dec) * sin(o_dec) * cos(d_ra));
        fg.dn = 1 - sqrt(1 - fg.dl * fg.dl - fg.dm * fg.dm);
    }
    return fg;
}

// Checks of the arguments of a plan that need no device work.
int validate(const LibView &lv, const gridding_parameters *p, int mode,
             const bullseye_batch_facet *facets, uint32_t n) {
    if (!lv.initialized) return fail("initLibrary has not been called");
    if (!mode) return fail("batch mode must be facet_single_pol, facet_duel_pol or facet_4_cor");
    if (!p || !facets || n == 0) return fail("empty batch");
    if (n > MAX_FACETS) return fail("a batch holds at most %u facets", MAX_FACETS);
    const bool any_rows = p->row_count > 0;   // per-row arrays may be NULL in an empty set
    if (any_rows && (!p->antenna_1_ids || !p->antenna_2_ids))
        return fail("antenna_1_ids/antenna_2_ids are required");
    if (p->row_count > 0xffffffffull) return fail("row_count too large");
    const size_t npol = p->number_of_polarization_terms;
    if (mode == MODE_1 && p->polarization_index >= npol) return fail("polarization_index out of range");
    if (mode == MODE_2 &&
        (p->polarization_index >= npol || p->second_polarization_index >= npol))
        return fail("polarization indexes out of range");
    if (mode == MODE_4 && npol < 4) return fail("4-correlation gridding needs 4 correlations");
    bool any_vis = false, any_psf = false;
    for (uint32_t f = 0; f < n; ++f) {
        const bullseye_batch_facet &fc = facets[f];
        if (fc.kernel_set >= MAX_SETS || !sets[fc.kernel_set].defined)
            return fail("facet %u: kernel set %u is not defined", f, fc.kernel_set);
        const KernelSet &ks = sets[fc.kernel_set];
        const uint32_t padded = 2 * ks.S + 3;
        if (fc.nx <= padded + 1 || fc.ny <= padded + 1 || fc.nx > 65536 || fc.ny > 65536)
            return fail("facet %u: grid %ux%u does not fit the kernel support", f, fc.nx, fc.ny);
        if (!(fc.cell_x > 0) || !(fc.cell_y > 0)) return fail("facet %u: cell size must be > 0", f);
        if (fc.sampling_function) any_psf = true;
        else any_vis = true;
    }
    if (any_rows && any_vis && (!p->visibilities || !p->visibility_weights || !p->flags))
        return fail("a required array pointer is NULL");
    if (any_vis && !p->channel_grid_indicies) return fail("channel_grid_indicies is NULL");
    if (any_psf && !p->sampling_function_channel_grid_indicies)
        return fail("sampling_function_channel_grid_indicies is NULL");
    if (!p->reference_wavelengths || !p->enabled_channels ||
        (any_rows && (!p->uvw_coords || !p->flagged_rows || !p->field_array || !p->spw_index_array)))
        return fail("a required array pointer is NULL");
    return 0;
}

// Streams of kernel set row k (in cls_host) by class, as chunked stream ids: for every class M
// the ascending ids (baseline, channel block, segment of `chunk_M` classification segments)
// holding at least one classification segment of class M.
void class_streams(const Plan &pl, uint32_t k, const bullseye_batch_tuning &t,
                   std::vector<std::vector<uint32_t>> &out) {
    out.assign(NCLS, {});
    const uint8_t *cls = pl.cls_host.data() + (size_t)k * pl.nbase;
    uint32_t nsegk[NCLS], chunk[NCLS];
    int64_t last[NCLS];
    for (uint32_t i = 0; i < NCLS; ++i) {
        chunk[i] = chunk_of(t, 2 * i + 1);
        nsegk[i] = (pl.nseg_base + chunk[i] - 1) / chunk[i];
        last[i] = -1;
    }
    const uint64_t nblcb = (uint64_t)pl.B * pl.ncb;
    for (uint64_t bc = 0; bc < nblcb; ++bc) {
        const uint8_t *c = cls + bc * pl.nseg_base;
        for (uint32_t s = 0; s < pl.nseg_base; ++s) {
            const uint32_t M = c[s];
            if (!M) continue;
            const uint32_t i = M / 2;
            const int64_t id = (int64_t)(bc * nsegk[i] + s / chunk[i]);
            if (id != last[i]) {
                out[i].push_back((uint32_t)id);
                last[i] = id;
            }
        }
    }
}

// The per-(kind, class) work lists and launches of the plan (facets merged), or per
// (facet, class) for the per-facet schedule.
int build_lists(Plan &pl, const bullseye_batch_tuning &t) {
    const LibView lv = lib_view();
    pl.launches.clear();
    pl.lists_host.clear();
    pl.lists64_host.clear();
    const uint32_t nsets_used = (uint32_t)(pl.cls_host.size() / std::max<uint64_t>(1, pl.nbase));
    std::vector<std::vector<std::vector<uint32_t>>> streams(nsets_used);
    for (uint32_t k = 0; k < nsets_used; ++k) class_streams(pl, k, t, streams[k]);
    for (uint32_t i = 0; i < NCLS; ++i) {
        const uint32_t chunk = chunk_of(t, 2 * i + 1);
        const uint64_t nsegk = (pl.nseg_base + chunk - 1) / chunk;
        if ((uint64_t)pl.B * pl.ncb * nsegk >= LIST_PAD)
            return fail("batch too large: %llu streams per facet exceed 2^32 - 1",
                        (unsigned long long)((uint64_t)pl.B * pl.ncb * nsegk));
    }
    // facets of each (kernel set row, kind)
    std::vector<std::vector<uint32_t>> of_vis(nsets_used), of_psf(nsets_used);
    for (uint32_t f = 0; f < pl.n; ++f) {
        const uint32_t k = pl.set_slot[pl.fset[f]];
        (pl.fkind[f] == MODE_PSF ? of_psf : of_vis)[k].push_back(f);
    }
    const bool per_facet = pl.schedule == BULLSEYE_BATCH_PER_FACET;
    // facet < 0: a launch over the merged list of every facet; facet >= 0: a launch of the
    // single-facet kernels with that facet's own arguments (plain stream ids, its normalisation
    // block as the base of the sums)
    auto add_launch = [&](int kind, uint32_t i, uint64_t off, uint64_t cnt, int facet) {
        if (!cnt) return 0;
        if (cnt > 0xffffffffull) return fail("work list too long");
        const uint32_t M = 2 * i + 1;
        const uint32_t chunk = chunk_of(t, M);
        Launch L;
        memset(&L.w, 0, sizeof L.w);
        if (facet >= 0) {
            L.w = pl.fargs_host[facet];
            L.w.g.norm = pl.norm.as<double>() + (size_t)facet * pl.cubeB * pl.NC;
        } else {
            L.w.g = pl.base;
            L.w.g.cube = pl.cubeB;
        }
        L.w.g.n_segments = (pl.nseg_base + chunk - 1) / chunk;
        L.w.seg_rows = chunk * pl.base_rows;
        L.w.nseg = L.w.g.n_segments;
        L.w.base_rows = pl.base_rows;
        L.w.nseg_base = pl.nseg_base;
        L.w.filter = chunk > 1 ? 1 : 0;
        L.w.M = M;
        // offsets into the plain (per-facet) or the merged lists until the upload
        if (facet >= 0)
            L.w.stream_list = (const uint32_t *)(uintptr_t)off;
        else
            L.w.stream_list64 = (const uint64_t *)(uintptr_t)off;
        L.w.n_list = (uint32_t)cnt;
        L.mode = kind;
        L.M = M;
        L.chunk = chunk;
        L.shape = (int)t.shape[i];
        if (!wproj_shape_valid((int)M, L.shape)) L.shape = WSHAPE_DEFAULT;
        L.work = (double)cnt * M * M;
        L.branch = 0;
        L.single = facet >= 0;
        pl.launches.push_back(L);
        return 0;
    };
    // Entries of one column-kernel item (the streams one wavefront walks, GPW of them) share a
    // facet, so the kernel reads that facet's arguments once per wave with scalar loads; a
    // short run is padded with no-op entries (stream PAD).
    auto emit_run = [&](uint32_t f, const uint32_t *ids, size_t n, uint32_t g) {
        for (size_t q = 0; q < n; ++q) pl.lists64_host.push_back((uint64_t)f << 32 | ids[q]);
        for (size_t q = n; q % g; ++q) pl.lists64_host.push_back((uint64_t)f << 32 | LIST_PAD);
    };
    for (int kind : {pl.mode, (int)MODE_PSF}) {
        auto &of = kind == MODE_PSF ? of_psf : of_vis;
        if (per_facet) {
            for (uint32_t f = 0; f < pl.n; ++f) {
                if (pl.fkind[f] != kind) continue;
                const uint32_t k = pl.set_slot[pl.fset[f]];
                for (uint32_t i = 0; i < NCLS; ++i) {
                    const uint64_t off = pl.lists_host.size();
                    pl.lists_host.insert(pl.lists_host.end(), streams[k][i].begin(),
                                         streams[k][i].end());
                    if (add_launch(kind, i, off, pl.lists_host.size() - off, (int)f)) return 1;
                }
            }
            continue;
        }
        for (uint32_t i = 0; i < NCLS; ++i) {
            const uint64_t off = pl.lists64_host.size();
            const uint32_t g = column_run((int)(2 * i + 1), kind, pl.path);
            if (g > 1 && !t.facet_major) {
                // runs of g consecutive streams of a kernel set, every facet of the set in turn
                // (the run's visibilities shared in L2), runs of all sets by first stream
                std::vector<std::pair<uint32_t, uint32_t>> runs;   // (first stream, set)
                for (uint32_t k = 0; k < nsets_used; ++k) {
                    if (of[k].empty()) continue;
                    for (size_t q = 0; q < streams[k][i].size(); q += g)
                        runs.push_back({streams[k][i][q], k});
                }
                std::sort(runs.begin(), runs.end());
                for (const auto &r : runs) {
                    const std::vector<uint32_t> &sv = streams[r.second][i];
                    const size_t q0 = std::lower_bound(sv.begin(), sv.end(), r.first) - sv.begin();
                    const size_t n = std::min<size_t>(g, sv.size() - q0);
                    for (uint32_t f : of[r.second]) emit_run(f, sv.data() + q0, n, g);
                }
                if (add_launch(kind, i, off, pl.lists64_host.size() - off, -1)) return 1;
                continue;
            }
            if (t.facet_major) {  // each facet's streams together
                for (uint32_t f = 0; f < pl.n; ++f) {
                    if (pl.fkind[f] != kind) continue;
                    const std::vector<uint32_t> &sv = streams[pl.set_slot[pl.fset[f]]][i];
                    emit_run(f, sv.data(), sv.size(), g);
                }
                if (add_launch(kind, i, off, pl.lists64_host.size() - off, -1)) return 1;
                continue;
            }
            // merge the kernel sets' ascending stream lists; each stream's facets together
            std::vector<size_t> pos(nsets_used, 0);
            for (;;) {
                uint32_t best = 0xffffffffu;
                for (uint32_t k = 0; k < nsets_used; ++k)
                    if (!of[k].empty() && pos[k] < streams[k][i].size())
                        best = std::min(best, streams[k][i][pos[k]]);
                if (best == 0xffffffffu) break;
                for (uint32_t k = 0; k < nsets_used; ++k) {
                    if (of[k].empty() || pos[k] >= streams[k][i].size() ||
                        streams[k][i][pos[k]] != best)
                        continue;
                    for (uint32_t f : of[k]) pl.lists64_host.push_back((uint64_t)f << 32 | best);
                    ++pos[k];
                }
            }
            if (add_launch(kind, i, off, pl.lists64_host.size() - off, -1)) return 1;
        }
    }
    // device lists, queues and branch assignment (longest first onto the least loaded)
    HIP_TRY(pl.lists.ensure(std::max<size_t>(4, pl.lists_host.size() * 4)));
    if (!pl.lists_host.empty())
        HIP_TRY(hipMemcpyAsync(pl.lists.p, pl.lists_host.data(), pl.lists_host.size() * 4,
                               hipMemcpyHostToDevice, lv.stream));
    HIP_TRY(pl.lists64.ensure(std::max<size_t>(8, pl.lists64_host.size() * 8)));
    if (!pl.lists64_host.empty())
        HIP_TRY(hipMemcpyAsync(pl.lists64.p, pl.lists64_host.data(), pl.lists64_host.size() * 8,
                               hipMemcpyHostToDevice, lv.stream));
    HIP_TRY(pl.queues.ensure(std::max<size_t>(4, pl.launches.size() * 4)));
    std::vector<size_t> order(pl.launches.size());
    for (size_t j = 0; j < order.size(); ++j) order[j] = j;
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return pl.launches[a].work > pl.launches[b].work; });
    double load[BRANCHES] = {0, 0, 0, 0};
    for (size_t j : order) {
        Launch &L = pl.launches[j];
        int b = 0;
        for (int q = 1; q < BRANCHES; ++q)
            if (load[q] < load[b]) b = q;
        L.branch = per_facet ? 0 : b;
        load[L.branch] += L.work;
    }
    for (size_t j = 0; j < pl.launches.size(); ++j) {
        Launch &L = pl.launches[j];
        if (L.single)
            L.w.stream_list = pl.lists.as<uint32_t>() + (uintptr_t)L.w.stream_list;
        else
            L.w.stream_list64 = pl.lists64.as<uint64_t>() + (uintptr_t)L.w.stream_list64;
        L.w.facet_args = L.single ? nullptr : pl.fargs.as<wp::WArgs>();
        L.w.queue = per_facet ? nullptr : pl.queues.as<uint32_t>() + j;
    }
    HIP_TRY(hipStreamSynchronize(lv.stream));  // lists_host stays valid until the upload is done
    std::vector<uint32_t>().swap(pl.lists_host);
    std::vector<uint64_t>().swap(pl.lists64_host);
    return 0;
}

// Issue one class launch (queue reset by the caller for persistent launches).
hipError_t issue(const Plan &pl, const Launch &L, int shape, hipStream_t st) {
    wp::WArgs w = L.w;
    if (w.queue) {
        const uint32_t per = wproj_streams_per_item((int)L.M, L.mode, pl.path, shape);
        w.n_items = (w.n_list + per - 1) / per;
    }
    return launch_wproj_class(w, L.mode, pl.path, shape, st);
}

// Clears and launches of one execution: on the library stream, branches forked off it.
int enqueue_all(Plan &pl, hipStream_t st) {
    // the row table is gathered from the row arrays on every pass: with device-resident buffers
    // the caller's arrays are read in place, so changes after planning are seen
    HIP_TRY(build_row_tables(pl.base, pl.rows, false, pl.rowrec.as<RowRec>(), nullptr, st));
    HIP_TRY(fill_f64(pl.acc.as<double>(), pl.acc_elems, 0.0, st));
    HIP_TRY(fill_f64(pl.norm.as<double>(), pl.norm_stride * pl.base.nslots, 0.0, st));
    if (pl.schedule != BULLSEYE_BATCH_PER_FACET)
        HIP_TRY(fill_u32(pl.queues.as<uint32_t>(), (uint32_t)pl.launches.size(), 0u, st));
    bool used[BRANCHES] = {false, false, false, false};
    for (const Launch &L : pl.launches) used[L.branch] = true;
    HIP_TRY(hipEventRecord(pl.fork, st));
    for (int b = 1; b < BRANCHES; ++b)
        if (used[b]) HIP_TRY(hipStreamWaitEvent(pl.side[b - 1], pl.fork, 0));
    for (const Launch &L : pl.launches) {
        hipStream_t s = L.branch ? pl.side[L.branch - 1] : st;
        HIP_TRY(issue(pl, L, L.shape, s));
    }
    for (int b = 1; b < BRANCHES; ++b) {
        if (!used[b]) continue;
        HIP_TRY(hipEventRecord(pl.join[b - 1], pl.side[b - 1]));
        HIP_TRY(hipStreamWaitEvent(st, pl.join[b - 1], 0));
    }
    return 0;
}

int capture(Plan &pl, hipStream_t st) {
    pl.drop_graph();
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_all(pl, st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return 1;
    }
    HIP_TRY(e);
    pl.graph = graph;
    HIP_TRY(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
    return 0;
}

// Everything of a plan up to the lists: staging, sort, geometry, facet table, classes.
int build_plan(gridding_parameters *p, int mode, const bullseye_batch_facet *facets, uint32_t n,
               int schedule, const bullseye_batch_tuning &t, Plan &pl) {
    const LibView lv = lib_view();
    if (validate(lv, p, mode, facets, n)) return 1;
    if (schedule < 0 || schedule > BULLSEYE_BATCH_PER_FACET) return fail("unknown batch schedule");
    HIP_TRY(hipStreamSynchronize(lv.stream));
    pl.drop_graph();
    pl.ready = false;
    pl.mode = mode;
    pl.schedule = schedule;
    pl.path = lv.path;
    pl.n = n;
    pl.NC = mode == MODE_1 ? 1 : mode == MODE_2 ? 2 : 4;
    pl.facets.assign(facets, facets + n);
    const size_t rows = p->row_count, C = p->channel_count, npol = p->number_of_polarization_terms;
    const size_t nspw = std::max<size_t>(1, p->spw_count);
    pl.C = (uint32_t)C;
    pl.B = (uint32_t)(p->baseline_count ? p->baseline_count
                                        : p->antenna_count * (p->antenna_count + 1) / 2);
    pl.cube = (uint32_t)std::max<size_t>(1, p->cube_channel_dim_size);
    pl.sfC = (uint32_t)std::max<size_t>(1, p->sampling_function_channel_count);
    bool any_psf = false;
    for (uint32_t f = 0; f < n; ++f) any_psf |= facets[f].sampling_function != 0;
    pl.cubeB = std::max(pl.cube, any_psf ? pl.sfC : 1u);

    // staging of the visibility set
    GridArgs &a = pl.base;
    memset(&a, 0, sizeof a);
    const void *d;
    HIP_TRY(stage(pl.in_vis, p->visibilities, rows * C * npol * 2 * sizeof(visibility_base_type), &d));
    a.vis = (const visibility_base_type *)d;
    HIP_TRY(stage(pl.in_uvw, p->uvw_coords, rows * 3 * sizeof(uvw_base_type), &d));
    a.uvw = (const uvw_base_type *)d;
    HIP_TRY(stage(pl.in_wl, p->reference_wavelengths,
                  nspw * C * sizeof(reference_wavelengths_base_type), &d));
    a.wavelengths = (const reference_wavelengths_base_type *)d;
    HIP_TRY(stage(pl.in_w, p->visibility_weights,
                  rows * C * npol * sizeof(visibility_weights_base_type), &d));
    a.weights = (const visibility_weights_base_type *)d;
    HIP_TRY(stage(pl.in_flags, p->flags, rows * C * npol, &d));
    a.flags = (const uint8_t *)d;
    HIP_TRY(stage(pl.in_rowflags, p->flagged_rows, rows, &d));
    a.row_flags = (const uint8_t *)d;
    HIP_TRY(stage(pl.in_field, p->field_array, rows * 4, &d));
    a.field = (const uint32_t *)d;
    HIP_TRY(stage(pl.in_spw, p->spw_index_array, rows * 4, &d));
    a.spw = (const uint32_t *)d;
    HIP_TRY(stage(pl.in_a1, p->antenna_1_ids, rows * 4, &d));
    a.ant1 = (const uint32_t *)d;
    HIP_TRY(stage(pl.in_a2, p->antenna_2_ids, rows * 4, &d));
    a.ant2 = (const uint32_t *)d;
    HIP_TRY(stage(pl.in_enabled, p->enabled_channels, nspw * C, &d));
    a.enabled = (const uint8_t *)d;
    HIP_TRY(stage(pl.in_cgi, p->channel_grid_indicies, nspw * C * sizeof(size_t), &d));
    const uint64_t *cgi_vis = (const uint64_t *)d;
    HIP_TRY(stage(pl.in_sfcgi, p->sampling_function_channel_grid_indicies,
                  nspw * C * sizeof(size_t), &d));
    const uint64_t *cgi_psf = (const uint64_t *)d;
    a.cgi = cgi_vis ? cgi_vis : cgi_psf;

    // ingest: stable sort of rows by baseline
    HIP_TRY(pl.perm.ensure(rows * 4));
    HIP_TRY(pl.keys.ensure(rows * 4));
    HIP_TRY(pl.vals.ensure(rows * 4));
    HIP_TRY(pl.keys_out.ensure(rows * 4));
    HIP_TRY(pl.counts.ensure(((size_t)pl.B + 1) * 4));
    HIP_TRY(pl.blstart.ensure(((size_t)pl.B + 1) * 4));
    HIP_TRY(pl.ws.ensure(ingest_workspace_bytes((uint32_t)rows, pl.B)));
    HIP_TRY(ingest_sort(a.ant1, a.ant2, (uint32_t)rows, (uint32_t)p->antenna_count, pl.B,
                        pl.keys.as<uint32_t>(), pl.vals.as<uint32_t>(), pl.keys_out.as<uint32_t>(),
                        pl.counts.as<uint32_t>(), pl.perm.as<uint32_t>(), pl.blstart.as<uint32_t>(),
                        pl.ws.p, pl.ws.bytes, lv.errflag, lv.stream));

    // channels per stream: several only when every enabled channel maps to one slice
    bool same_slice = true;
    if (!lv.device_resident) {
        for (const size_t *cg : {p->channel_grid_indicies,
                                 any_psf ? p->sampling_function_channel_grid_indicies : nullptr}) {
            if (!cg) continue;
            long first = -1;
            for (size_t i = 0; i < nspw * C; ++i) {
                if (!p->enabled_channels[i]) continue;
                if (first < 0) first = (long)cg[i];
                else if ((long)cg[i] != first) same_slice = false;
            }
        }
    } else {
        same_slice = pl.cube == 1 && (!any_psf || pl.sfC == 1);
    }
    uint32_t cps = t.channels_per_stream ? t.channels_per_stream : 8;
    if (!same_slice) cps = 1;
    pl.cps = std::max<uint32_t>(1, std::min<uint32_t>(cps, (uint32_t)std::max<size_t>(1, C)));
    pl.ncb = (uint32_t)((C + pl.cps - 1) / pl.cps);
    pl.base_rows = t.base_rows ? t.base_rows : 32;

    std::vector<uint32_t> bls((size_t)pl.B + 1);
    HIP_TRY(hipMemcpyAsync(bls.data(), pl.blstart.p, bls.size() * 4, hipMemcpyDeviceToHost,
                           lv.stream));
    HIP_TRY(hipStreamSynchronize(lv.stream));
    uint32_t max_nb = 0;
    for (uint32_t b = 0; b < pl.B; ++b) max_nb = std::max(max_nb, bls[b + 1] - bls[b]);
    pl.nseg_base = std::max<uint32_t>(1, (max_nb + pl.base_rows - 1) / pl.base_rows);
    pl.nbase = (uint64_t)pl.B * pl.ncb * pl.nseg_base;

    a.perm = pl.perm.as<uint32_t>();
    a.bl_start = pl.blstart.as<uint32_t>();
    a.counters = lv.counters;
    a.error_flag = lv.errflag;
    a.nslots = lv.nslots;
    a.baselines = pl.B;
    a.chan = (uint32_t)C;
    a.npol = (uint32_t)npol;
    a.spw_count = (uint32_t)nspw;
    a.antenna_count = (uint32_t)p->antenna_count;
    a.imaging_field = p->imaging_field;
    a.facets = 1;
    a.chan_per_stream = pl.cps;
    a.n_chan_blocks = pl.ncb;
    a.n_segments = pl.nseg_base;
    a.n_streams = pl.nbase;
    for (int k = 0; k < 4; ++k) a.pol_idx[k] = (uint32_t)k;
    if (mode == MODE_1) a.pol_idx[0] = (uint32_t)p->polarization_index;
    if (mode == MODE_2) {
        a.pol_idx[0] = (uint32_t)p->polarization_index;
        a.pol_idx[1] = (uint32_t)p->second_polarization_index;
    }
    a.phase = 1;
    a.transform = 1;
    // the walk order compares footprint movement per time step and per channel step in grid
    // cells; the ratio does not depend on the cell scale, so the first facet's scale stands in
    // for all (the shared arguments carry no scale of their own)
    a.u_scale = (double)((double)facets[0].nx * facets[0].cell_x * (double)ARCSEC_TO_RAD);
    a.v_scale = -(double)((double)facets[0].ny * facets[0].cell_y * (double)ARCSEC_TO_RAD);
    HIP_TRY(pl.walk.ensure(walk_order_bytes()));
    a.time_inner = pl.walk.as<int>();
    HIP_TRY(launch_walk_order(a, lv.walk_order, pl.walk.as<int>(), lv.stream));
    // sorted row table for the class kernels' sample preparation (rebuilt by every pass, see
    // enqueue_all; this copy serves the autotuner's single launches)
    pl.rows = (uint32_t)rows;
    HIP_TRY(pl.rowrec.ensure(std::max<size_t>(1, rows) * sizeof(RowRec)));
    a.rowrec = pl.rowrec.as<RowRec>();
    HIP_TRY(build_row_tables(a, pl.rows, false, pl.rowrec.as<RowRec>(), nullptr, lv.stream));

    // per-facet accumulation grids, normalisation slots and geometry
    pl.fkind.resize(n);
    pl.fcube.resize(n);
    pl.fnc.resize(n);
    pl.fset.resize(n);
    pl.acc_off.resize(n);
    pl.geom_host.resize(n);
    size_t tot = 0;
    for (uint32_t f = 0; f < n; ++f) {
        const bullseye_batch_facet &fc = facets[f];
        const bool psf = fc.sampling_function != 0;
        pl.fkind[f] = psf ? MODE_PSF : mode;
        pl.fcube[f] = psf ? pl.sfC : pl.cube;
        pl.fnc[f] = psf ? 1 : pl.NC;
        pl.fset[f] = fc.kernel_set;
        pl.acc_off[f] = tot;
        tot += (size_t)pl.fcube[f] * pl.fnc[f] * fc.ny * fc.nx * 2;
        pl.geom_host[f] = facet_geometry(p->phase_centre_ra, p->phase_centre_dec, fc.ra, fc.dec,
                                         !psf);
    }
    pl.acc_elems = tot;
    HIP_TRY(pl.acc.ensure(tot * 8));
    pl.norm_stride = (size_t)n * pl.cubeB * pl.NC;
    HIP_TRY(pl.norm.ensure(pl.norm_stride * lv.nslots * 8));
    a.norm = pl.norm.as<double>();
    a.norm_stride = pl.norm_stride;
    HIP_TRY(pl.geom.ensure(n * sizeof(FacetGeom)));
    HIP_TRY(hipMemcpyAsync(pl.geom.p, pl.geom_host.data(), n * sizeof(FacetGeom),
                           hipMemcpyHostToDevice, lv.stream));

    // classes per kernel set in use
    pl.set_slot.assign(MAX_SETS, 0xffffffffu);
    uint32_t nused = 0;
    for (uint32_t f = 0; f < n; ++f)
        if (pl.set_slot[pl.fset[f]] == 0xffffffffu) pl.set_slot[pl.fset[f]] = nused++;
    HIP_TRY(pl.cls_all.ensure(std::max<size_t>(16, (size_t)nused * pl.nbase)));
    for (uint32_t id = 0; id < MAX_SETS; ++id) {
        const uint32_t k = pl.set_slot[id];
        if (k == 0xffffffffu) continue;
        const KernelSet &ks = sets[id];
        HIP_TRY(launch_stream_class(a, ks.halfd.as<uint32_t>(), ks.wplanes, ks.wmax,
                                    pl.nseg_base, pl.base_rows,
                                    pl.cls_all.as<uint8_t>() + (size_t)k * pl.nbase, lv.stream));
    }
    pl.cls_host.resize((size_t)nused * pl.nbase);
    HIP_TRY(hipMemcpyAsync(pl.cls_host.data(), pl.cls_all.p, pl.cls_host.size(),
                           hipMemcpyDeviceToHost, lv.stream));

    // facet argument table
    pl.fargs_host.resize(n);
    for (uint32_t f = 0; f < n; ++f) {
        const bullseye_batch_facet &fc = facets[f];
        const KernelSet &ks = sets[fc.kernel_set];
        wp::WArgs &w = pl.fargs_host[f];
        memset(&w, 0, sizeof w);
        GridArgs &fa = w.g;
        fa = a;
        fa.grid = pl.acc.as<double>() + pl.acc_off[f];
        fa.geom = pl.geom.as<FacetGeom>() + f;
        fa.nx = fc.nx;
        fa.ny = fc.ny;
        fa.S = ks.S;
        fa.W = 2 * ks.S + 1;
        fa.OS = ks.OS;
        fa.csz = ks.csz;
        fa.cube = pl.fcube[f];
        fa.pstride = pl.fnc[f];
        fa.cgi = pl.fkind[f] == MODE_PSF ? cgi_psf : cgi_vis;
        fa.phase = pl.fkind[f] == MODE_PSF ? 0 : 1;
        // templated_gridder.h:23-27 with this facet's size and cell
        fa.u_scale = (double)((double)fc.nx * fc.cell_x * (double)ARCSEC_TO_RAD);
        fa.v_scale = -(double)((double)fc.ny * fc.cell_y * (double)ARCSEC_TO_RAD);
        fa.off_x = (double)(fc.nx / 2 - ks.S);
        fa.off_y = (double)(fc.ny / 2 - ks.S);
        w.taps = lv.path == PATH_FP64 ? ks.taps_d.p : ks.taps_f.p;
        w.plane_off = ks.off.as<uint64_t>();
        w.ctaps = lv.path == PATH_FP64 ? ks.ctaps_d.p : ks.ctaps_f.p;
        w.cplane_off = ks.coff.as<uint64_t>();
        w.plane_half = ks.halfd.as<uint32_t>();
        w.plane_norm = ks.norm.as<double>();
        w.wplanes = ks.wplanes;
        w.nf = ks.nf;
        w.fbase = ks.fbase;
        w.wmax_f = ks.wmax;
        w.seg_cls = pl.cls_all.as<uint8_t>() + (size_t)pl.set_slot[fc.kernel_set] * pl.nbase;
    }
    HIP_TRY(pl.fargs.ensure(n * sizeof(wp::WArgs)));
    HIP_TRY(hipMemcpyAsync(pl.fargs.p, pl.fargs_host.data(), n * sizeof(wp::WArgs),
                           hipMemcpyHostToDevice, lv.stream));
    HIP_TRY(hipStreamSynchronize(lv.stream));

    if (!pl.fork) {
        HIP_TRY(hipEventCreateWithFlags(&pl.fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreate(&pl.t0));
        HIP_TRY(hipEventCreate(&pl.t1));
        for (int b = 0; b < BRANCHES - 1; ++b) {
            HIP_TRY(hipStreamCreateWithFlags(&pl.side[b], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&pl.join[b], hipEventDisableTiming));
        }
    }
    return build_lists(pl, t);
}

int plan_impl(gridding_parameters *p, int mode, const bullseye_batch_facet *facets, uint32_t n,
              int schedule) {
    if (build_plan(p, mode, facets, n, schedule, tuning, plan)) return 1;
    if (schedule == BULLSEYE_BATCH_PERSISTENT_GRAPH && capture(plan, lib_view().stream)) return 1;
    plan.ready = true;
    return 0;
}

unsigned long long counters_total(const LibView &lv, unsigned long long out[2]) {
    std::vector<unsigned long long> cs(2 * (size_t)lv.nslots);
    if (hipMemcpy(cs.data(), lv.counters, cs.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return 1;
    out[0] = out[1] = 0;
    for (size_t i = 0; i < lv.nslots; ++i) {
        out[0] += cs[2 * i];
        out[1] += cs[2 * i + 1];
    }
    return 0;
}

int execute_impl() {
    if (!plan.ready) return fail("no batch plan (bullseye_batch_plan)");
    const LibView lv = lib_view();
    Plan &pl = plan;
    const hipStream_t st = lv.stream;
    HIP_TRY(hipStreamSynchronize(st));
    unsigned long long c0[2];
    if (counters_total(lv, c0)) return fail("reading counters");
    HIP_TRY(hipEventRecord(pl.t0, st));
    if (pl.exec) {
        HIP_TRY(hipGraphLaunch(pl.exec, st));
    } else if (enqueue_all(pl, st)) {
        return 1;
    }
    HIP_TRY(hipEventRecord(pl.t1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, pl.t0, pl.t1));
    if (check_device_errors()) return 1;
    unsigned long long c1[2];
    if (counters_total(lv, c1)) return fail("reading counters");
    lv.stats->samples_gridded = c1[0] - c0[0];
    lv.stats->taps = c1[1] - c0[1];
    lv.stats->kernel_ms = ms;
    lv.stats->prep_ms = 0;
    lv.stats->copy_ms = 0;
    int ti = 0;
    HIP_TRY(hipMemcpy(&ti, pl.walk.p, sizeof ti, hipMemcpyDeviceToHost));
    lv.stats->time_inner = (uint32_t)ti;
    lv.last_counters[0] = c1[0];
    lv.last_counters[1] = c1[1];
    *lv.grid_seconds += ms * 1e-3;

    // copy out: grids in the grid type, normalisation sums per facet
    HIP_TRY(fold_slots(pl.norm.as<double>(), pl.norm_stride, lv.nslots, st));
    pl.norm_host.resize(pl.norm_stride);
    HIP_TRY(hipMemcpyAsync(pl.norm_host.data(), pl.norm.p, pl.norm_stride * 8,
                           hipMemcpyDeviceToHost, st));
    for (uint32_t f = 0; f < pl.n; ++f) {
        const bullseye_batch_facet &fc = pl.facets[f];
        if (!fc.output) continue;
        const uint64_t npix = (uint64_t)fc.nx * fc.ny;
        const size_t bytes = (size_t)pl.fcube[f] * pl.fnc[f] * npix * 2 * sizeof(grid_base_type);
        void *dst = fc.output;
        if (!lv.device_resident) {
            HIP_TRY(pl.out_tmp.ensure(bytes));
            dst = pl.out_tmp.p;
        }
        HIP_TRY(grid_to_abi(pl.acc.as<double>() + pl.acc_off[f], dst,
                            pl.fcube[f], pl.fnc[f], pl.fnc[f], npix, st));
        if (!lv.device_resident) {
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(hipMemcpy(fc.output, pl.out_tmp.p, bytes, hipMemcpyDeviceToHost));
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t f = 0; f < pl.n; ++f) {
        double *nout = pl.facets[f].normalization;
        if (!nout) continue;
        const uint32_t nc = pl.fnc[f];
        for (uint32_t c = 0; c < pl.fcube[f]; ++c)
            for (uint32_t k = 0; k < nc; ++k)
                nout[c * nc + k] = pl.fkind[f] == MODE_PSF
                                       ? 0.0
                                       : pl.norm_host[((size_t)f * pl.cubeB + c) * pl.NC + k];
    }
    return 0;
}

// Device time of one class launch (queue cleared first), best of two.
int time_launch(Plan &pl, size_t j, int shape, double &ms_out) {
    const LibView lv = lib_view();
    const hipStream_t st = lv.stream;
    float best = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
        HIP_TRY(fill_u32(pl.queues.as<uint32_t>() + j, 1, 0u, st));
        HIP_TRY(hipEventRecord(pl.t0, st));
        HIP_TRY(issue(pl, pl.launches[j], shape, st));
        HIP_TRY(hipEventRecord(pl.t1, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, pl.t0, pl.t1));
        best = std::min(best, ms);
    }
    ms_out = best;
    return 0;
}

int autotune_impl(gridding_parameters *p, int mode, const bullseye_batch_facet *facets,
                  uint32_t n, bullseye_batch_tuning *result) {
    const LibView lv = lib_view();
    if (validate(lv, p, mode, facets, n)) return 1;
    plan.release();
    const std::vector<uint32_t> cps_list =
        tuning.channels_per_stream ? std::vector<uint32_t>{tuning.channels_per_stream}
                                   : std::vector<uint32_t>{4, 8, 16};
    const std::vector<uint32_t> base_list =
        tuning.base_rows ? std::vector<uint32_t>{tuning.base_rows}
                         : std::vector<uint32_t>{16, 32, 64};
    const int shapes[] = {WSHAPE_COLUMN, WSHAPE_PAIRS1, WSHAPE_PAIRS2, WSHAPE_PAIRS4};
    bullseye_batch_tuning best_t{};
    double best_total = 1e300, default_total = 0;
    for (uint32_t cps : cps_list) {
        for (uint32_t base : base_list) {
            // per class (index i): best time and choice over chunks and shapes, summed over kinds
            double best_ms[NCLS], def_ms[NCLS] = {0};
            uint32_t best_shape[NCLS] = {0}, best_chunk[NCLS] = {0};
            for (uint32_t i = 0; i < NCLS; ++i) best_ms[i] = 1e300;
            for (uint32_t chunk : {1u, 2u, 4u}) {
                bullseye_batch_tuning t{};
                t.struct_size = sizeof t;
                t.channels_per_stream = cps;
                t.base_rows = base;
                for (uint32_t i = 0; i < NCLS; ++i) t.chunk[i] = chunk;
                if (build_plan(p, mode, facets, n, BULLSEYE_BATCH_PERSISTENT, t, plan)) return 1;
                double ms_cls[4][NCLS];
                for (auto &row : ms_cls)
                    for (double &v : row) v = 0;
                bool valid[4][NCLS];
                for (int s = 0; s < 4; ++s)
                    for (uint32_t i = 0; i < NCLS; ++i)
                        valid[s][i] = wproj_shape_valid((int)(2 * i + 1), shapes[s]);
                for (size_t j = 0; j < plan.launches.size(); ++j) {
                    const uint32_t i = plan.launches[j].M / 2;
                    for (int s = 0; s < 4; ++s) {
                        if (!valid[s][i]) continue;
                        double ms = 0;
                        if (time_launch(plan, j, shapes[s], ms)) return 1;
                        ms_cls[s][i] += ms;
                    }
                }
                for (uint32_t i = 0; i < NCLS; ++i) {
                    for (int s = 0; s < 4; ++s) {
                        if (!valid[s][i]) continue;
                        if (ms_cls[s][i] < best_ms[i]) {
                            best_ms[i] = ms_cls[s][i];
                            best_shape[i] = (uint32_t)shapes[s];
                            best_chunk[i] = chunk;
                        }
                    }
                    if (chunk == 1) {  // default tile: column kernel up to 15, fewest pairs above
                        const int s_def = 2 * i + 1 <= 15 ? 0 : 1;
                        int s = s_def;
                        while (s < 4 && !valid[s][i]) ++s;
                        def_ms[i] = s < 4 ? ms_cls[s][i] : 0;
                    }
                }
                if (check_device_errors()) return 1;
            }
            double total = 0, dtotal = 0;
            for (uint32_t i = 0; i < NCLS; ++i) {
                if (best_ms[i] < 1e299) total += best_ms[i];
                dtotal += def_ms[i];
            }
            if (total < best_total) {
                best_total = total;
                best_t.struct_size = sizeof best_t;
                best_t.channels_per_stream = cps;
                best_t.base_rows = base;
                for (uint32_t i = 0; i < NCLS; ++i) {
                    best_t.shape[i] = best_ms[i] < 1e299 ? best_shape[i] : 0;
                    best_t.chunk[i] = best_ms[i] < 1e299 ? best_chunk[i] : 1;
                }
            }
            if (cps == 8 || cps_list.size() == 1)
                if (base == 32 || base_list.size() == 1) default_total = dtotal;
        }
    }
    // work-list order at the chosen tiles
    double order_ms[2] = {0, 0};
    for (uint32_t fm = 0; fm < 2; ++fm) {
        bullseye_batch_tuning t = best_t;
        t.facet_major = fm;
        if (build_plan(p, mode, facets, n, BULLSEYE_BATCH_PERSISTENT, t, plan)) return 1;
        for (size_t j = 0; j < plan.launches.size(); ++j) {
            double ms = 0;
            if (time_launch(plan, j, plan.launches[j].shape, ms)) return 1;
            order_ms[fm] += ms;
        }
    }
    if (check_device_errors()) return 1;
    best_t.facet_major = order_ms[1] < order_ms[0] ? 1 : 0;
    best_total = std::min(order_ms[0], order_ms[1]);
    plan.release();
    best_t.tuned = 1;
    best_t.tuned_ms = best_total;
    best_t.default_ms = default_total;
    tuning = best_t;
    if (result) *result = best_t;
    return 0;
}

}  // namespace

void batch_release_all() {
    plan.release();
    for (KernelSet &ks : sets) ks.release();
}

}  // namespace bam

using namespace bam;

extern "C" {

int bullseye_batch_kernel_set(uint32_t id, const void *conv, uint32_t wplanes,
                              uint32_t conv_support, uint32_t conv_oversample,
                              const uint32_t *half_support, float wmax) {
    const LibView lv = lib_view();
    if (!lv.initialized) return fail("initLibrary has not been called");
    if (id >= MAX_SETS) return fail("kernel set id must be < %u", MAX_SETS);
    if (!conv || wplanes < 2 || wplanes > 255) return fail("a kernel set needs 2..255 w-planes");
    if (conv_support > 31 || conv_oversample < 1 || conv_oversample > 128)
        return fail("kernel set support must be <= 31 and oversampling 1..128");
    if (!(wmax > 0)) return fail("wmax must be > 0");
    plan.release();
    KernelSet &ks = sets[id];
    ks.release();
    ks.wplanes = wplanes;
    ks.S = conv_support;
    ks.OS = conv_oversample;
    ks.csz = (2 * ks.S + 3) + (2 * ks.S + 2) * (ks.OS - 1);
    ks.wmax = wmax;
    ks.half.assign(wplanes, ks.S);
    if (half_support)
        for (uint32_t q = 0; q < wplanes; ++q) {
            if (half_support[q] > ks.S)
                return fail("plane %u: half support %u exceeds conv_support", q, half_support[q]);
            ks.half[q] = half_support[q];
        }
    if (build_wtaps(conv, lv.device_resident, wplanes, ks.S, ks.OS, ks.half, ks.taps_f, ks.taps_d,
                    ks.off, ks.halfd, ks.norm, ks.ctaps_f, ks.ctaps_d, ks.coff, ks.nf, ks.fbase))
        return 1;
    ks.defined = true;
    return 0;
}

int bullseye_batch_set_tuning(const bullseye_batch_tuning *t) {
    if (!t || t->struct_size != sizeof(bullseye_batch_tuning))
        return fail("bullseye_batch_tuning.struct_size mismatch");
    if (t->base_rows && (t->base_rows < 4 || t->base_rows > 4096))
        return fail("base_rows must be 4..4096");
    if (t->channels_per_stream > 4096) return fail("channels_per_stream too large");
    for (uint32_t i = 0; i < NCLS; ++i) {
        if (t->shape[i] > WSHAPE_PAIRS4) return fail("unknown lane shape %u", t->shape[i]);
        if (t->chunk[i] > 4 || t->chunk[i] == 3) return fail("chunk must be 0, 1, 2 or 4");
    }
    tuning = *t;
    plan.release();
    return 0;
}

int bullseye_batch_get_tuning(bullseye_batch_tuning *t) {
    if (!t || t->struct_size != sizeof(bullseye_batch_tuning))
        return fail("bullseye_batch_tuning.struct_size mismatch");
    *t = tuning;
    return 0;
}

int bullseye_batch_plan(gridding_parameters *p, int mode, const bullseye_batch_facet *facets,
                        uint32_t n, int schedule) {
    const int rc = plan_impl(p, mode_of(mode), facets, n, schedule);
    if (rc) plan.release();
    return rc;
}

int bullseye_batch_execute(void) { return execute_impl(); }

int bullseye_batch_autotune(gridding_parameters *p, int mode, const bullseye_batch_facet *facets,
                            uint32_t n, bullseye_batch_tuning *result) {
    const int rc = autotune_impl(p, mode_of(mode), facets, n, result);
    if (rc) plan.release();
    return rc;
}

int bullseye_batch_release(void) {
    (void)hipStreamSynchronize(lib_view().stream);
    batch_release_all();
    return 0;
}

}  // extern "C"
