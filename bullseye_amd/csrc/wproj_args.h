// wproj_args.h -- arguments of the w-projection class kernels (csrc/wproj.hip), shared with the
// host code that builds batched launches.
#pragma once
#include "device_common.h"

namespace bam {
namespace wp {

constexpr int WCOL_MAX = 15;                           // widest class of the column kernel
constexpr int WCOL_CLASSES = (WCOL_MAX + 1) / 2;       // its classes 1, 3, .., WCOL_MAX

struct WArgs {
    GridArgs g;                        // shared fields (vis, uvw, flags, grid, norm, ...)
    const void *taps;                  // re-laid taps, complex acc_t
    const uint64_t *plane_off;         // [wplanes] element offset of plane p's blocks
    const uint32_t *plane_half;        // [wplanes] S_p
    const double *plane_norm;          // [wplanes][nf][nf] sum of Re(taps) per block
    // Column tables of the classes M = 1, 3, .., WCOL_MAX (column kernel): cplane_off[(M >> 1)
    // * wplanes + p] is the first tap of plane p's table in class M (planes with W_p <= M).
    // Per (frac_v, frac_u) block 2M rows of M taps: entry (j, t) holds window row j mod M of
    // column t, zero where j mod M >= W_p or t >= W_p.
    const void *ctaps;                 // complex acc_t
    const uint64_t *cplane_off;
    const uint32_t *stream_list;       // streams of this class
    // batched facets: entries (facet << 32 | stream), stream 0xffffffff a no-op padding entry
    const uint64_t *stream_list64;
    uint32_t n_list;
    uint32_t wplanes, nf, fbase;       // fraction index range [fbase, fbase + nf)
    uint32_t M;                        // class width (odd)
    uint32_t lanes_per_stream, streams_per_block;
    float wmax_f;
    // Stream decomposition of the launch (read from the launch's own arguments, not a batched
    // facet's): rows per segment and segments per (baseline, channel block).  A launch whose
    // segments span several classification segments of base_rows rows walks only the rows of
    // the segments whose class (seg_cls, nseg_base per (baseline, channel block)) is M.
    uint32_t seg_rows, nseg;
    uint32_t base_rows, nseg_base;
    uint32_t filter;                   // 1: skip rows of other classes (seg_rows > base_rows)
    const uint8_t *seg_cls;            // per facet: the classes of its kernel set
    // Batched facets: list entries are (facet << 24 | stream) and facet_args[facet] holds that
    // facet's arguments (grid, geometry, kernel set); nullptr for a single-facet-set call.
    const WArgs *facet_args;
    // Persistent launch: waves (column kernel) or blocks (pair kernel) take items (packs of
    // streams) from this counter until n_items are taken; nullptr: one item per wave/block.
    uint32_t *queue;
    uint32_t n_items;
    // Single-facet-set launches whose class lists were built on the device (nullptr: the host
    // built stream_list): the class's entries are stream_list[*list_start .. *list_end), entry
    // e = segment * n_chan_blocks + channel block, and list index li is entry li / facets with
    // facet li % facets; segment s belongs to baseline seg_bl[s], whose first segment is
    // seg_first[baseline].  n_list and n_items then follow from the device counts.
    const uint32_t *list_start, *list_end;
    const uint32_t *seg_bl, *seg_first;
    // Device lists holding several classes (pair kernel, lanes shaped for the widest, M): the
    // class of sorted list position i is list_cls[i], and each workgroup's stream takes its own
    // width; nullptr: every stream has width M.
    const uint8_t *list_cls;
    // Pair kernel: on-chip window of the grid around each stream's track, win x win cells per
    // residue pair and axis (a (win M)^2 window), set by the launch from the LDS it has room for;
    // 0: every flush goes to the grid.
    uint32_t win;
};

// Stream of list index li of a launch without batched facets: facet, baseline, channel block
// and segment.
__device__ __forceinline__ void single_stream(const WArgs &w, const GridArgs &a, uint64_t li,
                                              uint32_t &facet, uint32_t &bl, uint32_t &cb,
                                              uint32_t &seg) {
    if (w.seg_bl) {
        const uint32_t e = w.stream_list[*w.list_start + (uint32_t)(li / a.facets)];
        facet = (uint32_t)(li % a.facets);
        cb = e % a.n_chan_blocks;
        const uint32_t s = e / a.n_chan_blocks;
        bl = w.seg_bl[s];
        seg = s - w.seg_first[bl];
        return;
    }
    uint64_t t = w.stream_list[li];
    facet = (uint32_t)(t % a.facets);
    t /= a.facets;
    seg = (uint32_t)(t % w.nseg);
    t /= w.nseg;
    cb = (uint32_t)(t % a.n_chan_blocks);
    bl = (uint32_t)(t / a.n_chan_blocks);
}

// Streams in the launch's list, and work items of `per` streams.
__device__ __forceinline__ uint64_t list_len(const WArgs &w) {
    return w.seg_bl ? (uint64_t)(*w.list_end - *w.list_start) * w.g.facets : w.n_list;
}
__device__ __forceinline__ uint64_t list_items(const WArgs &w, uint32_t per) {
    return w.seg_bl ? (list_len(w) + per - 1) / per : w.n_items;
}

}  // namespace wp

// Lane shapes of a class launch: the column-ownership kernel (M <= 15), or the residue-pair
// kernel with the fewest pairs per lane that fit (1 up to M = 31, 2 up to 45, else 4) or with
// at least 2 or 4 pairs per lane (fewer work-items per footprint).
enum { WSHAPE_DEFAULT = 0, WSHAPE_COLUMN = 1, WSHAPE_PAIRS1 = 2, WSHAPE_PAIRS2 = 3,
       WSHAPE_PAIRS4 = 4 };

// One class launch (width w.M, odd, <= 63) in the given shape; persistent when w.queue is set.
hipError_t launch_wproj_class(const wp::WArgs &w, int mode, int path, int shape, hipStream_t st);
// Residue-pair kernel shape of width M with at least min_pairs pairs per lane.
void wproj_shape(int M, int min_pairs, uint32_t &block, uint32_t &lanes, uint32_t &spb,
                 int &pairs);
// Whether a shape applies to width M (and differs from a smaller pairs count).
bool wproj_shape_valid(int M, int shape);
// Streams per work item of a class launch: GPW for the column kernel (one wave), the block's
// streams for the pair kernel.
uint32_t wproj_streams_per_item(int M, int mode, int path, int shape);

// Device-built class lists of a single-facet-set w-projection call (no host round trip): the
// stream segments of SEG rows per baseline, each (segment, channel block) entry classified by
// the widest plane its samples use and sorted stably by class; class M's entries start at
// class_start[M] and end at class_start[M + 1].  Workspace layout and sizes from
// wproj_list_plan; every array lives in the workspace.
struct WListPlan {
    uint64_t nseg_max, entries;          // segment bound, entries = nseg_max * n_chan_blocks
    size_t off_seg_first, off_seg_bl, off_cls, off_cls_sorted, off_val, off_val_sorted,
        off_class_start, off_queues, off_sort, sort_bytes, total;
};
WListPlan wproj_list_plan(uint32_t rows, uint32_t baselines, uint32_t n_chan_blocks);
hipError_t wproj_build_lists(const GridArgs &a, const WListPlan &pl, void *ws,
                             const uint32_t *plane_half, uint32_t wplanes, float wmax_f,
                             hipStream_t st);
// One class launch (persistent) over the device-built list of class M.
hipError_t launch_wproj_device_list(const GridArgs &a, const WListPlan &pl, void *ws,
                                    const void *taps, const uint64_t *plane_off,
                                    const void *ctaps, const uint64_t *cplane_off,
                                    const uint32_t *plane_half, const double *plane_norm,
                                    uint32_t wplanes, uint32_t nf, uint32_t fbase, uint32_t M,
                                    uint32_t M_hi, float wmax_f, int mode, int path,
                                    hipStream_t st);

}  // namespace bam
