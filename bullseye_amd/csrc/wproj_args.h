// wproj_args.h -- arguments of the w-projection class kernels (csrc/wproj.hip), shared with the
// host code that builds batched launches.
#pragma once
#include "device_common.h"

namespace bam {
namespace wp {

struct WArgs {
    GridArgs g;                        // shared fields (vis, uvw, flags, grid, norm, ...)
    const void *taps;                  // re-laid taps, complex acc_t
    const uint64_t *plane_off;         // [wplanes] element offset of plane p's blocks
    const uint32_t *plane_half;        // [wplanes] S_p
    const double *plane_norm;          // [wplanes][nf][nf] sum of Re(taps) per block
    const uint32_t *stream_list;       // streams of this class
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
};

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

}  // namespace bam
