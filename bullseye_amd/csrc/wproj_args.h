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
    uint32_t seg_rows;                 // rows per stream segment
    // Batched facets: list entries are (facet << 24 | stream) and facet_args[facet] holds that
    // facet's arguments (grid, geometry, kernel set); nullptr for a single-facet-set call.
    const WArgs *facet_args;
    // Persistent launch: waves (column kernel) or blocks (pair kernel) take items (packs of
    // streams) from this counter until n_items are taken; nullptr: one item per wave/block.
    uint32_t *queue;
    uint32_t n_items;
};

// The arguments and stream index of list entry li (see WArgs::facet_args).
__device__ __forceinline__ const WArgs &entry_args(const WArgs &w, uint64_t li, uint32_t &stream,
                                                   uint32_t &fidx) {
    const uint32_t e = w.stream_list[li];
    if (w.facet_args) {
        fidx = e >> 24;
        stream = e & 0xffffffu;
        return w.facet_args[fidx];
    }
    fidx = 0;
    stream = e;
    return w;
}

}  // namespace wp

// One class launch (width w.M): the column-ownership kernel for M <= 15, residue pairs above;
// persistent when w.queue is set.
hipError_t launch_wproj_class(const wp::WArgs &w, int mode, int path, hipStream_t st);
// Lanes per stream, streams per item and block size of a class of width M.
void wproj_shape(int M, uint32_t &block, uint32_t &lanes, uint32_t &spb, int &pairs);
// Streams per work item of a class launch: GPW for the column kernel (one wave), the block's
// streams for the pair kernel.
uint32_t wproj_streams_per_item(int M, int mode, int path);

}  // namespace bam
