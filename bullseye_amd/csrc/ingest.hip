// ingest.hip -- device-side Romein ingest: stable sort of rows by baseline.
//
// Replaces the host counting sort of bullseye's GPU library (gpu_algorithm/wrapper.cu:358-454,
// repack_input_data): instead of physically reordering every row array, the rows are
// radix-sorted by baseline index (stable, so time order inside a baseline is kept) into a
// permutation the gridding kernel gathers through, and the per-baseline row counts are
// prefix-summed into baseline starting indexes.  The baseline index is the reference's
// compute_baseline_index (wrapper.cu:358-366): s*(2N+1-s)/2 + |a1-a2| with s = min(a1, a2),
// autocorrelations included, N(N+1)/2 baselines.
#include <hipcub/hipcub.hpp>

#include "device_common.h"

namespace bam {

__global__ void k_baseline_keys(const uint32_t *a1, const uint32_t *a2, uint32_t rows,
                                uint32_t n_ant, uint32_t n_bl, uint32_t *keys, uint32_t *vals,
                                uint32_t *counts, int *err) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t x = a1[r], y = a2[r];
    uint32_t b = 0;
    if (x >= n_ant || y >= n_ant) {
        atomicOr(err, 2);
    } else {
        const uint64_t s = x < y ? x : y;
        const uint64_t d = x < y ? y - x : x - y;
        b = (uint32_t)((s * (2 * (uint64_t)n_ant + 1 - s)) / 2 + d);
        if (b >= n_bl) {
            atomicOr(err, 2);
            b = 0;
        }
    }
    keys[r] = b;
    vals[r] = r;
    atomicAdd(counts + b, 1u);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static int bits_for(uint32_t n) {
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

// Workspace bytes for ingest of `rows` rows and `n_bl` baselines.
size_t ingest_workspace_bytes(uint32_t rows, uint32_t n_bl) {
    size_t sort_bytes = 0, scan_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)nullptr,
                                       (uint32_t *)nullptr, (uint32_t *)nullptr,
                                       (uint32_t *)nullptr, (int)rows, 0, bits_for(n_bl + 1));
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr,
                                     (uint32_t *)nullptr, (int)n_bl + 1);
    return (sort_bytes > scan_bytes ? sort_bytes : scan_bytes) + 256;
}

// keys/vals/keys_out: [rows] scratch; counts: [n_bl+1] scratch; perm: [rows] out;
// bl_start: [n_bl+1] out.
hipError_t ingest_sort(const uint32_t *a1, const uint32_t *a2, uint32_t rows, uint32_t n_ant,
                       uint32_t n_bl, uint32_t *keys, uint32_t *vals, uint32_t *keys_out,
                       uint32_t *counts, uint32_t *perm, uint32_t *bl_start, void *ws,
                       size_t ws_bytes, int *err, hipStream_t st) {
    // cleared with a kernel, not hipMemsetAsync, so the sequence stays graph-capturable
    k_fill_u32<<<(n_bl + 1 + 255) / 256, 256, 0, st>>>(counts, n_bl + 1, 0u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (rows) {
        k_baseline_keys<<<(rows + 255) / 256, 256, 0, st>>>(a1, a2, rows, n_ant, n_bl, keys,
                                                              vals, counts, err);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        size_t b = ws_bytes;
        e = hipcub::DeviceRadixSort::SortPairs(ws, b, keys, keys_out, vals, perm, (int)rows, 0,
                                               bits_for(n_bl + 1), st);
        if (e != hipSuccess) return e;
    }
    size_t b2 = ws_bytes;
    return hipcub::DeviceScan::ExclusiveSum(ws, b2, counts, bl_start, (int)n_bl + 1, st);
}

}  // namespace bam
