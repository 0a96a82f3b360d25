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

// Walk order of the Romein streams: the mean footprint movement, |du| + |dv| in grid cells,
// per time step (consecutive rows of a baseline at the middle channel) and per channel step
// (the middle channel pair), over up to 4096 baselines x 32 rows.  Time inner when a time step
// moves the footprint less: that order flushes fewer grid points.  WALK_BLOCKS workgroups each
// sum a share of the baselines into their own partial (no atomics); one thread then adds the
// partials in block order, so the result is the same on every run.  The flag stays on the
// device for the gridding kernels.  override_order: 1 channels inner, 2 time inner, anything
// else measured.
constexpr int WALK_BLOCKS = 64;
constexpr int WALK_THREADS = 256;

__global__ __launch_bounds__(WALK_THREADS) void k_walk_partial(GridArgs a, double *part) {
    __shared__ double s_d[2][WALK_THREADS];
    __shared__ uint32_t s_n[2][WALK_THREADS];
    double dt = 0, dc = 0;
    uint32_t nt = 0, nc = 0;
    if (a.chan > 1) {
        const uint32_t B = a.baselines;
        const uint32_t nbs = B < 4096u ? B : 4096u;
        const uint32_t cm = (a.chan - 1) / 2, cn = cm + 1;
        for (uint32_t i = blockIdx.x * WALK_THREADS + threadIdx.x; i < nbs;
             i += WALK_BLOCKS * WALK_THREADS) {
            const uint32_t b = (uint32_t)((uint64_t)i * B / nbs);
            const uint32_t r0 = a.bl_start[b];
            const uint32_t r1 = min(a.bl_start[b + 1], r0 + 33);
            double pu = 0, pv = 0;
            bool have = false;
            for (uint32_t r = r0; r < r1; ++r) {
                const uint32_t row = a.perm[r];
                const uint32_t s = a.spw[row];
                if (s >= a.spw_count) continue;
                const double u = (double)a.uvw[3 * (size_t)row] * a.u_scale;
                const double v = (double)a.uvw[3 * (size_t)row + 1] * a.v_scale;
                const double rl = 1.0 / (double)a.wavelengths[s * a.chan + cm];
                const double rl2 = 1.0 / (double)a.wavelengths[s * a.chan + cn];
                if (have) {
                    dt += fabs(u * rl - pu) + fabs(v * rl - pv);
                    ++nt;
                }
                pu = u * rl;
                pv = v * rl;
                have = true;
                dc += (fabs(u) + fabs(v)) * fabs(rl2 - rl);
                ++nc;
            }
        }
    }
    s_d[0][threadIdx.x] = dt;
    s_d[1][threadIdx.x] = dc;
    s_n[0][threadIdx.x] = nt;
    s_n[1][threadIdx.x] = nc;
    __syncthreads();
    for (int h = WALK_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_d[0][threadIdx.x] += s_d[0][threadIdx.x + h];
            s_d[1][threadIdx.x] += s_d[1][threadIdx.x + h];
            s_n[0][threadIdx.x] += s_n[0][threadIdx.x + h];
            s_n[1][threadIdx.x] += s_n[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = part + 4 * blockIdx.x;
        o[0] = s_d[0][0];
        o[1] = s_d[1][0];
        o[2] = (double)s_n[0][0];
        o[3] = (double)s_n[1][0];
    }
}

__global__ void k_walk_decide(const double *part, int override_order, int *time_inner) {
    if (threadIdx.x != 0) return;
    int ti = 0;
    if (override_order == 2) {
        ti = 1;
    } else if (override_order != 1) {
        double dt = 0, dc = 0, nt = 0, nc = 0;
        for (int b = 0; b < WALK_BLOCKS; ++b) {
            dt += part[4 * b];
            dc += part[4 * b + 1];
            nt += part[4 * b + 2];
            nc += part[4 * b + 3];
        }
        if (nt > 0 && nc > 0) ti = dt / nt < dc / nc ? 1 : 0;
    }
    *time_inner = ti;
}

// Sorted row table: row i of the pass is row perm[i] of the arrays.  Jones modes also range-check
// the time and antenna ids here and precompute each antenna's Jones row offset
// (t * antennas + ant) * facets * spw_count * chan; the gridding kernel raises the error for a
// row with bad ids when one of its samples reaches the Jones correction, as it would have
// reading the ids itself.
__global__ void k_row_table(GridArgs a, uint32_t rows, int jones, uint64_t fsc, RowRec *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const uint32_t row = a.perm[i];
    RowRec r;
    r.u = (double)a.uvw[3 * (size_t)row];
    r.v = (double)a.uvw[3 * (size_t)row + 1];
    r.w = (double)a.uvw[3 * (size_t)row + 2];
    r.row = row;
    r.spw = a.spw[row];
    r.state = (!a.row_flags[row] && a.field[row] == a.imaging_field) ? ROW_USABLE : 0u;
    r.pad = 0;
    r.j1 = r.j2 = 0;
    if (jones) {
        const uint64_t t = a.time_ids[row];
        const uint32_t x = a.ant1[row], y = a.ant2[row];
        if (t >= a.ntime || x >= a.antenna_count || y >= a.antenna_count) {
            r.state |= ROW_BAD_IDS;
        } else {
            r.j1 = (t * a.antenna_count + x) * fsc;
            r.j2 = (t * a.antenna_count + y) * fsc;
        }
    }
    out[i] = r;
}

__global__ void k_chan_table(GridArgs a, uint32_t n, ChanRec *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ChanRec c;
    c.rw = 1.0 / (double)a.wavelengths[i];
    const uint64_t g = a.cgi[i];
    c.cgi = g > 0xffffffffull ? 0xffffffffu : (uint32_t)g;
    c.enabled = a.enabled[i] ? 1u : 0u;
    out[i] = c;
}

hipError_t build_row_tables(const GridArgs &a, uint32_t rows, bool jones, RowRec *rowrec,
                            ChanRec *chanrec, hipStream_t st) {
    const uint32_t nfc = a.spw_count * a.chan;
    if (nfc && chanrec) k_chan_table<<<(nfc + 255) / 256, 256, 0, st>>>(a, nfc, chanrec);
    if (rows) {
        const uint64_t fsc = (uint64_t)a.facets * a.spw_count * a.chan;
        k_row_table<<<(rows + 255) / 256, 256, 0, st>>>(a, rows, jones ? 1 : 0, fsc, rowrec);
    }
    return hipGetLastError();
}

// Per-facet geometry in FP64 on the device, from the facet centres (arcsec, as given):
// baseline_transform_policies.h:120-143 (planar uvw transform) and phase_transform_policies.h:
// 123-133 (phase offset); arcsec are converted with the reference's float constant
// (cu_common.h:63).  grid_* modes use the phase centre: every term stays zero.
__global__ void k_facet_geom(const uvw_base_type *fc, double ra0, double dec0, uint32_t F,
                             int facet, int phase, FacetGeom *out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    constexpr double A2R = (double)(float)(3.14159265358979323846 / (180.0 * 3600.0));
    FacetGeom g = {};
    if (facet) {
        const double ra = (double)fc[2 * f], dec = (double)fc[2 * f + 1];
        const double d_ra = (ra - ra0) * A2R;
        const double n_dec = dec * A2R, o_dec = dec0 * A2R;
        const double li0 = cos(n_dec) * sin(d_ra);
        const double mi0 = sin(n_dec) * cos(o_dec) - cos(n_dec) * sin(o_dec) * cos(d_ra);
        const double ni0 = sin(n_dec) * sin(o_dec) + cos(n_dec) * cos(o_dec) * cos(d_ra);
        g.u_term = li0 / ni0;
        g.v_term = mi0 / ni0;
        if (phase) {
            g.dl = -cos(n_dec) * sin(d_ra);
            g.dm = -(sin(n_dec) * cos(o_dec) - cos(n_dec) * sin(o_dec) * cos(d_ra));
            g.dn = 1 - sqrt(1 - g.dl * g.dl - g.dm * g.dm);
        }
    }
    out[f] = g;
}

hipError_t launch_facet_geom(const uvw_base_type *fc, double ra0, double dec0, uint32_t F,
                             bool facet, bool phase, FacetGeom *out, hipStream_t st) {
    if (!F) return hipSuccess;
    if (facet && !fc) return hipErrorInvalidValue;
    k_facet_geom<<<(F + 63) / 64, 64, 0, st>>>(fc, ra0, dec0, F, facet ? 1 : 0, phase ? 1 : 0,
                                               out);
    return hipGetLastError();
}

// A kernel, not hipMemsetAsync, so that it can be captured into a graph and replayed.
hipError_t fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_u32<<<(n + 255) / 256, 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

size_t walk_order_bytes() { return 256 + 4 * WALK_BLOCKS * sizeof(double); }

hipError_t launch_walk_order(const GridArgs &a, int override_order, int *time_inner,
                             hipStream_t st) {
    double *part = reinterpret_cast<double *>(reinterpret_cast<char *>(time_inner) + 256);
    if (override_order != 1 && override_order != 2)
        k_walk_partial<<<WALK_BLOCKS, WALK_THREADS, 0, st>>>(a, part);
    k_walk_decide<<<1, 64, 0, st>>>(part, override_order, time_inner);
    return hipGetLastError();
}

}  // namespace bam
