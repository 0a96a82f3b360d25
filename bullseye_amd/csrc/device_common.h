// device_common.h -- shared device-side types for the gfx950 imaging kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bullseye_amd.h"

namespace bam {

// Compute paths of the gridding kernels (enum bullseye_compute_path without DEFAULT).
enum Path : int { PATH_FP64 = 1, PATH_F32K = 2, PATH_F32 = 3 };

// Correlation handling modes of the entry points (cpu_algorithm/wrapper.cpp:182-425).
enum CorrMode : int { MODE_1 = 1, MODE_2 = 2, MODE_4 = 4, MODE_4J = 5, MODE_PSF = 6 };

template <int MODE> struct ModeTraits { static constexpr int NCORR = 1; };
template <> struct ModeTraits<MODE_2> { static constexpr int NCORR = 2; };
template <> struct ModeTraits<MODE_4> { static constexpr int NCORR = 4; };
template <> struct ModeTraits<MODE_4J> { static constexpr int NCORR = 4; };

template <int PATH> struct PathTraits {  // FP32 paths
    typedef float acc_t;
    static constexpr bool KAHAN = (PATH == PATH_F32K);
};
template <> struct PathTraits<PATH_FP64> {
    typedef double acc_t;
    static constexpr bool KAHAN = false;
};

// Per-facet geometry computed on the host in FP64 (phase_transform_policies.h:110-134 and
// baseline_transform_policies.h:117-143).
struct FacetGeom {
    double u_term, v_term;   // planar uvw transform: u -= w*u_term, v -= w*v_term
    double dl, dm, dn;       // phase offset; rotor exp(i 2 pi (u dl + v dm + w dn))
    double pad[3];
};

// One row of a Romein pass in baseline-sorted order, gathered once per call (through the sort
// permutation) from the row arrays, so that a gridding kernel reads a row with one 64-byte load
// instead of a chain of dependent gathers.  uvw is widened exactly to double.
constexpr uint32_t ROW_USABLE = 1;    // not flagged and in the imaging field
constexpr uint32_t ROW_BAD_IDS = 2;   // time or antenna id out of range (Jones modes)
struct __align__(16) RowRec {
    double u, v, w;
    uint32_t row;          // original row index (visibility, weight and flag addressing)
    uint32_t spw;          // spectral window as given (range-checked by the kernel)
    uint32_t state;        // ROW_* bits
    uint32_t pad;
    uint64_t j1, j2;       // Jones modes: (t * antennas + ant) * facets * spw_count * chan
};

// One (spw, channel) of a pass: reciprocal wavelength (the same FP64 division a kernel would
// do per sample), grid slice (clipped to 0xffffffff) and the enabled flag.
struct __align__(16) ChanRec {
    double rw;
    uint32_t cgi;
    uint32_t enabled;
};
static_assert(sizeof(RowRec) == 64 && sizeof(ChanRec) == 16, "table record sizes");

// The whole row record in one round trip: the empty asm takes the four 16-byte words as
// operands, so the compiler cannot sink the uvw and Jones-offset loads below the caller's
// checks (which would make them a second dependent load).
__device__ __forceinline__ RowRec load_row_record(const RowRec *p) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const u4 *q = reinterpret_cast<const u4 *>(p);
    u4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    asm volatile("" : "+v"(q0), "+v"(q1), "+v"(q2), "+v"(q3));
    RowRec r;
    __builtin_memcpy(reinterpret_cast<char *>(&r), &q0, 16);
    __builtin_memcpy(reinterpret_cast<char *>(&r) + 16, &q1, 16);
    __builtin_memcpy(reinterpret_cast<char *>(&r) + 32, &q2, 16);
    __builtin_memcpy(reinterpret_cast<char *>(&r) + 48, &q3, 16);
    return r;
}
constexpr uint32_t CHAN_LDS_MAX = 1024;  // channel tables up to this size are staged in LDS

// One Romein stream: a lane group walks the rows [row_lo, row_hi) of baseline `baseline`
// (in time order) for channels [chan_lo, chan_hi) of facet `facet`, serpentine over channels.
// Streams are enumerated implicitly: index -> (facet, baseline, channel block, time segment).
struct GridArgs {
    // inputs (device pointers; element types are the ABI's)
    const visibility_base_type *vis;           // complex [row][chan][npol]
    const uvw_base_type *uvw;                  // [row][3]
    const reference_wavelengths_base_type *wavelengths;  // [spw][chan]
    const visibility_weights_base_type *weights;          // [row][chan][npol]
    const uint8_t *flags;                      // [row][chan][npol]
    const uint8_t *row_flags;                  // [row]
    const uint32_t *field;                     // [row]
    const uint32_t *spw;                       // [row]
    const uint32_t *ant1, *ant2;               // [row]
    const uint64_t *time_ids;                  // [row]
    const visibility_base_type *jones;         // complex [t][ant][facet][spw][chan][4]
    const void *jones_inv;                     // the same, inverted, in the path's type (acc_t)
    const uint8_t *enabled;                    // [spw*chan]
    const uint64_t *cgi;                       // [spw*chan] grid slice per channel
    const uint32_t *perm;                      // [row] rows sorted by baseline, time order kept
    const uint32_t *bl_start;                  // [baseline_count + 1]
    const FacetGeom *geom;                     // [facet]
    const void *conv;                          // AA taps (acc_t) [csz]
    const double *tapsum;                      // [OS+1] sum of the 2S+1 taps per fraction
    const int *time_inner;                     // [1] walk order chosen on the device
    const RowRec *rowrec;                      // [row] sorted rows (kernels that use them)
    const ChanRec *chanrec;                    // [spw*chan]
    // outputs
    double *grid;                              // [facet][cube][ny][nx][pstride][2]
    double *norm;                              // [nslots][facet*cube*ncorr]
    unsigned long long *counters;              // [nslots][2]: samples, taps
    int *error_flag;   // bit 0: singular Jones, 1: antenna id, 2: grid slice / spw / time id
    // sizes
    uint32_t nx, ny, S, W, OS, csz;
    uint32_t chan, npol, spw_count, antenna_count, facets, cube, pstride, ntime;
    uint32_t pol_idx[4];
    uint32_t imaging_field;
    uint32_t baselines, chan_per_stream, n_chan_blocks, n_segments;
    uint64_t n_streams;
    uint64_t norm_stride;                      // facet*cube*ncorr (elements of one slot)
    uint32_t nslots;                           // copies of norm and counters (>= 1)
    double u_scale, v_scale, off_x, off_y;
    int phase, transform;
    uint32_t chan_lds;     // chanrec is staged in LDS (spw*chan <= CHAN_LDS_MAX)
    uint32_t pol_contig;   // 4 correlations: a sample's vis/weights/flags are contiguous, aligned
};

// ---------- small complex helpers --------------------------------------------------------
template <typename T> struct cpx { T re, im; };

template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T> __device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return {a.re, -a.im}; }
// a * conj(b)
template <typename T> __device__ __forceinline__ cpx<T> cmul_conj(cpx<T> a, cpx<T> b) {
    return {fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -(a.re * b.im))};
}
// acc + a * b and acc + a * conj(b), chained by FMA
template <typename T> __device__ __forceinline__ cpx<T> cmac(cpx<T> acc, cpx<T> a, cpx<T> b) {
    return {fma(a.re, b.re, fma(-a.im, b.im, acc.re)), fma(a.re, b.im, fma(a.im, b.re, acc.im))};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmac_conj(cpx<T> acc, cpx<T> a, cpx<T> b) {
    return {fma(a.re, b.re, fma(a.im, b.im, acc.re)), fma(a.im, b.re, fma(-a.re, b.im, acc.im))};
}

// Kahan-compensated running sum of one real component.  The true sum is s - c.
template <typename T, bool KAHAN> struct acc1 {
    T s, c;
    __device__ __forceinline__ void zero() { s = 0; c = 0; }
    __device__ __forceinline__ void add(T x) {
        if (KAHAN) {
            T y = x - c;
            T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += x;
        }
    }
    // Kahan step with a pre-formed term y = x - c (the compensation already subtracted).
    __device__ __forceinline__ void add_pre(T y) {
        T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const {
        return KAHAN ? (double)s - (double)c : (double)s;
    }
};

// Sample s of a stream of nrows rows x ncb channels -> (row offset tr, channel offset j).
// Channels inner: rows in order, channels serpentine; time inner: channels in order, rows
// serpentine.  Either way consecutive samples are neighbours in (time, frequency).
// Written with selects only: swapping the outputs through branches makes the compiler build a
// dynamically indexed stack array (scratch memory) for them.
__device__ __forceinline__ void stream_sample(uint32_t s, uint32_t nrows, uint32_t ncb,
                                              bool time_inner, uint32_t &tr, uint32_t &j) {
    const uint32_t inner = time_inner ? nrows : ncb;
    const uint32_t o = s / inner;
    const uint32_t i0 = s - o * inner;
    const uint32_t i = (o & 1) ? inner - 1 - i0 : i0;
    tr = time_inner ? i : o;
    j = time_inner ? o : i;
}

// Workgroups are dispatched to the XCDs round-robin (block b runs on XCD b % NXCD).  This
// renumbering is a bijection of [0, gridDim.x) under which consecutive logical blocks run on one
// XCD, so work items that read the same rows (the facets of one baseline segment) share that
// XCD's L2 instead of each XCD fetching them.  Results do not depend on it.
constexpr uint32_t NXCD = 8;
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t b = blockIdx.x, n = gridDim.x;
    const uint32_t x = b % NXCD, i = b / NXCD, q = n / NXCD, r = n % NXCD;
    return x * q + (x < r ? x : r) + i;
}

// ---------- publication of per-lane partial sums ---------------------------------------------
// Block b adds its normalisation sums and counters into copy (slot) b % nslots; the host folds
// the slots.  Within a wave, lanes holding the same key are summed before one lane issues the
// atomic.  Without both, every stream of a launch adds into the same few addresses and the
// atomics serialise at the cache.  Every lane of the wave must reach these calls.
__device__ __forceinline__ uint32_t block_slot(const GridArgs &a) { return blockIdx.x % a.nslots; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds v[k] of every lane into norm[slot][key * NC + k]; key < 0 means nothing to add.
template <int NC>
__device__ __forceinline__ void publish_norm(const GridArgs &a, int key, const double (&v)[NC]) {
    const int lane = threadIdx.x & 63;
    double *base = a.norm + (size_t)block_slot(a) * a.norm_stride;
    int k = key;
    for (;;) {
        const unsigned long long pend = __ballot(k >= 0);
        if (!pend) break;
        const int lead = __ffsll((unsigned long long)pend) - 1;
        const int k0 = __shfl(k, lead, 64);
        const bool mine = (k == k0);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double s = wave_sum(mine ? v[c] : 0.0);
            if (lane == lead) unsafeAtomicAdd(base + (size_t)k0 * NC + c, s);
        }
        if (mine) k = -1;
    }
}

__device__ __forceinline__ void publish_counts(const GridArgs &a, unsigned long long samples,
                                               unsigned long long taps) {
    const unsigned long long s = wave_sum(samples), t = wave_sum(taps);
    if ((threadIdx.x & 63) == 0 && s) {
        unsigned long long *c = a.counters + 2 * (size_t)block_slot(a);
        atomicAdd(c, s);
        atomicAdd(c + 1, t);
    }
}

// In-stream flush of one lane's sums (a stream's channel moved to another grid slice).
template <int NC>
__device__ __forceinline__ void flush_norm_lane(const GridArgs &a, int key, double (&v)[NC]) {
    double *base = a.norm + (size_t)block_slot(a) * a.norm_stride + (size_t)key * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        unsafeAtomicAdd(base + c, v[c]);
        v[c] = 0;
    }
}

}  // namespace bam
