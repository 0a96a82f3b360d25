// wproj.hip -- Romein gridding with w-projection kernels (2D, complex, support varying by plane).
//
// Semantics: bullseye's CPU w-projection policy (cpu_algorithm/convolution_policies.h:231-289):
// for w < 0 the visibility is conjugated and (u, v, w) negated; the tap origin is
// lrint(u + nx/2 - S) (round to nearest even), the fraction index size_t((1 - u + lrint(u)) * OS),
// the plane lrint(|w| / (float)wmax * (wplanes - 1)); a sample beyond the last plane or within
// 2S+3 of the high grid edge is dropped; tap (j, k) of the footprint is
// conv[plane][frac_v + j*OS][frac_u + k*OS]; the normalisation adds weight * sum Re(taps).
//
// Variable support.  Each plane p carries its own half support S_p <= S (S = conv_support):
// its taps outside [S - S_p, S + S_p] are taken as zero (the kernel generator truncates them), so
// only the W_p x W_p centre of the (2S+1)^2 footprint is walked.  At initLibrary the taps are
// re-laid out into one contiguous W_p x W_p block per (plane, frac_v, frac_u).
//
// Decomposition.  A stream is (facet, baseline, channel block, time segment of SEG rows).  A
// pre-pass computes each stream's support class -- the widest plane any of its samples uses --
// and a counting sort lists the streams of each class.  One launch per class with modulus
// M = class width: lane pair (a, b) owns the grid point x = a mod M, y = b mod M of the
// stream's current footprint (Romein 2012); a sample narrower than M touches only the pairs
// inside its own W_p window.  Small classes pack several streams into a 256-lane workgroup;
// classes wider than 16 use 1024-lane workgroups with up to PAIRS residue pairs per lane.
#include <algorithm>

#include "device_common.h"

namespace bam {

namespace wp {

constexpr int SEG = 32;        // rows per stream segment
constexpr int REC = 64;        // samples prepared per stream per step

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
};

template <typename T, int NC> struct WRec {
    int x0, y0, xm, ym;     // class-footprint origin and its residues mod M
    int off, wp, cgi;       // off = S_M - S_p (first active tap), plane width W_p, grid slice
    uint32_t tb;            // first tap of the sample's W_p x W_p block
    cpx<T> vis[NC];
};

__device__ __forceinline__ int mod_m(int v, int M) {
    v = v < 0 ? v + M : v;
    return v >= M ? v - M : v;
}

// Support class of every stream: the widest plane (2 S_p + 1) any of its samples can use.
__global__ void k_stream_class(GridArgs a, const uint32_t *plane_half, uint32_t wplanes,
                               float wmax_f, uint32_t nseg_max, uint8_t *cls) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)a.baselines * a.n_chan_blocks * nseg_max;
    if (s >= n) return;
    const uint32_t seg = (uint32_t)(s % nseg_max);
    const uint64_t t = s / nseg_max;
    const uint32_t cb = (uint32_t)(t % a.n_chan_blocks);
    const uint32_t bl = (uint32_t)(t / a.n_chan_blocks);
    const uint32_t b0 = a.bl_start[bl], nb = a.bl_start[bl + 1] - b0;
    const uint32_t r_lo = seg * SEG, r_hi = min(nb, r_lo + SEG);
    const uint32_t c_lo = cb * a.chan_per_stream, c_hi = min(a.chan, c_lo + a.chan_per_stream);
    int best = 0;
    for (uint32_t r = r_lo; r < r_hi; ++r) {
        const uint32_t row = a.perm[b0 + r];
        const uint32_t spw = a.spw[row];
        if (spw >= a.spw_count) continue;
        const double w = fabs((double)a.uvw[3 * (size_t)row + 2]);
        for (uint32_t c = c_lo; c < c_hi; ++c) {
            const double wl = w / (double)a.wavelengths[spw * a.chan + c];
            const double pl = rint(wl / (double)wmax_f * (double)(wplanes - 1));
            if (pl < (double)wplanes) best = max(best, (int)(2 * plane_half[(int)pl] + 1));
        }
    }
    cls[s] = (uint8_t)(r_lo < r_hi ? best : 0);
}

template <int MODE, int PATH, int PAIRS> struct WGridder {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    typedef typename PathTraits<PATH>::acc_t T;
    static constexpr bool KAHAN = PathTraits<PATH>::KAHAN;
    typedef WRec<T, NC> R_t;

    static __device__ __forceinline__ bool prep(const WArgs &w, uint32_t facet, uint32_t row,
                                                uint32_t c, R_t &r, double nw[NC]) {
        const GridArgs &a = w.g;
        const uint32_t s = a.spw[row];
        if (s >= a.spw_count) {
            atomicOr(a.error_flag, 4);
            return false;
        }
        const uint32_t fc = s * a.chan + c;
        if (!a.enabled[fc]) return false;
        const uint64_t cgi = a.cgi[fc];
        if (cgi >= a.cube) {
            atomicOr(a.error_flag, 4);
            return false;
        }
        if (a.row_flags[row] || a.field[row] != a.imaging_field) return false;
        const double rw = 1.0 / (double)a.wavelengths[fc];
        double u = (double)a.uvw[3 * (size_t)row] * rw;
        double v = (double)a.uvw[3 * (size_t)row + 1] * rw;
        double ww = (double)a.uvw[3 * (size_t)row + 2] * rw;
        cpx<T> vis[NC];
        T wt[NC];
        if constexpr (MODE == MODE_PSF) {
            vis[0] = {T(1), T(0)};
            wt[0] = T(1);
        } else {
            const size_t base = ((size_t)row * a.chan + c) * a.npol;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const size_t idx = base + a.pol_idx[k];
                vis[k] = {(T)a.vis[2 * idx], (T)a.vis[2 * idx + 1]};
                wt[k] = a.flags[idx] ? T(0) : (T)a.weights[idx];
            }
        }
        if constexpr (MODE == MODE_4J) {
            const size_t t = a.time_ids[row];
            if (t >= a.ntime || a.ant1[row] >= a.antenna_count || a.ant2[row] >= a.antenna_count) {
                atomicOr(a.error_flag, 4);
                return false;
            }
            const size_t j1 = ((((size_t)t * a.antenna_count + a.ant1[row]) * a.facets + facet) *
                                   a.spw_count + s) * a.chan + c;
            const size_t j2 = ((((size_t)t * a.antenna_count + a.ant2[row]) * a.facets + facet) *
                                   a.spw_count + s) * a.chan + c;
            const cpx<T> *Pi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * j1;
            const cpx<T> *Qi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * j2;
            const cpx<T> h0 = conjg(Qi[0]), h1 = conjg(Qi[2]), h2 = conjg(Qi[1]), h3 = conjg(Qi[3]);
            const cpx<T> t0 = cadd(cmul(vis[0], h0), cmul(vis[1], h2));
            const cpx<T> t1 = cadd(cmul(vis[0], h1), cmul(vis[1], h3));
            const cpx<T> t2 = cadd(cmul(vis[2], h0), cmul(vis[3], h2));
            const cpx<T> t3 = cadd(cmul(vis[2], h1), cmul(vis[3], h3));
            const cpx<T> p0 = Pi[0], p1 = Pi[1], p2 = Pi[2], p3 = Pi[3];
            vis[0] = cadd(cmul(p0, t0), cmul(p1, t2));
            vis[1] = cadd(cmul(p0, t1), cmul(p1, t3));
            vis[2] = cadd(cmul(p2, t0), cmul(p3, t2));
            vis[3] = cadd(cmul(p2, t1), cmul(p3, t3));
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            any |= (wt[k] != T(0));
            vis[k].re *= wt[k];
            vis[k].im *= wt[k];
        }
        if (!any) return false;
        const FacetGeom *gp = a.geom + facet;
        if (a.phase) {
            const double turns = u * gp->dl + v * gp->dm + ww * gp->dn;
            const double fr = 2.0 * (turns - rint(turns));
            T sn, cs;
            if constexpr (PATH == PATH_FP64) {
                double s_, c_;
                sincospi(fr, &s_, &c_);
                sn = s_;
                cs = c_;
            } else {
                float s_, c_;
                sincospif((float)fr, &s_, &c_);
                sn = s_;
                cs = c_;
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) vis[k] = cmul(vis[k], cpx<T>{cs, sn});
        }
        if (a.transform) {
            u = u - ww * gp->u_term;
            v = v - ww * gp->v_term;
        }
        u *= a.u_scale;
        v *= a.v_scale;
        if (ww < 0) {  // convolution_policies.h:248-253
#pragma unroll
            for (int k = 0; k < NC; ++k) vis[k] = conjg(vis[k]);
            u = -u;
            v = -v;
            ww = -ww;
        }
        const double ru = rint(u), rv = rint(v);
        const int dx = (int)rint(u + a.off_x), dy = (int)rint(v + a.off_y);
        const int fu = (int)((1.0 - u + ru) * a.OS);
        const int fv = (int)((1.0 - v + rv) * a.OS);
        const double pl = rint(ww / (double)w.wmax_f * (double)(w.wplanes - 1));
        const int padded = 2 * (int)a.S + 3;
        if (dx < 0 || dy < 0 || dx + padded >= (int)a.nx || dy + padded >= (int)a.ny ||
            pl >= (double)w.wplanes)
            return false;
        const int plane = (int)pl;
        const int sp = (int)w.plane_half[plane];
        const int sm = ((int)w.M - 1) / 2;
        // class footprint [x0, x0 + M) is centred where the reference's taps are
        r.x0 = dx + (int)a.S - sm;
        r.y0 = dy + (int)a.S - sm;
        r.xm = r.x0 % (int)w.M;
        r.ym = r.y0 % (int)w.M;
        r.off = sm - sp;
        const int ru_ = fu - (int)w.fbase, rv_ = fv - (int)w.fbase;
        r.cgi = (int)cgi;
        r.wp = 2 * sp + 1;
        r.tb = (uint32_t)(w.plane_off[plane] + ((uint64_t)rv_ * w.nf + ru_) * (uint64_t)(r.wp * r.wp));
#pragma unroll
        for (int k = 0; k < NC; ++k) r.vis[k] = vis[k];
        const double ts = w.plane_norm[((size_t)plane * w.nf + rv_) * w.nf + ru_];
#pragma unroll
        for (int k = 0; k < NC; ++k) nw[k] = (double)wt[k] * ts;
        return true;
    }

    static __device__ __forceinline__ void flush(const GridArgs &a, uint32_t facet, int cgi,
                                                 int x, int y, acc1<T, KAHAN> (&acc)[NC][2]) {
        bool nz = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) nz |= (acc[k][0].s != T(0)) | (acc[k][1].s != T(0));
        if (nz) {
            double *cell = a.grid + ((((size_t)facet * a.cube + cgi) * a.ny + y) * a.nx + x) *
                                        (size_t)a.pstride * 2;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                unsafeAtomicAdd(cell + 2 * k, acc[k][0].value());
                unsafeAtomicAdd(cell + 2 * k + 1, acc[k][1].value());
            }
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            acc[k][0].zero();
            acc[k][1].zero();
        }
    }

    static __device__ void run(const WArgs &w, unsigned char *smem) {
        const GridArgs &a = w.g;
        const int M = (int)w.M;
        const int L = (int)w.lanes_per_stream;
        const int slot = threadIdx.x / L;            // stream slot in the block
        const int sl = threadIdx.x - slot * L;       // lane within the stream
        const bool in_slot = slot < (int)w.streams_per_block;
        const uint64_t li = (uint64_t)blockIdx.x * w.streams_per_block + slot;
        const bool live = in_slot && li < w.n_list;
        R_t *recs = reinterpret_cast<R_t *>(smem) + slot * REC;
        __shared__ uint32_t s_steps;
        if (threadIdx.x == 0) s_steps = 0;
        __syncthreads();
        // stream geometry
        uint32_t facet = 0, b0 = 0, r_lo = 0, c_lo = 0, ncb = 1, nsamp = 0;
        if (live) {
            uint64_t t = w.stream_list[li];
            facet = (uint32_t)(t % a.facets);
            t /= a.facets;
            const uint32_t seg = (uint32_t)(t % a.n_segments);
            t /= a.n_segments;
            const uint32_t cb = (uint32_t)(t % a.n_chan_blocks);
            const uint32_t bl = (uint32_t)(t / a.n_chan_blocks);
            b0 = a.bl_start[bl];
            const uint32_t nb = a.bl_start[bl + 1] - b0;
            r_lo = seg * SEG;
            const uint32_t r_hi = min(nb, r_lo + SEG);
            c_lo = cb * a.chan_per_stream;
            ncb = min(a.chan, c_lo + a.chan_per_stream) - c_lo;
            nsamp = r_hi > r_lo ? (r_hi - r_lo) * ncb : 0;
            if (sl == 0) atomicMax(&s_steps, (nsamp + REC - 1) / REC);
        }
        __syncthreads();
        const uint32_t steps = s_steps;
        // residue pairs owned by this lane
        int pa[PAIRS], pb[PAIRS], px[PAIRS], py[PAIRS], pc[PAIRS];
        bool own[PAIRS];
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) {
            const int pair = sl + p * L;
            own[p] = live && pair < M * M;
            pa[p] = pair % M;
            pb[p] = pair / M;
            px[p] = -1;
            py[p] = -1;
            pc[p] = -1;
        }
        acc1<T, KAHAN> acc[PAIRS][NC][2];
#pragma unroll
        for (int p = 0; p < PAIRS; ++p)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                acc[p][k][0].zero();
                acc[p][k][1].zero();
            }
        int ncgi = -1;
        double nacc[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) nacc[k] = 0;
        unsigned long long cnt = 0, taps = 0;
        const cpx<T> *tapbase = reinterpret_cast<const cpx<T> *>(w.taps);

        for (uint32_t st = 0; st < steps; ++st) {
            const uint32_t step0 = st * REC;
            if (live && sl < REC) {
                R_t rec;
                rec.cgi = -1;
                const uint32_t s = step0 + sl;
                if (s < nsamp) {
                    const uint32_t tr = s / ncb;
                    const uint32_t j = s - tr * ncb;
                    const uint32_t c = c_lo + ((tr & 1) ? (ncb - 1 - j) : j);
                    const uint32_t row = a.perm[b0 + r_lo + tr];
                    double nw[NC];
                    if (prep(w, facet, row, c, rec, nw)) {
                        ++cnt;
                        taps += (unsigned long long)(rec.wp * rec.wp);
                        if constexpr (MODE != MODE_PSF) {
                            if (rec.cgi != ncgi) {
                                if (ncgi >= 0) flush_norm_lane<NC>(a, (int)(facet * a.cube + ncgi), nacc);
                                ncgi = rec.cgi;
                            }
#pragma unroll
                            for (int k = 0; k < NC; ++k) nacc[k] += nw[k];
                        }
                    } else {
                        rec.cgi = -1;
                    }
                }
                recs[sl] = rec;
            }
            __syncthreads();
            const int kmax = live ? (int)min((uint32_t)REC, nsamp > step0 ? nsamp - step0 : 0u) : 0;
            for (int kk = 0; kk < kmax; ++kk) {
                const R_t &r = recs[kk];
                if (r.cgi < 0) continue;
#pragma unroll
                for (int p = 0; p < PAIRS; ++p) {
                    if (!own[p]) continue;
                    const int ku = mod_m(pa[p] - r.xm, M);
                    const int kv = mod_m(pb[p] - r.ym, M);
                    const int x = r.x0 + ku, y = r.y0 + kv;
                    if (x != px[p] || y != py[p] || r.cgi != pc[p]) {
                        if (pc[p] >= 0) flush(a, facet, pc[p], px[p], py[p], acc[p]);
                        px[p] = x;
                        py[p] = y;
                        pc[p] = r.cgi;
                    }
                    const int tu = ku - r.off, tv = kv - r.off;
                    if ((unsigned)tu >= (unsigned)r.wp || (unsigned)tv >= (unsigned)r.wp) continue;
                    const cpx<T> tap = tapbase[r.tb + (uint32_t)(tv * r.wp + tu)];
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        const cpx<T> m = cmul(r.vis[k], tap);
                        acc[p][k][0].add(m.re);
                        acc[p][k][1].add(m.im);
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int p = 0; p < PAIRS; ++p)
            if (own[p] && pc[p] >= 0) flush(a, facet, pc[p], px[p], py[p], acc[p]);
        // every thread of the block gets here
        if constexpr (MODE != MODE_PSF)
            publish_norm<NC>(a, ncgi >= 0 ? (int)(facet * a.cube + ncgi) : -1, nacc);
        publish_counts(a, cnt, taps);
    }
};

template <int MODE, int PATH, int PAIRS>
__global__ __launch_bounds__(1024) void wproj_kernel(WArgs w) {
    extern __shared__ __align__(16) unsigned char smem[];
    WGridder<MODE, PATH, PAIRS>::run(w, smem);
}

template <int MODE, int PATH, int PAIRS>
static hipError_t launch_t(const WArgs &w, uint32_t block, hipStream_t st) {
    typedef typename WGridder<MODE, PATH, PAIRS>::R_t R_t;
    const uint64_t blocks = (w.n_list + w.streams_per_block - 1) / w.streams_per_block;
    if (!blocks) return hipSuccess;
    const size_t smem = (size_t)w.streams_per_block * REC * sizeof(R_t);
    wproj_kernel<MODE, PATH, PAIRS><<<dim3((uint32_t)blocks), dim3(block), smem, st>>>(w);
    return hipGetLastError();
}

template <int MODE, int PATH>
static hipError_t launch_p(const WArgs &w, int pairs, uint32_t block, hipStream_t st) {
    switch (pairs) {
        case 1: return launch_t<MODE, PATH, 1>(w, block, st);
        case 2: return launch_t<MODE, PATH, 2>(w, block, st);
        case 4: return launch_t<MODE, PATH, 4>(w, block, st);
    }
    return hipErrorInvalidValue;
}

template <int MODE>
static hipError_t launch_m(const WArgs &w, int path, int pairs, uint32_t block, hipStream_t st) {
    switch (path) {
        case PATH_FP64: return launch_p<MODE, PATH_FP64>(w, pairs, block, st);
        case PATH_F32K: return launch_p<MODE, PATH_F32K>(w, pairs, block, st);
        case PATH_F32: return launch_p<MODE, PATH_F32>(w, pairs, block, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace wp

// Lane shape of a class of width M: (block size, lanes per stream, streams per block, pairs).
void wproj_shape(int M, uint32_t &block, uint32_t &lanes, uint32_t &spb, int &pairs) {
    const int m2 = M * M;
    block = m2 <= 256 ? 256 : 1024;
    pairs = m2 <= 1024 ? 1 : m2 <= 2048 ? 2 : 4;
    const int per = (m2 + pairs - 1) / pairs;
    // lanes per stream: whole wavefronts, at least one (REC preparers)
    lanes = (uint32_t)std::max(64, (per + 63) / 64 * 64);
    spb = block / lanes;
}

hipError_t launch_stream_class(const GridArgs &a, const uint32_t *plane_half, uint32_t wplanes,
                               float wmax_f, uint32_t nseg_max, uint8_t *cls, hipStream_t st) {
    const uint64_t n = (uint64_t)a.baselines * a.n_chan_blocks * nseg_max;
    if (!n) return hipSuccess;
    wp::k_stream_class<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, plane_half, wplanes,
                                                                     wmax_f, nseg_max, cls);
    return hipGetLastError();
}

size_t wproj_record_bytes(int mode, int path) {
    const int nc = (mode == MODE_2) ? 2 : (mode == MODE_4 || mode == MODE_4J) ? 4 : 1;
    const size_t t = path == PATH_FP64 ? 8 : 4;
    return 32 + 2 * t * nc;
}

hipError_t launch_wproj(const GridArgs &a, const void *taps, const uint64_t *plane_off,
                        const uint32_t *plane_half, const double *plane_norm,
                        const uint32_t *stream_list, uint32_t n_list, uint32_t wplanes,
                        uint32_t nf, uint32_t fbase, uint32_t M, float wmax_f, int mode, int path,
                        hipStream_t st) {
    wp::WArgs w;
    w.g = a;
    w.taps = taps;
    w.plane_off = plane_off;
    w.plane_half = plane_half;
    w.plane_norm = plane_norm;
    w.stream_list = stream_list;
    w.n_list = n_list;
    w.wplanes = wplanes;
    w.nf = nf;
    w.fbase = fbase;
    w.M = M;
    w.wmax_f = wmax_f;
    uint32_t block;
    int pairs;
    wproj_shape((int)M, block, w.lanes_per_stream, w.streams_per_block, pairs);
    switch (mode) {
        case MODE_1: return wp::launch_m<MODE_1>(w, path, pairs, block, st);
        case MODE_2: return wp::launch_m<MODE_2>(w, path, pairs, block, st);
        case MODE_4: return wp::launch_m<MODE_4>(w, path, pairs, block, st);
        case MODE_4J: return wp::launch_m<MODE_4J>(w, path, pairs, block, st);
        case MODE_PSF: return wp::launch_m<MODE_PSF>(w, path, pairs, block, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace bam
