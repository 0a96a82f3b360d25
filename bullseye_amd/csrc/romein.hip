// romein.hip -- Romein-strategy convolutional gridding for gfx950.
//
// Work decomposition.  A *stream* is (facet, baseline, channel block, time segment).  A lane
// group of G = W * ceil(W/R) lanes walks one stream's samples in time order, serpentine over
// the block's channels (channel order reverses every row, so consecutive samples stay close
// in uv).  Lane (a, q) owns, for R v-residues j = qR..qR+R-1, the grid points (x, y) of the
// current footprint with x = a (mod W) and y = j (mod W) (Romein 2012): it accumulates the
// taps for those points in registers and writes them to HBM (FP64 atomics) only when the
// footprint moves so that its point changes, or the grid slice changes.
//
// Per step every lane of a group prepares one sample (row/flag/weight reads, Jones,
// FP64 phase argument, uvw transform, tap selection) and publishes the result as a record in
// LDS; then all lanes of the group walk the G records.  Records never cross a wavefront
// (G <= 64), so no workgroup barrier is needed after the kernel table is staged.
//
// Semantics follow bullseye's CPU gridder (cpu_algorithm/templated_gridder.h:44-101,
// convolution_policies.h:76-117 and :231-289): AA tap choice floor(u) with fraction index
// size_t((1 - u + floor(u)) * OS), w-projection nearest, a sample is dropped when its
// footprint comes within 2S+3 of the high edge, normalisation += weight * sum(taps).
#include "device_common.h"

namespace bam {

constexpr int R = 8;          // v-residues owned per lane
constexpr int BLOCK = 256;    // threads per workgroup (4 wavefronts)

template <typename T, int NC> struct Rec {
    cpx<T> vis[NC];
    int x0, y0, xm, ym, fu, fv, cgi, plane;
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


template <int MODE, int PATH, bool WPROJ> struct Gridder {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    typedef typename PathTraits<PATH>::acc_t T;
    static constexpr bool KAHAN = PathTraits<PATH>::KAHAN;

    // Prepare one (facet, row, channel) sample.  Returns a record with cgi = -1 when the sample
    // contributes nothing (disabled channel, flagged or out-of-field row, all weights zero,
    // footprint off the grid).
    static __device__ __forceinline__ void prep(const GridArgs &a, const FacetGeom &g,
                                                uint32_t facet, uint32_t row, uint32_t c,
                                                Rec<T, NC> &r, double nw[NC], int &nval) {
        r.cgi = -1;
        nval = 0;
        const uint32_t s = a.spw[row];
        if (s >= a.spw_count) {
            atomicOr(a.error_flag, 4);
            return;
        }
        const uint32_t fc = s * a.chan + c;
        if (!a.enabled[fc]) return;
        const uint64_t cgi = a.cgi[fc];
        if (cgi >= a.cube) {
            atomicOr(a.error_flag, 4);
            return;
        }
        if (a.row_flags[row] || a.field[row] != a.imaging_field) return;
        const double rw = 1.0 / (double)a.wavelengths[fc];
        double u = (double)a.uvw[3 * (size_t)row] * rw;
        double v = (double)a.uvw[3 * (size_t)row + 1] * rw;
        double w = (double)a.uvw[3 * (size_t)row + 2] * rw;
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
        if constexpr (MODE == MODE_4J) {  // V = P^-1 V (Q^-1)^H (correlation_gridding_policies.h:523-542)
            const size_t t = a.time_ids[row];
            if (t >= a.ntime || a.ant1[row] >= a.antenna_count || a.ant2[row] >= a.antenna_count) {
                atomicOr(a.error_flag, 4);
                return;
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
            vis[0] = cadd(cmul(Pi[0], t0), cmul(Pi[1], t2));
            vis[1] = cadd(cmul(Pi[0], t1), cmul(Pi[1], t3));
            vis[2] = cadd(cmul(Pi[2], t0), cmul(Pi[3], t2));
            vis[3] = cadd(cmul(Pi[2], t1), cmul(Pi[3], t3));
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            any |= (wt[k] != T(0));
            vis[k].re *= wt[k];
            vis[k].im *= wt[k];
        }
        if (!any) return;
        if (a.phase) {  // phase_transform_policies.h:138-146, argument in turns, FP64
            const double turns = u * g.dl + v * g.dm + w * g.dn;
            const double fr = 2.0 * (turns - rint(turns));
            T sn, cs;
            if constexpr (PATH == PATH_FP64) {
                double s_, c_;
                sincospi(fr, &s_, &c_);
                sn = (T)s_;
                cs = (T)c_;
            } else {
                float s_, c_;
                sincospif((float)fr, &s_, &c_);
                sn = (T)s_;
                cs = (T)c_;
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) vis[k] = cmul(vis[k], cpx<T>{cs, sn});
        }
        if (a.transform) {  // baseline_transform_policies.h:145-148
            u = u - w * g.u_term;
            v = v - w * g.v_term;
        }
        u *= a.u_scale;
        v *= a.v_scale;
        int x0, y0, fu, fv, plane = 0;
        if constexpr (!WPROJ) {
            // lrint under FE_DOWNWARD: floor(u + off) == floor(u) + off for integral off
            const double flu = floor(u), flv = floor(v);
            x0 = (int)flu + (int)a.off_x;
            y0 = (int)flv + (int)a.off_y;
            fu = (int)((1.0 - u + flu) * a.OS);
            fv = (int)((1.0 - v + flv) * a.OS);
        } else {
            // convolution_policies.h:248-269 (FE_TONEAREST)
            if (w < 0) {
#pragma unroll
                for (int k = 0; k < NC; ++k) vis[k] = conjg(vis[k]);
                u = -u;
                v = -v;
                w = -w;
            }
            x0 = (int)rint(u + a.off_x);
            y0 = (int)rint(v + a.off_y);
            fu = (int)((1.0 - u + rint(u)) * a.OS);
            fv = (int)((1.0 - v + rint(v)) * a.OS);
            plane = 0;  // w-projection kernels are handled by a separate entry (not this path)
        }
        const int padded = 2 * (int)a.S + 3;
        if (x0 < 0 || y0 < 0 || x0 + padded >= (int)a.nx || y0 + padded >= (int)a.ny) return;
        r.x0 = x0;
        r.y0 = y0;
        r.xm = x0 % (int)a.W;
        r.ym = y0 % (int)a.W;
        r.fu = fu;
        r.fv = fv;
        r.plane = plane;
        r.cgi = (int)cgi;
#pragma unroll
        for (int k = 0; k < NC; ++k) r.vis[k] = vis[k];
        const double ts = a.tapsum[fu] * a.tapsum[fv];
#pragma unroll
        for (int k = 0; k < NC; ++k) nw[k] = (double)wt[k] * ts;
        nval = 1;
    }

    static __device__ __forceinline__ void flush(const GridArgs &a, uint32_t facet, int cgi, int x,
                                                 int y, acc1<T, KAHAN> (&acc)[NC][2]) {
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

    // Publication of the lanes' partial sums needs every lane of the wave: walk() may return
    // early for lanes without a stream, run() then publishes for all of them.
    static __device__ void run(const GridArgs &a, const T *sconv, Rec<T, NC> *srec) {
        int nkey = -1;
        double nacc[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) nacc[k] = 0;
        unsigned long long cnt = 0;
        walk(a, sconv, srec, nkey, nacc, cnt);
        if constexpr (MODE != MODE_PSF) publish_norm<NC>(a, nkey, nacc);
        publish_counts(a, cnt, cnt * (unsigned long long)(a.W * a.W));
    }

    static __device__ __forceinline__ void walk(const GridArgs &a, const T *sconv,
                                                Rec<T, NC> *srec, int &nkey, double (&nacc)[NC],
                                                unsigned long long &cnt_samples) {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int W = (int)a.W;
        const int nchunk = (W + R - 1) / R;
        const int G = W * nchunk;
        const int gpw = 64 / G;
        const int gw = lane / G;
        const int gl = lane - gw * G;
        if (gw >= gpw) return;
        const uint64_t stream =
            ((uint64_t)xcd_block() * (BLOCK / 64) + wave) * (uint64_t)gpw + (uint64_t)gw;
        if (stream >= a.n_streams) return;
        // stream -> (baseline, channel block, segment, facet); facet fastest so that the
        // facets of one baseline share visibility reads through L2
        uint64_t t = stream;
        const uint32_t facet = (uint32_t)(t % a.facets);
        t /= a.facets;
        const uint32_t seg = (uint32_t)(t % a.n_segments);
        t /= a.n_segments;
        const uint32_t cb = (uint32_t)(t % a.n_chan_blocks);
        const uint32_t bl = (uint32_t)(t / a.n_chan_blocks);
        const uint32_t b0 = a.bl_start[bl];
        const uint32_t nb = a.bl_start[bl + 1] - b0;
        const uint32_t r_lo = (uint32_t)(((uint64_t)nb * seg) / a.n_segments);
        const uint32_t r_hi = (uint32_t)(((uint64_t)nb * (seg + 1)) / a.n_segments);
        const uint32_t c_lo = cb * a.chan_per_stream;
        const uint32_t c_hi = min(a.chan, c_lo + a.chan_per_stream);
        const uint32_t ncb = c_hi - c_lo;
        const uint32_t nrows = r_hi - r_lo;
        const uint32_t nsamp = nrows * ncb;
        const bool tin = *a.time_inner != 0;
        const FacetGeom g = a.geom[facet];

        const int ua = gl % W;   // u residue of this lane
        const int q = gl / W;    // which R v-residues
        Rec<T, NC> *grp = srec + wave * 64 + gw * G;

        acc1<T, KAHAN> acc[R][NC][2];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                acc[r][k][0].zero();
                acc[r][k][1].zero();
            }
        int py[R];
#pragma unroll
        for (int r = 0; r < R; ++r) py[r] = -1;
        int pcgi = -1, px = -1;
        int ncgi = -1;

        for (uint32_t step0 = 0; step0 < nsamp; step0 += G) {
            // ---- prepare one sample per lane and publish it -------------------------------
            const uint32_t s = step0 + gl;
            Rec<T, NC> rec;
            rec.cgi = -1;
            if (s < nsamp) {
                uint32_t tr, j;
                stream_sample(s, nrows, ncb, tin, tr, j);
                const uint32_t c = c_lo + j;
                const uint32_t row = a.perm[b0 + r_lo + tr];
                double nw[NC];
                int nval;
                prep(a, g, facet, row, c, rec, nw, nval);
                if (nval && MODE != MODE_PSF) {
                    if (rec.cgi != ncgi) {
                        if (ncgi >= 0) flush_norm_lane<NC>(a, (int)(facet * a.cube + ncgi), nacc);
                        ncgi = rec.cgi;
                    }
#pragma unroll
                    for (int k = 0; k < NC; ++k) nacc[k] += nw[k];
                }
                cnt_samples += nval;
            }
            grp[gl] = rec;
            wave_sync();
            // ---- walk the group's records ------------------------------------------------
            const int kmax = (int)min((uint32_t)G, nsamp - step0);
            for (int kk = 0; kk < kmax; ++kk) {
                const Rec<T, NC> rr = grp[kk];
                if (rr.cgi < 0) continue;
                int ku = ua - rr.xm;
                ku += (ku < 0) ? W : 0;
                const int x = rr.x0 + ku;
                const T cu = sconv[rr.fu + ku * (int)a.OS];
                if (rr.cgi != pcgi || x != px) {
                    if (pcgi >= 0) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (py[r] >= 0) flush(a, facet, pcgi, px, py[r], acc[r]);
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) py[r] = -1;
                    pcgi = rr.cgi;
                    px = x;
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int jj = q * R + r;
                    if (jj < W) {
                        int kv = jj - rr.ym;
                        kv += (kv < 0) ? W : 0;
                        const int y = rr.y0 + kv;
                        if (y != py[r]) {
                            if (py[r] >= 0) flush(a, facet, pcgi, px, py[r], acc[r]);
                            py[r] = y;
                        }
                        const T wgt = cu * sconv[rr.fv + kv * (int)a.OS];
#pragma unroll
                        for (int k = 0; k < NC; ++k) {
                            acc[r][k][0].add(rr.vis[k].re * wgt);
                            acc[r][k][1].add(rr.vis[k].im * wgt);
                        }
                    }
                }
            }
            wave_sync();
        }
        // ---- final flush ------------------------------------------------------------------
        if (pcgi >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (py[r] >= 0) flush(a, facet, pcgi, px, py[r], acc[r]);
        }
        nkey = ncgi >= 0 ? (int)(facet * a.cube + ncgi) : -1;
    }
};

template <int MODE, int PATH>
__global__ __launch_bounds__(BLOCK) void romein_aa_kernel(GridArgs a) {
    typedef typename PathTraits<PATH>::acc_t T;
    extern __shared__ __align__(16) unsigned char smem[];
    T *sconv = reinterpret_cast<T *>(smem);
    const uint32_t conv_bytes = (a.csz * sizeof(T) + 15) & ~15u;
    auto *srec = reinterpret_cast<Rec<T, ModeTraits<MODE>::NCORR> *>(smem + conv_bytes);
    const T *gconv = reinterpret_cast<const T *>(a.conv);
    for (uint32_t i = threadIdx.x; i < a.csz; i += BLOCK) sconv[i] = gconv[i];
    __syncthreads();
    Gridder<MODE, PATH, false>::run(a, sconv, srec);
}

// ---- host-side launch ---------------------------------------------------------------------
template <int MODE, int PATH> static size_t smem_bytes(const GridArgs &a) {
    typedef typename PathTraits<PATH>::acc_t T;
    return ((a.csz * sizeof(T) + 15) & ~15u) + BLOCK * sizeof(Rec<T, ModeTraits<MODE>::NCORR>);
}

template <int MODE, int PATH>
static hipError_t launch_t(const GridArgs &a, hipStream_t st) {
    const int W = (int)a.W;
    const int G = W * ((W + R - 1) / R);
    const int gpw = 64 / G;
    const uint64_t waves = (a.n_streams + gpw - 1) / gpw;
    const uint64_t blocks = (waves + BLOCK / 64 - 1) / (BLOCK / 64);
    if (blocks == 0) return hipSuccess;
    const size_t smem = smem_bytes<MODE, PATH>(a);
    romein_aa_kernel<MODE, PATH><<<dim3((uint32_t)blocks), dim3(BLOCK), smem, st>>>(a);
    return hipGetLastError();
}

template <int MODE> static hipError_t launch_m(const GridArgs &a, int path, hipStream_t st) {
    switch (path) {
        case PATH_FP64: return launch_t<MODE, PATH_FP64>(a, st);
        case PATH_F32K: return launch_t<MODE, PATH_F32K>(a, st);
        case PATH_F32: return launch_t<MODE, PATH_F32>(a, st);
    }
    return hipErrorInvalidValue;
}

// Largest support this kernel accepts: one lane group must fit a wavefront.
int romein_max_width() { return 21; }

size_t romein_conv_elem_bytes(int path) { return path == PATH_FP64 ? 8 : 4; }

hipError_t launch_romein_aa(const GridArgs &a, int mode, int path, hipStream_t st) {
    switch (mode) {
        case MODE_1: return launch_m<MODE_1>(a, path, st);
        case MODE_2: return launch_m<MODE_2>(a, path, st);
        case MODE_4: return launch_m<MODE_4>(a, path, st);
        case MODE_4J: return launch_m<MODE_4J>(a, path, st);
        case MODE_PSF: return launch_m<MODE_PSF>(a, path, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace bam
