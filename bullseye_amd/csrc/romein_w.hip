// romein_w.hip -- Romein gridding specialised on the kernel's full support W (3..11).
//
// Same decomposition and semantics as romein.hip, with these changes for the common case:
//  * W is a compile-time constant, so residues and ownership are constant-folded;
//  * the taps are staged in LDS *rotated to residue order*: row f of the table holds
//    conv[f + (k mod W) * OS] for k in [0, 2W), so the W taps of a sample with fraction f and
//    origin residue m, in residue order (tap j for the grid point congruent to j mod W), are
//    the contiguous run starting at f * 2W + W - m.  The lane that prepares a sample writes the
//    two run offsets into its LDS record; a walking lane reads its u-tap and the W v-taps from
//    the table, so the walk is loads + MACs;
//  * 4-correlation modes split the correlations over lanes (CPL = 1 per lane on the FP32
//    paths, 2 on the FP64 path), which cuts the accumulator registers per lane;
//  * FP32 compensated sums take one Kahan step per run of samples on an unchanged footprint:
//    the run's products are chained by FMA onto -c and the result is added with the Kahan
//    correction.  A run is at most one step of G samples long.  The leader lane of each quad
//    (QUAD consecutive samples) decides once per step whether the quad's valid samples share
//    one footprint, so the walking lanes read one flag per uniform quad; in a mixed quad each
//    sample either extends the open run (same footprint) or closes it and opens one on its
//    own footprint.  The error stays O(u) per term times the run length, independent of the
//    number of samples a register sums.
// A lane group is W * NSPLIT lanes: lane (h, a) owns u-residue a and correlations
// [h*CPL, (h+1)*CPL) of all W v-residues.  Each group's LDS records start on a different
// LDS bank so the four groups of a wavefront read without conflicts.
#include "device_common.h"

namespace bam {

namespace rw {

constexpr int BLOCK = 256;
constexpr int QUAD = 4;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// TR: the record carries the sample's residue-ordered taps (FP32 paths: the walk reads them with
// the record); otherwise it carries the runs' offsets in the tap table (FP64 path: half the
// record bytes, the taps read from the table by the walking lanes).
template <typename T, int NC, int W, bool TR> struct Rec {
    int x0, y0, cgi, qf;    // footprint; qf (first record of a quad): 1 = the quad is uniform
    int qx0, qy0, qcgi, pad;  // the uniform quad's footprint (qcgi < 0: no valid sample)
    cpx<T> vis[NC];
    T cu[W];
    T cv[W];
};
template <typename T, int NC, int W> struct Rec<T, NC, W, false> {
    int x0, y0, cgi, qf;
    int qx0, qy0, qcgi, pad;
    int tu, tv;             // first tap of the u and v runs in the rotated tap table
    cpx<T> vis[NC];
};

// Entries of the rotated tap table for oversampling OS and width W: OS + 1 fraction rows of 2W
// taps, then one row of zeros (the taps of a record that contributes nothing).
__host__ __device__ constexpr uint32_t rot_table_len(uint32_t os, int w) {
    return (os + 2) * 2u * (uint32_t)w;
}

template <int MODE, int PATH, int W> struct GridderW {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    // correlations per lane: one on the FP32 paths (fewest registers), two on the FP64 path
    // (its double-width records make per-sample LDS reads the larger cost)
    static constexpr int CPL = NC == 4 ? (PATH == PATH_FP64 ? 2 : 1) : NC;
    static constexpr int NSPLIT = NC / CPL;
    static constexpr int G = W * NSPLIT;               // lanes per group
    static constexpr int GPW = 64 / G;                 // groups per wavefront
    typedef typename PathTraits<PATH>::acc_t T;
    static constexpr bool KAHAN = PathTraits<PATH>::KAHAN;
    static constexpr bool TR = PATH != PATH_FP64;
    typedef Rec<T, NC, W, TR> R_t;
    // bytes between the record arrays of consecutive groups: 16-byte aligned and starting
    // 8 banks (of 4 bytes, 32 banks) after the previous group
    static constexpr int GRAW = (G * (int)sizeof(R_t) + 15) & ~15;
    static constexpr int GSTRIDE = GRAW + ((8 - (GRAW / 4) % 32 + 32) % 32) * 4;
    static constexpr int WAVE_BYTES = GPW * GSTRIDE;

    // groups start on a multiple of 4 lanes, so the QUAD samples of a quad are a lane quad
    static constexpr bool QUAD_LANES = (G % QUAD) == 0;

    // v of lane q of this lane's quad (DPP quad permutation, no LDS)
    static __device__ __forceinline__ int quad_lane(int v, int q) {
        switch (q) {
            case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xf, 0xf, false);
            case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xf, 0xf, false);
            case 2: return __builtin_amdgcn_mov_dpp(v, 0xaa, 0xf, 0xf, false);
        }
        return __builtin_amdgcn_mov_dpp(v, 0xff, 0xf, 0xf, false);
    }

    static __device__ __forceinline__ int mod_w(int v) {  // v in (-W, 2W)
        v = v < 0 ? v + W : v;
        return v >= W ? v - W : v;
    }

    // 16-byte vector loads of N consecutive values (the caller guarantees the alignment).
    template <typename V, int N>
    static __device__ __forceinline__ void load_vec(const V *p, V (&o)[N]) {
        static_assert((N * sizeof(V)) % 16 == 0, "whole 16-byte words");
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u4 *q = reinterpret_cast<const u4 *>(p);
#pragma unroll
        for (int i = 0; i < (int)(N * sizeof(V) / 16); ++i) {
            const u4 x = q[i];
            __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * i, &x, 16);
        }
    }

    // Prepare sample (sorted row ri, channel c); returns false when it contributes nothing.
    // The checks and their order (and so the errors raised) are those of a gather straight from
    // the row arrays: spw range, enabled channel, grid slice range, row flag and field, then the
    // Jones ids.
    static __device__ __forceinline__ bool prep(const GridArgs &a, const FacetGeom *__restrict__ gp,
                                                const T *stap, const double *sts,
                                                const ChanRec *schan, uint32_t ri,
                                                uint64_t jfacet, uint32_t c, R_t &r,
                                                double nw[NC]) {
        const RowRec rr = load_row_record(a.rowrec + ri);
        const uint32_t s = rr.spw;
        if (s >= a.spw_count) {
            atomicOr(a.error_flag, 4);
            return false;
        }
        const uint32_t fc = s * a.chan + c;
        const ChanRec ch = schan[fc];   // the channel table is staged in LDS
        if (!ch.enabled) return false;
        if (ch.cgi >= a.cube) {
            atomicOr(a.error_flag, 4);
            return false;
        }
        if (!(rr.state & ROW_USABLE)) return false;
        if constexpr (MODE == MODE_4J) {
            // raised here, before the visibility loads, so that the Jones loads issue with them
            // (nothing between this point and the Jones correction can drop the sample)
            if (rr.state & ROW_BAD_IDS) {
                atomicOr(a.error_flag, 4);
                return false;
            }
        }
        const uint32_t row = rr.row;
        const double rw = ch.rw;
        double u = rr.u * rw;
        double v = rr.v * rw;
        const double w = rr.w * rw;
        cpx<T> P[4], Q[4];
        if constexpr (MODE == MODE_4J) {
            const cpx<T> *Pi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * (rr.j1 + jfacet + fc);
            const cpx<T> *Qi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * (rr.j2 + jfacet + fc);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                P[k] = Pi[k];
                Q[k] = Qi[k];
            }
        }
        cpx<T> vis[NC];
        T wt[NC];
        if constexpr (MODE == MODE_PSF) {
            vis[0] = {T(1), T(0)};
            wt[0] = T(1);
        } else {
            const size_t base = ((size_t)row * a.chan + c) * a.npol;
            if constexpr (NC == 4) {
                if (a.pol_contig) {
                    visibility_base_type vv[8];
                    visibility_weights_base_type ww[4];
                    load_vec(a.vis + 2 * base, vv);
                    load_vec(a.weights + base, ww);
                    const uint32_t fl = *reinterpret_cast<const uint32_t *>(a.flags + base);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        vis[k] = {(T)vv[2 * k], (T)vv[2 * k + 1]};
                        wt[k] = ((fl >> (8 * k)) & 0xffu) ? T(0) : (T)ww[k];
                    }
                }
            }
            if (NC != 4 || !a.pol_contig) {
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    const size_t idx = base + a.pol_idx[k];
                    vis[k] = {(T)a.vis[2 * idx], (T)a.vis[2 * idx + 1]};
                    wt[k] = a.flags[idx] ? T(0) : (T)a.weights[idx];
                }
            }
        }
        if constexpr (MODE == MODE_4J) {  // V = P^-1 V (Q^-1)^H (correlation_gridding_policies.h:523-542)
            // (Q^-1)^H = [conj q0, conj q2; conj q1, conj q3]; products chained by FMA
            const cpx<T> t0 = cmac_conj(cmul_conj(vis[0], Q[0]), vis[1], Q[1]);
            const cpx<T> t1 = cmac_conj(cmul_conj(vis[0], Q[2]), vis[1], Q[3]);
            const cpx<T> t2 = cmac_conj(cmul_conj(vis[2], Q[0]), vis[3], Q[1]);
            const cpx<T> t3 = cmac_conj(cmul_conj(vis[2], Q[2]), vis[3], Q[3]);
            vis[0] = cmac(cmul(P[0], t0), P[1], t2);
            vis[1] = cmac(cmul(P[0], t1), P[1], t3);
            vis[2] = cmac(cmul(P[2], t0), P[3], t2);
            vis[3] = cmac(cmul(P[2], t1), P[3], t3);
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            any |= (wt[k] != T(0));
            vis[k].re *= wt[k];
            vis[k].im *= wt[k];
        }
        if (!any) return false;
        if (a.phase) {  // phase_transform_policies.h:138-146, argument in turns, FP64
            const double turns = u * gp->dl + v * gp->dm + w * gp->dn;
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
        if (a.transform) {  // baseline_transform_policies.h:145-148
            u = u - w * gp->u_term;
            v = v - w * gp->v_term;
        }
        u *= a.u_scale;
        v *= a.v_scale;
        // AA tap choice (convolution_policies.h:88-96 under FE_DOWNWARD)
        const double flu = floor(u), flv = floor(v);
        const int x0 = (int)flu + (int)a.off_x;
        const int y0 = (int)flv + (int)a.off_y;
        const int fu = (int)((1.0 - u + flu) * a.OS);
        const int fv = (int)((1.0 - v + flv) * a.OS);
        if (x0 < 0 || y0 < 0 || x0 + W + 2 >= (int)a.nx || y0 + W + 2 >= (int)a.ny) return false;
        r.x0 = x0;
        r.y0 = y0;
        r.cgi = (int)ch.cgi;
#pragma unroll
        for (int k = 0; k < NC; ++k) r.vis[k] = vis[k];
        // residue-ordered tap runs: tap j of a run is conv[f + ((j - m) mod W) * OS]
        const int tu = fu * (2 * W) + W - x0 % W;
        const int tv = fv * (2 * W) + W - y0 % W;
        if constexpr (TR) {
#pragma unroll
            for (int j = 0; j < W; ++j) {
                r.cu[j] = stap[tu + j];
                r.cv[j] = stap[tv + j];
            }
        } else {
            r.tu = tu;
            r.tv = tv;
        }
        const double ts = sts[fu] * sts[fv];
#pragma unroll
        for (int k = 0; k < NC; ++k) nw[k] = (double)wt[k] * ts;
        return true;
    }

    typedef acc1<T, KAHAN> Acc[W][CPL][2];

    // Adds one grid point's sums into the FP64 grid (skipped when they are zero) and clears them.
    static __device__ __forceinline__ void flush1(double *cell, acc1<T, KAHAN> (&acc)[CPL][2]) {
        bool nz = false;
#pragma unroll
        for (int k = 0; k < CPL; ++k) nz |= (acc[k][0].s != T(0)) | (acc[k][1].s != T(0));
        if (nz) {
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                unsafeAtomicAdd(cell + 2 * k, acc[k][0].value());
                unsafeAtomicAdd(cell + 2 * k + 1, acc[k][1].value());
            }
        }
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
            acc[k][0].zero();
            acc[k][1].zero();
        }
    }

    // Grid address of (row 0, column xo, correlation c0) of slice (facet, cgi); row y is
    // y * nx * pstride * 2 doubles further.
    static __device__ __forceinline__ double *column(const GridArgs &a, uint32_t facet, int cgi,
                                                     int xo, int c0) {
        return a.grid + ((((size_t)facet * a.cube + cgi) * a.ny * a.nx + xo) * (size_t)a.pstride +
                         c0) * 2;
    }

    // Move this lane's ownership to footprint (x0, y0, cgi): flush the points it gives up.  The
    // lane owns, for v-residue r, the row of the footprint's window [py0, py0 + W) congruent to
    // r mod W, and the column of [px0, px0 + W) congruent to its u-residue; a point stays owned
    // exactly when it lies inside the new window too.
    static __device__ __forceinline__ void move_to(const GridArgs &a, uint32_t facet, int ua,
                                                   int c0, int x0, int y0, int cgi, int &px0,
                                                   int &py0, int &pcgi, Acc &acc) {
        if (x0 == px0 && y0 == py0 && cgi == pcgi) return;
        if (pcgi >= 0) {
            const int xo = px0 + mod_w(ua - px0 % W);
            const bool all = (cgi != pcgi) || ((unsigned)(xo - x0) >= (unsigned)W);
            const int pym = py0 % W;
            const int ybase = py0 - pym;
            double *col = column(a, facet, pcgi, xo, c0);
            const uint32_t rs = a.nx * a.pstride * 2u;   // < 2^32: nx is limited at init
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const int yo = ybase + r + (r < pym ? W : 0);   // >= 0: footprints lie in the grid
                if (all || (unsigned)(yo - y0) >= (unsigned)W)
                    flush1(col + (size_t)(uint32_t)yo * rs, acc[r]);
            }
        }
        px0 = x0;
        py0 = y0;
        pcgi = cgi;
    }

    static __device__ __forceinline__ const T *v_taps(const R_t &rec, const T *tap) {
        if constexpr (TR)
            return rec.cv;
        else
            return tap + rec.tv;
    }
    static __device__ __forceinline__ T u_tap(const R_t &rec, const T *tap, int ua) {
        if constexpr (TR)
            return rec.cu[ua];
        else
            return tap[rec.tu + ua];
    }

    // One sample into the register sums (classic Kahan step on the compensated path).
    static __device__ __forceinline__ void add_one(const R_t &rec, const T *tap, int ua, int c0,
                                                   Acc &acc) {
        const T cu = u_tap(rec, tap, ua);
        const T *cvp = v_taps(rec, tap);
        cpx<T> vs[CPL];
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
            vs[k] = rec.vis[c0 + k];
            vs[k].re *= cu;
            vs[k].im *= cu;
        }
#pragma unroll
        for (int r = 0; r < W; ++r) {
            const T cv = cvp[r];
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                if constexpr (KAHAN) {
                    acc[r][k][0].add_pre(vs[k].re * cv - acc[r][k][0].c);
                    acc[r][k][1].add_pre(vs[k].im * cv - acc[r][k][1].c);
                } else {
                    acc[r][k][0].s = fma(vs[k].re, cv, acc[r][k][0].s);
                    acc[r][k][1].s = fma(vs[k].im, cv, acc[r][k][1].s);
                }
            }
        }
    }

    typedef T Run[W][CPL][2];

    // A run starts from -c of every register sum and ends with one Kahan step.
    static __device__ __forceinline__ void open_run(const Acc &acc, Run &y) {
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                y[r][k][0] = -acc[r][k][0].c;
                y[r][k][1] = -acc[r][k][1].c;
            }
    }

    static __device__ __forceinline__ void close_run(Acc &acc, const Run &y) {
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                acc[r][k][0].add_pre(y[r][k][0]);
                acc[r][k][1].add_pre(y[r][k][1]);
            }
    }

    // One sample on the run's footprint: its products chained by FMA onto the run.
    static __device__ __forceinline__ void run_one(const R_t &rec, const T *tap, int ua, int c0,
                                                   Run &y) {
        const T cu = u_tap(rec, tap, ua);
        const T *cvp = v_taps(rec, tap);
        cpx<T> vs[CPL];
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
            vs[k] = rec.vis[c0 + k];
            vs[k].re *= cu;
            vs[k].im *= cu;
        }
#pragma unroll
        for (int r = 0; r < W; ++r) {
            const T cv = cvp[r];
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                y[r][k][0] = fma(vs[k].re, cv, y[r][k][0]);
                y[r][k][1] = fma(vs[k].im, cv, y[r][k][1]);
            }
        }
    }

    // QUAD samples on the run's footprint.
    static __device__ __forceinline__ void run_quad(const R_t *recs, const T *tap, int ua, int c0,
                                                    Run &y) {
#pragma unroll
        for (int q = 0; q < QUAD; ++q) run_one(recs[q], tap, ua, c0, y);
    }

    // Publication of the lanes' partial sums needs every lane of the wave: walk() may return
    // early for lanes without a stream, run() then publishes for all of them.
    static __device__ void run(const GridArgs &a, const T *stap, const double *sts,
                               const ChanRec *schan, unsigned char *swave) {
        int nkey = -1;
        double nacc[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) nacc[k] = 0;
        unsigned long long cnt = 0;
        walk(a, stap, sts, schan, swave, nkey, nacc, cnt);
        if constexpr (MODE != MODE_PSF) publish_norm<NC>(a, nkey, nacc);
        publish_counts(a, cnt, cnt * (unsigned long long)(W * W));
    }

    static __device__ __forceinline__ void walk(const GridArgs &a, const T *stap,
                                                const double *sts, const ChanRec *schan,
                                                unsigned char *swave,
                                                int &nkey,
                                                double (&nacc)[NC], unsigned long long &cnt) {
        const int lane = threadIdx.x & 63;
        const int gw = lane / G;
        const int gl = lane - gw * G;
        if (gw >= GPW) return;
        const uint64_t stream =
            ((uint64_t)xcd_block() * (BLOCK / 64) + (threadIdx.x >> 6)) * (uint64_t)GPW + gw;
        if (stream >= a.n_streams) return;
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
        const FacetGeom *gp = a.geom + facet;
        const uint64_t jfacet = (uint64_t)facet * a.spw_count * a.chan;

        const int ua = gl % W;          // u residue of this lane
        const int c0 = (gl / W) * CPL;  // first correlation of this lane
        const uint32_t zero_run = (a.OS + 1) * 2u * W;   // the table's row of zeros
        R_t *grp = reinterpret_cast<R_t *>(swave + gw * GSTRIDE);

        Acc acc;
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                acc[r][k][0].zero();
                acc[r][k][1].zero();
            }
        int px0 = -1, py0 = -1, pcgi = -1;   // footprint of the accumulated samples
        int ncgi = -1;

        // this lane's sample s = step0 + gl as (outer, inner) indexes of the stream's
        // serpentine order (stream_sample), advanced by G samples a step without a division
        const uint32_t inner = max(1u, tin ? nrows : ncb);
        uint32_t so = (uint32_t)gl / inner, si = (uint32_t)gl - so * inner;
        const uint32_t dso = (uint32_t)G / inner, dsi = (uint32_t)G - dso * inner;
        for (uint32_t step0 = 0; step0 < nsamp; step0 += G) {
            const uint32_t s = step0 + gl;
            // issue priority: low while preparing (mostly waiting on loads), high while walking
            // (a dense VALU/LDS stream that the other waves' waits can then interleave with)
            __builtin_amdgcn_s_setprio(0);
            {
                R_t rec;
                rec.cgi = -1;
                if (s < nsamp) {
                    const uint32_t i = (so & 1) ? inner - 1 - si : si;
                    const uint32_t tr = tin ? i : so, j = tin ? so : i;
                    const uint32_t c = c_lo + j;
                    double nw[NC];
                    if (prep(a, gp, stap, sts, schan, b0 + r_lo + tr, jfacet, c, rec, nw)) {
                        ++cnt;
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
                if (rec.cgi < 0) {
                    // contributes nothing; zero taps so it can ride inside a QUAD run
#pragma unroll
                    for (int k = 0; k < NC; ++k) rec.vis[k] = {T(0), T(0)};
                    if constexpr (TR) {
#pragma unroll
                        for (int j = 0; j < W; ++j) {
                            rec.cu[j] = T(0);
                            rec.cv[j] = T(0);
                        }
                    } else {
                        rec.tu = rec.tv = (int)zero_run;
                    }
                }
                if constexpr (KAHAN && QUAD_LANES) {
                    // the quad's lanes are a hardware lane quad: each reads the quad's records
                    // from its neighbours' registers, and the first writes the quad's decision
                    // (below) into its own record
                    int fc = -1, fx = 0, fy = 0;
                    bool same = true;
#pragma unroll
                    for (int q = 0; q < QUAD; ++q) {
                        const int cq = quad_lane(rec.cgi, q);
                        const int xq = quad_lane(rec.x0, q), yq = quad_lane(rec.y0, q);
                        if (cq < 0) continue;
                        if (fc < 0) {
                            fc = cq;
                            fx = xq;
                            fy = yq;
                        } else {
                            same = same && cq == fc && xq == fx && yq == fy;
                        }
                    }
                    rec.qf = same ? 1 : 0;
                    rec.qx0 = fx;
                    rec.qy0 = fy;
                    rec.qcgi = fc;
                }
                grp[gl] = rec;
            }
            __builtin_amdgcn_s_setprio(3);
            wave_sync();
            const int kmax = (int)min((uint32_t)G, nsamp - step0);
            int kk = 0;
            if constexpr (KAHAN) {
                const int nq = kmax / QUAD;
                // otherwise the leader of each full quad decides, once, whether the quad's valid
                // samples share one footprint
                if (!QUAD_LANES && (gl % QUAD) == 0 && gl / QUAD < nq) {
                    int cg = -1, x0 = 0, y0 = 0;
                    bool same = true;
#pragma unroll
                    for (int q = 0; q < QUAD; ++q) {
                        const int c1 = grp[gl + q].cgi;
                        if (c1 < 0) continue;
                        const int x1 = grp[gl + q].x0, y1 = grp[gl + q].y0;
                        if (cg < 0) {
                            cg = c1;
                            x0 = x1;
                            y0 = y1;
                        } else {
                            same = same && c1 == cg && x1 == x0 && y1 == y0;
                        }
                    }
                    grp[gl].qf = same ? 1 : 0;
                    grp[gl].qx0 = x0;
                    grp[gl].qy0 = y0;
                    grp[gl].qcgi = cg;
                }
                if (!QUAD_LANES) wave_sync();
                Run y;
                bool open = false;
                for (int q = 0; q < nq; ++q) {
                    const R_t &L = grp[q * QUAD];
                    if (L.qf) {
                        if (L.qcgi < 0) continue;   // no valid sample in the quad
                        if (open && (L.qx0 != px0 || L.qy0 != py0 || L.qcgi != pcgi)) {
                            close_run(acc, y);
                            open = false;
                        }
                        if (!open) {
                            move_to(a, facet, ua, c0, L.qx0, L.qy0, L.qcgi, px0, py0, pcgi, acc);
                            open_run(acc, y);
                            open = true;
                        }
                        run_quad(grp + q * QUAD, stap, ua, c0, y);
                    } else {
                        // a mixed quad: a sample on the open run's footprint extends the run;
                        // one elsewhere closes it, moves the ownership and opens a new run
#pragma unroll
                        for (int k2 = 0; k2 < QUAD; ++k2) {
                            const R_t &rr = grp[q * QUAD + k2];
                            if (rr.cgi < 0) continue;
                            if (!open || rr.x0 != px0 || rr.y0 != py0 || rr.cgi != pcgi) {
                                if (open) close_run(acc, y);
                                move_to(a, facet, ua, c0, rr.x0, rr.y0, rr.cgi, px0, py0, pcgi,
                                        acc);
                                open_run(acc, y);
                                open = true;
                            }
                            run_one(rr, stap, ua, c0, y);
                        }
                    }
                }
                if (open) close_run(acc, y);
                kk = nq * QUAD;
            }
            for (; kk < kmax; ++kk) {
                const int cg = grp[kk].cgi;
                if (cg < 0) continue;
                move_to(a, facet, ua, c0, grp[kk].x0, grp[kk].y0, cg, px0, py0, pcgi, acc);
                add_one(grp[kk], stap, ua, c0, acc);
            }
            wave_sync();
            si += dsi;
            so += dso;
            if (si >= inner) {
                si -= inner;
                ++so;
            }
        }
        if (pcgi >= 0) {
            const int xo = px0 + mod_w(ua - px0 % W);
            const int pym = py0 % W;
            double *col = column(a, facet, pcgi, xo, c0);
            const uint32_t rs = a.nx * a.pstride * 2u;
#pragma unroll
            for (int r = 0; r < W; ++r)
                flush1(col + (size_t)(uint32_t)(py0 + mod_w(r - pym)) * rs, acc[r]);
        }
        nkey = ncgi >= 0 ? (int)(facet * a.cube + ncgi) : -1;
    }
};

template <int MODE, int PATH, int W>
__global__ __launch_bounds__(BLOCK) void romein_w_kernel(GridArgs a) {
    typedef GridderW<MODE, PATH, W> Gd;
    typedef typename Gd::T T;
    extern __shared__ __align__(16) unsigned char smem[];
    // rotated tap table: row f < OS + 1 holds conv[f + (k mod W) * OS], k < 2W; row OS + 1 zeros
    T *stap = reinterpret_cast<T *>(smem);
    const uint32_t ntap = rot_table_len(a.OS, W);
    const uint32_t conv_bytes = (ntap * sizeof(T) + 15) & ~15u;
    const T *gconv = reinterpret_cast<const T *>(a.conv);
    for (uint32_t i = threadIdx.x; i < ntap; i += BLOCK) {
        const uint32_t f = i / (2 * W), k = i - f * (2 * W);
        stap[i] = f <= a.OS ? gconv[f + (k < W ? k : k - W) * a.OS] : T(0);
    }
    // tap sums per fraction (OS + 1 doubles), then the channel table when it is staged
    double *sts = reinterpret_cast<double *>(smem + conv_bytes);
    const uint32_t ts_bytes = ((a.OS + 1) * 8 + 15) & ~15u;
    for (uint32_t i = threadIdx.x; i <= a.OS; i += BLOCK) sts[i] = a.tapsum[i];
    ChanRec *schan = reinterpret_cast<ChanRec *>(smem + conv_bytes + ts_bytes);
    const uint32_t nfc = a.spw_count * a.chan;   // <= CHAN_LDS_MAX (launch_romein_w checks)
    for (uint32_t i = threadIdx.x; i < nfc; i += BLOCK) schan[i] = a.chanrec[i];
    __syncthreads();
    Gd::run(a, stap, sts, schan,
            smem + conv_bytes + ts_bytes + ((nfc * sizeof(ChanRec) + 15) & ~15u) +
                (threadIdx.x >> 6) * Gd::WAVE_BYTES);
}

template <int MODE, int PATH, int W> static hipError_t launch_t(const GridArgs &a, hipStream_t st) {
    typedef GridderW<MODE, PATH, W> Gd;
    const uint64_t waves = (a.n_streams + Gd::GPW - 1) / Gd::GPW;
    const uint64_t blocks = (waves + BLOCK / 64 - 1) / (BLOCK / 64);
    if (blocks == 0) return hipSuccess;
    const size_t smem = ((rot_table_len(a.OS, W) * sizeof(typename Gd::T) + 15) & ~15u) +
                        (((a.OS + 1) * 8 + 15) & ~15u) +
                        (((size_t)a.spw_count * a.chan * sizeof(ChanRec) + 15) & ~(size_t)15) +
                        (BLOCK / 64) * (size_t)Gd::WAVE_BYTES;
    if (!a.rowrec || !a.chanrec) return hipErrorInvalidValue;
    romein_w_kernel<MODE, PATH, W><<<dim3((uint32_t)blocks), dim3(BLOCK), smem, st>>>(a);
    return hipGetLastError();
}

template <int MODE, int W> static hipError_t launch_p(const GridArgs &a, int path, hipStream_t st) {
    switch (path) {
        case PATH_FP64: return launch_t<MODE, PATH_FP64, W>(a, st);
        case PATH_F32K: return launch_t<MODE, PATH_F32K, W>(a, st);
        case PATH_F32: return launch_t<MODE, PATH_F32, W>(a, st);
    }
    return hipErrorInvalidValue;
}

template <int W> static hipError_t launch_m(const GridArgs &a, int mode, int path, hipStream_t st) {
    switch (mode) {
        case MODE_1: return launch_p<MODE_1, W>(a, path, st);
        case MODE_2: return launch_p<MODE_2, W>(a, path, st);
        case MODE_4: return launch_p<MODE_4, W>(a, path, st);
        case MODE_4J: return launch_p<MODE_4J, W>(a, path, st);
        case MODE_PSF: return launch_p<MODE_PSF, W>(a, path, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace rw

hipError_t launch_romein_aa(const GridArgs &a, int mode, int path, hipStream_t st);

// Lane groups per wavefront for a support width, or 0 when no specialisation exists.
int romein_w_groups_per_wave(int W, int mode, int path) {
    const int split = (mode == MODE_4 || mode == MODE_4J) ? (path == PATH_FP64 ? 2 : 4) : 1;
    switch (W) {
        case 3: case 5: case 7: case 9: case 11: return 64 / (W * split);
    }
    return 0;
}

hipError_t launch_romein_w(const GridArgs &a, int mode, int path, hipStream_t st) {
    // channel tables too large for LDS go to the generic kernel
    if (!a.chan_lds) return launch_romein_aa(a, mode, path, st);
    switch (a.W) {
        case 3: return rw::launch_m<3>(a, mode, path, st);
        case 5: return rw::launch_m<5>(a, mode, path, st);
        case 7: return rw::launch_m<7>(a, mode, path, st);
        case 9: return rw::launch_m<9>(a, mode, path, st);
        case 11: return rw::launch_m<11>(a, mode, path, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace bam
