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
// The 1024-lane workgroups (one per CU) hold an on-chip FP64 window of the grid around each
// stream's track in their spare LDS: the flushes that land in it are added there, and the
// window goes to the grid once per stream, whole grid-row runs per atomic instruction (on the
// 4096^2 w-projection workload 2.7x fewer flush requests from these kernels).
#include <algorithm>
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "wproj_args.h"

namespace bam {

namespace wp {

constexpr int SEG = 32;        // default rows per stream segment
constexpr int REC = 128;       // samples prepared per stream per step (at most the stream's lanes)

template <typename T, int NC> struct WRec {
    int x0, y0, xm, ym;     // class-footprint origin and its residues mod M
    int off, wp, cgi;       // off = S_M - S_p (first active tap), plane width W_p, grid slice
    // pair kernels: first tap of the sample's W_p x W_p block; column kernel: the tap of
    // window column 0 for v residue 0 in the class's column table
    uint32_t tb;
    int ru0;                // column kernel: u residue of the window's first column
    cpx<T> vis[NC];
};

__device__ __forceinline__ int mod_m(int v, int M) {
    v = v < 0 ? v + M : v;
    return v >= M ? v - M : v;
}

// Support class of every stream: the widest plane (2 S_p + 1) any of its samples can use.
__global__ void k_stream_class(GridArgs a, const uint32_t *plane_half, uint32_t wplanes,
                               float wmax_f, uint32_t nseg_max, uint32_t seg_rows, uint8_t *cls) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)a.baselines * a.n_chan_blocks * nseg_max;
    if (s >= n) return;
    const uint32_t seg = (uint32_t)(s % nseg_max);
    const uint64_t t = s / nseg_max;
    const uint32_t cb = (uint32_t)(t % a.n_chan_blocks);
    const uint32_t bl = (uint32_t)(t / a.n_chan_blocks);
    const uint32_t b0 = a.bl_start[bl], nb = a.bl_start[bl + 1] - b0;
    const uint32_t r_lo = seg * seg_rows, r_hi = min(nb, r_lo + seg_rows);
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

// Preparation of one sample for a class of width M: the record and the normalisation terms.
template <int MODE, int PATH> struct WPrep {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    typedef typename PathTraits<PATH>::acc_t T;
    typedef WRec<T, NC> R_t;

    // Sample (sorted row ri, channel c).  The checks and their order (and so the errors raised)
    // are those of a gather straight from the row arrays: spw range, enabled channel, grid
    // slice range, row flag and field, then the Jones ids.
    // col: the record addresses the class's column table (column kernel), else the plane's
    // W_p x W_p blocks.
    static __device__ __forceinline__ bool prep(const WArgs &w, int M, bool col, uint32_t facet,
                                                uint32_t ri, uint32_t c, R_t &r, double nw[NC]) {
        const GridArgs &a = w.g;
        const RowRec rr = load_row_record(a.rowrec + ri);
        const uint32_t s = rr.spw;
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
        if (!(rr.state & ROW_USABLE)) return false;
        const uint32_t row = rr.row;
        const double rw = 1.0 / (double)a.wavelengths[fc];
        double u = rr.u * rw;
        double v = rr.v * rw;
        double ww = rr.w * rw;
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
            if (rr.state & ROW_BAD_IDS) {
                atomicOr(a.error_flag, 4);
                return false;
            }
            const uint64_t jf = ((uint64_t)facet * a.spw_count + s) * a.chan + c;
            const cpx<T> *Pi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * (rr.j1 + jf);
            const cpx<T> *Qi = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * (rr.j2 + jf);
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
        if (2 * sp + 1 > M) {  // class lists older than the data: never walk outside the class
            atomicOr(a.error_flag, 8);
            return false;
        }
        const int sm = (M - 1) / 2;
        // class footprint [x0, x0 + M) is centred where the reference's taps are
        r.x0 = dx + (int)a.S - sm;
        r.y0 = dy + (int)a.S - sm;
        r.xm = r.x0 % M;
        r.ym = r.y0 % M;
        r.off = sm - sp;
        const int ru_ = fu - (int)w.fbase, rv_ = fv - (int)w.fbase;
        r.cgi = (int)cgi;
        r.wp = 2 * sp + 1;
        const uint64_t q = (uint64_t)rv_ * w.nf + ru_;
        if (col) {
            // window row tv of v residue r is (r - rv0) mod M: the 2M table rows of a block
            // read from row M - rv0 on are the rows of residues 0 .. M-1 in order
            const int rv0 = (r.ym + r.off) % M;
            r.ru0 = (r.xm + r.off) % M;
            r.tb = (uint32_t)(w.cplane_off[(M >> 1) * w.wplanes + plane] +
                              (q * 2 + 1) * (uint64_t)(M * M) - (uint64_t)(rv0 * M));
        } else {
            r.tb = (uint32_t)(w.plane_off[plane] + q * (uint64_t)(r.wp * r.wp));
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) r.vis[k] = vis[k];
        const double ts = w.plane_norm[((size_t)plane * w.nf + rv_) * w.nf + ru_];
#pragma unroll
        for (int k = 0; k < NC; ++k) nw[k] = (double)wt[k] * ts;
        return true;
    }
};

// BATCH: list entries name their facet (WArgs::facet_args); without it every argument comes
// from the kernel's own (uniform) arguments.
// WIN: the launch gives the stream slots an on-chip window of the grid (WArgs::win).
template <int MODE, int PATH, int PAIRS, bool BATCH, bool WIN> struct WGridder {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    static constexpr int UNR = PAIRS == 1 ? 4 : 2;   // records whose tap loads issue together
    typedef typename PathTraits<PATH>::acc_t T;
    static constexpr bool KAHAN = PathTraits<PATH>::KAHAN;
    typedef WRec<T, NC> R_t;

    // On-chip window of the grid around a stream's track.  Cell (x, y) is always owned by the
    // residue pair (x mod M, y mod M), so each window cell has one owning lane: the lane keeps
    // its cells in private LDS slots (slot (p, j, k) of lane sl at ((p * win^2 + j) * NC + k) * L
    // + sl, j = jy * win + jx the cell's index among the pair's cells of the window) and adds to
    // them without atomics, in FP64 like the grid.  The window is flushed to the grid once per
    // stream, a row of cells per instruction.
    struct Win {
        cpx<double> *lane;   // slot (0, 0, 0) of this lane; nullptr: no window
        int L, win, x0, y0, X, Y, cgi;
        float inv_m;
        // This lane's slot of correlation 0 for cell (x, y) of slice cgi, residue pair p; nullptr
        // when the cell lies outside the window.
        __device__ __forceinline__ cpx<double> *slot(int p, int c, int x, int y) const {
            const int dx = x - x0, dy = y - y0;
            if (!lane || c != cgi || (unsigned)dx >= (unsigned)X || (unsigned)dy >= (unsigned)Y)
                return nullptr;
            // dx = r + jx M with r < M: (dx + 1/2) / M lies at least 1/(2M) from an integer
            const int jx = (int)(((float)dx + 0.5f) * inv_m);
            const int jy = (int)(((float)dy + 0.5f) * inv_m);
            return lane + ((p * win * win + jy * win + jx) * NC) * L;
        }
    };

    static __device__ __forceinline__ void flush(const GridArgs &a, uint32_t facet, int cgi,
                                                 int x, int y, acc1<T, KAHAN> (&acc)[NC][2],
                                                 const Win &wn, int p) {
        bool nz = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) nz |= (acc[k][0].s != T(0)) | (acc[k][1].s != T(0));
        if (nz) {
            cpx<double> *sl = WIN ? wn.slot(p, cgi, x, y) : nullptr;
            if (sl) {
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    cpx<double> v = sl[k * wn.L];
                    v.re += acc[k][0].value();
                    v.im += acc[k][1].value();
                    sl[k * wn.L] = v;
                }
            } else {
                double *cell = a.grid + ((((size_t)facet * a.cube + cgi) * a.ny + y) * a.nx + x) *
                                            (size_t)a.pstride * 2;
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    unsafeAtomicAdd(cell + 2 * k, acc[k][0].value());
                    unsafeAtomicAdd(cell + 2 * k + 1, acc[k][1].value());
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            acc[k][0].zero();
            acc[k][1].zero();
        }
    }

    // Adds the nonzero values of the window's sub-box [rx0, rx0 + SX) x [ry0, ry0 + SY) (window
    // coordinates) into the grid: lane sl of the stream takes one double (correlation, re/im) of
    // one cell per row, R rows per pass, so an instruction adds whole runs of a grid row (one
    // request per 64-byte segment).
    static __device__ __forceinline__ void flush_window(const GridArgs &a, uint32_t facet,
                                                        const Win &wn, const cpx<double> *wslot,
                                                        int M, int sl, int rx0, int ry0, int SX,
                                                        int SY) {
        const int D = 2 * NC * SX;     // doubles per row (SX <= X and the launch keeps 2 NC X <= L)
        const int R = wn.L / D;        // rows per pass
        const int r = sl / D;
        if (r >= R) return;
        const int e = sl - r * D;
        const int cx = rx0 + e / (2 * NC), kc = e - (e / (2 * NC)) * (2 * NC);
        const int k = kc >> 1, comp = kc & 1;
        const int x = wn.x0 + cx;
        const int jx = cx / M;
        const int am = ((x % M) + M) % M;           // u residue of the column
        const int ym = ((wn.y0 % M) + M) % M;       // v residue of the window's first row
        const double *ws = reinterpret_cast<const double *>(wslot);
        int jy = (ry0 + r) / M, rm = (ry0 + r) - jy * M;   // window row ry = jy * M + rm
        for (int ry = ry0 + r; ry < ry0 + SY; ry += R) {
            int b = ym + rm;
            if (b >= M) b -= M;
            int pair = b * M + am, p = 0;           // owner: residue pair b * M + am
            while (pair >= wn.L) {
                pair -= wn.L;
                ++p;
            }
            const double v =
                ws[((size_t)((p * wn.win * wn.win + jy * wn.win + jx) * NC + k) * wn.L + pair) * 2 +
                   comp];
            if (v != 0.0)
                unsafeAtomicAdd(a.grid + ((((size_t)facet * a.cube + wn.cgi) * a.ny +
                                           (size_t)(wn.y0 + ry)) * a.nx + (size_t)x) *
                                             (size_t)a.pstride * 2 + 2 * k + comp,
                                v);
            rm += R;
            while (rm >= M) {
                rm -= M;
                ++jy;
            }
        }
    }

    // Every thread of the block runs this the same number of times (workgroup barriers).
    static __device__ void run(const WArgs &w, unsigned char *smem) {
        if (!w.queue) {
            run_item(w, xcd_block(), smem);
            return;
        }
        __shared__ uint32_t s_item;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) s_item = atomicAdd(w.queue, 1u);
            __syncthreads();
            const uint32_t item = s_item;
            if (item >= list_items(w, w.streams_per_block)) break;
            run_item(w, item, smem);
        }
    }

    static __device__ void run_item(const WArgs &w, uint64_t item, unsigned char *smem) {
        const int L = (int)w.lanes_per_stream;
        const int slot = threadIdx.x / L;            // stream slot in the block
        const int sl = threadIdx.x - slot * L;       // lane within the stream
        const bool in_slot = slot < (int)w.streams_per_block;
        const uint64_t li = item * w.streams_per_block + slot;
        bool live = in_slot && li < list_len(w);
        // width of this stream: the launch's, or its own class in a merged list
        const int M = (!BATCH && w.list_cls && live)
                          ? (int)w.list_cls[*w.list_start + (uint32_t)(li / w.g.facets)]
                          : (int)w.M;
        uint32_t stream = 0, fidx = 0;
        const WArgs *wfp = &w;
        if constexpr (BATCH) {
            // a stream slot is whole wavefronts, so the entry is uniform across the wave
            const uint64_t e = live ? w.stream_list64[li] : 0ull;
            fidx = __builtin_amdgcn_readfirstlane((uint32_t)(e >> 32));
            stream = __builtin_amdgcn_readfirstlane((uint32_t)e);
            if (stream == 0xffffffffu) stream = 0, live = false;   // padding entry
            wfp = w.facet_args + fidx;
        }
        const WArgs &wf = *wfp;
        const GridArgs &a = wf.g;
        const uint32_t nrec = min((uint32_t)REC, (uint32_t)L);   // records per step
        R_t *recs = reinterpret_cast<R_t *>(smem) + slot * nrec;
        __shared__ uint32_t s_steps;
        if (threadIdx.x == 0) s_steps = 0;
        __syncthreads();
        // stream geometry
        uint32_t facet = 0, b0 = 0, r_lo = 0, c_lo = 0, ncb = 1, nsamp = 0, nrows = 0;
        uint64_t cls_base = 0;
        const bool tin = *a.time_inner != 0;
        if (live) {
            uint32_t seg, cb, bl;
            if constexpr (BATCH) {
                uint64_t t = stream;
                facet = (uint32_t)(t % a.facets);
                t /= a.facets;
                seg = (uint32_t)(t % w.nseg);
                t /= w.nseg;
                cb = (uint32_t)(t % a.n_chan_blocks);
                bl = (uint32_t)(t / a.n_chan_blocks);
            } else {
                single_stream(w, a, li, facet, bl, cb, seg);
            }
            cls_base = ((uint64_t)bl * a.n_chan_blocks + cb) * w.nseg_base;
            b0 = a.bl_start[bl];
            const uint32_t nb = a.bl_start[bl + 1] - b0;
            r_lo = seg * w.seg_rows;
            const uint32_t r_hi = min(nb, r_lo + w.seg_rows);
            c_lo = cb * a.chan_per_stream;
            ncb = min(a.chan, c_lo + a.chan_per_stream) - c_lo;
            nrows = r_hi > r_lo ? r_hi - r_lo : 0;
            nsamp = nrows * ncb;
            if (sl == 0) atomicMax(&s_steps, (nsamp + nrec - 1) / nrec);
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
        // the stream slot's window (placed after the first step's records are prepared)
        Win wn;
        wn.lane = nullptr;
        wn.L = L;
        wn.win = WIN ? (int)w.win : 0;
        wn.X = wn.Y = wn.win * M;
        wn.x0 = wn.y0 = 0;
        wn.cgi = -1;
        wn.inv_m = 1.0f / (float)M;
        cpx<double> *wslot = nullptr;   // the slot's window area
        if (wn.win) {
            const size_t rec_bytes =
                ((size_t)w.streams_per_block * nrec * sizeof(R_t) + 15) & ~(size_t)15;
            wslot = reinterpret_cast<cpx<double> *>(smem + rec_bytes) +
                    (size_t)slot * L * PAIRS * wn.win * wn.win * NC;
            if (live) {
                // each lane clears its own slots (the previous stream's window flush, which
                // reads every lane's slots, ended before this item's first barrier)
#pragma unroll
                for (int p = 0; p < PAIRS; ++p)
                    if (own[p])
                        for (int j = 0; j < wn.win * wn.win * NC; ++j)
                            wslot[(p * wn.win * wn.win * NC + j) * L + sl] = {0.0, 0.0};
            }
        }
        int ncgi = -1;
        double nacc[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) nacc[k] = 0;
        unsigned long long cnt = 0, taps = 0;
        const cpx<T> *tapbase = reinterpret_cast<const cpx<T> *>(wf.taps);
        // normalisation key: facet slice, or the global facet of a batched entry
        const int kbase = BATCH ? (int)(fidx * w.g.cube) : (int)(facet * a.cube);

        __shared__ int s_box[16][5];   // per wavefront: its records' footprint box and slice
        // box of the footprints of every record this lane prepared (the window flush scans only
        // the part of the window inside the stream's box)
        int tlx = 0x7fffffff, thx = -0x7fffffff, tly = 0x7fffffff, thy = -0x7fffffff;
        for (uint32_t st = 0; st < steps; ++st) {
            const uint32_t step0 = st * nrec;
            int bx = 0, by = 0, bc = -1;    // this lane's record: footprint origin and slice
            if (live && sl < (int)nrec) {
                R_t rec;
                rec.cgi = -1;
                const uint32_t s = step0 + sl;
                if (s < nsamp) {
                    uint32_t tr, j;
                    stream_sample(s, nrows, ncb, tin, tr, j);
                    const uint32_t c = c_lo + j;
                    double nw[NC];
                    const bool mine =
                        !w.filter || wf.seg_cls[cls_base + (r_lo + tr) / w.base_rows] == (uint8_t)M;
                    if (mine &&
                        WPrep<MODE, PATH>::prep(wf, M, false, facet, b0 + r_lo + tr, c, rec, nw)) {
                        ++cnt;
                        taps += (unsigned long long)(rec.wp * rec.wp);
                        if constexpr (MODE != MODE_PSF) {
                            if (rec.cgi != ncgi) {
                                if (ncgi >= 0) flush_norm_lane<NC>(w.g, kbase + ncgi, nacc);
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
                bx = rec.x0;
                by = rec.y0;
                bc = rec.cgi;
                if (bc >= 0) {
                    tlx = min(tlx, bx);
                    thx = max(thx, bx);
                    tly = min(tly, by);
                    thy = max(thy, by);
                }
            }
            if (wn.win && st == 0) {
                // box of the first step's footprints (the stream's first samples in walk order),
                // reduced per wavefront
                int lx = bc >= 0 ? bx : 0x7fffffff, hx = bc >= 0 ? bx : -0x7fffffff;
                int ly = bc >= 0 ? by : 0x7fffffff, hy = bc >= 0 ? by : -0x7fffffff;
                int lc = bc >= 0 ? bc : 0x7fffffff;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    lx = min(lx, __shfl_xor(lx, o));
                    hx = max(hx, __shfl_xor(hx, o));
                    ly = min(ly, __shfl_xor(ly, o));
                    hy = max(hy, __shfl_xor(hy, o));
                    lc = min(lc, __shfl_xor(lc, o));
                }
                if ((threadIdx.x & 63) == 0) {
                    int *b = s_box[threadIdx.x >> 6];
                    b[0] = lx;
                    b[1] = hx;
                    b[2] = ly;
                    b[3] = hy;
                    b[4] = lc;
                }
            }
            __syncthreads();
            if (wn.win && st == 0 && live) {
                // the window is centred on the box of the slot's preparing wavefronts
                int lx = 0x7fffffff, hx = -0x7fffffff, ly = 0x7fffffff, hy = -0x7fffffff;
                int lc = 0x7fffffff;
                const int w0 = slot * (L >> 6);
                for (int i = 0; i < (int)(nrec >> 6); ++i) {
                    const int *b = s_box[w0 + i];
                    lx = min(lx, b[0]);
                    hx = max(hx, b[1]);
                    ly = min(ly, b[2]);
                    hy = max(hy, b[3]);
                    lc = min(lc, b[4]);
                }
                if (lc != 0x7fffffff) {
                    wn.cgi = lc;
                    wn.x0 = (lx + hx + M) / 2 - wn.X / 2;
                    wn.y0 = (ly + hy + M) / 2 - wn.Y / 2;
                    wn.lane = wslot + sl;
                }
            }
            const int kmax = live ? (int)min(nrec, nsamp > step0 ? nsamp - step0 : 0u) : 0;
            // records in chunks of UNR: the chunk's tap loads are issued together, so their
            // latency overlaps instead of one load in flight per lane at a time
            for (int kk = 0; kk < kmax; kk += UNR) {
                cpx<T> tq[UNR][PAIRS];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const R_t &r = recs[min(kk + u, kmax - 1)];
#pragma unroll
                    for (int p = 0; p < PAIRS; ++p) {
                        const int tu = mod_m(pa[p] - r.xm, M) - r.off;
                        const int tv = mod_m(pb[p] - r.ym, M) - r.off;
                        const bool ok = kk + u < kmax && r.cgi >= 0 && own[p] &&
                                        (unsigned)tu < (unsigned)r.wp &&
                                        (unsigned)tv < (unsigned)r.wp;
                        tq[u][p] = ok ? tapbase[r.tb + (uint32_t)(tv * r.wp + tu)]
                                      : cpx<T>{T(0), T(0)};
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    if (kk + u >= kmax) break;
                    const R_t &r = recs[kk + u];
                    if (r.cgi < 0) continue;
#pragma unroll
                    for (int p = 0; p < PAIRS; ++p) {
                        if (!own[p]) continue;
                        const int ku = mod_m(pa[p] - r.xm, M);
                        const int kv = mod_m(pb[p] - r.ym, M);
                        const int x = r.x0 + ku, y = r.y0 + kv;
                        if (x != px[p] || y != py[p] || r.cgi != pc[p]) {
                            if (pc[p] >= 0) flush(a, facet, pc[p], px[p], py[p], acc[p], wn, p);
                            px[p] = x;
                            py[p] = y;
                            pc[p] = r.cgi;
                        }
                        const int tu = ku - r.off, tv = kv - r.off;
                        if ((unsigned)tu >= (unsigned)r.wp || (unsigned)tv >= (unsigned)r.wp)
                            continue;
                        const cpx<T> tap = tq[u][p];
#pragma unroll
                        for (int k = 0; k < NC; ++k) {
                            const cpx<T> m = cmul(r.vis[k], tap);
                            acc[p][k][0].add(m.re);
                            acc[p][k][1].add(m.im);
                        }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int p = 0; p < PAIRS; ++p)
            if (own[p] && pc[p] >= 0) flush(a, facet, pc[p], px[p], py[p], acc[p], wn, p);
        if (wn.win) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                tlx = min(tlx, __shfl_xor(tlx, o));
                thx = max(thx, __shfl_xor(thx, o));
                tly = min(tly, __shfl_xor(tly, o));
                thy = max(thy, __shfl_xor(thy, o));
            }
            if ((threadIdx.x & 63) == 0) {
                int *b = s_box[threadIdx.x >> 6];
                b[0] = tlx;
                b[1] = thx;
                b[2] = tly;
                b[3] = thy;
            }
            __syncthreads();   // every lane's sums are in the window, the boxes in s_box
            if (live && wn.lane) {
                int lx = 0x7fffffff, hx = -0x7fffffff, ly = 0x7fffffff, hy = -0x7fffffff;
                const int w0 = slot * (L >> 6);
                for (int i = 0; i < (int)(nrec >> 6); ++i) {
                    const int *b = s_box[w0 + i];
                    lx = min(lx, b[0]);
                    hx = max(hx, b[1]);
                    ly = min(ly, b[2]);
                    hy = max(hy, b[3]);
                }
                // cells flushed into the window lie in the footprints' box, [lx, hx + M) x
                // [ly, hy + M); scan its part inside the window
                const int sx0 = max(wn.x0, lx), sx1 = min(wn.x0 + wn.X, hx + M);
                const int sy0 = max(wn.y0, ly), sy1 = min(wn.y0 + wn.Y, hy + M);
                if (sx1 > sx0 && sy1 > sy0)
                    flush_window(a, facet, wn, wslot, M, sl, sx0 - wn.x0, sy0 - wn.y0,
                                 sx1 - sx0, sy1 - sy0);
            }
        }
        // every thread of the block gets here; sums go to the launch's slots (batched facets
        // share one normalisation array, keyed by global facet)
        if constexpr (MODE != MODE_PSF) publish_norm<NC>(w.g, ncgi >= 0 ? kbase + ncgi : -1, nacc);
        publish_counts(w.g, cnt, taps);
    }
};

// ---------- column ownership for narrow classes ----------------------------------------------
// For M <= WCOL_MAX a group of M * NSPLIT lanes walks one stream inside a wavefront, the AA
// gridder's layout: lane (h, a) owns u-residue a and correlations [h*CPL, (h+1)*CPL) of all M
// v-residues.  Per sample a lane issues its M tap loads together and does M complex MACs, so
// the per-sample bookkeeping (record read, ownership move) is shared by M taps and the loads'
// latency overlaps.  FP32 compensated sums take one Kahan step per run of QUAD samples on one
// footprint (products chained by FMA onto -c).  The taps of a lane's column come from the
// class's column table at fixed strides from one address, so the loads need no per-tap
// address, window test or select.
constexpr int WCOL_BLOCK = 256;
constexpr int QUAD = 4;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int MODE, int PATH, int M, bool BATCH> struct WColGridder {
    static constexpr int NC = ModeTraits<MODE>::NCORR;
    static constexpr int CPL = NC == 4 ? (PATH == PATH_FP64 ? 2 : 1) : NC;
    static constexpr int NSPLIT = NC / CPL;
    static constexpr int G = M * NSPLIT;     // lanes per group
    static constexpr int GPW = 64 / G;       // groups per wavefront
    typedef typename PathTraits<PATH>::acc_t T;
    static constexpr bool KAHAN = PathTraits<PATH>::KAHAN;
    typedef WRec<T, NC> R_t;
    // record arrays of consecutive groups start 8 LDS banks apart
    static constexpr int GRAW = (G * (int)sizeof(R_t) + 15) & ~15;
    static constexpr int GSTRIDE = GRAW + ((8 - (GRAW / 4) % 32 + 32) % 32) * 4;
    // LDS staging of a column that leaves the footprint: M rows x NC correlations x (re, im)
    static constexpr int STAGE = 2 * M * NC;
    static constexpr int WAVE_BYTES = GPW * GSTRIDE + GPW * STAGE * 8;
    static_assert(G < 64, "a group's lanes form a 64-bit ballot slice");
    static constexpr uint64_t GMASK = (1ull << G) - 1;
    typedef acc1<T, KAHAN> Acc[M][CPL][2];

    static __device__ __forceinline__ int mod_m(int v) {  // v in (-M, 2M)
        v = v < 0 ? v + M : v;
        return v >= M ? v - M : v;
    }

    static __device__ __forceinline__ void flush1(const GridArgs &a, uint32_t facet, int cgi,
                                                  int x, int y, int c0,
                                                  acc1<T, KAHAN> (&acc)[CPL][2]) {
        bool nz = false;
#pragma unroll
        for (int k = 0; k < CPL; ++k) nz |= (acc[k][0].s != T(0)) | (acc[k][1].s != T(0));
        if (nz) {
            double *cell = a.grid + (((((size_t)facet * a.cube + cgi) * a.ny + y) * a.nx + x) *
                                         (size_t)a.pstride + c0) * 2;
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

    // Move this lane's ownership to footprint (x0, y0, cgi): flush the points it gives up.
    // Rows that leave (and every row when the slice changes or the footprint jumps by more
    // than two columns) are flushed by all lanes of the group, adjacent cells of a grid row
    // together.  The other rows of a column that leaves go through the group's LDS staging
    // area, so that the real and imaginary parts (and the correlations) of a cell are added by
    // adjacent lanes of one instruction instead of one lane issuing a request per value.
    static __device__ __forceinline__ void move_to(const GridArgs &a, uint32_t facet, int gw,
                                                   int gl, int ua, int c0, int x0, int y0,
                                                   int cgi, int &px0, int &py0, int &pcgi,
                                                   Acc &acc, double *stage) {
        if (x0 == px0 && y0 == py0 && cgi == pcgi) return;
        if (pcgi >= 0) {
            const int pxm = px0 % M, pym = py0 % M, ym = y0 % M;
            const int xo = px0 + mod_m(ua - pxm);
            const int xn = x0 + mod_m(ua - x0 % M);
            const bool every = cgi != pcgi || x0 - px0 > 2 || px0 - x0 > 2;
#pragma unroll
            for (int r = 0; r < M; ++r) {
                const int yo = py0 + mod_m(r - pym);
                const int yn = y0 + mod_m(r - ym);
                if (every || yn != yo) flush1(a, facet, pcgi, xo, yo, c0, acc[r]);
            }
            bool pend = !every && xn != xo;
            for (;;) {
                const uint64_t gm = (__ballot(pend) >> (gw * G)) & GMASK;
                if (!__ballot(gm != 0)) break;
                const int lead = gm ? __ffsll((unsigned long long)gm) - 1 : -1;
                const int lua = lead >= 0 ? lead % M : 0;
                if (pend && ua == lua) {   // the NSPLIT lanes of the group's first leaving column
#pragma unroll
                    for (int r = 0; r < M; ++r)
#pragma unroll
                        for (int k = 0; k < CPL; ++k) {
                            stage[(r * NC + c0 + k) * 2] = acc[r][k][0].value();
                            stage[(r * NC + c0 + k) * 2 + 1] = acc[r][k][1].value();
                            acc[r][k][0].zero();
                            acc[r][k][1].zero();
                        }
                    pend = false;
                }
                wave_sync();
                if (lead >= 0) {
                    const int lxo = px0 + mod_m(lua - pxm);
                    const size_t col = (((size_t)facet * a.cube + pcgi) * a.ny) * a.nx + lxo;
#pragma unroll
                    for (int i = 0; i < 2 * CPL; ++i) {
                        const int j = gl + i * G;
                        const double v = stage[j];
                        if (v != 0.0) {
                            const int yo = py0 + mod_m(j / (2 * NC) - pym);
                            unsafeAtomicAdd(a.grid + ((col + (size_t)yo * a.nx) * a.pstride) * 2 +
                                                (j % (2 * NC)),
                                            v);
                        }
                    }
                }
                wave_sync();
            }
        }
        px0 = x0;
        py0 = y0;
        pcgi = cgi;
    }

    // The lane's column of the sample's taps in v-residue order (zero outside the plane's
    // W_p x W_p window); false when the lane's column is outside the window.
    // The column's taps are M entries M apart in the class's column table, window column tu
    // from the sample's v rotation on (zeros outside the window rows); the lanes of a group
    // read adjacent entries.
    static __device__ __forceinline__ bool column(const R_t &rec, const cpx<T> *taps, int ua,
                                                  cpx<T> (&tp)[M]) {
        const int tu = ua - rec.ru0 + (ua < rec.ru0 ? M : 0);
        if (tu >= rec.wp) return false;
        const cpx<T> *col = taps + (rec.tb + (uint32_t)tu);
#pragma unroll
        for (int r = 0; r < M; ++r) tp[r] = col[r * M];
        return true;
    }

    static __device__ __forceinline__ void add_one(const R_t &rec, const cpx<T> *taps, int ua,
                                                   int c0, Acc &acc) {
        cpx<T> tp[M];
        if (!column(rec, taps, ua, tp)) return;
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                const cpx<T> v = rec.vis[c0 + k];
                const T re = fma(v.re, tp[r].re, -v.im * tp[r].im);
                const T im = fma(v.re, tp[r].im, v.im * tp[r].re);
                if constexpr (KAHAN) {
                    acc[r][k][0].add_pre(re - acc[r][k][0].c);
                    acc[r][k][1].add_pre(im - acc[r][k][1].c);
                } else {
                    acc[r][k][0].s += re;
                    acc[r][k][1].s += im;
                }
            }
    }

    static __device__ __forceinline__ void add_quad(const R_t *recs, const cpx<T> *taps, int ua,
                                                    int c0, Acc &acc) {
        T y[M][CPL][2];
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                y[r][k][0] = -acc[r][k][0].c;
                y[r][k][1] = -acc[r][k][1].c;
            }
        bool any = false;
#pragma unroll
        for (int q = 0; q < QUAD; ++q) {
            const R_t &rec = recs[q];
            cpx<T> tp[M];
            if (!column(rec, taps, ua, tp)) continue;
            any = true;
#pragma unroll
            for (int r = 0; r < M; ++r)
#pragma unroll
                for (int k = 0; k < CPL; ++k) {
                    const cpx<T> v = rec.vis[c0 + k];
                    y[r][k][0] = fma(v.re, tp[r].re, fma(-v.im, tp[r].im, y[r][k][0]));
                    y[r][k][1] = fma(v.re, tp[r].im, fma(v.im, tp[r].re, y[r][k][1]));
                }
        }
        if (!any) return;
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                acc[r][k][0].add_pre(y[r][k][0]);
                acc[r][k][1].add_pre(y[r][k][1]);
            }
    }

    // Publication needs every lane of the wave: walk() returns early for lanes without a
    // stream.  With a queue each wave takes items (GPW streams) until the list is exhausted.
    static __device__ void run(const WArgs &w, unsigned char *swave) {
        if (!w.queue) {
            item(w, (uint64_t)xcd_block() * (WCOL_BLOCK / 64) + (threadIdx.x >> 6), swave);
            return;
        }
        const int lane = threadIdx.x & 63;
        for (;;) {
            uint32_t it = 0;
            if (lane == 0) it = atomicAdd(w.queue, 1u);
            it = __shfl(it, 0, 64);
            if (it >= list_items(w, GPW)) break;
            item(w, it, swave);
        }
    }

    static __device__ __forceinline__ void item(const WArgs &w, uint64_t it, unsigned char *swave) {
        int nkey = -1;
        double nacc[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) nacc[k] = 0;
        unsigned long long cnt = 0, taps = 0;
        walk(w, it, swave, nkey, nacc, cnt, taps);
        if constexpr (MODE != MODE_PSF) publish_norm<NC>(w.g, nkey, nacc);
        publish_counts(w.g, cnt, taps);
    }

    static __device__ __forceinline__ void walk(const WArgs &w, uint64_t it, unsigned char *swave,
                                                int &nkey, double (&nacc)[NC],
                                                unsigned long long &cnt,
                                                unsigned long long &ntaps) {
        const int lane = threadIdx.x & 63;
        const int gw = lane / G;
        const int gl = lane - gw * G;
        if (gw >= GPW) return;
        const uint64_t li = it * (uint64_t)GPW + gw;
        if (li >= list_len(w)) return;
        uint32_t facet, seg, cb, bl, fidx = 0;
        const WArgs *wfp = &w;
        if constexpr (BATCH) {
            // the host lays out an item's GPW entries with one facet: read it once per wave
            fidx = __builtin_amdgcn_readfirstlane(
                (uint32_t)(w.stream_list64[it * (uint64_t)GPW] >> 32));
            const uint32_t stream = (uint32_t)w.stream_list64[li];
            if (stream == 0xffffffffu) return;   // padding
            wfp = w.facet_args + fidx;
            uint64_t t = stream;
            facet = (uint32_t)(t % wfp->g.facets);
            t /= wfp->g.facets;
            seg = (uint32_t)(t % w.nseg);
            t /= w.nseg;
            cb = (uint32_t)(t % wfp->g.n_chan_blocks);
            bl = (uint32_t)(t / wfp->g.n_chan_blocks);
        } else {
            single_stream(w, w.g, li, facet, bl, cb, seg);
        }
        const WArgs &wf = *wfp;
        const GridArgs &a = wf.g;
        const uint64_t cls_base = ((uint64_t)bl * a.n_chan_blocks + cb) * w.nseg_base;
        const uint32_t b0 = a.bl_start[bl];
        const uint32_t nb = a.bl_start[bl + 1] - b0;
        const uint32_t r_lo = seg * w.seg_rows;
        const uint32_t r_hi = min(nb, r_lo + w.seg_rows);
        const uint32_t c_lo = cb * a.chan_per_stream;
        const uint32_t ncb = min(a.chan, c_lo + a.chan_per_stream) - c_lo;
        const uint32_t nrows = r_hi > r_lo ? r_hi - r_lo : 0;
        const uint32_t nsamp = nrows * ncb;
        const bool tin = *a.time_inner != 0;
        const int kbase = BATCH ? (int)(fidx * w.g.cube) : (int)(facet * a.cube);

        const int ua = gl % M;          // u residue of this lane
        const int c0 = (gl / M) * CPL;  // first correlation of this lane
        R_t *grp = reinterpret_cast<R_t *>(swave + gw * GSTRIDE);
        double *stage = reinterpret_cast<double *>(swave + GPW * GSTRIDE) + gw * STAGE;
        const cpx<T> *taps = reinterpret_cast<const cpx<T> *>(wf.ctaps);

        Acc acc;
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                acc[r][k][0].zero();
                acc[r][k][1].zero();
            }
        int px0 = -1, py0 = -1, pcgi = -1;
        int ncgi = -1;

        for (uint32_t step0 = 0; step0 < nsamp; step0 += G) {
            const uint32_t s = step0 + gl;
            {
                R_t rec;
                rec.cgi = -1;
                if (s < nsamp) {
                    uint32_t tr, j;
                    stream_sample(s, nrows, ncb, tin, tr, j);
                    const uint32_t c = c_lo + j;
                    double nw[NC];
                    const bool mine =
                        !w.filter || wf.seg_cls[cls_base + (r_lo + tr) / w.base_rows] == (uint8_t)M;
                    if (mine &&
                        WPrep<MODE, PATH>::prep(wf, M, true, facet, b0 + r_lo + tr, c, rec, nw)) {
                        ++cnt;
                        ntaps += (unsigned long long)(rec.wp * rec.wp);
                        if constexpr (MODE != MODE_PSF) {
                            if (rec.cgi != ncgi) {
                                if (ncgi >= 0) flush_norm_lane<NC>(w.g, kbase + ncgi, nacc);
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
                    // contributes nothing: no column of it is inside its (empty) window
                    rec.x0 = rec.y0 = rec.xm = rec.ym = 0;
                    rec.off = 0;
                    rec.wp = 0;
                    rec.tb = 0;
                    rec.ru0 = 0;
#pragma unroll
                    for (int k = 0; k < NC; ++k) rec.vis[k] = {T(0), T(0)};
                }
                grp[gl] = rec;
            }
            wave_sync();
            const int kmax = (int)min((uint32_t)G, nsamp - step0);
            int kk = 0;
            if constexpr (KAHAN) {
                while (kk + QUAD <= kmax) {
                    int cg = -1, x0 = 0, y0 = 0;
                    bool same = true;
#pragma unroll
                    for (int q = 0; q < QUAD; ++q) {
                        const int c1 = grp[kk + q].cgi;
                        if (c1 < 0) continue;
                        const int x1 = grp[kk + q].x0, y1 = grp[kk + q].y0;
                        if (cg < 0) {
                            cg = c1;
                            x0 = x1;
                            y0 = y1;
                        } else {
                            same = same && c1 == cg && x1 == x0 && y1 == y0;
                        }
                    }
                    if (!same) break;
                    if (cg >= 0) {
                        move_to(a, facet, gw, gl, ua, c0, x0, y0, cg, px0, py0, pcgi, acc,
                                stage);
                        add_quad(grp + kk, taps, ua, c0, acc);
                    }
                    kk += QUAD;
                }
            }
            for (; kk < kmax; ++kk) {
                const int cg = grp[kk].cgi;
                if (cg < 0) continue;
                move_to(a, facet, gw, gl, ua, c0, grp[kk].x0, grp[kk].y0, cg, px0, py0, pcgi, acc,
                        stage);
                add_one(grp[kk], taps, ua, c0, acc);
            }
            wave_sync();
        }
        if (pcgi >= 0) {
            const int xo = px0 + mod_m(ua - px0 % M);
            const int pym = py0 % M;
#pragma unroll
            for (int r = 0; r < M; ++r) flush1(a, facet, pcgi, xo, py0 + mod_m(r - pym), c0, acc[r]);
        }
        nkey = ncgi >= 0 ? kbase + ncgi : -1;
    }
};

template <int MODE, int PATH, int M, bool BATCH>
__global__ __launch_bounds__(WCOL_BLOCK) void wcol_kernel(WArgs w) {
    extern __shared__ __align__(16) unsigned char smem[];
    typedef WColGridder<MODE, PATH, M, BATCH> Gd;
    Gd::run(w, smem + (threadIdx.x >> 6) * Gd::WAVE_BYTES);
}

// Blocks of a persistent launch: as many as fit on the device at once (occupancy), no more
// than the items need.
template <typename K>
static uint64_t persistent_blocks(K kernel, uint32_t block, size_t smem, uint64_t needed) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 1;
    }
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)block, smem) != hipSuccess ||
        per <= 0)
        per = 1;
    return std::min<uint64_t>(needed, (uint64_t)per * (uint64_t)cus);
}

template <int MODE, int PATH, int M, bool BATCH>
static hipError_t launch_col(const WArgs &w, hipStream_t st) {
    typedef WColGridder<MODE, PATH, M, BATCH> Gd;
    // device-built lists: the length is on the device, the launch persistent at full occupancy
    const uint64_t n_list = w.seg_bl ? (uint64_t)1 << 40 : w.n_list;
    const uint64_t waves = (n_list + Gd::GPW - 1) / Gd::GPW;
    uint64_t blocks = (waves + WCOL_BLOCK / 64 - 1) / (WCOL_BLOCK / 64);
    if (w.seg_bl && !w.queue) return hipErrorInvalidValue;
    if (!blocks) return hipSuccess;
    const size_t smem = (size_t)(WCOL_BLOCK / 64) * Gd::WAVE_BYTES;
    if (w.queue)
        blocks = persistent_blocks(wcol_kernel<MODE, PATH, M, BATCH>, WCOL_BLOCK, smem, blocks);
    wcol_kernel<MODE, PATH, M, BATCH><<<dim3((uint32_t)blocks), dim3(WCOL_BLOCK), smem, st>>>(w);
    return hipGetLastError();
}

template <int MODE, int PATH, bool BATCH>
static hipError_t launch_col_m(const WArgs &w, hipStream_t st) {
    switch (w.M) {
        case 1: return launch_col<MODE, PATH, 1, BATCH>(w, st);
        case 3: return launch_col<MODE, PATH, 3, BATCH>(w, st);
        case 5: return launch_col<MODE, PATH, 5, BATCH>(w, st);
        case 7: return launch_col<MODE, PATH, 7, BATCH>(w, st);
        case 9: return launch_col<MODE, PATH, 9, BATCH>(w, st);
        case 11: return launch_col<MODE, PATH, 11, BATCH>(w, st);
        case 13: return launch_col<MODE, PATH, 13, BATCH>(w, st);
        case 15: return launch_col<MODE, PATH, 15, BATCH>(w, st);
    }
    return hipErrorInvalidValue;
}

template <int MODE, int PATH, int PAIRS, bool BATCH, bool WIN>
__global__ __launch_bounds__(1024) void wproj_kernel(WArgs w) {
    extern __shared__ __align__(16) unsigned char smem[];
    WGridder<MODE, PATH, PAIRS, BATCH, WIN>::run(w, smem);
}

// LDS a workgroup may allocate on this device.
static size_t lds_per_block() {
    static int bytes = 0;
    if (!bytes) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess ||
            bytes <= 0)
            bytes = 65536;
    }
    return (size_t)bytes;
}

// Window size (cells per residue pair and axis) of a pair-kernel launch: the largest whose slots
// fit the LDS left beside the records, at least 2, with a window row's doubles no more than a
// stream's lanes (the window flush takes a row per lane set).  Only 1024-lane workgroups (one per
// CU) take a window; smaller ones keep their LDS for more resident workgroups.
static uint32_t window_cells(uint32_t block, uint32_t lanes, uint32_t spb, int pairs, int nc,
                             uint32_t m_max, size_t rec_bytes) {
    if (block != 1024) return 0;
    const size_t reserve = 2048;   // static LDS of the kernel
    const size_t avail = lds_per_block();
    if (avail <= rec_bytes + reserve) return 0;
    const size_t per_cell = (size_t)spb * lanes * pairs * nc * 16;
    uint32_t wn = (uint32_t)std::sqrt((double)((avail - rec_bytes - reserve) / per_cell));
    while (wn >= 2 && (size_t)2 * nc * wn * m_max > lanes) --wn;
    return wn >= 2 ? wn : 0;
}

template <int MODE, int PATH, int PAIRS, bool BATCH>
static hipError_t launch_t(const WArgs &w0, uint32_t block, hipStream_t st) {
    typedef typename WGridder<MODE, PATH, PAIRS, BATCH, false>::R_t R_t;
    WArgs w = w0;
    const uint64_t n_list = w.seg_bl ? (uint64_t)1 << 40 : w.n_list;
    uint64_t blocks = (n_list + w.streams_per_block - 1) / w.streams_per_block;
    if (w.seg_bl && !w.queue) return hipErrorInvalidValue;
    if (!blocks) return hipSuccess;
    const size_t rec_bytes =
        ((size_t)w.streams_per_block * std::min<uint32_t>(REC, w.lanes_per_stream) * sizeof(R_t) +
         15) & ~(size_t)15;
    const int nc = ModeTraits<MODE>::NCORR;
    // Windows pay where a stream's track crosses many cells.  Measured: a 4096^2 grid gains
    // (w-projection pass 55.8 -> 50.6 ms), while the 128^2-2048^2 facets of the batched workload
    // lose (200-facet batch 372 -> 354 Gtaps/s, one launch per facet 344 -> 316, though the pair
    // kernels' requests fell 3.3x): their tracks are short and the window's per-stream clear and
    // flush outweigh the requests saved.  So batched launches and grids under 4096^2 cells flush
    // straight to the grid, in kernels built without the window code (its registers would
    // spill there for nothing).
    const bool big = (uint64_t)w.g.nx * w.g.ny >= ((uint64_t)1 << 24);
    w.win = (!BATCH && big) ? window_cells(block, w.lanes_per_stream, w.streams_per_block, PAIRS,
                                           nc, w.M, rec_bytes)
                            : 0u;
    const size_t smem = rec_bytes + (size_t)w.streams_per_block * w.lanes_per_stream * PAIRS *
                                        w.win * w.win * nc * 16;
    if constexpr (!BATCH) {
        if (w.win) {
            if (w.queue)
                blocks = persistent_blocks(wproj_kernel<MODE, PATH, PAIRS, false, true>, block,
                                           smem, blocks);
            wproj_kernel<MODE, PATH, PAIRS, false, true>
                <<<dim3((uint32_t)blocks), dim3(block), smem, st>>>(w);
            return hipGetLastError();
        }
    }
    if (w.queue)
        blocks = persistent_blocks(wproj_kernel<MODE, PATH, PAIRS, BATCH, false>, block, smem,
                                   blocks);
    wproj_kernel<MODE, PATH, PAIRS, BATCH, false><<<dim3((uint32_t)blocks), dim3(block), smem, st>>>(w);
    return hipGetLastError();
}

template <int MODE, int PATH, bool BATCH>
static hipError_t launch_p(const WArgs &w, int pairs, uint32_t block, hipStream_t st) {
    switch (pairs) {
        case 1: return launch_t<MODE, PATH, 1, BATCH>(w, block, st);
        case 2: return launch_t<MODE, PATH, 2, BATCH>(w, block, st);
        case 4: return launch_t<MODE, PATH, 4, BATCH>(w, block, st);
    }
    return hipErrorInvalidValue;
}

template <int MODE, bool BATCH>
static hipError_t launch_m(const WArgs &w, int path, bool column, int pairs, uint32_t block,
                           hipStream_t st) {
    if (column) {
        switch (path) {
            case PATH_FP64: return launch_col_m<MODE, PATH_FP64, BATCH>(w, st);
            case PATH_F32K: return launch_col_m<MODE, PATH_F32K, BATCH>(w, st);
            case PATH_F32: return launch_col_m<MODE, PATH_F32, BATCH>(w, st);
        }
        return hipErrorInvalidValue;
    }
    switch (path) {
        case PATH_FP64: return launch_p<MODE, PATH_FP64, BATCH>(w, pairs, block, st);
        case PATH_F32K: return launch_p<MODE, PATH_F32K, BATCH>(w, pairs, block, st);
        case PATH_F32: return launch_p<MODE, PATH_F32, BATCH>(w, pairs, block, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace wp

// Lane shape of a class of width M under the residue-pair kernel: (block size, lanes per
// stream, streams per block, pairs per lane).  min_pairs raises the pairs per lane above the
// smallest that fits (fewer lanes per footprint, more streams per block).
void wproj_shape(int M, int min_pairs, uint32_t &block, uint32_t &lanes, uint32_t &spb,
                 int &pairs) {
    const int m2 = M * M;
    block = m2 <= 256 ? 256 : 1024;
    pairs = m2 <= 1024 ? 1 : m2 <= 2048 ? 2 : 4;
    pairs = std::max(pairs, min_pairs == 2 ? 2 : min_pairs >= 4 ? 4 : 1);
    const int per = (m2 + pairs - 1) / pairs;
    // lanes per stream: whole wavefronts, at least one (the first REC lanes prepare samples)
    lanes = (uint32_t)std::max(64, (per + 63) / 64 * 64);
    spb = block / lanes;
}

hipError_t launch_stream_class(const GridArgs &a, const uint32_t *plane_half, uint32_t wplanes,
                               float wmax_f, uint32_t nseg_max, uint32_t seg_rows, uint8_t *cls,
                               hipStream_t st) {
    const uint64_t n = (uint64_t)a.baselines * a.n_chan_blocks * nseg_max;
    if (!n) return hipSuccess;
    wp::k_stream_class<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, plane_half, wplanes,
                                                                     wmax_f, nseg_max, seg_rows,
                                                                     cls);
    return hipGetLastError();
}

bool wproj_shape_valid(int M, int shape) {
    if (shape == WSHAPE_COLUMN) return M <= wp::WCOL_MAX;
    if (shape == WSHAPE_DEFAULT) return true;
    if (shape < WSHAPE_PAIRS1 || shape > WSHAPE_PAIRS4) return false;
    uint32_t block, lanes, spb;
    int pairs;
    const int req = shape == WSHAPE_PAIRS1 ? 1 : shape == WSHAPE_PAIRS2 ? 2 : 4;
    wproj_shape(M, req, block, lanes, spb, pairs);
    return pairs == req;   // a width too large for the requested pairs is not a distinct shape
}

static bool use_column(int M, int shape) {
    return shape == WSHAPE_COLUMN || (shape == WSHAPE_DEFAULT && M <= wp::WCOL_MAX);
}

static int shape_pairs(int shape) {
    return shape == WSHAPE_PAIRS2 ? 2 : shape == WSHAPE_PAIRS4 ? 4 : 1;
}

uint32_t wproj_streams_per_item(int M, int mode, int path, int shape) {
    if (use_column(M, shape)) {
        const int nc = (mode == MODE_2) ? 2 : (mode == MODE_4 || mode == MODE_4J) ? 4 : 1;
        const int cpl = nc == 4 ? (path == PATH_FP64 ? 2 : 1) : nc;
        return (uint32_t)(64 / (M * (nc / cpl)));
    }
    uint32_t block, lanes, spb;
    int pairs;
    wproj_shape(M, shape_pairs(shape), block, lanes, spb, pairs);
    return spb;
}

hipError_t launch_wproj_class(const wp::WArgs &w0, int mode, int path, int shape,
                              hipStream_t st) {
    wp::WArgs w = w0;
    if (!wproj_shape_valid((int)w.M, shape) || w.M > 63 || !(w.M & 1)) return hipErrorInvalidValue;
    uint32_t block;
    int pairs;
    const bool column = use_column((int)w.M, shape);
    wproj_shape((int)w.M, shape_pairs(shape), block, w.lanes_per_stream, w.streams_per_block,
                pairs);
    switch (mode) {
        case MODE_1:
            return w.facet_args ? wp::launch_m<MODE_1, true>(w, path, column, pairs, block, st)
                                : wp::launch_m<MODE_1, false>(w, path, column, pairs, block, st);
        case MODE_2:
            return w.facet_args ? wp::launch_m<MODE_2, true>(w, path, column, pairs, block, st)
                                : wp::launch_m<MODE_2, false>(w, path, column, pairs, block, st);
        case MODE_4:
            return w.facet_args ? wp::launch_m<MODE_4, true>(w, path, column, pairs, block, st)
                                : wp::launch_m<MODE_4, false>(w, path, column, pairs, block, st);
        case MODE_4J:
            return w.facet_args ? wp::launch_m<MODE_4J, true>(w, path, column, pairs, block, st)
                                : wp::launch_m<MODE_4J, false>(w, path, column, pairs, block, st);
        case MODE_PSF:
            return w.facet_args ? wp::launch_m<MODE_PSF, true>(w, path, column, pairs, block, st)
                                : wp::launch_m<MODE_PSF, false>(w, path, column, pairs, block, st);
    }
    return hipErrorInvalidValue;
}

size_t wproj_record_bytes(int mode, int path) {
    const int nc = (mode == MODE_2) ? 2 : (mode == MODE_4 || mode == MODE_4J) ? 4 : 1;
    const size_t t = path == PATH_FP64 ? 8 : 4;
    return 32 + 2 * t * nc;
}

// ---------- device-built class lists of a single-facet-set call ------------------------------

hipError_t fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st);

namespace wp {

// Segments of each baseline (SEG rows each, the last one partial).
__global__ void k_seg_count(const uint32_t *bl_start, uint32_t B, uint32_t *cnt) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > B) return;
    cnt[b] = b < B ? (bl_start[b + 1] - bl_start[b] + SEG - 1) / SEG : 0u;
}

__global__ void k_seg_fill(const uint32_t *seg_first, uint32_t B, uint32_t *seg_bl) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    for (uint32_t s = seg_first[b]; s < seg_first[b + 1]; ++s) seg_bl[s] = b;
}

// Class of every (segment, channel block) entry: the widest plane (2 S_p + 1) any of its samples
// can use (k_stream_class for the segment table); entries beyond the segments get class 0.
__global__ void k_entry_class(GridArgs a, const uint32_t *seg_first, const uint32_t *seg_bl,
                              uint64_t entries, const uint32_t *plane_half, uint32_t wplanes,
                              float wmax_f, uint8_t *cls, uint32_t *val) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries) return;
    val[e] = (uint32_t)e;
    const uint32_t s = (uint32_t)(e / a.n_chan_blocks);
    const uint32_t cb = (uint32_t)(e % a.n_chan_blocks);
    int best = 0;
    if (s < seg_first[a.baselines]) {
        const uint32_t bl = seg_bl[s];
        const uint32_t b0 = a.bl_start[bl], nb = a.bl_start[bl + 1] - b0;
        const uint32_t r_lo = (s - seg_first[bl]) * SEG, r_hi = min(nb, r_lo + SEG);
        const uint32_t c_lo = cb * a.chan_per_stream;
        const uint32_t c_hi = min(a.chan, c_lo + a.chan_per_stream);
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
    }
    cls[e] = (uint8_t)best;
}

// First sorted entry of each class value c (0..256).
__global__ void k_class_start(const uint8_t *sorted, uint64_t entries, uint32_t *start) {
    const uint32_t c = threadIdx.x + blockIdx.x * blockDim.x;
    if (c > 256) return;
    uint64_t lo = 0, hi = entries;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((uint32_t)sorted[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[c] = (uint32_t)lo;
}

}  // namespace wp

static size_t wl_align(size_t n) { return (n + 255) & ~(size_t)255; }

WListPlan wproj_list_plan(uint32_t rows, uint32_t baselines, uint32_t n_chan_blocks) {
    WListPlan p{};
    p.nseg_max = ((uint64_t)rows + wp::SEG - 1) / wp::SEG + baselines;
    p.entries = p.nseg_max * n_chan_blocks;
    size_t scan_bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr,
                                           (uint32_t *)nullptr, (int)baselines + 1);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, p.sort_bytes, (uint8_t *)nullptr,
                                             (uint8_t *)nullptr, (uint32_t *)nullptr,
                                             (uint32_t *)nullptr, (int)p.entries, 0, 8);
    p.sort_bytes = std::max(p.sort_bytes, scan_bytes);
    size_t o = 0;
    p.off_seg_first = o;  o += wl_align(((size_t)baselines + 1) * 4);
    p.off_seg_bl = o;     o += wl_align(p.nseg_max * 4);
    p.off_cls = o;        o += wl_align(p.entries);
    p.off_cls_sorted = o; o += wl_align(p.entries);
    // (val first holds the per-baseline segment counts that seg_first is scanned from)
    p.off_val = o;        o += wl_align(std::max<uint64_t>(p.entries, (uint64_t)baselines + 1) * 4);
    p.off_val_sorted = o; o += wl_align(p.entries * 4);
    p.off_class_start = o; o += wl_align(257 * 4);
    p.off_queues = o;     o += wl_align(256 * 4);
    p.off_sort = o;       o += wl_align(p.sort_bytes);
    p.total = o;
    return p;
}

hipError_t wproj_build_lists(const GridArgs &a, const WListPlan &pl, void *ws,
                             const uint32_t *plane_half, uint32_t wplanes, float wmax_f,
                             hipStream_t st) {
    unsigned char *b = static_cast<unsigned char *>(ws);
    uint32_t *seg_first = reinterpret_cast<uint32_t *>(b + pl.off_seg_first);
    uint32_t *seg_bl = reinterpret_cast<uint32_t *>(b + pl.off_seg_bl);
    uint8_t *cls = b + pl.off_cls, *cls_sorted = b + pl.off_cls_sorted;
    uint32_t *val = reinterpret_cast<uint32_t *>(b + pl.off_val);
    uint32_t *val_sorted = reinterpret_cast<uint32_t *>(b + pl.off_val_sorted);
    uint32_t *cstart = reinterpret_cast<uint32_t *>(b + pl.off_class_start);
    uint32_t *queues = reinterpret_cast<uint32_t *>(b + pl.off_queues);
    void *tmp = b + pl.off_sort;
    const uint32_t B = a.baselines;
    // seg_first (counts, then their exclusive sum in place through val as the scan's input)
    wp::k_seg_count<<<(B + 1 + 255) / 256, 256, 0, st>>>(a.bl_start, B, val);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = pl.sort_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, val, seg_first, (int)B + 1, st);
    if (e != hipSuccess) return e;
    if (B) wp::k_seg_fill<<<(B + 255) / 256, 256, 0, st>>>(seg_first, B, seg_bl);
    if (pl.entries) {
        wp::k_entry_class<<<(unsigned)((pl.entries + 255) / 256), 256, 0, st>>>(
            a, seg_first, seg_bl, pl.entries, plane_half, wplanes, wmax_f, cls, val);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        bytes = pl.sort_bytes;
        e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, cls, cls_sorted, val, val_sorted,
                                               (int)pl.entries, 0, 8, st);
        if (e != hipSuccess) return e;
    }
    wp::k_class_start<<<2, 256, 0, st>>>(cls_sorted, pl.entries, cstart);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return fill_u32(queues, 256, 0u, st);
}

hipError_t launch_wproj_device_list(const GridArgs &a, const WListPlan &pl, void *ws,
                                    const void *taps, const uint64_t *plane_off,
                                    const void *ctaps, const uint64_t *cplane_off,
                                    const uint32_t *plane_half, const double *plane_norm,
                                    uint32_t wplanes, uint32_t nf, uint32_t fbase, uint32_t M,
                                    uint32_t M_hi, float wmax_f, int mode, int path,
                                    hipStream_t st) {
    unsigned char *b = static_cast<unsigned char *>(ws);
    if (M_hi < M || (M_hi > M && M_hi <= (uint32_t)wp::WCOL_MAX)) return hipErrorInvalidValue;
    wp::WArgs w;
    w.g = a;
    w.taps = taps;
    w.plane_off = plane_off;
    w.ctaps = ctaps;
    w.cplane_off = cplane_off;
    w.plane_half = plane_half;
    w.plane_norm = plane_norm;
    w.stream_list = reinterpret_cast<const uint32_t *>(b + pl.off_val_sorted);
    w.stream_list64 = nullptr;
    w.n_list = 0;
    w.wplanes = wplanes;
    w.nf = nf;
    w.fbase = fbase;
    w.M = M_hi;
    w.wmax_f = wmax_f;
    w.seg_rows = wp::SEG;
    w.nseg = 0;
    w.base_rows = wp::SEG;
    w.nseg_base = 0;
    w.filter = 0;
    w.seg_cls = nullptr;
    w.facet_args = nullptr;
    w.queue = reinterpret_cast<uint32_t *>(b + pl.off_queues) + M;
    w.n_items = 0;
    const uint32_t *cstart = reinterpret_cast<const uint32_t *>(b + pl.off_class_start);
    w.list_start = cstart + M;
    w.list_end = cstart + M_hi + 1;
    w.seg_bl = reinterpret_cast<const uint32_t *>(b + pl.off_seg_bl);
    w.seg_first = reinterpret_cast<const uint32_t *>(b + pl.off_seg_first);
    w.list_cls = M_hi > M ? b + pl.off_cls_sorted : nullptr;
    // classes 19..31 walk two residue pairs per lane: two to five streams per workgroup instead
    // of one or two (their windows are smaller, the workgroups' lanes busier; 4096^2
    // w-projection pass 53.6 -> 50.6 ms, against 50.8 with class 17 too and 51.8 with four
    // pairs from 25)
    const int shape = (M_hi == M && M >= 19 && M <= 31) ? WSHAPE_PAIRS2 : WSHAPE_DEFAULT;
    return launch_wproj_class(w, mode, path, shape, st);
}

}  // namespace bam
