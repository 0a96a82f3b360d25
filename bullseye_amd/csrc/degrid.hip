// degrid.hip -- the degridding adjoint of the gridding kernels.
//
// bullseye has no degridder in C++ (writing_and_wiki/notebooks/Resampling.ipynb sketches the
// gather in Python).  For the visibility modes this kernel computes the exact transpose, over
// the reals, of the map the gridding kernels apply from visibilities to the unnormalised grids,
// summed over facets.  Per (row, channel) and facet, with the forward's geometry, tap choice,
// edge drop and plane choice:
//   s_k = sum over the footprint of conj(C) * model_k      (C real for the AA kernel)
//   s_k = conj(s_k) if the w-projection sample had w < 0  (the forward conjugated it)
//   s_k *= exp(-i x) (facet phase), s_k *= weight_k (zero when flagged),
//   S = Pinv^H S Qinv in the Jones mode (the forward applies V -> Pinv V Qinv^H),
// and out[row][chan][corr_k] = sum over facets of s_k; every other correlation is zero.
//
// A gather: no atomics.  A group of GL lanes shares one (row, channel) sample and splits the
// footprint's taps; a butterfly over the group's lanes combines the partial sums.  Neighbouring
// samples of a row read overlapping grid rows, which stay in L2.
#include "device_common.h"

namespace bam {
namespace dg {

constexpr int GL = 16;      // lanes per sample
constexpr int BLOCK = 256;

struct DArgs {
    GridArgs g;                      // inputs and geometry (g.grid and g.norm unused)
    const grid_base_type *model;     // complex [facet][cube][P][ny][nx]
    visibility_base_type *out;       // complex [row][chan][npol]
    const void *taps;                // w-projection taps per (plane, frac_v, frac_u), acc_t
    const uint64_t *plane_off;
    const uint32_t *plane_half;
    uint32_t wplanes, nf, fbase, wproj;
    float wmax_f;
    uint32_t P;                      // correlation slices per (facet, cube) group
    uint64_t rows;
};

template <typename T> __device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = GL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, GL);
    return v;
}

template <int MODE, int PATH>
__global__ __launch_bounds__(BLOCK) void degrid_kernel(DArgs d) {
    constexpr int NC = ModeTraits<MODE>::NCORR;
    typedef typename PathTraits<PATH>::acc_t T;
    const GridArgs &a = d.g;
    const uint64_t sample = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / GL;
    const int gl = (int)(threadIdx.x % GL);
    const bool in_range = sample < d.rows * a.chan;
    const uint32_t row = in_range ? (uint32_t)(sample / a.chan) : 0;
    const uint32_t c = in_range ? (uint32_t)(sample % a.chan) : 0;
    const size_t base = ((size_t)row * a.chan + c) * a.npol;

    // per-sample checks, in the forward's order (every lane of the group agrees)
    bool live = in_range;
    uint32_t s = 0;
    uint64_t cgi = 0;
    if (live) {
        s = a.spw[row];
        if (s >= a.spw_count) {
            if (gl == 0) atomicOr(a.error_flag, 4);
            live = false;
        }
    }
    if (live && !a.enabled[s * a.chan + c]) live = false;
    if (live) {
        cgi = a.cgi[s * a.chan + c];
        if (cgi >= a.cube) {
            if (gl == 0) atomicOr(a.error_flag, 4);
            live = false;
        }
    }
    if (live && (a.row_flags[row] || a.field[row] != a.imaging_field)) live = false;
    T wt[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) wt[k] = T(0);
    if (live) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const size_t idx = base + a.pol_idx[k];
            wt[k] = a.flags[idx] ? T(0) : (T)a.weights[idx];
            any |= wt[k] != T(0);
        }
        live = any;
    }
    size_t jb1 = 0, jb2 = 0;
    if constexpr (MODE == MODE_4J) {
        if (live) {
            const size_t t = a.time_ids[row];
            if (t >= a.ntime || a.ant1[row] >= a.antenna_count ||
                a.ant2[row] >= a.antenna_count) {
                if (gl == 0) atomicOr(a.error_flag, 4);
                live = false;
            } else {
                jb1 = ((size_t)t * a.antenna_count + a.ant1[row]) * a.facets;
                jb2 = ((size_t)t * a.antenna_count + a.ant2[row]) * a.facets;
            }
        }
    }
    double u0 = 0, v0 = 0, w0 = 0;
    if (live) {
        const double rw = 1.0 / (double)a.wavelengths[s * a.chan + c];
        u0 = (double)a.uvw[3 * (size_t)row] * rw;
        v0 = (double)a.uvw[3 * (size_t)row + 1] * rw;
        w0 = (double)a.uvw[3 * (size_t)row + 2] * rw;
    }

    const size_t slice = (size_t)a.nx * a.ny;
    const T *conv = reinterpret_cast<const T *>(a.conv);
    const cpx<T> *wtaps = reinterpret_cast<const cpx<T> *>(d.taps);
    const int padded = 2 * (int)a.S + 3;
    cpx<T> acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = {T(0), T(0)};

    for (uint32_t f = 0; f < a.facets; ++f) {
        bool on = live;
        const FacetGeom *gp = a.geom + f;
        double u = u0, v = v0, w = w0;
        T cs = T(1), sn = T(0);
        if (on && a.phase) {  // the forward's rotor, in turns, FP64 argument
            const double turns = u * gp->dl + v * gp->dm + w * gp->dn;
            const double fr = 2.0 * (turns - rint(turns));
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
        }
        if (a.transform) {
            u = u - w * gp->u_term;
            v = v - w * gp->v_term;
        }
        u *= a.u_scale;
        v *= a.v_scale;
        bool conj_out = false;
        int x0 = 0, y0 = 0, fu = 0, fv = 0, wp = 0;
        uint32_t tb = 0;
        if (!d.wproj) {  // AA tap choice under the forward's floor rounding
            const double flu = floor(u), flv = floor(v);
            x0 = (int)flu + (int)a.off_x;
            y0 = (int)flv + (int)a.off_y;
            fu = (int)((1.0 - u + flu) * a.OS);
            fv = (int)((1.0 - v + flv) * a.OS);
            wp = (int)a.W;
            if (x0 < 0 || y0 < 0 || x0 + wp + 2 >= (int)a.nx || y0 + wp + 2 >= (int)a.ny)
                on = false;
        } else {  // w-projection: w < 0 conjugation, nearest rounding, plane choice
            if (w < 0) {
                conj_out = true;
                u = -u;
                v = -v;
                w = -w;
            }
            const double ru = rint(u), rv = rint(v);
            const int dx = (int)rint(u + a.off_x), dy = (int)rint(v + a.off_y);
            fu = (int)((1.0 - u + ru) * a.OS);
            fv = (int)((1.0 - v + rv) * a.OS);
            const double pl = rint(w / (double)d.wmax_f * (double)(d.wplanes - 1));
            if (dx < 0 || dy < 0 || dx + padded >= (int)a.nx || dy + padded >= (int)a.ny ||
                pl >= (double)d.wplanes)
                on = false;
            if (on) {
                const int plane = (int)pl;
                const int sp = (int)d.plane_half[plane];
                wp = 2 * sp + 1;
                x0 = dx + (int)a.S - sp;
                y0 = dy + (int)a.S - sp;
                tb = (uint32_t)(d.plane_off[plane] +
                                ((uint64_t)(fv - (int)d.fbase) * d.nf + (uint64_t)(fu - (int)d.fbase)) *
                                    (uint64_t)(wp * wp));
            }
        }
        cpx<T> sk[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) sk[k] = {T(0), T(0)};
        if (on) {
            const grid_base_type *gb = d.model + ((size_t)f * a.cube + cgi) * d.P * slice * 2;
            const int ntap = wp * wp;
            for (int t = gl; t < ntap; t += GL) {
                const int j = t / wp, k = t - j * wp;
                cpx<T> tap;
                if (!d.wproj) {
                    tap = {conv[fu + k * (int)a.OS] * conv[fv + j * (int)a.OS], T(0)};
                } else {
                    const cpx<T> tw = wtaps[tb + (uint32_t)t];
                    tap = {tw.re, -tw.im};
                }
                const size_t pos = (size_t)(y0 + j) * a.nx + (size_t)(x0 + k);
#pragma unroll
                for (int kk = 0; kk < NC; ++kk) {
                    const grid_base_type *cell = gb + 2 * ((size_t)kk * slice + pos);
                    const cpx<T> gv = {(T)cell[0], (T)cell[1]};
                    sk[kk] = cadd(sk[kk], cmul(tap, gv));
                }
            }
        }
        // the whole group takes part in the butterfly, samples that are off included
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            sk[k].re = group_sum(sk[k].re);
            sk[k].im = group_sum(sk[k].im);
        }
        if (on) {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                if (conj_out) sk[k] = conjg(sk[k]);
                sk[k] = cmul(sk[k], cpx<T>{cs, -sn});
                sk[k].re *= wt[k];
                sk[k].im *= wt[k];
            }
            if constexpr (MODE == MODE_4J) {  // S = Pinv^H S Qinv
                const size_t j1 = ((jb1 + f) * a.spw_count + s) * a.chan + c;
                const size_t j2 = ((jb2 + f) * a.spw_count + s) * a.chan + c;
                const cpx<T> *A = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * j1;
                const cpx<T> *B = reinterpret_cast<const cpx<T> *>(a.jones_inv) + 4 * j2;
                const cpx<T> a0 = conjg(A[0]), a1 = conjg(A[1]), a2 = conjg(A[2]),
                             a3 = conjg(A[3]);
                const cpx<T> t0 = cadd(cmul(a0, sk[0]), cmul(a2, sk[2]));
                const cpx<T> t1 = cadd(cmul(a0, sk[1]), cmul(a2, sk[3]));
                const cpx<T> t2 = cadd(cmul(a1, sk[0]), cmul(a3, sk[2]));
                const cpx<T> t3 = cadd(cmul(a1, sk[1]), cmul(a3, sk[3]));
                sk[0] = cadd(cmul(t0, B[0]), cmul(t1, B[2]));
                sk[1] = cadd(cmul(t0, B[1]), cmul(t1, B[3]));
                sk[2] = cadd(cmul(t2, B[0]), cmul(t3, B[2]));
                sk[3] = cadd(cmul(t2, B[1]), cmul(t3, B[3]));
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) acc[k] = cadd(acc[k], sk[k]);
        }
    }
    if (!in_range) return;
    for (uint32_t cc = (uint32_t)gl; cc < a.npol; cc += GL) {
        cpx<T> val = {T(0), T(0)};
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (a.pol_idx[k] == cc) val = cadd(val, acc[k]);
        d.out[2 * (base + cc)] = (visibility_base_type)val.re;
        d.out[2 * (base + cc) + 1] = (visibility_base_type)val.im;
    }
}

template <int MODE, int PATH> static hipError_t launch_t(const DArgs &d, hipStream_t st) {
    const uint64_t threads = d.rows * d.g.chan * GL;
    const uint64_t blocks = (threads + BLOCK - 1) / BLOCK;
    if (!blocks) return hipSuccess;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    degrid_kernel<MODE, PATH><<<dim3((uint32_t)blocks), dim3(BLOCK), 0, st>>>(d);
    return hipGetLastError();
}

template <int MODE> static hipError_t launch_m(const DArgs &d, int path, hipStream_t st) {
    switch (path) {
        case PATH_FP64: return launch_t<MODE, PATH_FP64>(d, st);
        case PATH_F32K: return launch_t<MODE, PATH_F32K>(d, st);
        case PATH_F32: return launch_t<MODE, PATH_F32>(d, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace dg

hipError_t launch_degrid(const GridArgs &a, const void *model, void *out, uint64_t rows,
                         uint32_t P, const void *taps, const uint64_t *plane_off,
                         const uint32_t *plane_half, uint32_t wplanes, uint32_t nf,
                         uint32_t fbase, float wmax_f, int mode, int path, hipStream_t st) {
    dg::DArgs d;
    d.g = a;
    d.model = (const grid_base_type *)model;
    d.out = (visibility_base_type *)out;
    d.taps = taps;
    d.plane_off = plane_off;
    d.plane_half = plane_half;
    d.wplanes = wplanes;
    d.nf = nf;
    d.fbase = fbase;
    d.wproj = wplanes > 1 ? 1u : 0u;
    d.wmax_f = wmax_f;
    d.P = P;
    d.rows = rows;
    switch (mode) {
        case MODE_1: return dg::launch_m<MODE_1>(d, path, st);
        case MODE_2: return dg::launch_m<MODE_2>(d, path, st);
        case MODE_4: return dg::launch_m<MODE_4>(d, path, st);
        case MODE_4J: return dg::launch_m<MODE_4J>(d, path, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace bam
