/*
 * gridder_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of bullseye's CPU imaging library (the reference's own FP64 oracle and
 * CPU baseline), written from the reference's source text.  It is the checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library never links, calls or falls back to it.
 *
 * Built twice by oracle/Makefile with the reference's CPU flags (-O3 -fopenmp, no -ffast-math,
 * bullseye/mo/cpu_algorithm/{single,double}/CMakeLists.txt:17-18):
 *   liboracle_imaging64.so  -DBULLSEYE_DOUBLE   (the FP64 oracle)
 *   liboracle_imaging32.so  -DBULLSEYE_SINGLE   (the single-precision CPU baseline)
 *
 * Each function cites the reference lines it restates.  Semantics kept on purpose:
 *  - AA gridding runs under FE_DOWNWARD (cpu_algorithm/convolution_policies.h:73-75), so lrint
 *    is floor and every FP op of the gridding thread rounds toward -inf; w-projection runs under
 *    FE_TONEAREST (convolution_policies.h:235).  OpenMP workers set the mode themselves.
 *  - kernel-sample choice truncates: size_t((1 - u + lrint(u)) * OS) (convolution_policies.h:92-93)
 *  - ARCSEC_TO_RAD is a float constant even in the FP64 build (cpu_gpu_common/cu_common.h:63)
 *  - a sample whose footprint comes within 2S+3 of the high edge is dropped
 *    (convolution_policies.h:95-96)
 *  - PSF gridding ignores per-sample flags and weights (correlation_gridding_policies.h:194-205)
 *  - normalisation accumulates weight * sum(conv taps); w-projection uses Re(conv)
 *    (convolution_policies.h:283)
 */
#define _GNU_SOURCE
#include <complex.h>
#include <fenv.h>
#include <math.h>
/* type-generic sin/cos/sqrt/lrint: float arguments use the float functions, as the
 * reference's C++ <cmath> overloads do in the single-precision build */
#include <tgmath.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/bullseye_amd.h"

#ifdef BULLSEYE_SINGLE
typedef float complex vis_c;
typedef float complex grid_c;
typedef float complex conv_c;
#define CONJ conjf
#define CREAL crealf
#define CIMAG cimagf
#define MKC(re, im) __builtin_complex((float)(re), (float)(im))
#else
typedef double complex vis_c;
typedef double complex grid_c;
typedef double complex conv_c;
#define CONJ conj
#define CREAL creal
#define CIMAG cimag
#define MKC(re, im) __builtin_complex((double)(re), (double)(im))
#endif

/* cpu_gpu_common/cu_common.h:63 */
static const float ARCSEC_TO_RAD = (float)(M_PI / (180.0 * 3600.0));

enum oracle_mode {
    GRID_SINGLE = 0, FACET_SINGLE, GRID_DUEL, FACET_DUEL, GRID_4, FACET_4, FACET_4_JONES,
    GRID_PSF, FACET_PSF
};

static double g_gridding_seconds = 0;
static normalization_base_type *g_norm = NULL;
static char g_error[256] = "";

const char *oracle_last_error(void) { return g_error; }
double oracle_gridding_walltime(void) { return g_gridding_seconds; }

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* cpu_algorithm/wrapper.cpp:79-106: allocate normalization_terms[F*C*P] and publish it */
int oracle_init(gridding_parameters *p) {
    free(g_norm);
    size_t n = p->num_facet_centres * p->cube_channel_dim_size *
               p->number_of_polarization_terms_being_gridded;
    g_norm = (normalization_base_type *)calloc(n ? n : 1, sizeof(normalization_base_type));
    p->normalization_terms = g_norm;
    g_gridding_seconds = 0;
    return 0;
}

void oracle_release(void) {
    free(g_norm);
    g_norm = NULL;
}

/* jones_2x2.h:88-106 (invert) and :113-126 (invert_all, throws on a singular matrix) */
int oracle_invert_all(vis_c *jones, size_t count) {
    int singular = 0;
#pragma omp parallel for reduction(| : singular)
    for (size_t j = 0; j < count; ++j) {
        vis_c *m = jones + 4 * j;
        vis_c det = m[0] * m[3] - m[1] * m[2];
        if (det == 0) {
            singular = 1;
            continue;
        }
        vis_c det_inv = 1 / det;
        m[1] *= -det_inv;
        m[2] *= -det_inv;
        vis_c swp = m[0];
        m[0] = m[3] * det_inv;
        m[3] = swp * det_inv;
    }
    if (singular) {
        snprintf(g_error, sizeof g_error, "JONES MATRIX SET CONTAINS SINGULAR MATRICIES");
        return 1;
    }
    return 0;
}

/* Per-facet geometry: templated_gridder.h:36-42 (facet centre), baseline_transform_policies.h:
 * 120-143 (planar uvw transform with w) and phase_transform_policies.h:110-134 (phase offset). */
typedef struct {
    uvw_base_type u_term, v_term, dl, dm, dn;
} facet_geom;

static facet_geom facet_setup(const gridding_parameters *p, size_t f, int facet, int phase) {
    facet_geom g = {0, 0, 0, 0, 0};
    uvw_base_type ra = p->phase_centre_ra, dec = p->phase_centre_dec;
    if (facet) {
        ra = p->facet_centres[2 * f];
        dec = p->facet_centres[2 * f + 1];
    }
    if (facet) {
        uvw_base_type d_ra = (ra - p->phase_centre_ra) * ARCSEC_TO_RAD,
                      n_dec = dec * ARCSEC_TO_RAD, o_dec = p->phase_centre_dec * ARCSEC_TO_RAD,
                      c_d_ra = cos(d_ra), s_d_ra = sin(d_ra), c_new_dec = cos(n_dec),
                      c_old_dec = cos(o_dec), s_new_dec = sin(n_dec), s_old_dec = sin(o_dec),
                      li0 = c_new_dec * s_d_ra,
                      mi0 = s_new_dec * c_old_dec - c_new_dec * s_old_dec * c_d_ra,
                      ni0 = s_new_dec * s_old_dec + c_new_dec * c_old_dec * c_d_ra;
        g.u_term = li0 / ni0;
        g.v_term = mi0 / ni0;
    }
    if (phase) {
        uvw_base_type d_ra = (ra - p->phase_centre_ra) * ARCSEC_TO_RAD,
                      d_dec = dec * ARCSEC_TO_RAD, d_decp = p->phase_centre_dec * ARCSEC_TO_RAD,
                      c_d_dec = cos(d_dec), s_d_dec = sin(d_dec), s_d_ra = sin(d_ra),
                      c_d_ra = cos(d_ra), c_d_decp = cos(d_decp), s_d_decp = sin(d_decp);
        g.dl = -c_d_dec * s_d_ra;
        g.dm = -(s_d_dec * c_d_decp - c_d_dec * s_d_decp * c_d_ra);
        g.dn = 1 - sqrt(1 - g.dl * g.dl - g.dm * g.dm);
    }
    return g;
}

/* One facet of cpu_algorithm/templated_gridder.h:12-104 with the policies chosen by
 * cpu_algorithm/wrapper.cpp:182-425 (mode) and wplanes (AA 1D vs 2D w-projection). */
static void grid_facet(const gridding_parameters *p, int mode, size_t f) {
    const int wproj = p->wplanes > 1;
    fesetround(wproj ? FE_TONEAREST : FE_DOWNWARD);
    const int facet = (mode == FACET_SINGLE || mode == FACET_DUEL || mode == FACET_4 ||
                       mode == FACET_4_JONES || mode == FACET_PSF);
    const int phase = facet && mode != FACET_PSF; /* wrapper.cpp:411-412 */
    const int psf = (mode == GRID_PSF || mode == FACET_PSF);
    const int jones = (mode == FACET_4_JONES);
    size_t ncorr = 1;
    if (mode == GRID_DUEL || mode == FACET_DUEL) ncorr = 2;
    if (mode == GRID_4 || mode == FACET_4 || mode == FACET_4_JONES) ncorr = 4;
    const size_t P = psf ? 1 : p->number_of_polarization_terms_being_gridded;

    const size_t S = p->conv_support, OS = p->conv_oversample;
    const size_t full = 2 * S + 1, padded = full + 2;
    /* templated_gridder.h:23-27 */
    const uvw_base_type u_scale = p->nx * p->cell_size_x * ARCSEC_TO_RAD;
    const uvw_base_type v_scale = -(p->ny * p->cell_size_y * ARCSEC_TO_RAD);
    const uvw_base_type off_x = p->nx / 2 - p->conv_support;
    const uvw_base_type off_y = p->ny / 2 - p->conv_support;
    const size_t slice = p->nx * p->ny; /* complex elements per grid slice */

    /* correlation_gridding_policies.h:114-118 / :188-192 (compute_facet_grid_ptr) */
    grid_c *facet_grid = psf ? (grid_c *)p->sampling_function_buffer +
                                   slice * p->sampling_function_channel_count * f
                             : (grid_c *)p->output_buffer + slice * P * p->cube_channel_dim_size * f;

    const facet_geom fg = facet_setup(p, f, facet, phase);
    const uvw_base_type u_term = fg.u_term, v_term = fg.v_term, dl = fg.dl, dm = fg.dm,
                        dn = fg.dn;
    const size_t conv_dim = padded + (padded - 1) * (OS - 1);
    const convolution_base_type *conv_r = p->conv;
    const conv_c *conv_w = (const conv_c *)p->conv;
    const vis_c *vis_all = (const vis_c *)p->visibilities;
    const vis_c *jones_all = (const vis_c *)p->jones_terms;

    for (size_t row = 0; row < p->row_count; ++row) {
        const uvw_base_type *uvw = p->uvw_coords + 3 * row;
        const int row_flagged = p->flagged_rows[row];
        const int in_field = p->field_array[row] == p->imaging_field;
        const size_t spw = p->spw_index_array[row];
        for (size_t c = 0; c < p->channel_count; ++c) {
            const size_t fc = spw * p->channel_count + c;
            if (!p->enabled_channels[fc]) continue;
            const size_t cgi = psf ? p->sampling_function_channel_grid_indicies[fc]
                                   : p->channel_grid_indicies[fc];
            reference_wavelengths_base_type rw = 1 / p->reference_wavelengths[fc];
            /* correlation_gridding_policies.h:97-109, 194-205, 230-240, 367-376 */
            vis_c vis[4];
            int flag[4];
            visibility_weights_base_type weight[4];
            const size_t base = (row * p->channel_count + c) * p->number_of_polarization_terms;
            for (size_t k = 0; k < ncorr; ++k) {
                size_t corr = k;
                if (ncorr == 1) corr = p->polarization_index;
                if (ncorr == 2) corr = k == 0 ? p->polarization_index : p->second_polarization_index;
                if (psf) {
                    vis[k] = 1;
                    flag[k] = 0;
                    weight[k] = 1;
                } else {
                    vis[k] = vis_all[base + corr];
                    flag[k] = p->flags[base + corr];
                    weight[k] = p->visibility_weights[base + corr];
                }
            }
            /* templated_gridder.h:63-67 */
            uvw_base_type u = uvw[0] * rw, v = uvw[1] * rw, w = uvw[2] * rw;
            if (jones) { /* correlation_gridding_policies.h:523-542: V = Pinv (V Qinv^H) */
                size_t t = p->timestamp_ids[row];
                size_t a1 = p->antenna_1_ids[row], a2 = p->antenna_2_ids[row];
                size_t j1 = (((t * p->antenna_count + a1) * p->num_facet_centres + f) * p->spw_count +
                             spw) * p->channel_count + c;
                size_t j2 = (((t * p->antenna_count + a2) * p->num_facet_centres + f) * p->spw_count +
                             spw) * p->channel_count + c;
                const vis_c *P1 = jones_all + 4 * j1;
                const vis_c *Q = jones_all + 4 * j2;
                /* jones_2x2.h:60-71 hermitian transpose */
                vis_c qh[4] = {CONJ(Q[0]), CONJ(Q[2]), CONJ(Q[1]), CONJ(Q[3])};
                /* correlation_gridding_traits.h:148-159 (vis * jones) */
                vis_c t0 = vis[0] * qh[0] + vis[1] * qh[2];
                vis_c t1 = vis[0] * qh[1] + vis[1] * qh[3];
                vis_c t2 = vis[2] * qh[0] + vis[3] * qh[2];
                vis_c t3 = vis[2] * qh[1] + vis[3] * qh[3];
                /* correlation_gridding_traits.h:137-146 (jones * vis) */
                vis[0] = P1[0] * t0 + P1[1] * t2;
                vis[1] = P1[0] * t1 + P1[1] * t3;
                vis[2] = P1[2] * t0 + P1[3] * t2;
                vis[3] = P1[2] * t1 + P1[3] * t3;
            }
            /* templated_gridder.h:78-82: weight * int(!(flag || row_flag) && in_field) */
            visibility_weights_base_type cw[4];
            for (size_t k = 0; k < ncorr; ++k) {
                int keep = !(flag[k] || row_flagged) && in_field;
                cw[k] = weight[k] * (visibility_base_type)keep;
                vis[k] = MKC(CREAL(vis[k]) * cw[k], CIMAG(vis[k]) * cw[k]);
            }
            if (phase) { /* phase_transform_policies.h:138-146 */
                uvw_base_type x = 2 * M_PI * (u * dl + v * dm + w * dn);
                uvw_base_type s_, c_;
#ifdef BULLSEYE_SINGLE
                sincosf(x, &s_, &c_);
#else
                sincos(x, &s_, &c_);
#endif
                for (size_t k = 0; k < ncorr; ++k) {
                    visibility_base_type re = CREAL(vis[k]), im = CIMAG(vis[k]);
                    vis[k] = MKC(re * c_ - im * s_, re * s_ + im * c_);
                }
            }
            if (facet) { /* baseline_transform_policies.h:145-148 */
                u = u - w * u_term;
                v = v - w * v_term;
            }
            u *= u_scale; /* templated_gridder.h:89-90 */
            v *= v_scale;

            normalization_base_type norm_acc = 0;
            if (!wproj) { /* convolution_policies.h:76-117 (AA 1D precomputed) */
                uvw_base_type tu = u + off_x, tv = v + off_y;
                size_t disc_u = (size_t)lrint(tu), disc_v = (size_t)lrint(tv);
                size_t frac_u = (size_t)((1 - u + lrint(u)) * OS);
                size_t frac_v = (size_t)((1 - v + lrint(v)) * OS);
                if (disc_v + padded >= p->ny || disc_u + padded >= p->nx || disc_v >= p->ny ||
                    disc_u >= p->nx)
                    goto store;
                grid_c *g = facet_grid + (psf ? cgi : cgi * P) * slice;
                size_t cv_i = frac_v;
                for (size_t sv = 0; sv < full; ++sv) {
                    convolution_base_type cvw = conv_r[cv_i];
                    size_t cu_i = frac_u;
                    for (size_t su = 0; su < full; ++su) {
                        convolution_base_type cww = conv_r[cu_i] * cvw;
                        size_t pos = (disc_v + sv) * p->nx + disc_u + su;
                        for (size_t k = 0; k < ncorr; ++k) {
                            grid_base_type re = CREAL(vis[k]) * cww, im = CIMAG(vis[k]) * cww;
                            grid_base_type *cell = (grid_base_type *)(g + k * slice + pos);
                            cell[0] += re;
                            cell[1] += im;
                        }
                        norm_acc += cww;
                        cu_i += OS;
                    }
                    cv_i += OS;
                }
            } else { /* convolution_policies.h:231-289 (2D w-projection precomputed) */
                if (w < 0) {
                    for (size_t k = 0; k < ncorr; ++k) vis[k] = CONJ(vis[k]);
                    u *= -1;
                    v *= -1;
                    w *= -1;
                }
                uvw_base_type tu = u + off_x, tv = v + off_y;
                size_t disc_u = (size_t)lrint(tu), disc_v = (size_t)lrint(tv);
                size_t frac_u = (size_t)((1 - u + lrint(u)) * OS);
                size_t frac_v = (size_t)((1 - v + lrint(v)) * OS);
                size_t plane = (size_t)lrint(fabs(w) / (float)p->wmax_est * (p->wplanes - 1));
                size_t filter_offset = plane * conv_dim * conv_dim;
                if (disc_v + padded >= p->ny || disc_u + padded >= p->nx || disc_v >= p->ny ||
                    disc_u >= p->nx || plane >= p->wplanes)
                    goto store;
                grid_c *g = facet_grid + (psf ? cgi : cgi * P) * slice;
                size_t cv_i = filter_offset + frac_v * conv_dim;
                for (size_t sv = 0; sv < full; ++sv) {
                    size_t cu_i = frac_u;
                    for (size_t su = 0; su < full; ++su) {
                        conv_c cww = conv_w[cv_i + cu_i];
                        size_t pos = (disc_v + sv) * p->nx + disc_u + su;
                        for (size_t k = 0; k < ncorr; ++k) {
                            /* basic_complex product (cu_basic_complex.h:73-75) */
                            grid_base_type vr = CREAL(vis[k]), vi = CIMAG(vis[k]);
                            grid_base_type cr = CREAL(cww), ci = CIMAG(cww);
                            grid_base_type *cell = (grid_base_type *)(g + k * slice + pos);
                            cell[0] += vr * cr - vi * ci;
                            cell[1] += vr * ci + vi * cr;
                        }
                        norm_acc += CREAL(cww);
                        cu_i += OS;
                    }
                    cv_i += OS * conv_dim;
                }
            }
        store:
            /* templated_gridder.h:97-99 + store_normalization_term (P=1,2,4) */
            if (!psf) {
                size_t nidx = (f * p->cube_channel_dim_size + cgi) * ncorr;
                for (size_t k = 0; k < ncorr; ++k)
                    p->normalization_terms[nidx + k] +=
                        (normalization_base_type)(cw[k] * (visibility_weights_base_type)norm_acc);
            }
        }
    }
    fesetround(FE_TONEAREST);
}

/* cpu_algorithm/wrapper.cpp:182-425: one entry point per mode; OpenMP over facets
 * (templated_gridder.h:29).  Synchronous here; the reference runs it in std::async. */
int oracle_grid(gridding_parameters *p, int mode) {
    if (mode < GRID_SINGLE || mode > FACET_PSF) {
        snprintf(g_error, sizeof g_error, "bad mode %d", mode);
        return 1;
    }
    double t0 = now_seconds();
    if (mode == FACET_4_JONES) { /* wrapper.cpp:356-360: invert every Jones term in place */
        size_t n = p->no_timestamps_read * p->antenna_count * p->num_facet_centres *
                   p->spw_count * p->channel_count;
        if (oracle_invert_all((vis_c *)p->jones_terms, n)) return 1;
    }
    /* grid_* modes also loop over num_facet_centres, each at the phase centre
     * (phase_transform_policies.h:60-65) */
#pragma omp parallel for schedule(static)
    for (size_t f = 0; f < p->num_facet_centres; ++f) grid_facet(p, mode, f);
    g_gridding_seconds += now_seconds() - t0;
    return 0;
}

/* The degridding adjoint (no reference code exists; bullseye's Resampling.ipynb cell 17 is a
 * Python sketch of the gather).  For the visibility modes (not PSF) this is the exact
 * transpose, over the reals, of grid_facet's map from visibilities to the unnormalised grids,
 * summed over facets: per (row, channel) and facet, with the forward's geometry, tap choice,
 * edge drop and plane choice,
 *   s_k = sum over the footprint of conj(C) * model_k       (C real for the AA kernel)
 *   s_k = conj(s_k) when the w-projection sample had w < 0  (the forward conjugated it)
 *   s_k *= exp(-i x) (facet phase), s_k *= weight_k * keep_k,
 *   S = Pinv^H S Qinv for the Jones mode (the forward is V -> Pinv V Qinv^H),
 * and out[row][chan][corr_k] = sum over facets of s_k; every other correlation is zero.
 * model has output_buffer's layout; out has the visibilities' layout.  The Jones terms are
 * inverted on a private copy. */
int oracle_degrid(const gridding_parameters *p, int mode, const grid_c *model, vis_c *out) {
    if (mode < GRID_SINGLE || mode > FACET_4_JONES) {
        snprintf(g_error, sizeof g_error, "degridding is defined for the visibility modes only");
        return 1;
    }
    const int wproj = p->wplanes > 1;
    const int facet = (mode == FACET_SINGLE || mode == FACET_DUEL || mode == FACET_4 ||
                       mode == FACET_4_JONES);
    const int phase = facet;
    const int jones = (mode == FACET_4_JONES);
    size_t ncorr = 1;
    if (mode == GRID_DUEL || mode == FACET_DUEL) ncorr = 2;
    if (mode == GRID_4 || mode == FACET_4 || mode == FACET_4_JONES) ncorr = 4;
    const size_t P = p->number_of_polarization_terms_being_gridded;
    const size_t F = p->num_facet_centres, npol = p->number_of_polarization_terms;
    const size_t S = p->conv_support, OS = p->conv_oversample;
    const size_t full = 2 * S + 1, padded = full + 2;
    const size_t slice = p->nx * p->ny;
    const size_t conv_dim = padded + (padded - 1) * (OS - 1);
    vis_c *jinv = NULL;
    if (jones) {
        size_t n = p->no_timestamps_read * p->antenna_count * F * p->spw_count * p->channel_count;
        jinv = (vis_c *)malloc(4 * n * sizeof(vis_c));
        memcpy(jinv, p->jones_terms, 4 * n * sizeof(vis_c));
        if (oracle_invert_all(jinv, n)) {
            free(jinv);
            return 1;
        }
    }
    facet_geom *geo = (facet_geom *)malloc((F ? F : 1) * sizeof(facet_geom));
    for (size_t f = 0; f < F; ++f) geo[f] = facet_setup(p, f, facet, phase);
    memset(out, 0, p->row_count * p->channel_count * npol * sizeof(vis_c));
    const convolution_base_type *conv_r = p->conv;
    const conv_c *conv_w = (const conv_c *)p->conv;

#pragma omp parallel for schedule(static)
    for (size_t row = 0; row < p->row_count; ++row) {
        fesetround(wproj ? FE_TONEAREST : FE_DOWNWARD);
        const uvw_base_type u_scale = p->nx * p->cell_size_x * ARCSEC_TO_RAD;
        const uvw_base_type v_scale = -(p->ny * p->cell_size_y * ARCSEC_TO_RAD);
        const uvw_base_type off_x = p->nx / 2 - p->conv_support;
        const uvw_base_type off_y = p->ny / 2 - p->conv_support;
        const uvw_base_type *uvw = p->uvw_coords + 3 * row;
        const int row_flagged = p->flagged_rows[row];
        const int in_field = p->field_array[row] == p->imaging_field;
        const size_t spw = p->spw_index_array[row];
        for (size_t c = 0; c < p->channel_count; ++c) {
            const size_t fc = spw * p->channel_count + c;
            if (!p->enabled_channels[fc]) continue;
            const size_t cgi = p->channel_grid_indicies[fc];
            reference_wavelengths_base_type rw = 1 / p->reference_wavelengths[fc];
            const size_t base = (row * p->channel_count + c) * npol;
            size_t corr[4];
            visibility_weights_base_type cw[4];
            for (size_t k = 0; k < ncorr; ++k) {
                corr[k] = k;
                if (ncorr == 1) corr[k] = p->polarization_index;
                if (ncorr == 2) corr[k] = k == 0 ? p->polarization_index
                                                 : p->second_polarization_index;
                int keep = !(p->flags[base + corr[k]] || row_flagged) && in_field;
                cw[k] = p->visibility_weights[base + corr[k]] * (visibility_base_type)keep;
            }
            vis_c acc[4] = {0, 0, 0, 0};
            for (size_t f = 0; f < F; ++f) {
                uvw_base_type u = uvw[0] * rw, v = uvw[1] * rw, w = uvw[2] * rw;
                uvw_base_type s_ = 0, c_ = 1;
                if (phase) {
                    uvw_base_type x = 2 * M_PI * (u * geo[f].dl + v * geo[f].dm + w * geo[f].dn);
#ifdef BULLSEYE_SINGLE
                    sincosf(x, &s_, &c_);
#else
                    sincos(x, &s_, &c_);
#endif
                }
                if (facet) {
                    u = u - w * geo[f].u_term;
                    v = v - w * geo[f].v_term;
                }
                u *= u_scale;
                v *= v_scale;
                const grid_c *g = model + (f * p->cube_channel_dim_size + cgi) * P * slice;
                vis_c sk[4] = {0, 0, 0, 0};
                int conj_out = 0;
                if (!wproj) {
                    uvw_base_type tu = u + off_x, tv = v + off_y;
                    size_t disc_u = (size_t)lrint(tu), disc_v = (size_t)lrint(tv);
                    size_t frac_u = (size_t)((1 - u + lrint(u)) * OS);
                    size_t frac_v = (size_t)((1 - v + lrint(v)) * OS);
                    if (disc_v + padded >= p->ny || disc_u + padded >= p->nx ||
                        disc_v >= p->ny || disc_u >= p->nx)
                        continue;
                    size_t cv_i = frac_v;
                    for (size_t sv = 0; sv < full; ++sv) {
                        convolution_base_type cvw = conv_r[cv_i];
                        size_t cu_i = frac_u;
                        for (size_t su = 0; su < full; ++su) {
                            convolution_base_type cww = conv_r[cu_i] * cvw;
                            size_t pos = (disc_v + sv) * p->nx + disc_u + su;
                            for (size_t k = 0; k < ncorr; ++k) sk[k] += cww * g[k * slice + pos];
                            cu_i += OS;
                        }
                        cv_i += OS;
                    }
                } else {
                    if (w < 0) {
                        conj_out = 1;
                        u *= -1;
                        v *= -1;
                        w *= -1;
                    }
                    uvw_base_type tu = u + off_x, tv = v + off_y;
                    size_t disc_u = (size_t)lrint(tu), disc_v = (size_t)lrint(tv);
                    size_t frac_u = (size_t)((1 - u + lrint(u)) * OS);
                    size_t frac_v = (size_t)((1 - v + lrint(v)) * OS);
                    size_t plane = (size_t)lrint(fabs(w) / (float)p->wmax_est * (p->wplanes - 1));
                    if (disc_v + padded >= p->ny || disc_u + padded >= p->nx ||
                        disc_v >= p->ny || disc_u >= p->nx || plane >= p->wplanes)
                        continue;
                    size_t cv_i = plane * conv_dim * conv_dim + frac_v * conv_dim;
                    for (size_t sv = 0; sv < full; ++sv) {
                        size_t cu_i = frac_u;
                        for (size_t su = 0; su < full; ++su) {
                            conv_c cww = CONJ(conv_w[cv_i + cu_i]);
                            size_t pos = (disc_v + sv) * p->nx + disc_u + su;
                            for (size_t k = 0; k < ncorr; ++k) sk[k] += cww * g[k * slice + pos];
                            cu_i += OS;
                        }
                        cv_i += OS * conv_dim;
                    }
                }
                for (size_t k = 0; k < ncorr; ++k) {
                    if (conj_out) sk[k] = CONJ(sk[k]);
                    sk[k] *= MKC(c_, -s_);
                    sk[k] *= cw[k];
                }
                if (jones) {
                    size_t t = p->timestamp_ids[row];
                    size_t a1 = p->antenna_1_ids[row], a2 = p->antenna_2_ids[row];
                    size_t j1 = (((t * p->antenna_count + a1) * F + f) * p->spw_count + spw) *
                                    p->channel_count + c;
                    size_t j2 = (((t * p->antenna_count + a2) * F + f) * p->spw_count + spw) *
                                    p->channel_count + c;
                    const vis_c *A = jinv + 4 * j1, *B = jinv + 4 * j2;
                    /* T = A^H S, then T B */
                    vis_c t0 = CONJ(A[0]) * sk[0] + CONJ(A[2]) * sk[2];
                    vis_c t1 = CONJ(A[0]) * sk[1] + CONJ(A[2]) * sk[3];
                    vis_c t2 = CONJ(A[1]) * sk[0] + CONJ(A[3]) * sk[2];
                    vis_c t3 = CONJ(A[1]) * sk[1] + CONJ(A[3]) * sk[3];
                    sk[0] = t0 * B[0] + t1 * B[2];
                    sk[1] = t0 * B[1] + t1 * B[3];
                    sk[2] = t2 * B[0] + t3 * B[2];
                    sk[3] = t2 * B[1] + t3 * B[3];
                }
                for (size_t k = 0; k < ncorr; ++k) acc[k] += sk[k];
            }
            for (size_t k = 0; k < ncorr; ++k) out[base + corr[k]] = acc[k];
        }
        fesetround(FE_TONEAREST);
    }
    free(geo);
    free(jinv);
    return 0;
}

/* cpu_algorithm/wrapper.cpp:136-163 */
void oracle_normalize(gridding_parameters *p) {
    const size_t P = p->number_of_polarization_terms_being_gridded;
    for (size_t f = 0; f < p->num_facet_centres; ++f)
        for (size_t ch = 0; ch < p->cube_channel_dim_size; ++ch)
            for (size_t corr = 0; corr < P; ++corr) {
                normalization_base_type nv =
                    p->normalization_terms[(f * p->cube_channel_dim_size + ch) * P + corr];
                grid_c *g = (grid_c *)p->output_buffer +
                            ((f * p->cube_channel_dim_size + ch) * P + corr) * p->ny * p->nx;
                grid_base_type d = (grid_base_type)nv;
                for (size_t i = 0; i < p->ny * p->nx; ++i) {
                    grid_base_type *cell = (grid_base_type *)(g + i);
                    cell[0] /= d;
                    cell[1] /= d;
                }
            }
}

/* cpu_algorithm/wrapper.cpp:116-135 (channel_grid_indicies indexed by the sampling-function
 * channel, as in the reference) */
void oracle_weight_uniformly(gridding_parameters *p) {
    const size_t P = p->number_of_polarization_terms_being_gridded;
    const grid_c *sf = (const grid_c *)p->sampling_function_buffer;
    grid_c *out = (grid_c *)p->output_buffer;
    for (size_t f = 0; f < p->num_facet_centres; ++f)
        for (size_t g = 0; g < p->cube_channel_dim_size; ++g)
            for (size_t y = 0; y < p->ny; ++y)
                for (size_t x = 0; x < p->nx; ++x) {
                    grid_base_type count = 0.0000001f;
                    for (size_t c = 0; c < p->sampling_function_channel_count; ++c)
                        count += (int)(p->channel_grid_indicies[c] == g) *
                                 CREAL(sf[((f * p->sampling_function_channel_count + c) * p->ny + y) *
                                              p->nx + x]);
                    count = 1 / count;
                    for (size_t corr = 0; corr < P; ++corr) {
                        grid_base_type *cell =
                            (grid_base_type *)(out + (((f * p->cube_channel_dim_size + g) * P + corr) *
                                                          p->ny + y) * p->nx + x);
                        cell[0] *= count;
                        cell[1] *= count;
                    }
                }
}

/* gpu_algorithm/wrapper.cu:358-366 */
long oracle_baseline_index(long a1, long a2, long n_ant) {
    long s = a1 < a2 ? a1 : a2;
    return (s * (-s + (2 * n_ant + 1))) / 2 + labs(a1 - a2);
}

/* gpu_algorithm/wrapper.cu:367-381: counts per baseline and the prefix sum only (the row
 * permutation is the stable order of rows within each baseline) */
void oracle_baseline_starts(const unsigned int *a1, const unsigned int *a2, size_t rows,
                            size_t n_ant, size_t n_bl, size_t *starts, size_t *perm) {
    memset(starts, 0, sizeof(size_t) * (n_bl + 1));
    for (size_t r = 0; r < rows; ++r) ++starts[oracle_baseline_index(a1[r], a2[r], n_ant) + 1];
    for (size_t b = 0; b < n_bl; ++b) starts[b + 1] += starts[b];
    size_t *cur = (size_t *)calloc(n_bl, sizeof(size_t));
    for (size_t r = 0; r < rows; ++r) {
        size_t b = oracle_baseline_index(a1[r], a2[r], n_ant);
        perm[starts[b] + cur[b]++] = r;
    }
    free(cur);
}

/* fft_shift_utils.cpp:41-74 (in place, odd sizes safe) */
void oracle_fftshift(grid_c *grid, size_t nx, size_t ny, size_t slices) {
    for (size_t s = 0; s < slices; ++s) {
        grid_c *g = grid + s * nx * ny;
        size_t hx = nx / 2, hy = ny / 2, ox = nx % 2, oy = ny % 2;
        for (size_t iy = 0; iy < ny; ++iy) {
            grid_c mid = g[iy * nx + hx];
            for (size_t ix = 0; ix < hx; ++ix) {
                grid_c swp = g[iy * nx + ix];
                g[iy * nx + ix] = g[iy * nx + hx + ix + ox];
                g[iy * nx + hx + ix] = swp;
            }
            if (ox) g[iy * nx + nx - 1] = mid;
        }
        for (size_t ix = 0; ix < nx; ++ix) {
            grid_c mid = g[hy * nx + ix];
            for (size_t iy = 0; iy < hy; ++iy) {
                grid_c swp = g[iy * nx + ix];
                g[iy * nx + ix] = g[(iy + hy + oy) * nx + ix];
                g[(hy + iy) * nx + ix] = swp;
            }
            if (oy) g[(ny - 1) * nx + ix] = mid;
        }
    }
}

/* fft_shift_utils.cpp:75-112 */
void oracle_ifftshift(grid_c *grid, size_t nx, size_t ny, size_t slices) {
    for (size_t s = 0; s < slices; ++s) {
        grid_c *g = grid + s * nx * ny;
        size_t hx = nx / 2, hy = ny / 2, ox = nx % 2, oy = ny % 2;
        for (size_t iy = 0; iy < ny; ++iy) {
            grid_c mid = g[iy * nx + hx];
            for (size_t ix = 0; ix < hx; ++ix) {
                size_t r = hx - 1 - ix;
                grid_c swp = g[iy * nx + hx + r + ox];
                g[iy * nx + hx + r + ox] = g[iy * nx + r];
                g[iy * nx + r + ox] = swp;
            }
            g[iy * nx] = mid;
        }
        for (size_t ix = 0; ix < nx; ++ix) {
            grid_c mid = g[hy * nx + ix];
            for (size_t iy = 0; iy < hy; ++iy) {
                size_t r = hy - 1 - iy;
                grid_c swp = g[(hy + r + oy) * nx + ix];
                g[(hy + r + oy) * nx + ix] = g[r * nx + ix];
                g[(r + oy) * nx + ix] = swp;
            }
            g[ix] = mid;
        }
    }
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* offsetof every gridding_parameters field, in declaration order (ABI layout check) */
#include <stddef.h>
#define OFF(f) offsetof(gridding_parameters, f)
size_t oracle_struct_offsets(size_t *out) {
    const size_t o[] = {
        OFF(visibilities), OFF(uvw_coords), OFF(reference_wavelengths), OFF(visibility_weights),
        OFF(flags), OFF(flagged_rows), OFF(field_array), OFF(spw_index_array), OFF(imaging_field),
        OFF(baseline_count), OFF(row_count), OFF(chunk_max_row_count), OFF(channel_count),
        OFF(number_of_polarization_terms), OFF(number_of_polarization_terms_being_gridded),
        OFF(spw_count), OFF(no_timestamps_read), OFF(nx), OFF(ny), OFF(cell_size_x),
        OFF(cell_size_y), OFF(conv), OFF(conv_support), OFF(conv_oversample),
        OFF(polarization_index), OFF(second_polarization_index), OFF(output_buffer),
        OFF(phase_centre_ra), OFF(phase_centre_dec), OFF(facet_centres), OFF(num_facet_centres),
        OFF(jones_terms), OFF(should_invert_jones_terms), OFF(antenna_1_ids), OFF(antenna_2_ids),
        OFF(timestamp_ids), OFF(antenna_count), OFF(enabled_channels), OFF(channel_grid_indicies),
        OFF(cube_channel_dim_size), OFF(should_grid_sampling_function),
        OFF(sampling_function_buffer), OFF(sampling_function_channel_grid_indicies),
        OFF(sampling_function_channel_count), OFF(is_final_data_chunk), OFF(wplanes),
        OFF(wmax_est), OFF(baseline_starting_indexes), OFF(antenna_jones_starting_indexes),
        OFF(jones_time_indicies_per_antenna), OFF(normalization_terms), sizeof(gridding_parameters)};
    const size_t n = sizeof(o) / sizeof(o[0]);
    if (out) memcpy(out, o, sizeof(o));
    return n;
}
