// finalize.hip -- grid layout conversion, normalisation, uniform weighting and the facet
// FFT with its shifts and the grid-taper correction.
//
// finalize restates ifft_machine::repack_and_ifft_uv_grids (fft_and_repacking_routines.cpp:
// 83-118): per facet the first `n` slices are ifftshifted, transformed with an unnormalised
// backward c2c FFT (exp(+i), FFTW_BACKWARD), fftshifted and their real parts stored as float32
// from the start of the facet block.  Here the shifts are gathers fused around one batched
// hipFFT call, and the optional taper correction divides each image pixel by T(x) T(y), the
// image-plane response of the sampled AA kernel (no counterpart in the reference).
#include <hipfft/hipfft.h>

#include "device_common.h"

namespace bam {

typedef visibility_base_type vis_t;

// internal FP64 accumulation grid [F][C][ny][nx][pstride][2] <-> ABI grid
// complex<grid_base_type>[F][C][P][ny][nx]
__global__ void k_grid_to_abi(const double *g, grid_base_type *out, uint64_t slices_fc, uint32_t P,
                              uint32_t pstride, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = slices_fc * npix;
    if (i >= total) return;
    const uint64_t fc = i / npix, pix = i - fc * npix;
    const double *src = g + (fc * npix + pix) * pstride * 2;
    for (uint32_t p = 0; p < P; ++p) {
        grid_base_type *dst = out + ((fc * P + p) * npix + pix) * 2;
        dst[0] = (grid_base_type)src[2 * p];
        dst[1] = (grid_base_type)src[2 * p + 1];
    }
}

__global__ void k_abi_to_grid(const grid_base_type *in, double *g, uint64_t slices_fc, uint32_t P,
                              uint32_t pstride, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = slices_fc * npix;
    if (i >= total) return;
    const uint64_t fc = i / npix, pix = i - fc * npix;
    double *dst = g + (fc * npix + pix) * pstride * 2;
    for (uint32_t p = 0; p < pstride; ++p) {
        if (p < P && in) {
            const grid_base_type *src = in + ((fc * P + p) * npix + pix) * 2;
            dst[2 * p] = (double)src[0];
            dst[2 * p + 1] = (double)src[1];
        } else {
            dst[2 * p] = 0;
            dst[2 * p + 1] = 0;
        }
    }
}

__global__ void k_fill_f64(double *p, uint64_t n, double v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// cpu_algorithm/wrapper.cpp:136-163: slice (f, ch, corr) /= (grid type)norm[(f*C+ch)*P+corr]
__global__ void k_normalize(grid_base_type *out, const double *norm, uint64_t slices, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slices * npix) return;
    const grid_base_type d = (grid_base_type)norm[i / npix];
    out[2 * i] /= d;
    out[2 * i + 1] /= d;
}

// cpu_algorithm/wrapper.cpp:116-135 (channel_grid_indicies indexed by the sampling-function
// channel, as in the reference)
__global__ void k_weight_uniformly(grid_base_type *out, const grid_base_type *sf,
                                   const uint64_t *cgi, uint32_t F, uint32_t C, uint32_t P,
                                   uint32_t sfC, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * C * npix) return;
    const uint64_t fg = i / npix, pix = i - fg * npix;
    const uint64_t f = fg / C, g = fg - f * C;
    grid_base_type count = 0.0000001f;
    for (uint32_t c = 0; c < sfC; ++c)
        count += (int)(cgi[c] == g) * sf[((f * sfC + c) * npix + pix) * 2];
    count = 1 / count;
    for (uint32_t corr = 0; corr < P; ++corr) {
        grid_base_type *cell = out + (((f * C + g) * P + corr) * npix + pix) * 2;
        cell[0] *= count;
        cell[1] *= count;
    }
}

// Gather the first n slices of every facet block (facet stride `fstride` slices) into a
// contiguous [F*n][ny][nx] buffer, applying ifftshift (numpy: out[i] = in[(i + N/2) % N]).
__global__ void k_ifftshift_gather(const grid_base_type *in, grid_base_type *out, uint32_t F,
                                   uint32_t n, uint32_t fstride, uint32_t nx, uint32_t ny) {
    const uint64_t npix = (uint64_t)nx * ny;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * n * npix) return;
    const uint64_t s = i / npix, pix = i - s * npix;
    const uint32_t f = (uint32_t)(s / n), k = (uint32_t)(s - (uint64_t)f * n);
    const uint32_t y = (uint32_t)(pix / nx), x = (uint32_t)(pix - (uint64_t)y * nx);
    const uint32_t sy = (y + ny / 2) % ny, sx = (x + nx / 2) % nx;
    const grid_base_type *src = in + (((uint64_t)f * fstride + k) * npix + (uint64_t)sy * nx + sx) * 2;
    out[2 * i] = src[0];
    out[2 * i + 1] = src[1];
}

// Scatter back with fftshift (out[i] = in[(i + (N+1)/2) % N]) and the optional taper division.
__global__ void k_fftshift_scatter(const grid_base_type *in, grid_base_type *out, uint32_t F,
                                   uint32_t n, uint32_t fstride, uint32_t nx, uint32_t ny,
                                   const double *tx, const double *ty) {
    const uint64_t npix = (uint64_t)nx * ny;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * n * npix) return;
    const uint64_t s = i / npix, pix = i - s * npix;
    const uint32_t f = (uint32_t)(s / n), k = (uint32_t)(s - (uint64_t)f * n);
    const uint32_t y = (uint32_t)(pix / nx), x = (uint32_t)(pix - (uint64_t)y * nx);
    const uint32_t sy = (y + (ny + 1) / 2) % ny, sx = (x + (nx + 1) / 2) % nx;
    const grid_base_type *src = in + (s * npix + (uint64_t)sy * nx + sx) * 2;
    grid_base_type re = src[0], im = src[1];
    if (tx) {
        const double t = 1.0 / (tx[x] * ty[y]);
        re = (grid_base_type)(re * t);
        im = (grid_base_type)(im * t);
    }
    grid_base_type *dst = out + (((uint64_t)f * fstride + k) * npix + pix) * 2;
    dst[0] = re;
    dst[1] = im;
}

// Real parts of the first n slices of each facet block -> float32 [F][n][npix]
__global__ void k_extract_real(const grid_base_type *in, float *tmp, uint32_t F, uint32_t n,
                               uint32_t fstride, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * n * npix) return;
    const uint64_t fn = (uint64_t)n * npix;
    const uint64_t f = i / fn, j = i - f * fn;
    tmp[i] = (float)in[(f * fstride * npix + j) * 2];
}

// float32 images written from float index f*fstride*npix (the in-place layout of
// fft_and_repacking_routines.cpp:106-116)
__global__ void k_place_real(const float *tmp, float *out, uint32_t F, uint32_t n,
                             uint32_t fstride, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * n * npix) return;
    const uint64_t fn = (uint64_t)n * npix;
    const uint64_t f = i / fn, j = i - f * fn;
    out[f * fstride * npix + j] = tmp[i];
}

static inline unsigned nblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t grid_to_abi(const double *g, void *out, uint64_t slices_fc, uint32_t P,
                       uint32_t pstride, uint64_t npix, hipStream_t st) {
    if (!slices_fc || !npix) return hipSuccess;
    k_grid_to_abi<<<nblk(slices_fc * npix), 256, 0, st>>>(g, (grid_base_type *)out, slices_fc, P,
                                                         pstride, npix);
    return hipGetLastError();
}

hipError_t abi_to_grid(const void *in, double *g, uint64_t slices_fc, uint32_t P,
                       uint32_t pstride, uint64_t npix, hipStream_t st) {
    if (!slices_fc || !npix) return hipSuccess;
    k_abi_to_grid<<<nblk(slices_fc * npix), 256, 0, st>>>((const grid_base_type *)in, g,
                                                         slices_fc, P, pstride, npix);
    return hipGetLastError();
}

// Sums the slot copies of an array into slot 0 and clears the others.
__global__ void k_fold_slots(double *p, uint64_t stride, uint32_t nslots) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= stride) return;
    double s = p[i];
    for (uint32_t j = 1; j < nslots; ++j) {
        s += p[(uint64_t)j * stride + i];
        p[(uint64_t)j * stride + i] = 0.0;
    }
    p[i] = s;
}

hipError_t fold_slots(double *p, uint64_t stride, uint32_t nslots, hipStream_t st) {
    if (!stride || nslots < 2) return hipSuccess;
    k_fold_slots<<<nblk(stride), 256, 0, st>>>(p, stride, nslots);
    return hipGetLastError();
}

hipError_t fill_f64(double *p, uint64_t n, double v, hipStream_t st) {
    if (!n) return hipSuccess;
    k_fill_f64<<<nblk(n), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

// Stokes formation (bullseye/data/stokes.py:90-172): for every (facet, channel) the value the
// pipeliner stores in correlation slot 0, (g[a] + sign * g[b]) / 2 in the grid type, or slot 0
// itself when first < 0.  tmp: [facet][cube][npix] complex.
__global__ void k_stokes_combine(const grid_base_type *g, grid_base_type *tmp, uint32_t F,
                                 uint32_t C, uint32_t P, uint64_t npix, int first, int second,
                                 int sign) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)F * C * npix) return;
    const uint64_t fc = i / npix, px = i - fc * npix;
    const grid_base_type *base = g + 2 * (fc * P * npix + px);
    grid_base_type re, im;
    if (first < 0) {
        re = base[0];
        im = base[1];
    } else {
        const grid_base_type *a = base + 2 * (uint64_t)first * npix;
        const grid_base_type *b = base + 2 * (uint64_t)second * npix;
        re = sign > 0 ? a[0] + b[0] : a[0] - b[0];
        im = sign > 0 ? a[1] + b[1] : a[1] - b[1];
        re /= (grid_base_type)2;
        im /= (grid_base_type)2;
    }
    tmp[2 * i] = re;
    tmp[2 * i + 1] = im;
}

// The pipeliner's result after Stokes formation and compaction (bullseye_pipeliner.py:406-416),
// decided per destination slot t of a facet's [cube][P] block: with compaction, slot (P-1)*c
// receives channel c's Stokes value (for P = 1 every channel targets slot 0 and the last one
// stays); otherwise slot 0 of channel c (t = P*c) receives it; every other slot is unchanged.
// contiguous = 1 places channel c at slot c instead.
__global__ void k_stokes_place(grid_base_type *g, const grid_base_type *tmp, uint32_t F,
                               uint32_t C, uint32_t P, uint64_t npix, int compact,
                               int contiguous) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t slots = (uint64_t)C * P;
    if (i >= (uint64_t)F * slots * npix) return;
    const uint64_t fs = i / npix, px = i - fs * npix;
    const uint64_t f = fs / slots, t = fs - f * slots;
    int64_t c = -1;
    if (compact) {
        if (contiguous) {
            if (t < C) c = (int64_t)t;
        } else if (P == 1) {
            if (t == 0) c = (int64_t)C - 1;
        } else if (t % (P - 1) == 0 && t / (P - 1) < C) {
            c = (int64_t)(t / (P - 1));
        }
    }
    if (c < 0 && t % P == 0) c = (int64_t)(t / P);
    if (c < 0) return;
    const uint64_t src = 2 * ((f * C + (uint64_t)c) * npix + px);
    g[2 * i] = tmp[src];
    g[2 * i + 1] = tmp[src + 1];
}

hipError_t stokes_device(void *g, void *tmp, uint32_t F, uint32_t C, uint32_t P, uint64_t npix,
                         int first, int second, int sign, int compact, int contiguous,
                         hipStream_t st) {
    const uint64_t n1 = (uint64_t)F * C * npix, n2 = n1 * P;
    if (!n1) return hipSuccess;
    k_stokes_combine<<<nblk(n1), 256, 0, st>>>((const grid_base_type *)g, (grid_base_type *)tmp,
                                               F, C, P, npix, first, second, sign);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_stokes_place<<<nblk(n2), 256, 0, st>>>((grid_base_type *)g, (const grid_base_type *)tmp, F,
                                             C, P, npix, compact, contiguous);
    return hipGetLastError();
}

hipError_t normalize_device(void *out, const double *norm, uint64_t slices, uint64_t npix,
                            hipStream_t st) {
    if (!slices || !npix) return hipSuccess;
    k_normalize<<<nblk(slices * npix), 256, 0, st>>>((grid_base_type *)out, norm, slices, npix);
    return hipGetLastError();
}

hipError_t weight_uniformly_device(void *out, const void *sf, const uint64_t *cgi, uint32_t F,
                                   uint32_t C, uint32_t P, uint32_t sfC, uint64_t npix,
                                   hipStream_t st) {
    const uint64_t n = (uint64_t)F * C * npix;
    if (!n) return hipSuccess;
    k_weight_uniformly<<<nblk(n), 256, 0, st>>>((grid_base_type *)out,
                                               (const grid_base_type *)sf, cgi, F, C, P, sfC,
                                               npix);
    return hipGetLastError();
}

// In place on the device buffer `buf` (complex<grid_base_type>, facet blocks of fstride
// slices): shifts + unnormalised backward FFT + optional taper, then (if place_real) the
// float32 real parts at the reference's positions.  scratch: F*n*npix complex;
// rtmp: F*n*npix float.
hipError_t finalize_device(void *buf, uint32_t F, uint32_t n, uint32_t fstride, uint32_t nx,
                           uint32_t ny, const double *tx, const double *ty, void *scratch,
                           float *rtmp, bool place_real, hipStream_t st) {
    const uint64_t npix = (uint64_t)nx * ny;
    const uint64_t total = (uint64_t)F * n * npix;
    if (!total) return hipSuccess;
    grid_base_type *b = (grid_base_type *)buf;
    grid_base_type *s = (grid_base_type *)scratch;
    k_ifftshift_gather<<<nblk(total), 256, 0, st>>>(b, s, F, n, fstride, nx, ny);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipfftHandle plan;
    int dims[2] = {(int)ny, (int)nx};
#if defined(BULLSEYE_DOUBLE)
    const hipfftType type = HIPFFT_Z2Z;
#else
    const hipfftType type = HIPFFT_C2C;
#endif
    if (hipfftPlanMany(&plan, 2, dims, nullptr, 1, (int)npix, nullptr, 1, (int)npix, type,
                       (int)(F * n)) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    hipfftSetStream(plan, st);
#if defined(BULLSEYE_DOUBLE)
    hipfftResult r = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)s, (hipfftDoubleComplex *)s,
                                   HIPFFT_BACKWARD);
#else
    hipfftResult r = hipfftExecC2C(plan, (hipfftComplex *)s, (hipfftComplex *)s, HIPFFT_BACKWARD);
#endif
    if (r != HIPFFT_SUCCESS) {
        hipfftDestroy(plan);
        return hipErrorUnknown;
    }
    k_fftshift_scatter<<<nblk(total), 256, 0, st>>>(s, b, F, n, fstride, nx, ny, tx, ty);
    e = hipGetLastError();
    if (e == hipSuccess && place_real) {
        k_extract_real<<<nblk(total), 256, 0, st>>>(b, rtmp, F, n, fstride, npix);
        k_place_real<<<nblk(total), 256, 0, st>>>(rtmp, (float *)buf, F, n, fstride, npix);
        e = hipGetLastError();
    }
    hipError_t es = hipStreamSynchronize(st);
    hipfftDestroy(plan);
    return e != hipSuccess ? e : es;
}

}  // namespace bam
