// dft.hip -- direct Fourier transforms in FP64: the brute-force cross-checks of gridding and
// degridding, and a point-component visibility predictor.
//
// bullseye ships an unused direct-FT kernel (gpu_algorithm/dft.h:51-83); these follow the
// conventions the gridder and finalize actually produce instead.  A sample at (u, v) in
// wavelengths lands at grid cell u*nx*cell (x axis) and -v*ny*cell (y axis), and finalize's
// unnormalised backward FFT turns grid cell offset g into exp(+2 pi i k g / n) at pixel offset
// k, so pixel (x, y) of a facet image is the direction
//   l_x = (x - nx/2) * cell_x,   m_y = -(y - ny/2) * cell_y   (radians; the gridder's float
//                                                               arcsecond constant)
// and the kernel-free image is sum_s weight_s * Re(V_s exp(2 pi i (u_s l_x + v_s m_y))).
// The predictor evaluates V = sum_k S_k exp(+2 pi i (u l_k + v m_k + w (n_k - 1))).
//
// Both loop over every (sample, pixel/component) pair: samples (components) are staged through
// LDS in tiles, one thread per pixel (sample).  FP64 rotors by sincospi of the argument in
// turns reduced to [-1, 1).
#include "device_common.h"

namespace bam {
namespace dft {

constexpr int BLOCK = 256;
constexpr int TILE = 256;

struct Sample {
    double u, v, re, im;   // u, v in wavelengths; weighted visibility (zero when dropped)
};

// One thread per (row, channel): the weighted single-correlation visibility and its uv.
__global__ void k_samples(GridArgs a, uint64_t rows, Sample *out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * a.chan) return;
    const uint32_t row = (uint32_t)(i / a.chan), c = (uint32_t)(i % a.chan);
    Sample s = {0, 0, 0, 0};
    const uint32_t spw = a.spw[row];
    if (spw < a.spw_count) {
        const uint32_t fc = spw * a.chan + c;
        const size_t idx = ((size_t)row * a.chan + c) * a.npol + a.pol_idx[0];
        const bool keep = a.enabled[fc] && !a.row_flags[row] && a.field[row] == a.imaging_field &&
                          !a.flags[idx];
        const double rw = 1.0 / (double)a.wavelengths[fc];
        s.u = (double)a.uvw[3 * (size_t)row] * rw;
        s.v = (double)a.uvw[3 * (size_t)row + 1] * rw;
        if (keep) {
            const double wt = (double)a.weights[idx];
            s.re = wt * (double)a.vis[2 * idx];
            s.im = wt * (double)a.vis[2 * idx + 1];
        }
    } else {
        atomicOr(a.error_flag, 4);
    }
    out[i] = s;
}

__device__ __forceinline__ void rotor(double turns, double &s, double &c) {
    sincospi(2.0 * (turns - rint(turns)), &s, &c);
}

// image[y][x] = sum_s Re(V_s exp(2 pi i (u_s l_x + v_s m_y)))
__global__ __launch_bounds__(BLOCK) void k_image(const Sample *smp, uint64_t n, uint32_t nx,
                                                 uint32_t ny, double cell_x, double cell_y,
                                                 double *image) {
    __shared__ Sample tile[TILE];
    const uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool on = p < (uint64_t)nx * ny;
    const uint32_t x = on ? (uint32_t)(p % nx) : 0, y = on ? (uint32_t)(p / nx) : 0;
    const double l = ((double)x - (double)(nx / 2)) * cell_x;
    const double m = -((double)y - (double)(ny / 2)) * cell_y;
    double acc = 0;
    for (uint64_t t0 = 0; t0 < n; t0 += TILE) {
        __syncthreads();
        if (t0 + threadIdx.x < n) tile[threadIdx.x] = smp[t0 + threadIdx.x];
        __syncthreads();
        const int cnt = (int)min((uint64_t)TILE, n - t0);
        for (int k = 0; k < cnt; ++k) {
            const Sample s = tile[k];
            double sn, cs;
            rotor(s.u * l + s.v * m, sn, cs);
            acc += s.re * cs - s.im * sn;
        }
    }
    if (on) image[p] = acc;
}

// vis = sum_k S_k exp(+2 pi i (u l_k + v m_k + w (n_k - 1))) into correlation pol of vis_out
__global__ __launch_bounds__(BLOCK) void k_predict(GridArgs a, uint64_t rows,
                                                   const double *comp, uint32_t ncomp,
                                                   visibility_base_type *vis_out) {
    __shared__ double tile[TILE][4];
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool on = i < rows * a.chan;
    const uint32_t row = on ? (uint32_t)(i / a.chan) : 0, c = on ? (uint32_t)(i % a.chan) : 0;
    double u = 0, v = 0, w = 0;
    bool ok = on;
    if (on) {
        const uint32_t spw = a.spw[row];
        if (spw >= a.spw_count) {
            atomicOr(a.error_flag, 4);
            ok = false;
        } else {
            const double rw = 1.0 / (double)a.wavelengths[spw * a.chan + c];
            u = (double)a.uvw[3 * (size_t)row] * rw;
            v = (double)a.uvw[3 * (size_t)row + 1] * rw;
            w = (double)a.uvw[3 * (size_t)row + 2] * rw;
        }
    }
    double re = 0, im = 0;
    for (uint32_t t0 = 0; t0 < ncomp; t0 += TILE) {
        __syncthreads();
        if (t0 + threadIdx.x < ncomp) {
            const double *q = comp + 3 * (size_t)(t0 + threadIdx.x);
            tile[threadIdx.x][0] = q[0];
            tile[threadIdx.x][1] = q[1];
            tile[threadIdx.x][2] = q[2];
            tile[threadIdx.x][3] = sqrt(1.0 - q[0] * q[0] - q[1] * q[1]) - 1.0;
        }
        __syncthreads();
        const int cnt = (int)min((uint32_t)TILE, ncomp - t0);
        for (int k = 0; k < cnt; ++k) {
            double sn, cs;
            rotor(u * tile[k][0] + v * tile[k][1] + w * tile[k][3], sn, cs);
            re += tile[k][2] * cs;
            im += tile[k][2] * sn;
        }
    }
    if (ok) {
        const size_t idx = ((size_t)row * a.chan + c) * a.npol + a.pol_idx[0];
        vis_out[2 * idx] = (visibility_base_type)re;
        vis_out[2 * idx + 1] = (visibility_base_type)im;
    }
}

static inline unsigned nblk(uint64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace dft

size_t dft_sample_bytes(uint64_t rows, uint32_t chan) { return rows * chan * sizeof(dft::Sample); }

hipError_t launch_dft_image(const GridArgs &a, uint64_t rows, void *samples, uint32_t nx,
                            uint32_t ny, double cell_x, double cell_y, double *image,
                            hipStream_t st) {
    const uint64_t n = rows * a.chan;
    dft::Sample *s = (dft::Sample *)samples;
    if (n) {
        dft::k_samples<<<dft::nblk(n), dft::BLOCK, 0, st>>>(a, rows, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    dft::k_image<<<dft::nblk((uint64_t)nx * ny), dft::BLOCK, 0, st>>>(s, n, nx, ny, cell_x,
                                                                     cell_y, image);
    return hipGetLastError();
}

hipError_t launch_dft_predict(const GridArgs &a, uint64_t rows, const double *comp,
                              uint32_t ncomp, void *vis_out, hipStream_t st) {
    const uint64_t n = rows * a.chan;
    if (!n) return hipSuccess;
    dft::k_predict<<<dft::nblk(n), dft::BLOCK, 0, st>>>(a, rows, comp, ncomp,
                                                       (visibility_base_type *)vis_out);
    return hipGetLastError();
}

}  // namespace bam
