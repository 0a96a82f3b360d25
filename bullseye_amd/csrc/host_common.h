// host_common.h -- host-side pieces shared by the library's C ABI translation units
// (wrapper.hip: the reference's entry points; batch.hip: batched facets).
#pragma once
#include <math.h>

#include <string>
#include <vector>

#include "wproj_args.h"

namespace bam {
hipError_t launch_romein_aa(const GridArgs &a, int mode, int path, hipStream_t st);
hipError_t launch_romein_w(const GridArgs &a, int mode, int path, hipStream_t st);
int romein_w_groups_per_wave(int W, int mode, int path);
hipError_t invert_jones(const visibility_base_type *in, void *out, uint64_t n, bool fp64,
                        int *err, hipStream_t st);
hipError_t launch_stream_class(const GridArgs &a, const uint32_t *plane_half, uint32_t wplanes,
                               float wmax_f, uint32_t nseg_max, uint32_t seg_rows, uint8_t *cls,
                               hipStream_t st);
int romein_max_width();
// time_inner: the flag followed by the measurement's partial sums (walk_order_bytes() in all)
size_t walk_order_bytes();
hipError_t launch_walk_order(const GridArgs &a, int override_order, int *time_inner,
                             hipStream_t st);
hipError_t launch_degrid(const GridArgs &a, const void *model, void *out, uint64_t rows,
                         uint32_t P, const void *taps, const uint64_t *plane_off,
                         const uint32_t *plane_half, uint32_t wplanes, uint32_t nf,
                         uint32_t fbase, float wmax_f, int mode, int path, hipStream_t st);
size_t ingest_workspace_bytes(uint32_t rows, uint32_t n_bl);
hipError_t ingest_sort(const uint32_t *a1, const uint32_t *a2, uint32_t rows, uint32_t n_ant,
                       uint32_t n_bl, uint32_t *keys, uint32_t *vals, uint32_t *keys_out,
                       uint32_t *counts, uint32_t *perm, uint32_t *bl_start, void *ws,
                       size_t ws_bytes, int *err, hipStream_t st);
hipError_t grid_to_abi(const double *g, void *out, uint64_t slices_fc, uint32_t P,
                       uint32_t pstride, uint64_t npix, hipStream_t st);
hipError_t abi_to_grid(const void *in, double *g, uint64_t slices_fc, uint32_t P,
                       uint32_t pstride, uint64_t npix, hipStream_t st);
hipError_t fill_f64(double *p, uint64_t n, double v, hipStream_t st);
hipError_t fold_slots(double *p, uint64_t stride, uint32_t nslots, hipStream_t st);
size_t dft_sample_bytes(uint64_t rows, uint32_t chan);
hipError_t launch_dft_image(const GridArgs &a, uint64_t rows, void *samples, uint32_t nx,
                            uint32_t ny, double cell_x, double cell_y, double *image,
                            hipStream_t st);
hipError_t launch_dft_predict(const GridArgs &a, uint64_t rows, const double *comp,
                              uint32_t ncomp, void *vis_out, hipStream_t st);
hipError_t stokes_device(void *g, void *tmp, uint32_t F, uint32_t C, uint32_t P, uint64_t npix,
                         int first, int second, int sign, int compact, int contiguous,
                         hipStream_t st);
hipError_t normalize_device(void *out, const double *norm, uint64_t slices, uint64_t npix,
                            hipStream_t st);
hipError_t weight_uniformly_device(void *out, const void *sf, const uint64_t *cgi, uint32_t F,
                                   uint32_t C, uint32_t P, uint32_t sfC, uint64_t npix,
                                   hipStream_t st);
hipError_t finalize_device(void *buf, uint32_t F, uint32_t n, uint32_t fstride, uint32_t nx,
                           uint32_t ny, const double *tx, const double *ty, void *scratch,
                           float *rtmp, bool place_real, hipStream_t st);
hipError_t fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st);
hipError_t build_row_tables(const GridArgs &a, uint32_t rows, bool jones, RowRec *rowrec,
                            ChanRec *chanrec, hipStream_t st);
hipError_t launch_facet_geom(const uvw_base_type *fc, double ra0, double dec0, uint32_t F,
                             bool facet, bool phase, FacetGeom *out, hipStream_t st);

// cpu_gpu_common/cu_common.h:63 -- a float constant in every precision build
static const float ARCSEC_TO_RAD = (float)(M_PI / (180.0 * 3600.0));

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n ? n : 16);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Record a formatted error message for bullseye_last_error; returns 1.
int fail(const char *fmt, ...);

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail("%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

// The library context as the batch code needs it.
struct LibView {
    bool initialized;
    hipStream_t stream;
    bool device_resident;
    int path;            // resolved compute path
    int walk_order;      // bullseye_options::walk_order
    int *errflag;        // device error bits
    unsigned long long *counters;  // device [nslots][2] samples, taps
    uint32_t nslots;
    unsigned long long *last_counters;  // host totals at the last stats update [2]
    bullseye_stats *stats;
    double *grid_seconds;               // get_gridding_walltime accumulator
};
LibView lib_view();
// Caller array -> device (a staging copy on the library stream, or the pointer itself when
// device resident).
hipError_t stage(DevBuf &b, const void *src, size_t bytes, const void **out);
// Read and clear the device error bits; 1 with a message when any was set.
int check_device_errors();
// Re-laid w-projection taps of a kernel stack (conv as initLibrary's).
int build_wtaps(const void *conv, bool conv_on_device, size_t P, size_t S, size_t OS,
                const std::vector<uint32_t> &halves, DevBuf &taps_f, DevBuf &taps_d,
                DevBuf &plane_off, DevBuf &plane_half, DevBuf &plane_norm, DevBuf &ctaps_f,
                DevBuf &ctaps_d, DevBuf &cplane_off, uint32_t &w_nf, uint32_t &w_fbase);
// Free batch plans and kernel sets (releaseLibrary).
void batch_release_all();

}  // namespace bam
