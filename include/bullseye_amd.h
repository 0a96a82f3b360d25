/*
 * bullseye_amd.h -- C ABI of the MI355X (gfx950) imaging library.
 *
 * The library is a drop-in for bullseye's imaging libraries
 * (bullseye/mo/cpu_gpu_common/wrapper.h:41-61): the same 19 extern "C" symbols taking the
 * same `gridding_parameters` struct (bullseye/mo/cpu_gpu_common/gridding_parameters.h:44-109),
 * field for field, so bullseye's ctypes mirror (bullseye/mo/gridding_parameters.py:46-112)
 * drives it unchanged.  Three builds come from one source tree:
 *
 *   libmi355x_imaging32.so     -DBULLSEYE_SINGLE  struct/pointee types of libgpu_imaging32.so
 *   libmi355x_imaging64.so     -DBULLSEYE_DOUBLE  struct/pointee types of libgpu_imaging64.so
 *   libmi355x_imaging_mixed.so -DBULLSEYE_MIXED   uvw/geometry float64, data float32
 *                                                  (struct layout identical to DOUBLE)
 *
 * Precision typedefs follow bullseye/mo/cpu_gpu_common/base_types.h:41-59; MIXED is new.
 *
 * Every entry point returns an int status (0 = success).  The reference declares them void;
 * its callers ignore the return value (bullseye/bin/bullseye_pipeliner.py:306-424 leaves
 * ctypes' default int restype), so an int return is invisible to them.  On failure the
 * message is available from bullseye_last_error(); nothing throws across the ABI.
 */
#ifndef BULLSEYE_AMD_H
#define BULLSEYE_AMD_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#if defined(BULLSEYE_SINGLE)
typedef float visibility_base_type;
typedef float uvw_base_type;
typedef float reference_wavelengths_base_type;
typedef float convolution_base_type;
typedef float visibility_weights_base_type;
typedef float grid_base_type;
typedef double normalization_base_type;
#elif defined(BULLSEYE_DOUBLE)
typedef double visibility_base_type;
typedef double uvw_base_type;
typedef double reference_wavelengths_base_type;
typedef double convolution_base_type;
typedef double visibility_weights_base_type;
typedef double grid_base_type;
typedef double normalization_base_type;
#elif defined(BULLSEYE_MIXED)
typedef float visibility_base_type;
typedef double uvw_base_type;
typedef double reference_wavelengths_base_type;
typedef float convolution_base_type;
typedef float visibility_weights_base_type;
typedef float grid_base_type;
typedef double normalization_base_type;
#else
#error "define one of BULLSEYE_SINGLE, BULLSEYE_DOUBLE, BULLSEYE_MIXED"
#endif

/*
 * Field order and widths are those of bullseye/mo/cpu_gpu_common/gridding_parameters.h:44-109.
 * Pointee layouts (C order):
 *   visibilities           complex<visibility_base_type>[row][channel][correlation]
 *   uvw_coords             uvw_base_type[row][3]          metres (u, v, w)
 *   reference_wavelengths  reference_wavelengths_base_type[spw][channel]   metres
 *   visibility_weights     visibility_weights_base_type[row][channel][correlation]
 *   flags                  bool[row][channel][correlation]
 *   flagged_rows, field_array, spw_index_array: per row
 *   conv  AA (wplanes <= 1): convolution_base_type[(2S+3)+(2S+2)(OS-1)]
 *         w-projection:      complex<convolution_base_type>[wplanes][csz][csz]
 *   output_buffer          complex<grid_base_type>[facet][cube channel][pol gridded][ny][nx]
 *   facet_centres          uvw_base_type[facet][2]        RA, Dec in arcsec
 *   jones_terms            complex<visibility_base_type>[time][antenna][facet][spw][channel][4]
 *   sampling_function_buffer complex<grid_base_type>[facet][sf channel][ny][nx]
 *   baseline_starting_indexes size_t[baseline_count + 1]
 *   normalization_terms    written by initLibrary: normalization_base_type[facet][cube][pol]
 */
typedef struct gridding_parameters {
    void *visibilities;
    uvw_base_type *uvw_coords;
    reference_wavelengths_base_type *reference_wavelengths;
    visibility_weights_base_type *visibility_weights;
    bool *flags;
    bool *flagged_rows;
    unsigned int *field_array;
    unsigned int *spw_index_array;
    unsigned int imaging_field;
    size_t baseline_count;
    size_t row_count;
    size_t chunk_max_row_count;
    size_t channel_count;
    size_t number_of_polarization_terms;
    size_t number_of_polarization_terms_being_gridded;
    size_t spw_count;
    size_t no_timestamps_read;
    size_t nx;
    size_t ny;
    uvw_base_type cell_size_x;
    uvw_base_type cell_size_y;
    convolution_base_type *conv;
    size_t conv_support;
    size_t conv_oversample;
    size_t polarization_index;
    size_t second_polarization_index;
    void *output_buffer;
    uvw_base_type phase_centre_ra;
    uvw_base_type phase_centre_dec;
    uvw_base_type *facet_centres;
    size_t num_facet_centres;
    void *jones_terms;
    bool should_invert_jones_terms;
    unsigned int *antenna_1_ids;
    unsigned int *antenna_2_ids;
    size_t *timestamp_ids;
    size_t antenna_count;
    bool *enabled_channels;
    size_t *channel_grid_indicies;
    size_t cube_channel_dim_size;
    bool should_grid_sampling_function;
    void *sampling_function_buffer;
    size_t *sampling_function_channel_grid_indicies;
    size_t sampling_function_channel_count;
    bool is_final_data_chunk;
    size_t wplanes;
    uvw_base_type wmax_est;
    size_t *baseline_starting_indexes;
    size_t *antenna_jones_starting_indexes;
    size_t *jones_time_indicies_per_antenna;
    normalization_base_type *normalization_terms;
} gridding_parameters;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the 19 entry points of bullseye/mo/cpu_gpu_common/wrapper.h:41-61 ---------------- */

/* wrapper.h:42 / cpu_algorithm/wrapper.cpp:66-69: seconds spent in grid_* and facet_* calls
 * (device time of the gridding stream, hipEvent-timed). */
double get_gridding_walltime(void);
/* wrapper.h:43 / cpu_algorithm/wrapper.cpp:70-73: seconds spent in finalize/finalize_psf. */
double get_inversion_walltime(void);
/* wrapper.h:44 / cpu_algorithm/wrapper.cpp:74-78: block until queued gridding has finished;
 * copies results to host buffers if the final chunk was gridded. */
void gridding_barrier(void);
/* wrapper.h:45 / cpu_algorithm/wrapper.cpp:79-106, gpu_algorithm/wrapper.cu:81-232:
 * one-time set-up; allocates normalization_terms and writes the pointer into params. */
int initLibrary(gridding_parameters *params);
/* wrapper.h:46 / cpu_algorithm/wrapper.cpp:107-115 */
int releaseLibrary(void);
/* wrapper.h:47 / cpu_algorithm/wrapper.cpp:116-135: uv-plane uniform weighting from the
 * gridded sampling function. */
int weight_uniformly(gridding_parameters *params);
/* wrapper.h:48 / cpu_algorithm/wrapper.cpp:136-163: grid /= normalization per slice. */
int normalize(gridding_parameters *params);
/* wrapper.h:49 / gpu_algorithm/wrapper.cu:367-454: in-place reorder of the row arrays by
 * baseline (time order kept) and prefix scan into baseline_starting_indexes. */
int repack_input_data(gridding_parameters *params);
/* wrapper.h:50 / fft_and_repacking_routines.cpp:83-118: ifftshift, unnormalised backward
 * c2c FFT, fftshift and float32 real part written in place. */
int finalize(gridding_parameters *params);
/* wrapper.h:51 / fft_and_repacking_routines.cpp:119-150: same for the sampling function. */
int finalize_psf(gridding_parameters *params);
/* wrapper.h:52-60 / cpu_algorithm/wrapper.cpp:182-425: grid one chunk of rows. grid_* grid
 * at the phase centre; facet_* rotate phase and uvw to each facet centre.  These return
 * once the work is queued; call gridding_barrier() before touching the buffers. */
int grid_single_pol(gridding_parameters *params);
int facet_single_pol(gridding_parameters *params);
int grid_duel_pol(gridding_parameters *params);
int facet_duel_pol(gridding_parameters *params);
int grid_4_cor(gridding_parameters *params);
int facet_4_cor(gridding_parameters *params);
int facet_4_cor_corrections(gridding_parameters *params);
int grid_sampling_function(gridding_parameters *params);
int facet_sampling_function(gridding_parameters *params);

/* ---- extensions (no counterpart in the reference) -------------------------------------- */

/* Message of the last failing call ("" if none). */
const char *bullseye_last_error(void);

/* Compute path of the gridding kernels. */
enum bullseye_compute_path {
    BULLSEYE_PATH_DEFAULT = 0, /* FP64 for the DOUBLE build, FP32_KAHAN otherwise */
    BULLSEYE_PATH_FP64 = 1,    /* FP64 geometry, taps, register accumulators and grid */
    BULLSEYE_PATH_FP32_KAHAN = 2, /* FP64 phase argument + uvw geometry, FP32 taps with
                                     compensated register sums, FP64 grid flushes */
    BULLSEYE_PATH_FP32 = 3     /* as FP32_KAHAN without the compensation term */
};

/* Library-wide options.  The reference struct has no spare fields, so options live beside
 * it; callers that never set them get the reference's behaviour. */
typedef struct bullseye_options {
    uint32_t struct_size;       /* sizeof(bullseye_options) */
    uint32_t device_resident;   /* 1: every array pointer in gridding_parameters (inputs,
                                   output_buffer, sampling_function_buffer) is device memory;
                                   no host<->device copies are made */
    uint32_t compute_path;      /* enum bullseye_compute_path */
    uint32_t apply_taper_correction; /* 1: finalize divides each facet image by the grid
                                   taper (DFT of the sampled AA kernel) */
    uint32_t channels_per_stream; /* Romein stream: channels walked together (0 = auto: up to
                                     32 AA / 8 w-projection channels when every enabled channel
                                     maps to one grid slice, else 1; a non-zero value is used
                                     as given, capped at channel_count, by both gridders) */
    uint32_t rows_per_stream;   /* Romein stream: time rows walked per work item (0 = auto) */
    uint32_t walk_order;        /* order of a stream's samples: 0 = whichever moves the kernel
                                   footprint less per step, measured from the data; 1 =
                                   channels inner; 2 = time inner */
    uint32_t reserved[9];       /* reserved[0] = 1 forces the generic (any-support) kernel */
} bullseye_options;

int bullseye_set_options(const bullseye_options *options);
int bullseye_get_options(bullseye_options *options);

/* Per-plane half supports of the w-projection kernels (plane p uses the central
 * 2*half_support[p]+1 taps of its (2S+1)^2 footprint; taps outside are taken as zero).  Call
 * before initLibrary; without it every plane uses conv_support, as in the reference
 * (gridding_parameters.h:71 has a single support for all planes). */
int bullseye_set_wplane_support(const uint32_t *half_support, uint32_t nplanes);

/* Zero the device accumulation grids and normalisation sums so a new imaging pass can
 * start without releaseLibrary/initLibrary. */
int bullseye_reset_grids(void);

/* Counters of the last gridding call (filled after gridding_barrier). */
typedef struct bullseye_stats {
    uint64_t samples_gridded;   /* (facet, row, channel) samples that were in-grid and unflagged */
    uint64_t taps;              /* sum over those samples of the tap count (2S+1)^2 */
    double kernel_ms;           /* device time of the gridding kernel(s) of the call */
    double prep_ms;             /* device time from the call's start to its gridding kernel:
                                   staging copies, Jones inversion, ingest (baseline sort),
                                   row tables and stream classification */
    double copy_ms;             /* device time of the call's host->device staging copies
                                   (0 with device_resident: nothing is copied) */
    uint32_t time_inner;        /* walk order of the call's streams: 1 time inner, 0 channels */
    uint32_t reserved;
} bullseye_stats;
int bullseye_get_stats(bullseye_stats *stats);

/* Measured rate of FP64 atomic adds to device memory on the current device, in lane-operations
 * per second: 2^27 adds of 1.0 at hashed addresses of a 16 MB buffer after one warm-up launch.
 * Every grid flush is such an add, and they execute on the memory side, so this is the ceiling
 * of a flush-bound gridder.  No counterpart in the reference (a measurement for the
 * benchmark's roofline).  Needs no initLibrary; allocates and frees its own buffer. */
int bullseye_fp64_atomic_rate(double *ops_per_second);

/* Gridding modes, numbered in the order of the entry points above. */
enum bullseye_grid_mode {
    BULLSEYE_GRID_SINGLE_POL = 0,
    BULLSEYE_FACET_SINGLE_POL = 1,
    BULLSEYE_GRID_DUEL_POL = 2,
    BULLSEYE_FACET_DUEL_POL = 3,
    BULLSEYE_GRID_4_COR = 4,
    BULLSEYE_FACET_4_COR = 5,
    BULLSEYE_FACET_4_COR_CORRECTIONS = 6,
    BULLSEYE_GRID_SAMPLING_FUNCTION = 7,
    BULLSEYE_FACET_SAMPLING_FUNCTION = 8
};

/* Allocation-free gridding.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Grids one chunk like the entry point numbered `mode` (enum
 * bullseye_grid_mode) with the AA kernel (wplanes <= 1) and device-resident buffers
 * (bullseye_options.device_resident = 1), into the grids initLibrary set up.  The call's own
 * tables (baseline sort, row and channel tables, facet geometry, inverted Jones terms) live in
 * a static 64 MiB device array per GPU that the library holds until the process ends; a call
 * whose tables exceed it is refused (600,000 rows per call fit without Jones terms).  It calls
 * no hipMalloc, hipFree or allocating library routine.  Calls that overlap on one GPU race on
 * that array; bullseye_stats is not updated. */
int bullseye_grid_noalloc(gridding_parameters *params, int mode);

/* Degridding: the adjoint of a gridding entry point.  No counterpart in bullseye's C++ library
 * (writing_and_wiki/notebooks/Resampling.ipynb sketches the gather in Python).  mode is one of
 * the visibility modes 0..6.  model has output_buffer's layout (complex [F][cube][P][ny][nx],
 * grid type); vis_out has the visibilities' layout (complex [rows][chan][npol], visibility
 * type) and is overwritten.  Each gridded correlation receives the transpose, over the reals,
 * of the gridding map applied to model: conjugated kernel taps, the w < 0 conjugation of
 * w-projection, the conjugate facet phase, weights and flags, and Pinv^H S Qinv for the Jones
 * mode; no normalisation.  Every other correlation is zero.  Uses the struct's uvw,
 * wavelengths, weights, flags, facets and Jones and the kernels given to initLibrary, on the
 * compute path of bullseye_options.  Synchronous; with device_resident, model and vis_out are
 * device pointers. */
int bullseye_degrid(gridding_parameters *params, int mode, const void *model, void *vis_out);

/* Stokes formation and compaction of the (normalised) grids in output_buffer, on the device: the
 * pipeliner's steps between normalize and finalize (bullseye/data/stokes.py:90-172 and
 * bullseye/bin/bullseye_pipeliner.py:406-416), which the reference runs on host arrays.  For
 * every (facet, channel), correlation slot 0 becomes (slot[first] + sign * slot[second]) / 2 in
 * the grid type (first < 0: unchanged, the wanted term was gridded directly).  With compact = 1,
 * channel c's slot-0 grid is then copied to flat slot (P-1)*c of its facet's [cube][P] block, the
 * reference's arithmetic (contiguous for P = 2; for P = 1 the last channel ends in slot 0); with
 * contiguous = 1 it goes to slot c instead.  Every other slot is left as it was.  Synchronous;
 * with device_resident, output_buffer is a device pointer and nothing crosses PCIe. */
int bullseye_form_stokes(gridding_parameters *params, int first, int second, int sign,
                         int compact, int contiguous);

/* Direct Fourier transforms in FP64 (bullseye ships an unused one, gpu_algorithm/dft.h:51-83;
 * these follow the conventions the gridder and finalize produce).  Pixel (x, y) of an nx x ny
 * image at cell_size_x/y is the direction l = (x - nx/2) cell_x, m = -(y - ny/2) cell_y (radians
 * through the gridder's float arcsecond constant).
 * bullseye_dft_image: image[y][x] = sum over the unflagged, enabled, in-field samples of
 *   correlation polarization_index of weight * Re(V exp(2 pi i (u l + v m))) -- the kernel-free
 *   image that gridding + finalize approximate (a brute-force cross-check).  image: double
 *   [ny][nx].
 * bullseye_dft_predict: V = sum_k flux_k exp(+2 pi i (u l_k + v m_k + w (n_k - 1))) for every
 *   (row, channel), written to correlation polarization_index of vis_out (visibility type,
 *   [rows][chan][npol]; other correlations untouched).  components: double [n][3] (l, m, flux).
 * Both need initLibrary and use the struct's uvw, wavelengths, weights, flags and selections;
 * with device_resident the array arguments are device pointers.  Synchronous. */
int bullseye_dft_image(gridding_parameters *params, double *image);
int bullseye_dft_predict(gridding_parameters *params, const double *components, uint32_t n,
                         void *vis_out);

/* ---- batched facets (w-projection) ----------------------------------------------------------
 * The reference grids one facet geometry per initLibrary: every facet of a call shares nx, ny,
 * cell size and kernels (gridding_parameters.h:45-72), and the pipeliner loops over calls
 * (bullseye/bin/bullseye_pipeliner.py:350-404).  These entry points grid up to 65536 facets of
 * different size, cell, w-kernel stack and kind (visibilities or sampling function) from one
 * visibility set in a single plan: the facets' streams are merged into one work queue per
 * support class and drained by persistent kernels, captured once in a HIP graph and replayed.
 * Each facet follows the semantics of facet_single_pol / facet_duel_pol / facet_4_cor (or
 * facet_sampling_function) with w-projection, as if initLibrary had been called with its own
 * nx, ny, cell size and kernels.  initLibrary must have been called (any geometry); its stream,
 * options (device_resident, compute_path, walk_order) and error reporting are used. */

/* A w-kernel stack: conv is [wplanes][csz][csz] complex convolution_base_type with
 * csz = (2S+3) + (2S+2)(OS-1), as initLibrary's conv (device memory with device_resident);
 * half_support[wplanes] gives each plane's half support (<= S <= 31), NULL: S for all; a
 * sample's plane is lrint(|w| / wmax * (wplanes - 1)) as in the reference.  id < 64; defining
 * an id again replaces it (and invalidates the plan). */
int bullseye_batch_kernel_set(uint32_t id, const void *conv, uint32_t wplanes,
                              uint32_t conv_support, uint32_t conv_oversample,
                              const uint32_t *half_support, float wmax);

typedef struct bullseye_batch_facet {
    uint32_t nx, ny;            /* facet grid size */
    uint32_t kernel_set;        /* id given to bullseye_batch_kernel_set */
    uint32_t sampling_function; /* 1: grid the facet's sampling function (PSF) */
    double cell_x, cell_y;      /* cell size, arcsec */
    double ra, dec;             /* facet centre, arcsec (as facet_centres) */
    void *output;               /* complex [cube][ncorr][ny][nx] grid_base_type, not normalised
                                   (cube = cube_channel_dim_size, or
                                   sampling_function_channel_count for a PSF facet; ncorr = 1, 2
                                   or 4 by mode, 1 for PSF); device memory with device_resident;
                                   NULL: not copied out */
    double *normalization;      /* host [cube][ncorr] normalisation sums, or NULL */
} bullseye_batch_facet;

enum bullseye_batch_schedule {
    BULLSEYE_BATCH_PERSISTENT_GRAPH = 0, /* one persistent launch per (kind, support class) over
                                            all facets, captured in a HIP graph and replayed */
    BULLSEYE_BATCH_PERSISTENT = 1,       /* the same launches issued on the stream each time */
    BULLSEYE_BATCH_PER_FACET = 2         /* one launch per (facet, support class), issued on the
                                            stream each time: the unbatched baseline */
};

/* Romein tile choices per support class M = 2i+1 (i < 32): shape[i] is the lane shape
 * (0 default, 1 column ownership (M <= 15), 2/3/4 residue pairs with at least 1/2/4 pairs per
 * lane) and chunk[i] the classification segments per work item (1, 2 or 4: the visibility
 * chunk length is chunk * base_rows rows).  channels_per_stream and base_rows are shared by all
 * classes (0: automatic, 32). */
typedef struct bullseye_batch_tuning {
    uint32_t struct_size;       /* sizeof(bullseye_batch_tuning) */
    uint32_t channels_per_stream;
    uint32_t base_rows;
    uint32_t tuned;             /* 1 when filled by bullseye_batch_autotune */
    uint32_t shape[32];
    uint32_t chunk[32];
    uint32_t facet_major;       /* work-list order: 0 = the facets of a stream together (the
                                   stream's visibilities shared in L2), 1 = each facet's streams
                                   together (its grid's atomics local in the memory-side cache) */
    uint32_t reserved;
    double tuned_ms;            /* autotune: summed device time of the chosen launches */
    double default_ms;          /* autotune: the same with the default tiles */
} bullseye_batch_tuning;
int bullseye_batch_set_tuning(const bullseye_batch_tuning *tuning);
int bullseye_batch_get_tuning(bullseye_batch_tuning *tuning);

/* Plan a batch: stage the visibility set of params (rows, uvw, vis, weights, flags, channels,
 * spw, antennas, selections) -- only the struct's visibility fields are read; its nx, ny, cell
 * sizes, facet list and conv are ignored -- sort rows by baseline, classify every stream per
 * kernel set, build the work lists and (schedule 0) capture the graph.  mode: one of
 * BULLSEYE_FACET_SINGLE_POL, BULLSEYE_FACET_DUEL_POL, BULLSEYE_FACET_4_COR for the
 * visibility facets.  The facet table and its output pointers are kept until the next plan.
 * With device_resident the caller's arrays are used in place and must stay allocated; without,
 * they are copied here and later changes are not seen. */
int bullseye_batch_plan(gridding_parameters *params, int mode, const bullseye_batch_facet *facets,
                        uint32_t n, int schedule);
/* Grid the planned batch from zero, copy each facet's grid and normalisation sums out, and fill
 * bullseye_stats (kernel_ms = device time of the gridding launches).  Synchronous. */
int bullseye_batch_execute(void);
/* Time every support class of the batch under each lane shape and chunk length, and each
 * channels-per-stream / base-rows pair (those left 0 in the current tuning), then the two
 * work-list orders at the chosen tiles; keep the fastest (as bullseye_batch_set_tuning) and
 * return it.  Leaves no plan: call bullseye_batch_plan afterwards. */
int bullseye_batch_autotune(gridding_parameters *params, int mode,
                            const bullseye_batch_facet *facets, uint32_t n,
                            bullseye_batch_tuning *result);
/* Free the plan and kernel sets (releaseLibrary does too). */
int bullseye_batch_release(void);

#ifdef __cplusplus
}
#endif
#endif
