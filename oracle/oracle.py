"""ctypes binding of the C oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; the
product library never imports it.  The oracle takes the same gridding_parameters struct as the
HIP library (double precision build for parity, single for the CPU baseline timing).
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBS = {"double": "liboracle_imaging64.so", "single": "liboracle_imaging32.so"}
MODES = {
    "grid_single_pol": 0, "facet_single_pol": 1, "grid_duel_pol": 2, "facet_duel_pol": 3,
    "grid_4_cor": 4, "facet_4_cor": 5, "facet_4_cor_corrections": 6,
    "grid_sampling_function": 7, "facet_sampling_function": 8,
}
_LOADED = {}


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def load(precision="double"):
    if precision not in _LOADED:
        path = os.path.join(ORACLE_DIR, LIBS[precision])
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.oracle_gridding_walltime.restype = ctypes.c_double
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_baseline_index.restype = ctypes.c_long
        lib.oracle_baseline_index.argtypes = [ctypes.c_long] * 3
        _LOADED[precision] = lib
    return _LOADED[precision]


def run(binding, mode, precision="double", normalize=True, weight_uniformly=False):
    """Grid binding (a bullseye_amd.binding.Binding built for `precision`) with the oracle.
    Returns (grid, normalization_terms) as NumPy arrays."""
    lib = load(precision)
    p = binding.params
    lib.oracle_init(binding.byref())
    nnorm = p.num_facet_centres * p.cube_channel_dim_size * \
        p.number_of_polarization_terms_being_gridded
    saved_jones = p.jones_terms
    if mode == "facet_4_cor_corrections":
        # the reference inverts jones_terms in place (cpu_algorithm/wrapper.cpp:356-360); do it
        # on a private copy so the caller's array stays usable as an input elsewhere
        jcopy = np.array(binding.keep["jones_terms"], copy=True)
        p.jones_terms = jcopy.ctypes.data
    rc = lib.oracle_grid(binding.byref(), MODES[mode])
    p.jones_terms = saved_jones
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    norm = np.ctypeslib.as_array(
        ctypes.cast(p.normalization_terms, ctypes.POINTER(ctypes.c_double)), (nnorm,)).copy()
    if weight_uniformly:
        lib.oracle_weight_uniformly(binding.byref())
    if normalize and not mode.endswith("sampling_function"):
        lib.oracle_normalize(binding.byref())
    return norm


def degrid(binding, mode, model, precision="double"):
    """The oracle's degridding adjoint of `mode` on binding's inputs (uvw, weights, flags,
    Jones, facets, kernel): model has output_buffer's shape [F, cube, P, ny, nx]; returns the
    complex visibilities [rows, chan, npol]."""
    lib = load(precision)
    p = binding.params
    dt = np.complex128 if precision == "double" else np.complex64
    m = np.ascontiguousarray(model, dtype=dt)
    want = (p.num_facet_centres * p.cube_channel_dim_size *
            p.number_of_polarization_terms_being_gridded * p.ny * p.nx)
    if m.size != want:
        raise ValueError("model has %d elements, the grids %d" % (m.size, want))
    out = np.zeros((p.row_count, p.channel_count, p.number_of_polarization_terms), dtype=dt)
    rc = lib.oracle_degrid(binding.byref(), MODES[mode], m.ctypes.data_as(ctypes.c_void_p),
                           out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return out


def pipeliner_stokes(gridded_vis, jones_corrections, data_correlations, gridded_correlations,
                     feeds, wanted):
    """The pipeliner's steps between normalize and finalize on a NumPy copy of the grids
    [F, cube, P, ny, nx]: bullseye/data/stokes.py:90-172 (slot 0 of every (facet, channel) becomes
    (slot a +- slot b) / 2 for the feed type and wanted term; nothing when the feed type is
    undetermined, as it is with Jones corrections) and the compaction of
    bullseye/bin/bullseye_pipeliner.py:406-416 (for I, Q, U, V: channel c's slot 0 is copied, in
    channel order, to [f, nc, nf] with nf, nc = divmod-like split of (P - 1) * c by P)."""
    g = np.array(gridded_vis, copy=True)
    codes = {"I": 1, "Q": 2, "U": 3, "V": 4, "RR": 5, "RL": 6, "LR": 7, "LL": 8,
             "XX": 9, "XY": 10, "YX": 11, "YY": 12}
    pairs = {("circular", "I"): ("RR", "LL", 1), ("circular", "V"): ("RR", "LL", -1),
             ("circular", "Q"): ("RL", "LR", 1), ("circular", "U"): ("RL", "LR", -1),
             ("linear", "I"): ("XX", "YY", 1), ("linear", "Q"): ("XX", "YY", -1),
             ("linear", "U"): ("XY", "YX", 1), ("linear", "V"): ("XY", "YX", -1)}
    order = list(data_correlations) if jones_corrections else list(gridded_correlations)
    if (feeds, wanted) in pairs:
        a, b, sign = pairs[(feeds, wanted)]
        ia, ib = order.index(codes[a]), order.index(codes[b])
        two = g.real.dtype.type(2)
        g[:, :, 0] = (g[:, :, ia] + g[:, :, ib]) / two if sign > 0 else \
            (g[:, :, ia] - g[:, :, ib]) / two
    if wanted in ("I", "Q", "U", "V"):
        P = len(gridded_correlations)
        for f in range(g.shape[0]):
            for c in range(g.shape[1]):
                shift = (P - 1) * c
                g[f, shift // P, shift % P] = g[f, c, 0]
    return g


def baseline_starts(ant1, ant2, n_ant):
    """Prefix scan of rows per baseline and the stable per-baseline row order
    (gpu_algorithm/wrapper.cu:358-454)."""
    lib = load("double")
    rows = ant1.size
    nbl = n_ant * (n_ant + 1) // 2
    a1 = np.ascontiguousarray(ant1, dtype=np.uint32)
    a2 = np.ascontiguousarray(ant2, dtype=np.uint32)
    starts = np.zeros(nbl + 1, dtype=np.uint64)
    perm = np.zeros(rows, dtype=np.uint64)
    lib.oracle_baseline_starts(a1.ctypes.data_as(ctypes.c_void_p),
                               a2.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(rows),
                               ctypes.c_size_t(n_ant), ctypes.c_size_t(nbl),
                               starts.ctypes.data_as(ctypes.c_void_p),
                               perm.ctypes.data_as(ctypes.c_void_p))
    return starts, perm


def fftshift_inplace(grid, inverse=False, precision="double"):
    """The reference's in-place (i)fftshift (fft_shift_utils.cpp:41-112) on a complex array
    [..., ny, nx]."""
    lib = load(precision)
    dt = np.complex128 if precision == "double" else np.complex64
    g = np.array(grid, dtype=dt, copy=True, order="C")
    ny, nx = g.shape[-2:]
    slices = g.size // (nx * ny)
    fn = lib.oracle_ifftshift if inverse else lib.oracle_fftshift
    fn(g.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nx), ctypes.c_size_t(ny),
       ctypes.c_size_t(slices))
    return g


def finalize_images(grid, n_slices_per_facet):
    """Restatement of ifft_machine::repack_and_ifft_uv_grids (fft_and_repacking_routines.cpp:
    83-118) in NumPy: per facet, the first n_slices_per_facet slices go through ifftshift,
    FFTW's unnormalised BACKWARD transform (= N * ifft2) and fftshift; the real parts are the
    float32 images.  Returns float32 [facet, n_slices, ny, nx]."""
    g = np.asarray(grid)
    F = g.shape[0]
    ny, nx = g.shape[-2:]
    g = g.reshape(F, -1, ny, nx)[:, :n_slices_per_facet]
    img = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g, axes=(-2, -1)), axes=(-2, -1)) *
                          (nx * ny), axes=(-2, -1))
    return img.real.astype(np.float32)
