"""Batched facets of uneven size, cell and w-kernel stack through the library's bullseye_batch_*
entry points.

bullseye's pipeliner images one facet geometry per initLibrary and loops over calls
(bullseye/bin/bullseye_pipeliner.py:350-404).  A FacetBatch grids many facets -- each with its
own grid size, cell size, centre, w-kernel stack and kind (visibilities or sampling function)
-- from the visibility set of one Binding in one plan: persistent class launches over a merged
work queue, captured once in a HIP graph and replayed.

    lib.initLibrary(binding.byref())            # any geometry: stream, options, errors
    batch = FacetBatch(lib, binding, "facet_single_pol")
    batch.kernel_set(0, conv, support=31, oversample=8, half_support=halves, wmax=wmax)
    batch.plan([BatchFacet(512, 1.5, (ra, dec), 0), BatchFacet(128, 1.5, (ra, dec), 0, True)])
    grids, norms = batch.execute()              # per facet [cube][ncorr][ny][nx], [cube][ncorr]
"""
import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from .base_types import force_precision
from .library_loader import (BATCH_PERSISTENT_GRAPH, bullseye_batch_facet,
                             bullseye_batch_tuning)

GRID_MODES = {"facet_single_pol": 1, "facet_duel_pol": 3, "facet_4_cor": 5}
NCORR = {"facet_single_pol": 1, "facet_duel_pol": 2, "facet_4_cor": 4}


@dataclass
class BatchFacet:
    npix: int                       # nx (and ny unless given)
    cell: float                     # arcsec
    centre: Tuple[float, float]     # (ra, dec) arcsec
    kernel_set: int
    sampling_function: bool = False
    ny: Optional[int] = None
    cell_y: Optional[float] = None


class FacetBatch:
    def __init__(self, lib, binding, mode="facet_single_pol"):
        if mode not in GRID_MODES:
            raise ValueError("batch mode must be one of %s" % sorted(GRID_MODES))
        self.lib, self.binding, self.mode = lib, binding, mode
        self.bt = force_precision(binding.precision)
        self.device = binding.device
        self.facets = []
        self.outputs, self.norms = [], []
        self._table = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.bullseye_last_error().decode())

    def kernel_set(self, kid, conv, support, oversample, half_support=None, wmax=1.0):
        """bullseye_batch_kernel_set: conv [wplanes][csz][csz] complex, in the build's w-kernel
        type (device memory when the binding is device resident)."""
        conv = np.ascontiguousarray(np.asarray(conv).astype(self.bt.w_fir_type))
        keep = conv
        ptr = conv.ctypes.data
        if self.device is not None:
            import torch
            keep = torch.from_numpy(conv).to(self.device)
            ptr = keep.data_ptr()
        halves = None
        if half_support is not None:
            halves = (ctypes.c_uint32 * len(half_support))(*[int(h) for h in half_support])
        self._check(self.lib.bullseye_batch_kernel_set(
            kid, ctypes.c_void_p(ptr), conv.shape[0], support, oversample, halves,
            ctypes.c_float(wmax)))
        del keep

    def _cube(self, psf):
        p = self.binding.params
        return max(1, int(p.sampling_function_channel_count if psf else p.cube_channel_dim_size))

    def plan(self, facets: Sequence[BatchFacet], schedule=BATCH_PERSISTENT_GRAPH):
        """Allocate each facet's output grid ([cube][ncorr][ny][nx], grid type) and normalisation
        sums, then bullseye_batch_plan."""
        self.facets = list(facets)
        n = len(self.facets)
        table = (bullseye_batch_facet * n)()
        self.outputs, self.norms = [], []
        for i, f in enumerate(self.facets):
            ny = f.ny or f.npix
            nc = 1 if f.sampling_function else NCORR[self.mode]
            cube = self._cube(f.sampling_function)
            shape = (cube, nc, ny, f.npix)
            if self.device is None:
                out = np.zeros(shape, self.bt.grid_type)
                ptr = out.ctypes.data
            else:
                import torch
                tdt = torch.complex64 if self.bt.grid_type == np.complex64 else torch.complex128
                out = torch.zeros(shape, dtype=tdt, device=self.device)
                ptr = out.data_ptr()
            norm = np.zeros((cube, nc), np.float64)
            self.outputs.append(out)
            self.norms.append(norm)
            e = table[i]
            e.nx, e.ny = f.npix, ny
            e.kernel_set = f.kernel_set
            e.sampling_function = 1 if f.sampling_function else 0
            e.cell_x = f.cell
            e.cell_y = f.cell_y if f.cell_y is not None else f.cell
            e.ra, e.dec = f.centre
            e.output = ptr
            e.normalization = norm.ctypes.data
        self._table = table
        self._check(self.lib.bullseye_batch_plan(self.binding.byref(), GRID_MODES[self.mode],
                                                 table, n, schedule))

    def execute(self):
        """bullseye_batch_execute; returns (grids, normalisation sums) per facet."""
        self._check(self.lib.bullseye_batch_execute())
        return self.outputs, self.norms

    def set_tuning(self, channels_per_stream=0, base_rows=0, shape=None, chunk=None,
                   facet_major=0):
        t = bullseye_batch_tuning(struct_size=ctypes.sizeof(bullseye_batch_tuning),
                                  channels_per_stream=channels_per_stream, base_rows=base_rows,
                                  facet_major=facet_major)
        for i in range(32):   # dict {class width M: value} or one value for every class
            t.shape[i] = _per_class(shape, i)
            t.chunk[i] = _per_class(chunk, i)
        self._check(self.lib.bullseye_batch_set_tuning(ctypes.byref(t)))

    def tuning(self):
        t = bullseye_batch_tuning(struct_size=ctypes.sizeof(bullseye_batch_tuning))
        self._check(self.lib.bullseye_batch_get_tuning(ctypes.byref(t)))
        return tuning_dict(t)

    def autotune(self, facets: Sequence[BatchFacet]):
        """bullseye_batch_autotune over the facets (outputs are not written); the choice stays
        in the library for later plans."""
        n = len(facets)
        table = (bullseye_batch_facet * n)()
        for i, f in enumerate(facets):
            e = table[i]
            e.nx, e.ny = f.npix, f.ny or f.npix
            e.kernel_set = f.kernel_set
            e.sampling_function = 1 if f.sampling_function else 0
            e.cell_x = f.cell
            e.cell_y = f.cell_y if f.cell_y is not None else f.cell
            e.ra, e.dec = f.centre
        t = bullseye_batch_tuning(struct_size=ctypes.sizeof(bullseye_batch_tuning))
        self._check(self.lib.bullseye_batch_autotune(self.binding.byref(), GRID_MODES[self.mode],
                                                     table, n, ctypes.byref(t)))
        return tuning_dict(t)

    def release(self):
        self.lib.bullseye_batch_release()


def _per_class(v, i):
    if isinstance(v, dict):
        return int(v.get(2 * i + 1, 0))
    return int(v or 0)


def tuning_dict(t):
    return {"channels_per_stream": int(t.channels_per_stream), "base_rows": int(t.base_rows),
            "tuned": int(t.tuned), "facet_major": int(t.facet_major),
            "shape": {2 * i + 1: int(t.shape[i]) for i in range(32)},
            "chunk": {2 * i + 1: int(t.chunk[i]) for i in range(32)},
            "tuned_ms": float(t.tuned_ms), "default_ms": float(t.default_ms)}
