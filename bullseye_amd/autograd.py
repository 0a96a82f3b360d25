"""Differentiable facet imaging: the library's gridding operator and its degridding adjoint as
torch.autograd functions.

bullseye itself has no gradients; this module layers them on the C ABI.  For one gridding mode
and a fixed observation (uvw, wavelengths, weights, flags, facets, Jones, kernels), let G be the
map from visibilities to the unnormalised uv-grids that the mode's entry point applies.  G is
linear over the reals (the w-projection path conjugates w < 0 samples), and bullseye_degrid
computes its transpose G^T over the reals.  With PyTorch's convention for complex tensors (the
gradient of a real loss is dL/dRe + i dL/dIm), the vector-Jacobian product of a map that is
linear over the reals is its real transpose, so

  GridFunction:    vis  -> G vis     backward: grad_vis  = G^T grad_grid  (degridding)
  DegridFunction:  grid -> G^T grid  backward: grad_grid = G grad_vis     (gridding)

Both run on the device the observation lives on; nothing is copied through the host.

The image <-> grid transforms are the transposes of each other as well: finalize maps a grid to
an image with ifftshift, the unnormalised backward FFT and fftshift (fft_and_repacking_routines.
cpp:83-118); its real transpose, from a real image to a grid, is fftshift, the unnormalised
forward FFT and ifftshift.  A chi-squared between degridded model visibilities and data,
  L(I) = sum |G^T F(I) - V|^2,
therefore has the gradient 2 Re finalize(G (G^T F(I) - V)) with respect to the real image I.
"""
import ctypes

import numpy as np
import torch

from .binding import Binding
from .library_loader import (PATH_FP32_KAHAN, bullseye_options, bullseye_stats, load_library)

VIS_MODES = ("grid_single_pol", "facet_single_pol", "grid_duel_pol", "facet_duel_pol",
             "grid_4_cor", "facet_4_cor", "facet_4_cor_corrections")


class FacetImager:
    """One device-resident imaging context on the library (the library holds one context per
    process, as bullseye's does: create one FacetImager at a time and close() it).

    obs, spec: bullseye_amd.synth.Observation and bullseye_amd.binding.ImagingSpec.
    mode: a visibility gridding mode.  precision: 'single', 'double' or 'mixed'.
    path: bullseye_compute_path.  wplane_support: per-plane half supports (w-projection).
    """

    def __init__(self, obs, spec, mode="grid_single_pol", precision="mixed",
                 path=PATH_FP32_KAHAN, device=None, wplane_support=None):
        if mode not in VIS_MODES:
            raise ValueError("%s is not a visibility gridding mode" % mode)
        self.mode, self.precision, self.path = mode, precision, path
        self.device = torch.device(device if device is not None else "cuda:0")
        self.lib = load_library("GPU", precision)
        opts = bullseye_options(struct_size=ctypes.sizeof(bullseye_options), device_resident=1,
                                compute_path=path)
        self._check(self.lib.bullseye_set_options(ctypes.byref(opts)), "bullseye_set_options")
        if wplane_support is not None:
            arr = (ctypes.c_uint32 * len(wplane_support))(*wplane_support)
            self._check(self.lib.bullseye_set_wplane_support(arr, len(wplane_support)),
                        "bullseye_set_wplane_support")
        else:
            self.lib.bullseye_set_wplane_support(None, 0)
        self.b = Binding(precision, obs, spec, device=self.device)
        torch.cuda.synchronize(self.device)
        self._check(self.lib.initLibrary(self.b.byref()), "initLibrary")
        self._open = True
        self.vis_shape = tuple(self.b.keep["visibilities"].shape)
        self.grid_shape = tuple(self.b.keep["output_buffer"].shape)
        self.vis_dtype = self.b.keep["visibilities"].dtype
        self.grid_dtype = self.b.keep["output_buffer"].dtype
        self._mode_id = VIS_MODES.index(mode)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, self.lib.bullseye_last_error().decode()))

    def close(self):
        if getattr(self, "_open", False):
            self.lib.releaseLibrary()
            self.lib.bullseye_set_wplane_support(None, 0)
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the two operators (no autograd) -------------------------------------------------
    def grid(self, vis):
        """G vis: the mode's unnormalised uv-grids [F, cube, P, ny, nx] of visibilities
        [rows, chan, npol] (weights, flags, phase, Jones and kernels of the observation)."""
        vis = self._as(vis, self.vis_shape, self.vis_dtype)
        p = self.b.params
        saved = p.visibilities
        p.visibilities = vis.data_ptr()
        torch.cuda.current_stream(self.device).synchronize()
        try:
            self._check(self.lib.bullseye_reset_grids(), "bullseye_reset_grids")
            # is_final_data_chunk is set: the call waits for the kernels, publishes the grids
            # and reports device-side errors in its status
            self._check(getattr(self.lib, self.mode)(self.b.byref()), self.mode)
        finally:
            p.visibilities = saved
        return self.b.keep["output_buffer"].clone()

    def degrid(self, grid):
        """G^T grid: visibilities [rows, chan, npol] of uv-grids [F, cube, P, ny, nx]."""
        grid = self._as(grid, self.grid_shape, self.grid_dtype)
        out = torch.empty(self.vis_shape, dtype=self.vis_dtype, device=self.device)
        torch.cuda.current_stream(self.device).synchronize()
        self._check(self.lib.bullseye_degrid(self.b.byref(), self._mode_id, grid.data_ptr(),
                                             out.data_ptr()), "bullseye_degrid")
        return out

    def stats(self):
        st = bullseye_stats()
        self.lib.bullseye_get_stats(ctypes.byref(st))
        return st

    def _as(self, t, shape, dtype):
        if not isinstance(t, torch.Tensor):
            raise TypeError("expected a torch tensor")
        if tuple(t.shape) != shape:
            raise ValueError("expected shape %s, got %s" % (shape, tuple(t.shape)))
        if t.dtype != dtype:
            raise TypeError("expected %s, got %s" % (dtype, t.dtype))
        if t.device != self.device:
            raise ValueError("expected a tensor on %s" % self.device)
        return t.contiguous()


class GridFunction(torch.autograd.Function):
    """vis -> G vis; backward degrids."""

    @staticmethod
    def forward(ctx, vis, imager):
        ctx.imager = imager
        return imager.grid(vis)

    @staticmethod
    def backward(ctx, grad_grid):
        return ctx.imager.degrid(grad_grid.contiguous()), None


class DegridFunction(torch.autograd.Function):
    """grid -> G^T grid; backward grids."""

    @staticmethod
    def forward(ctx, grid, imager):
        ctx.imager = imager
        return imager.degrid(grid)

    @staticmethod
    def backward(ctx, grad_vis):
        return ctx.imager.grid(grad_vis.contiguous()), None


def image_to_grid(image):
    """Real transpose of finalize's grid -> image chain: fftshift(fft2(ifftshift(image))) over
    the last two axes, unnormalised."""
    z = image.to(torch.complex128 if image.dtype == torch.float64 else torch.complex64)
    return torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(z, dim=(-2, -1))), dim=(-2, -1))


def grid_to_image(grid):
    """finalize's chain before the real part: fftshift(N * ifft2(ifftshift(grid)))."""
    n = grid.shape[-1] * grid.shape[-2]
    return torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(grid, dim=(-2, -1))) * n,
                              dim=(-2, -1))


def model_visibilities(image, imager):
    """Degridded model visibilities of a real model image [F, cube, P, ny, nx] (differentiable
    with respect to image)."""
    g = image_to_grid(image).to(imager.grid_dtype)
    return DegridFunction.apply(g, imager)


def chi_squared(image, data, imager):
    """sum |G^T F(image) - data|^2 over the gridded correlations, differentiable in image."""
    r = model_visibilities(image, imager) - data
    return (r.real ** 2 + r.imag ** 2).sum()
