"""Convolution kernels and the grid-taper correction.

Sampling layout follows bullseye's 1D AA filter array (bullseye/mo/convolution_filter.py:93-107
and bullseye/mo/cpu_gpu_common/gridding_parameters.h:69-72): for half support S and oversampling
OS the array has csz = (2S+3) + (2S+2)(OS-1) samples, centre index (S+1)*OS, sample spacing
1/OS grid cells, and is normalised to unit sum.

The reference only generates sinc x {unity, Kaiser-Bessel, Hamming}; the imaging configuration
asks for a prolate spheroidal (PSWF) anti-aliasing kernel.  It is generated here from the
zeroth-order prolate spheroidal angular function of the first kind (scipy.special.pro_ang1,
Flammer's normalisation) with bandwidth parameter c = pi * W / 2 over the full width W = 2S+1,
the classic gridding function of Schwab (1984) with alpha = 0.
"""
import numpy as np
from scipy import special


def aa_kernel_size(support, oversample):
    """Number of samples of the oversampled 1D kernel (convolution_filter.py:93-95)."""
    return (2 * support + 3) + (2 * support + 2) * (oversample - 1)


def _positions(support, oversample):
    csz = aa_kernel_size(support, oversample)
    centre = (support + 1) * oversample
    return (np.arange(csz) - centre) / float(oversample)


def pswf_aa_kernel(support, oversample, dtype=np.float64):
    """PSWF anti-aliasing kernel over 2S+1 cells, oversampled, unit sum."""
    width = 2 * support + 1
    x = _positions(support, oversample)
    nu = x / (width / 2.0)
    vals = np.zeros_like(x)
    inside = np.abs(nu) < 1.0
    vals[inside] = special.pro_ang1(0, 0, np.pi * width / 2.0, nu[inside])[0]
    vals /= vals.sum()
    return vals.astype(dtype)


def sinc_aa_kernel(support, oversample, dtype=np.float64):
    """bullseye's default 'sinc' AA filter (convolution_filter.py:96-100 with the unity window)."""
    csz = aa_kernel_size(support, oversample)
    centre = csz // 2
    x = np.arange(-centre, centre + 1)
    vals = np.sinc(x / float(oversample))
    vals /= vals.sum()
    return vals.astype(dtype)


def kernel_tap_sums(conv, support, oversample):
    """Sum of the 2S+1 taps used for each truncated fraction index f in [0, OS]."""
    idx = np.arange(oversample + 1)[:, None] + oversample * np.arange(2 * support + 1)[None, :]
    return np.asarray(conv, dtype=np.float64)[idx].sum(axis=1)


def grid_taper(conv, support, oversample, n):
    """Image-plane response of the gridding kernel along one axis.

    Returns T[k] for image pixel k = 0..n-1 in the order finalize produces (after fftshift,
    pixel k sits at l = k - n//2 cycles per grid): the DFT of the oversampled kernel samples
    taken at the pixel frequencies, T(l) = sum_i conv[i] cos(2 pi (i - c) l / (n OS)),
    normalised so that T at l = 0 is 1.  Dividing an image by T(l) T(m) removes the kernel's
    taper (the grid-taper correction).
    """
    conv = np.asarray(conv, dtype=np.float64)
    centre = (support + 1) * oversample
    off = (np.arange(conv.size) - centre) / float(oversample)
    l = np.arange(n) - n // 2
    t = np.cos(2.0 * np.pi * np.outer(l, off) / n) @ conv
    return t / conv.sum()


def linear_plane_supports(wplanes, s_min, s_max):
    """Half supports growing linearly with the plane index (|w|), s_min at w = 0."""
    p = np.arange(wplanes)
    return [int(x) for x in np.rint(s_min + (s_max - s_min) * p / max(wplanes - 1, 1))]


def wproj_kernels(support, oversample, wplanes, wmax, nx, cell_arcsec, aa_support=3,
                  half_supports=None, dtype=np.complex128, device=None):
    """2D w-projection kernels in the reference's layout complex[wplanes][csz][csz].

    Plane p is the kernel for w_p = p * wmax / (wplanes - 1) (the gridder picks
    lrint(|w| / wmax * (wplanes - 1)), convolution_policies.h:264): the inverse Fourier transform
    of T(l) T(m) exp(-2 pi i w_p (n - 1)) sampled on the image grid that the oversampled kernel's
    DFT spans (OS times the field of view), where T is the image-plane response of the PSWF AA
    kernel of half support aa_support.  Planes are normalised so plane 0 sums to one, then cut
    to their half supports (kernels.truncate_wkernels) when half_supports is given.  With a torch
    device the transforms run there (float64; equal to the NumPy result to rounding).
    """
    csz = aa_kernel_size(support, oversample)
    fov = nx * cell_arcsec * np.pi / 648000.0
    k = np.arange(csz) - csz // 2
    lm = k / float(csz) * fov * oversample
    aa = pswf_aa_kernel(aa_support, oversample)
    aa_pad = np.zeros(csz)
    a0 = csz // 2 - aa.size // 2
    aa_pad[a0:a0 + aa.size] = aa
    taper1 = np.real(np.fft.fftshift(np.fft.fft(np.fft.ifftshift(aa_pad))))
    taper2 = np.outer(taper1, taper1)
    lg, mg = np.meshgrid(lm, lm, indexing="ij")
    nm1 = np.sqrt(np.maximum(1.0 - lg ** 2 - mg ** 2, 0.0)) - 1.0
    step = wmax / float(max(wplanes - 1, 1))
    if device is not None:
        import torch
        t2 = torch.as_tensor(taper2, device=device)
        nm = torch.as_tensor(nm1, device=device)
        planes = []
        for p in range(wplanes):
            img = t2 * torch.exp(-2j * np.pi * (p * step) * nm)
            planes.append(torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(img))).cpu())
        out = torch.stack(planes).numpy()
    else:
        out = np.empty((wplanes, csz, csz), dtype=np.complex128)
        for p in range(wplanes):
            img = taper2 * np.exp(-2j * np.pi * (p * step) * nm1)
            out[p] = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(img)))
    out /= out[0].real.sum()
    if half_supports is not None:
        out = truncate_wkernels(out, half_supports, support, oversample)
    return out.astype(dtype)


def truncate_wkernels(conv, half_supports, support, oversample):
    """Zero every sample of plane p at |x| >= S_p + 1/2 cells (x measured from the kernel centre
    (S+1)*OS), in both axes, so plane p's footprint is exactly its central (2 S_p + 1)^2 taps for
    every fraction index the nearest-rounding w-projection gridder can pick."""
    conv = np.array(conv, copy=True)
    csz = conv.shape[-1]
    x = np.abs(np.arange(csz) / float(oversample) - (support + 1))
    for p, sp in enumerate(half_supports):
        keep = x < sp + 0.5
        conv[p] *= np.outer(keep, keep)
    return conv
