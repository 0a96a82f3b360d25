"""Differentiable facet imaging (bullseye_amd.autograd) on the GPU.

* The chi-squared gradient with respect to a sparse real model image, from autograd through the
  library (FP32 compensated path, float32 data), against a float64 reference gradient built
  from the oracle alone: 2 Re finalize(G r) with r = G^T F(I) - V from oracle_degrid, G from the
  oracle gridder and the FFT chain in NumPy.  Bar: rms <= 1e-5 and peak <= 1e-4 of the
  reference's peak.
* Central finite differences (step 1e-3) of the loss on 64 random pixels against autograd, on
  the FP64 path of the double build (the loss is quadratic in the image, so the central
  difference is exact up to rounding): relative error <= 1e-6.
* The dot-product test through autograd: the vector-Jacobian product of GridFunction is the
  degridding adjoint, Re<G v, g> = Re<v, vjp> to <= 1e-6 relative on the FP32 path.
"""
import numpy as np
import pytest
import torch

from common import (make_spec, oracle_chi2_gradient as reference_gradient, round_to, small_obs,
                    wproj_spec)
from bullseye_amd.autograd import DegridFunction, FacetImager, GridFunction, chi_squared
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def setup(mode, kernel, precision, seed=31, npix=128):
    obs = small_obs(mode, seed=seed, n_ant=10, n_time=24, n_chan=4)
    obs = round_to(obs, precision)
    if kernel == "aa":
        spec, halves = make_spec(obs, mode, precision=precision, npix=npix), None
    else:
        spec, halves = wproj_spec(obs, mode, precision, S=5, OS=4, planes=6,
                                  halves=[2, 3, 3, 4, 5, 5], npix=npix)
    return obs, spec, halves


def sparse_image(shape, n, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros(shape)
    flat = img.reshape(-1)
    idx = rng.choice(flat.size, n, replace=False)
    flat[idx] = rng.uniform(0, 1, n)
    return img


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
@pytest.mark.parametrize("mode", ["grid_single_pol", "facet_single_pol", "facet_4_cor_corrections"])
def test_chi2_gradient_matches_float64_reference(mode, kernel):
    print("case chi2", mode, kernel, flush=True)
    obs, spec, halves = setup(mode, kernel, "mixed")
    shape = (spec.n_facets, 1, spec.n_pol_gridded, spec.npix, spec.npix)
    image = sparse_image(shape, 50, seed=3)
    rng = np.random.default_rng(4)
    data = (rng.normal(size=obs.vis.shape) + 1j * rng.normal(size=obs.vis.shape)) * 0.5
    data = data.astype(np.complex64)
    # only the gridded correlations enter the loss; the others are zero in the model
    want, want_loss = reference_gradient(obs, spec, mode, image, mask_data(data, mode, spec))
    with FacetImager(obs, spec, mode=mode, precision="mixed", path=PATH_FP32_KAHAN,
                     wplane_support=halves) as im:
        img_t = torch.tensor(image, dtype=torch.float32, device=DEV, requires_grad=True)
        d = torch.from_numpy(mask_data(data, mode, spec)).to(DEV)
        loss = chi_squared(img_t, d, im)
        loss.backward()
        got = img_t.grad.double().cpu().numpy()
    peak = np.abs(want).max()
    rms = np.sqrt(np.mean((got - want) ** 2)) / peak
    pk = np.abs(got - want).max() / peak
    print("loss %.6e want %.6e grad rms %.3e peak %.3e" % (loss.item(), want_loss, rms, pk))
    assert abs(loss.item() - want_loss) <= 1e-5 * want_loss
    assert rms <= 1e-5 and pk <= 1e-4, (rms, pk)


def mask_data(data, mode, spec):
    """Zero the correlations a mode does not grid (the adjoint leaves them zero)."""
    out = np.zeros_like(data)
    if mode in ("grid_single_pol", "facet_single_pol"):
        out[..., spec.pol_index] = data[..., spec.pol_index]
    elif mode in ("grid_duel_pol", "facet_duel_pol"):
        out[..., spec.pol_index] = data[..., spec.pol_index]
        out[..., spec.second_pol_index] = data[..., spec.second_pol_index]
    else:
        out[:] = data
    return out


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
def test_chi2_gradient_central_differences(kernel):
    mode = "facet_single_pol"
    obs, spec, halves = setup(mode, kernel, "double", seed=32)
    shape = (spec.n_facets, 1, 1, spec.npix, spec.npix)
    image = sparse_image(shape, 50, seed=5)
    rng = np.random.default_rng(6)
    data = mask_data(rng.normal(size=obs.vis.shape) + 1j * rng.normal(size=obs.vis.shape),
                     mode, spec)
    with FacetImager(obs, spec, mode=mode, precision="double", path=PATH_FP64,
                     wplane_support=halves) as im:
        d = torch.from_numpy(data).to(DEV)
        img_t = torch.tensor(image, dtype=torch.float64, device=DEV, requires_grad=True)
        chi_squared(img_t, d, im).backward()
        grad = img_t.grad.cpu().numpy().reshape(-1)
        pix = np.random.default_rng(7).choice(image.size, 64, replace=False)
        h = 1e-3
        fd = np.empty(64)
        base = torch.tensor(image, dtype=torch.float64, device=DEV)
        with torch.no_grad():
            for i, p in enumerate(pix):
                e = torch.zeros(image.size, dtype=torch.float64, device=DEV)
                e[p] = h
                e = e.reshape(shape)
                lp = chi_squared(base + e, d, im).item()
                lm = chi_squared(base - e, d, im).item()
                fd[i] = (lp - lm) / (2 * h)
    err = np.abs(fd - grad[pix]).max() / np.abs(grad).max()
    print("central differences: max err %.3e of peak gradient" % err)
    assert err <= 1e-6, err


@pytest.mark.parametrize("kernel", ["aa", "wproj"])
def test_autograd_dot_product(kernel):
    mode = "facet_4_cor_corrections"
    obs, spec, halves = setup(mode, kernel, "mixed", seed=33)
    rng = np.random.default_rng(8)
    with FacetImager(obs, spec, mode=mode, precision="mixed", path=PATH_FP32_KAHAN,
                     wplane_support=halves) as im:
        v = torch.from_numpy((rng.normal(size=im.vis_shape) +
                              1j * rng.normal(size=im.vis_shape)).astype(np.complex64)).to(DEV)
        v.requires_grad_(True)
        gv = GridFunction.apply(v, im)
        noise = torch.from_numpy((rng.normal(size=im.grid_shape) +
                                  1j * rng.normal(size=im.grid_shape)).astype(np.complex64)).to(DEV)
        g = (gv.detach() / gv.detach().norm() + noise / noise.norm()).to(torch.complex64)
        lhs = torch.vdot(gv.reshape(-1), g.reshape(-1)).real
        (vjp,) = torch.autograd.grad(lhs, v)
        rhs = torch.vdot(v.detach().reshape(-1).to(torch.complex128),
                         vjp.reshape(-1).to(torch.complex128)).real.item()
        lhs = lhs.item()
        # the vjp of Re<G v, g> is G^T g: check it against the degridding call directly
        direct = im.degrid(g)
        assert torch.allclose(vjp, direct, rtol=0, atol=0)
        # and DegridFunction's backward is the gridding call
        gg = g.clone().requires_grad_(True)
        out = DegridFunction.apply(gg, im)
        (back,) = torch.autograd.grad(out, gg, v.detach())
        again = im.grid(v.detach())   # grid flushes are atomics: equal up to summation order
        assert (back - again).abs().max() <= 1e-6 * again.abs().max()
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("dot: lhs %.9e rhs %.9e rel %.3e" % (lhs, rhs, rel))
    assert rel <= 1e-6, rel
