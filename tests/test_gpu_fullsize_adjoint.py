"""The differentiable-imaging workload at full size: 1 M samples (a seeded row subset of the
w-projection observation, all 32 channels), one 2048^2 facet at 1.5" centred 0.2 deg from the
phase centre, 32 w-planes (supports 7 -> 63 by |w|, OS x8, complex64 kernels), and the
chi-squared of a 500-pixel model image against the data.

* Dot-product test of the gridding / degridding pair through autograd on the FP32-compensated
  path: Re<G v, g> = Re<v, G^T g> to <= 1e-6 relative.
* Central finite differences (step 1e-3) of the loss on 64 random pixels against the autograd
  gradient, FP64 path of the double build (the loss is quadratic in the image): <= 1e-6 of the
  peak gradient.
* The FP32-compensated autograd gradient against a float64 gradient built from the oracle
  alone (oracle_degrid for the residual, the oracle gridder for G r, the FFT chain in NumPy):
  rms <= 1e-5 and peak <= 1e-4 of its peak.  The oracle needs about half a minute for it.
"""
import numpy as np
import pytest
import torch

from common import oracle_chi2_gradient
from bullseye_amd import synth
from bullseye_amd.autograd import FacetImager, GridFunction, chi_squared
from bullseye_amd.library_loader import PATH_FP32_KAHAN, PATH_FP64

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
MODE = "facet_single_pol"


@pytest.fixture(scope="module")
def workload():
    print("case full-size adjoint: synthesising", flush=True)
    obs = synth.simulate(**synth.CONFIGS["c2"]["sim"], seed=3, device=DEV)
    obs, spec, halves, img = synth.differentiable_facet(obs, 10, device=DEV)
    # the oracle and the double build see the float32 values the mixed build grids
    spec.conv = spec.conv.astype(np.complex128)
    return obs, spec, halves, img


def test_dot_product(workload):
    obs, spec, halves, _ = workload
    print("case full-size adjoint: dot product", flush=True)
    rng = np.random.default_rng(8)
    with FacetImager(obs, spec, mode=MODE, precision="mixed", path=PATH_FP32_KAHAN,
                     wplane_support=halves) as im:
        v = torch.from_numpy(obs.vis).to(DEV).requires_grad_(True)
        gv = GridFunction.apply(v, im)
        noise = torch.from_numpy((rng.normal(size=im.grid_shape) +
                                  1j * rng.normal(size=im.grid_shape)).astype(np.complex64)).to(DEV)
        g = (gv.detach() / gv.detach().norm() + noise / noise.norm()).to(torch.complex64)
        lhs = torch.vdot(gv.reshape(-1).to(torch.complex128),
                         g.reshape(-1).to(torch.complex128)).real
        (vjp,) = torch.autograd.grad(lhs, v)
        rhs = torch.vdot(v.detach().reshape(-1).to(torch.complex128),
                         vjp.reshape(-1).to(torch.complex128)).real.item()
        lhs = lhs.item()
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("dot: lhs %.9e rhs %.9e rel %.3e" % (lhs, rhs, rel))
    assert rel <= 1e-6, rel


def test_central_differences(workload):
    obs, spec, halves, img = workload
    print("case full-size adjoint: central differences", flush=True)
    data = obs.vis.astype(np.complex128)
    with FacetImager(obs, spec, mode=MODE, precision="double", path=PATH_FP64,
                     wplane_support=halves) as im:
        d = torch.from_numpy(data).to(DEV)
        img_t = torch.tensor(img, dtype=torch.float64, device=DEV, requires_grad=True)
        chi_squared(img_t, d, im).backward()
        grad = img_t.grad.cpu().numpy().reshape(-1)
        pix = np.random.default_rng(7).choice(img.size, 64, replace=False)
        h = 1e-3
        fd = np.empty(64)
        base = torch.tensor(img, dtype=torch.float64, device=DEV)
        with torch.no_grad():
            for i, p in enumerate(pix):
                e = torch.zeros(img.size, dtype=torch.float64, device=DEV)
                e[p] = h
                e = e.reshape(img.shape)
                fd[i] = (chi_squared(base + e, d, im).item() -
                         chi_squared(base - e, d, im).item()) / (2 * h)
    err = np.abs(fd - grad[pix]).max() / np.abs(grad).max()
    print("central differences: max err %.3e of peak gradient" % err)
    assert err <= 1e-6, err


def test_gradient_matches_float64_oracle(workload):
    obs, spec, halves, img = workload
    print("case full-size adjoint: float64 oracle gradient", flush=True)
    want, want_loss = oracle_chi2_gradient(obs, spec, MODE, img, obs.vis.astype(np.complex128))
    with FacetImager(obs, spec, mode=MODE, precision="mixed", path=PATH_FP32_KAHAN,
                     wplane_support=halves) as im:
        img_t = torch.tensor(img, dtype=torch.float32, device=DEV, requires_grad=True)
        d = torch.from_numpy(obs.vis).to(DEV)
        loss = chi_squared(img_t, d, im)
        loss.backward()
        got = img_t.grad.double().cpu().numpy()
    peak = np.abs(want).max()
    rms = np.sqrt(np.mean((got - want) ** 2)) / peak
    pk = np.abs(got - want).max() / peak
    print("loss %.6e want %.6e grad rms %.3e peak %.3e" % (loss.item(), want_loss, rms, pk))
    assert abs(loss.item() - want_loss) <= 1e-5 * want_loss
    assert rms <= 1e-5 and pk <= 1e-4, (rms, pk)
