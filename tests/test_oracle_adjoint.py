"""Pins the oracle's degridding adjoint (oracle_degrid) without any GPU.

The reference has no degridder in C++ (bullseye's Resampling.ipynb sketches one in Python), so
the adjoint is pinned against the oracle gridder, itself pinned by test_oracle.py:
  * the dot-product test Re<G v, g> = Re<v, G^T g> over the reals with random complex v and g,
    for every visibility mode, the AA kernel and random complex w-projection kernels (a wrong
    conjugate, a transposed tap, a missing weight or a Jones product in the wrong order each
    break it by O(1));
  * closed forms: a delta model at the footprint centre returns the centre tap squared times
    the weight, and one cell to the right returns the neighbouring tap.
"""
import numpy as np
import pytest

from common import make_spec, small_obs, tiny_obs, wproj_spec
from bullseye_amd import kernels
from bullseye_amd.binding import Binding, ImagingSpec
from oracle import oracle

VIS_MODES = ["grid_single_pol", "facet_single_pol", "grid_duel_pol", "facet_duel_pol",
             "grid_4_cor", "facet_4_cor", "facet_4_cor_corrections"]


def random_obs(mode, seed):
    obs = small_obs(mode, seed=seed, n_ant=8, n_time=12, n_chan=4)
    rng = np.random.default_rng(seed + 100)
    obs.vis = rng.normal(size=obs.vis.shape) + 1j * rng.normal(size=obs.vis.shape)
    obs.weights = rng.uniform(0.5, 2.0, obs.weights.shape).astype(np.float32)
    return obs


def dot_test(obs, spec, mode, seed):
    b = Binding("double", obs, spec)
    oracle.run(b, mode, normalize=False)
    gv = b.output.copy()
    rng = np.random.default_rng(seed + 200)
    g = rng.normal(size=gv.shape) + 1j * rng.normal(size=gv.shape)
    gtg = oracle.degrid(b, mode, g)
    v = b.keep["visibilities"]
    lhs = np.vdot(gv, g).real
    rhs = np.vdot(v, gtg).real
    assert np.abs(gv).max() > 0 and np.abs(gtg).max() > 0
    # relative to |Gv| |g|: a wrong adjoint errs by about |Gv| |g| / sqrt(grid size) ~ 1e-2
    return abs(lhs - rhs) / (np.linalg.norm(gv) * np.linalg.norm(g))


@pytest.mark.parametrize("mode", VIS_MODES)
def test_adjoint_dot_product_aa(mode):
    obs = random_obs(mode, seed=11)
    spec = make_spec(obs, mode, npix=96)
    err = dot_test(obs, spec, mode, 11)
    assert err <= 1e-13, err


@pytest.mark.parametrize("mode", VIS_MODES)
def test_adjoint_dot_product_wprojection(mode):
    obs = random_obs(mode, seed=12)
    spec, _ = wproj_spec(obs, mode, "double", S=5, OS=4, planes=6, halves=[1, 2, 3, 4, 5, 5],
                         npix=96)
    err = dot_test(obs, spec, mode, 12)
    assert err <= 1e-13, err


def test_adjoint_rejects_psf_modes():
    obs = random_obs("grid_single_pol", seed=13)
    spec = make_spec(obs, "grid_single_pol", npix=64)
    b = Binding("double", obs, spec)
    with pytest.raises(RuntimeError, match="visibility modes"):
        oracle.degrid(b, "grid_sampling_function", b.output)


@pytest.mark.parametrize("dx,dy", [(0, 0), (1, 0), (0, -2), (-3, 1)])
def test_adjoint_delta_model_closed_form(dx, dy):
    """uv = 0, weight 1.5: the footprint is centred on (n/2, n/2) with taps conv[OS + k*OS]; a
    delta at (n/2 + dx, n/2 + dy) returns weight * conv[c + dx*OS] * conv[c + dy*OS]."""
    S, OS, n = 3, 63, 64
    conv = kernels.pswf_aa_kernel(S, OS)
    obs = tiny_obs([[0, 0, 0]], [[1.0 + 0.0j]])
    obs.weights[:] = 1.5
    spec = ImagingSpec(npix=n, cell=20.0, conv=conv, support=S, oversample=OS)
    b = Binding("double", obs, spec)
    model = np.zeros(b.output.shape, np.complex128)
    model[0, 0, 0, n // 2 + dy, n // 2 + dx] = 2.0 - 1.0j
    got = oracle.degrid(b, "grid_single_pol", model)[0, 0, 0]
    c = (S + 1) * OS
    want = 1.5 * conv[c + dx * OS] * conv[c + dy * OS] * (2.0 - 1.0j)
    assert got == pytest.approx(want, rel=1e-13, abs=1e-16)
