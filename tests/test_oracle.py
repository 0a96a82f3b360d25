"""Pins the C oracle (a restatement of bullseye's CPU gridder) without any GPU.

The reference ships no tests or golden vectors for the gridding path; the only known-answer
assertions in the repository are the fftshift/ifftshift checks of
writing_and_wiki/notebooks/Fourier Shifting.ipynb (cells 7 and 9).  Everything else here is
pinned by closed forms, invariants and brute force, each chosen so that a plausible mistake
(a dropped term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

from common import make_spec, small_obs, tiny_obs, u_scale
from bullseye_amd import kernels, synth
from bullseye_amd.binding import Binding, ImagingSpec, jones_reference_layout
from oracle import oracle

S, OS = 3, 63
CONV = kernels.pswf_aa_kernel(S, OS)


def grid_one(obs, mode="grid_single_pol", npix=64, cell=20.0, conv=CONV, support=S,
             oversample=OS, facets=None, normalize=False, **kw):
    spec = ImagingSpec(npix=npix, cell=cell, conv=conv, support=support, oversample=oversample,
                       facet_centres=facets, **kw)
    b = Binding("double", obs, spec, with_psf=mode.endswith("sampling_function"))
    norm = oracle.run(b, mode, normalize=normalize)
    return b, norm


# ---- known-answer tests from the reference's notebook ------------------------------------

@pytest.mark.parametrize("shape", [(6, 6), (5, 7), (7, 4), (1, 3), (8, 9)])
def test_fftshift_kat(shape):
    """Fourier Shifting.ipynb cells 7 and 9: the in-place shifts equal numpy's."""
    x = (np.arange(np.prod(shape)) + 1j * np.arange(np.prod(shape))[::-1]).reshape(shape)
    np.testing.assert_array_equal(oracle.fftshift_inplace(x), np.fft.fftshift(x))
    np.testing.assert_array_equal(oracle.fftshift_inplace(x, inverse=True), np.fft.ifftshift(x))


# ---- closed forms -----------------------------------------------------------------------

@pytest.mark.parametrize("support", [1, 3, 5])
def test_unit_visibility_at_origin_deposits_kernel_taps(support):
    """uv = 0: fraction index = OS, so the footprint is conv[OS + k*OS] outer conv[OS + j*OS]
    centred on (nx/2, ny/2); normalisation = weight * sum of those taps."""
    conv = kernels.pswf_aa_kernel(support, OS)
    obs = tiny_obs([[0, 0, 0]], [[2.0 - 1.0j]])
    b, norm = grid_one(obs, conv=conv, support=support)
    g = b.output[0, 0, 0]
    taps = conv[OS + OS * np.arange(2 * support + 1)]
    want = np.zeros_like(g)
    c = 32
    want[c - support:c + support + 1, c - support:c + support + 1] = np.outer(taps, taps) * (2 - 1j)
    np.testing.assert_allclose(g, want, rtol=1e-13, atol=1e-15)
    assert norm[0] == pytest.approx(taps.sum() ** 2, rel=1e-13)


def closed_form_footprint(u_cells, v_cells, n, conv=CONV, support=S, oversample=OS, vis=1.0):
    """The grid one sample deposits: origin floor(u) + n/2 - S, fraction index
    size_t((1 - u + floor(u)) * OS), taps conv[f + k*OS] (convolution_policies.h:86-117)."""
    g = np.zeros((n, n), dtype=np.complex128)
    x0 = int(np.floor(u_cells)) + n // 2 - support
    y0 = int(np.floor(v_cells)) + n // 2 - support
    fu = int((1 - u_cells + np.floor(u_cells)) * oversample)
    fv = int((1 - v_cells + np.floor(v_cells)) * oversample)
    k = np.arange(2 * support + 1)
    g[y0:y0 + 2 * support + 1, x0:x0 + 2 * support + 1] = \
        np.outer(conv[fv + oversample * k], conv[fu + oversample * k]) * vis
    return g


@pytest.mark.parametrize("du,dv", [(5, 0), (0, 4), (-3, 7), (11, -9)])
def test_footprint_position_and_axis_orientation(du, dv):
    """A sample at u = du cells (v = dv cells) lands du columns right of the centre and dv rows
    *above* it (v_scale < 0, templated_gridder.h:23-24), with the truncated fraction picking
    the kernel samples; fractional offsets exercise the floor rounding."""
    sc = u_scale(64, 20.0)
    # (exact integers are avoided: under FE_DOWNWARD the reference's product can land one ulp
    # below an integer, moving the origin one cell with fraction index ~0 -- the same physical
    # taps except the zero-valued padding samples)
    for frac in (0.125, 0.25, 0.75):
        u = (du + frac) / sc
        v = (dv + frac) / sc
        b, _ = grid_one(tiny_obs([[u, v, 0]], [[1.0 + 0.5j]]))
        want = closed_form_footprint(u * sc, -(v * sc), 64, vis=1.0 + 0.5j)
        np.testing.assert_allclose(b.output[0, 0, 0], want, rtol=1e-12, atol=1e-15)


def test_edge_drop_rule():
    """A sample is dropped when disc + 2S+3 >= n (convolution_policies.h:95-96): the highest
    kept origin is n - 2S - 4, the lowest is 0."""
    sc = u_scale(64, 20.0)
    hi_ok = (64 - 2 * S - 4) - (32 - S)       # origin n-2S-4
    for d, kept in [(hi_ok, True), (hi_ok + 1, False), (-(32 - S), True), (-(32 - S) - 1, False)]:
        b, norm = grid_one(tiny_obs([[(d + 0.5) / sc, 0, 0]], [[1.0]]))
        assert (np.abs(b.output).sum() > 0) == kept, d
        assert (norm[0] > 0) == kept


def test_linearity_and_weights():
    obs = small_obs("grid_single_pol", seed=11, n_ant=8, n_time=12, n_chan=4)
    b1, n1 = grid_one(obs, npix=128, cell=20.0)
    obs2 = small_obs("grid_single_pol", seed=11, n_ant=8, n_time=12, n_chan=4)
    obs2.vis = (obs2.vis * (2.0 - 3.0j)).astype(np.complex64)
    obs2.weights = obs2.weights * 0.5
    b2, n2 = grid_one(obs2, npix=128, cell=20.0)
    np.testing.assert_allclose(b2.output, b1.output * (2.0 - 3.0j) * 0.5,
                               rtol=1e-6, atol=1e-9 * np.abs(b1.output).max())
    np.testing.assert_allclose(n2, n1 * 0.5, rtol=1e-12)


def test_flags_rows_fields_channels_drop_samples():
    """Flagged samples, flagged rows, other fields and disabled channels contribute nothing
    (templated_gridder.h:51-55, 78-82); removing them gives the same grid."""
    obs = small_obs("grid_single_pol", seed=4, n_ant=8, n_time=10, n_chan=4, flag_fraction=0.2)
    obs.flagged_rows[::7] = True
    obs.field[::5] = 3
    en = np.array([True, False, True, True])
    b, norm = grid_one(obs, npix=128, enabled_channels=en)
    keep = ~obs.flags & ~obs.flagged_rows[:, None, None] & (obs.field == 0)[:, None, None]
    keep &= en[None, :, None]
    clean = small_obs("grid_single_pol", seed=4, n_ant=8, n_time=10, n_chan=4, flag_fraction=0.0)
    clean.vis = np.where(keep, obs.vis, 0).astype(np.complex64)
    clean.weights = keep.astype(np.float32)
    b2, norm2 = grid_one(clean, npix=128)
    np.testing.assert_allclose(b.output, b2.output, rtol=0, atol=1e-15)
    np.testing.assert_allclose(norm, norm2, rtol=1e-12)


def test_psf_ignores_sample_flags_and_weights_but_not_row_flags():
    obs = small_obs("grid_single_pol", seed=5, n_ant=6, n_time=6, n_chan=2, flag_fraction=0.5)
    obs.weights[:] = 3.0
    b, _ = grid_one(obs, mode="grid_sampling_function", npix=128, sf_count=1)
    obs.flags[:] = False
    obs.weights[:] = 1.0
    obs.vis[:] = 1.0
    b2, _ = grid_one(obs, mode="grid_single_pol", npix=128)
    np.testing.assert_allclose(b.psf[:, 0, 0], b2.output[:, 0, 0], rtol=1e-12, atol=1e-15)
    obs.flagged_rows[:] = True
    b3, _ = grid_one(obs, mode="grid_sampling_function", npix=128, sf_count=1)
    assert np.abs(b3.psf).max() == 0


def test_wprojection_conjugate_symmetry():
    """w-projection grids (u, v, w, V) and (-u, -v, -w, conj V) identically
    (convolution_policies.h:248-253) -- both are mapped to w >= 0."""
    wpl, S2, OS2 = 4, 3, 8
    rng = np.random.default_rng(0)
    conv = (rng.normal(size=(wpl, kernels.aa_kernel_size(S2, OS2), kernels.aa_kernel_size(S2, OS2))) +
            1j * rng.normal(size=(wpl, kernels.aa_kernel_size(S2, OS2), kernels.aa_kernel_size(S2, OS2))))
    uvw = np.array([[310.3, -120.7, 44.0]])
    vis = np.array([[0.3 + 0.8j]])
    kw = dict(conv=conv, support=S2, oversample=OS2, wplanes=wpl, wmax=100.0, npix=64)
    b1, n1 = grid_one(tiny_obs(uvw, vis), **kw)
    b2, n2 = grid_one(tiny_obs(-uvw, np.conj(vis)), **kw)
    assert np.abs(b1.output).max() > 0
    np.testing.assert_array_equal(b1.output, b2.output)
    np.testing.assert_array_equal(n1, n2)


def test_wprojection_plane_choice():
    """plane = lrint(|w| / wmax * (wplanes-1)) (convolution_policies.h:264): a kernel stack whose
    plane p is constant p+1 at every tap makes the gridded value reveal the plane used."""
    wpl, S2, OS2 = 5, 1, 4
    csz = kernels.aa_kernel_size(S2, OS2)
    conv = np.stack([np.full((csz, csz), p + 1.0 + 0j) for p in range(wpl)])
    for w, plane in [(0.0, 0), (12.4, 0), (12.6, 1), (50.0, 2), (99.0, 4), (100.0, 4)]:
        b, _ = grid_one(tiny_obs([[0, 0, w]], [[1.0]]), conv=conv, support=S2, oversample=OS2,
                        wplanes=wpl, wmax=100.0, npix=32)
        assert b.output.real.max() == pytest.approx(plane + 1.0), (w, plane)
    b, norm = grid_one(tiny_obs([[0, 0, 130.0]], [[1.0]]), conv=conv, support=S2, oversample=OS2,
                       wplanes=wpl, wmax=100.0, npix=32)
    assert np.abs(b.output).max() == 0 and norm[0] == 0   # plane >= wplanes: dropped


def wproj_closed_form(uc, vc, w, vis, conv, S2, OS2, wpl, wmax, n):
    """The grid and normalisation one w-projection sample deposits, from the reference's own
    arithmetic (cpu_algorithm/convolution_policies.h:247-287) written out independently: for
    w < 0 the visibility is conjugated and (u, v, w) negated; origin lrint(u + n/2 - S) (nearest,
    FE_TONEAREST), fraction size_t((1 - u + lrint(u)) * OS), plane lrint(|w|/wmax (W-1)); tap
    (j, k) of the footprint is conv[plane][frac_v + j OS][frac_u + k OS] (row = v, column = u),
    multiplied into the visibility unconjugated; the normalisation adds Re(tap)."""
    if w < 0:
        vis, uc, vc, w = np.conj(vis), -uc, -vc, -w
    x0 = int(np.rint(uc + (n // 2 - S2)))
    y0 = int(np.rint(vc + (n // 2 - S2)))
    fu = int((1 - uc + np.rint(uc)) * OS2)
    fv = int((1 - vc + np.rint(vc)) * OS2)
    plane = int(np.rint(abs(w) / float(np.float32(wmax)) * (wpl - 1)))
    k = np.arange(2 * S2 + 1)
    taps = conv[plane][np.ix_(fv + OS2 * k, fu + OS2 * k)]
    g = np.zeros((n, n), dtype=np.complex128)
    g[y0:y0 + 2 * S2 + 1, x0:x0 + 2 * S2 + 1] = vis * taps
    return g, taps.real.sum(), plane


@pytest.mark.parametrize("uvw,vis", [([210.3, -75.6, 31.0], 0.3 + 0.8j),
                                     ([-143.2, 96.45, 77.0], -1.1 + 0.25j),
                                     ([118.9, 40.7, -58.0], 0.6 - 0.4j),
                                     ([-96.1, -151.3, -12.0], 0.2 + 1.3j)])
def test_wprojection_single_sample_closed_form(uvw, vis):
    """One w-projection sample on a random asymmetric complex kernel stack lands exactly on the
    closed form: a transposed tap index, a conjugated kernel, an |conv| normalisation or a missing
    w < 0 conjugation in the oracle fails this."""
    wpl, S2, OS2, n, wmax = 6, 2, 8, 64, 100.0
    csz = kernels.aa_kernel_size(S2, OS2)
    rng = np.random.default_rng(17)
    conv = rng.normal(size=(wpl, csz, csz)) + 1j * rng.normal(size=(wpl, csz, csz))
    weight = 0.75
    obs = tiny_obs([uvw], [[vis]])
    obs.weights[:] = weight
    b, norm = grid_one(obs, conv=conv, support=S2, oversample=OS2, wplanes=wpl, wmax=wmax,
                       npix=n)
    sc = u_scale(n, 20.0)
    uc, vc, v0 = uvw[0] * sc, -(uvw[1] * sc), np.complex128(np.complex64(vis)) * weight
    want, tapsum, plane = wproj_closed_form(uc, vc, uvw[2], v0, conv, S2, OS2, wpl, wmax, n)
    assert 0 < plane < wpl - 1
    got = b.output[0, 0, 0]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)
    assert norm[0] == pytest.approx(weight * tapsum, rel=1e-13)
    # each plausible mistake gives a visibly different grid or normalisation on this stack
    wrong = [wproj_closed_form(uc, vc, uvw[2], v0, conv.transpose(0, 2, 1), S2, OS2, wpl, wmax,
                               n)[0],
             wproj_closed_form(uc, vc, uvw[2], v0, np.conj(conv), S2, OS2, wpl, wmax, n)[0]]
    if uvw[2] < 0:   # the w < 0 conjugation left out
        wrong.append(wproj_closed_form(uc, vc, uvw[2], np.conj(v0), conv, S2, OS2, wpl, wmax,
                                       n)[0])
    for g2 in wrong:
        assert np.abs(g2 - got).max() > 1e-3 * np.abs(got).max()
    k = np.arange(2 * S2 + 1)
    fu, fv = [int((1 - x + np.rint(x)) * OS2) for x in ((uc, vc) if uvw[2] > 0 else (-uc, -vc))]
    abs_sum = np.abs(conv[plane][np.ix_(fv + OS2 * k, fu + OS2 * k)]).sum()
    assert abs(weight * abs_sum - norm[0]) > 1e-3 * abs(norm[0])


# ---- facet phase rotation and uvw transform ---------------------------------------------

def facet_direction(ra0, dec0, ra, dec):
    """(l, m, n) of a facet centre seen from the phase centre (Synthesis Imaging II p. 388)."""
    a = np.pi / 648000.0
    dra = (ra - ra0) * a
    l = np.cos(dec * a) * np.sin(dra)
    m = np.sin(dec * a) * np.cos(dec0 * a) - np.cos(dec * a) * np.sin(dec0 * a) * np.cos(dra)
    return l, m, np.sqrt(1 - l * l - m * m)


def test_phase_rotation_centres_a_source_on_its_facet():
    """A point source at the facet centre has constant phase after the facet rotation: the
    normalised grid sums to the source flux (its image at the facet centre), while without
    rotation (grid_*) it does not."""
    obs = small_obs("facet_single_pol", seed=7, n_ant=10, n_time=20, n_chan=4, flag_fraction=0)
    fc = synth.facet_centres(1, 1.0, obs.phase_centre, np.random.default_rng(3))
    l, m, n = facet_direction(obs.phase_centre[0], obs.phase_centre[1], fc[0, 0], fc[0, 1])
    k = 2 * np.pi * obs.freqs / synth.C_LIGHT
    d = obs.uvw[:, 0] * l + obs.uvw[:, 1] * m + obs.uvw[:, 2] * (n - 1)
    obs.vis[:, :, 0] = (0.7 * np.exp(1j * d[:, None] * k[None, :])).astype(np.complex64)
    b, _ = grid_one(obs, mode="facet_single_pol", npix=128, facets=fc, normalize=True)
    assert b.output.sum() == pytest.approx(0.7, abs=2e-4)
    b2, _ = grid_one(obs, mode="grid_single_pol", npix=128, normalize=True)
    assert abs(b2.output.sum() - 0.7) > 0.1


def test_planar_uvw_transform_moves_footprint():
    """facet_* modes grid at u' = u - w l0/n0, v' = v - w m0/n0 (baseline_transform_policies.h:
    117-149); the footprint origin follows that closed form."""
    fc = np.array([[162000.0 + 3000.0, -108000.0 + 2000.0]])
    l, m, n = facet_direction(162000.0, -108000.0, fc[0, 0], fc[0, 1])
    sc = u_scale(256, 10.0)
    for uvw in ([120.0, -80.0, 3000.0], [-400.0, 250.0, -2500.0]):
        b, _ = grid_one(tiny_obs([uvw], [[1.0]]), mode="facet_sampling_function", npix=256,
                        cell=10.0, facets=fc, sf_count=1)
        u2 = (uvw[0] - uvw[2] * l / n) * sc
        v2 = -(uvw[1] - uvw[2] * m / n) * sc
        np.testing.assert_allclose(b.psf[0, 0, 0], closed_form_footprint(u2, v2, 256),
                                   rtol=1e-9, atol=1e-15)


# ---- Jones ------------------------------------------------------------------------------

def test_jones_identity_and_corruption_roundtrip():
    """Identity Jones leave facet_4_cor unchanged; corrupting with J_p V J_q^H and correcting
    with the same J gives back the uncorrupted grid (correlation_gridding_policies.h:523-542)."""
    obs = small_obs("facet_4_cor", seed=8, n_ant=6, n_time=8, n_chan=3)
    spec = make_spec(obs, "facet_4_cor", npix=128, n_facets=2)
    base = Binding("double", obs, spec)
    oracle.run(base, "facet_4_cor", normalize=False)
    rng = np.random.default_rng(1)
    J = np.eye(2) + 0.1 * (rng.normal(size=(obs.n_time, obs.n_ant, 2, 2)) +
                           1j * rng.normal(size=(obs.n_time, obs.n_ant, 2, 2)))
    for jones in (np.broadcast_to(np.eye(2), J.shape).astype(complex), J):
        o2 = small_obs("facet_4_cor", seed=8, n_ant=6, n_time=8, n_chan=3)
        Jp = jones[o2.time_idx, o2.ant1][:, None]
        Jq = jones[o2.time_idx, o2.ant2][:, None]
        V = o2.vis.astype(np.complex128).reshape(o2.rows, o2.n_chan, 2, 2)
        o2.vis = (Jp @ V @ np.conj(np.swapaxes(Jq, -1, -2))).reshape(o2.rows, o2.n_chan, 4)
        o2.jones = jones
        spec2 = make_spec(o2, "facet_4_cor", npix=128, n_facets=2)
        spec2.jones = jones_reference_layout(o2, 2, np.complex128)
        b = Binding("double", o2, spec2)
        oracle.run(b, "facet_4_cor_corrections", normalize=False)
        np.testing.assert_allclose(b.output, base.output, rtol=0,
                                   atol=2e-6 * np.abs(base.output).max())


def test_singular_jones_is_an_error():
    obs = small_obs("facet_4_cor_corrections", seed=2, n_ant=4, n_time=3, n_chan=2)
    spec = make_spec(obs, "facet_4_cor_corrections", npix=64, n_facets=1)
    spec.jones[0, 1] = 0
    b = Binding("double", obs, spec)
    with pytest.raises(RuntimeError, match="SINGULAR"):
        oracle.run(b, "facet_4_cor_corrections")


# ---- normalisation, uniform weighting, finalize ------------------------------------------

def test_normalize_and_weight_uniformly_closed_forms():
    obs = small_obs("grid_single_pol", seed=6, n_ant=6, n_time=6, n_chan=2)
    spec = ImagingSpec(npix=64, cell=20.0, conv=CONV, support=S, oversample=OS, sf_count=1)
    b = Binding("double", obs, spec, with_psf=True)
    oracle.run(b, "grid_sampling_function", normalize=False)
    norm = oracle.run(b, "grid_single_pol", normalize=False)
    raw = b.output.copy()
    psf = b.psf.copy()
    oracle.load("double").oracle_weight_uniformly(b.byref())
    count = np.float32(1e-7) + psf[:, 0, 0].real
    np.testing.assert_allclose(b.output[:, 0, 0], raw[:, 0, 0] / count, rtol=1e-12)
    weighted = b.output.copy()
    oracle.load("double").oracle_normalize(b.byref())
    np.testing.assert_allclose(b.output, weighted / norm[0], rtol=1e-14)


def weight_uniformly_closed_form(out, psf, cgi):
    """cpu_algorithm/wrapper.cpp:116-135 written out: for facet f, grid slice g and pixel,
    count = 1e-7f + sum over sampling-function channels c of [cgi[c] == g] Re(psf[f][c]) -- the
    visibility channel->grid map indexed by the sampling-function channel -- and every
    correlation of slice g is multiplied by 1/count."""
    F, G, P = out.shape[:3]
    sfc = psf.shape[1]
    res = out.copy()
    for f in range(F):
        for g in range(G):
            count = np.full(out.shape[3:], np.float32(1e-7), dtype=np.float64)
            for c in range(sfc):
                if cgi[c] == g:
                    count = count + psf[f, c, 0].real
            res[f, g] *= 1.0 / count
    return res


def test_weight_uniformly_multi_slice_closed_form():
    """Two facets, two grid slices, two correlations, four sampling-function channels whose
    channel_grid_indicies entries differ from their sampling_function_channel_grid_indicies and
    from the identity: the oracle must use channel_grid_indicies[c] for sampling-function
    channel c, as the reference does."""
    rng = np.random.default_rng(23)
    obs = tiny_obs([[10.0, 5.0, 0.0]], [[1.0, 1.0, 1.0, 1.0]])
    n = 16
    spec = ImagingSpec(npix=n, cell=20.0, conv=CONV, support=S, oversample=OS,
                       facet_centres=np.array([[162000.0, -108000.0], [162100.0, -108050.0]]),
                       n_pol_gridded=2, cube=2, sf_count=4,
                       channel_grid_index=np.array([1, 0, 1, 1], np.uint64),
                       sf_channel_grid_index=np.array([0, 0, 1, 1], np.uint64))
    b = Binding("double", obs, spec, with_psf=True)
    b.output[...] = rng.normal(size=b.output.shape) + 1j * rng.normal(size=b.output.shape)
    b.psf[...] = rng.uniform(0.5, 2.0, size=b.psf.shape) + 1j * rng.normal(size=b.psf.shape)
    raw = b.output.copy()
    oracle.load("double").oracle_weight_uniformly(b.byref())
    cgi = np.array([1, 0, 1, 1])
    np.testing.assert_allclose(b.output, weight_uniformly_closed_form(raw, b.psf, cgi),
                               rtol=1e-13)
    # the sampling-function map or the identity map in its place gives a different answer
    for wrong in (np.array([0, 0, 1, 1]), np.arange(4)):
        assert np.abs(weight_uniformly_closed_form(raw, b.psf, wrong) - b.output).max() > 1e-3


def test_baseline_index_and_stable_order():
    """compute_baseline_index (gpu_algorithm/wrapper.cu:358-366) enumerates the upper triangle
    with autocorrelations row by row; the repack keeps time order inside each baseline."""
    lib = oracle.load("double")
    n = 5
    want = 0
    for a in range(n):
        for b in range(a, n):
            assert lib.oracle_baseline_index(a, b, n) == want
            assert lib.oracle_baseline_index(b, a, n) == want
            want += 1
    a1 = np.array([0, 1, 0, 0, 1, 0], np.uint32)
    a2 = np.array([1, 2, 1, 2, 2, 1], np.uint32)
    starts, perm = oracle.baseline_starts(a1, a2, 3)
    np.testing.assert_array_equal(starts, [0, 0, 3, 4, 4, 6, 6])
    np.testing.assert_array_equal(perm, [0, 2, 5, 3, 1, 4])


def test_gridded_image_matches_direct_dft():
    """Brute force: image of gridded visibilities (shift + unnormalised backward FFT + shift)
    equals the direct Fourier sum of the same samples, each weighted by the image-plane response
    of the footprint it used, at every pixel of a small image."""
    obs = small_obs("grid_single_pol", seed=12, n_ant=5, n_time=4, n_chan=2, flag_fraction=0)
    npix, cell = 32, 60.0
    b, _ = grid_one(obs, npix=npix, cell=cell)
    img = oracle.finalize_images(b.output, 1)[0, 0].astype(np.float64)
    sc = u_scale(npix, cell)
    x = np.arange(npix) - npix // 2
    want = np.zeros((npix, npix))
    lam = obs.wavelengths[0]
    for r in range(obs.rows):
        for c in range(obs.n_chan):
            u = obs.uvw[r, 0] / lam[c] * sc
            v = -obs.uvw[r, 1] / lam[c] * sc
            x0, y0 = np.floor(u) + npix // 2 - S, np.floor(v) + npix // 2 - S
            if not (0 <= x0 and x0 + 2 * S + 3 < npix and 0 <= y0 and y0 + 2 * S + 3 < npix):
                continue
            fu = int((1 - u + np.floor(u)) * OS)
            fv = int((1 - v + np.floor(v)) * OS)
            gx = x0 + np.arange(2 * S + 1) - npix // 2
            gy = y0 + np.arange(2 * S + 1) - npix // 2
            cu = CONV[fu + OS * np.arange(2 * S + 1)]
            cv = CONV[fv + OS * np.arange(2 * S + 1)]
            ex = (cu[None, :] * np.exp(2j * np.pi * np.outer(x, gx) / npix)).sum(1)
            ey = (cv[None, :] * np.exp(2j * np.pi * np.outer(x, gy) / npix)).sum(1)
            want += (obs.vis[r, c, 0] * np.outer(ey, ex)).real
    np.testing.assert_allclose(img, want, rtol=0, atol=2e-5 * np.abs(want).max())


def test_taper_closed_form():
    """grid_taper is the DFT of the sampled kernel: 1 at the centre, symmetric, and equal to the
    direct sum over kernel samples at a few pixels."""
    n = 64
    t = kernels.grid_taper(CONV, S, OS, n)
    assert t[n // 2] == pytest.approx(1.0)
    np.testing.assert_allclose(t[n // 2 + 5], t[n // 2 - 5], rtol=1e-12)
    off = (np.arange(CONV.size) - (S + 1) * OS) / OS
    for k in (0, 7, 40):
        l = k - n // 2
        assert t[k] == pytest.approx((CONV * np.cos(2 * np.pi * off * l / n)).sum() / CONV.sum())
