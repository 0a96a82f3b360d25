"""Seeded synthetic observations with the array layouts bullseye's Measurement Set loader
produces (bullseye/data/data_set_loader.py:183-335): rows in time order, per row uvw in
metres, visibilities/weights/flags as [row][channel][correlation].

uvw follow the rotation used by bullseye's benchmark driver
(bullseye/mo/benchmarks/benchmark.cpp:192-202): for an equatorial baseline (Lx, Ly, Lz),
hour angle H and declination d,
    u = -sin H Lx + cos H Ly
    v = -sin d cos H Lx - sin d sin H Ly + cos d Lz
    w =  cos d cos H Lx + cos d sin H Ly + sin d Lz.
Model visibilities are V = sum_s S_s exp(+2 pi i (u l_s + v m_s + w (n_s - 1)) / lambda), the
sign for which bullseye's facet phase rotation (phase_transform_policies.h:122-146, rotor
exp(+i 2 pi (u dl + v dm + w dn)) with dl = -l_facet) brings a source at the facet centre to
zero phase.  (The reference's never-launched dft.h:74-79 assumes the opposite sign.)
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

C_LIGHT = 299792458.0
ARCSEC = np.pi / 648000.0
DEG = np.pi / 180.0


@dataclass
class Observation:
    n_ant: int
    n_time: int
    n_chan: int
    n_corr: int
    ant1: np.ndarray          # uint32 [rows]
    ant2: np.ndarray          # uint32 [rows]
    time_idx: np.ndarray      # uint64 [rows] index into the Jones time axis
    uvw: np.ndarray           # float64 [rows, 3] metres
    freqs: np.ndarray         # float64 [chan] Hz
    vis: np.ndarray           # complex64 [rows, chan, corr]
    weights: np.ndarray       # float32 [rows, chan, corr]
    flags: np.ndarray         # bool [rows, chan, corr]
    flagged_rows: np.ndarray  # bool [rows]
    field: np.ndarray         # uint32 [rows]
    spw: np.ndarray           # uint32 [rows]
    phase_centre: tuple       # (ra, dec) arcsec
    sources: np.ndarray       # float64 [nsrc, 3] (l, m, flux)
    jones: Optional[np.ndarray] = None   # complex128 [time, ant, 2, 2] (intrinsic set)
    meta: dict = field(default_factory=dict)
    wl: Optional[np.ndarray] = None      # explicit [spw, chan] wavelengths (else c / freqs)

    @property
    def rows(self):
        return self.uvw.shape[0]

    @property
    def baseline_count(self):
        return self.n_ant * (self.n_ant + 1) // 2

    @property
    def wavelengths(self):
        if self.wl is not None:
            return self.wl
        return (C_LIGHT / self.freqs)[None, :]


def antennas_disc(n_ant, diameter_m, rng):
    r = 0.5 * diameter_m * np.sqrt(rng.uniform(0, 1, n_ant))
    th = rng.uniform(0, 2 * np.pi, n_ant)
    return np.stack([r * np.cos(th), r * np.sin(th), rng.normal(0, 1.0, n_ant)], axis=1)


def antennas_log_radial(n_ant, max_radius_m, rng, arms=3):
    r = np.exp(rng.uniform(np.log(20.0), np.log(max_radius_m), n_ant))
    th = (np.arange(n_ant) % arms) * (2 * np.pi / arms) + rng.normal(0, 0.15, n_ant)
    return np.stack([r * np.cos(th), r * np.sin(th), rng.normal(0, 1.0, n_ant)], axis=1)


def enu_to_equatorial(enu, latitude_rad):
    e, n, u = enu[:, 0], enu[:, 1], enu[:, 2]
    sl, cl = np.sin(latitude_rad), np.cos(latitude_rad)
    return np.stack([-sl * n + cl * u, e, cl * n + sl * u], axis=1)


def random_directions(n, radius_rad, rng):
    """(l, m) uniform in a disc of the given radius."""
    r = radius_rad * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    return r * np.cos(th), r * np.sin(th)


def lm_to_radec(l, m, ra0, dec0):
    """Inverse of the direction-cosine relations of Synthesis Imaging II p. 388 (radians)."""
    n = np.sqrt(1 - l ** 2 - m ** 2)
    dec = np.arcsin(m * np.cos(dec0) + n * np.sin(dec0))
    dra = np.arctan2(l, n * np.cos(dec0) - m * np.sin(dec0))
    return ra0 + dra, dec


def simulate(n_ant=64, array="disc", array_size_m=3000.0, n_time=240, hours=8.0,
             dec_deg=-30.0, ra_deg=45.0, latitude_deg=-30.7, f_lo=1.40e9, f_hi=1.42e9,
             n_chan=16, n_corr=1, n_src=20, src_radius_deg=0.5, flux=(0.1, 1.0), noise=0.05,
             flag_fraction=0.05, jones_sigma=None, seed=0, predict_chunk=32768, device=None):
    """device: optional torch device on which the point-source prediction runs (float64); the
    result equals the NumPy prediction to rounding."""
    rng = np.random.default_rng(seed)
    if array == "disc":
        enu = antennas_disc(n_ant, array_size_m, rng)
    else:
        enu = antennas_log_radial(n_ant, array_size_m, rng)
    xyz = enu_to_equatorial(enu, latitude_deg * DEG)
    a1, a2 = np.triu_indices(n_ant, k=1)
    nbl = a1.size
    ha = np.linspace(-hours / 2, hours / 2, n_time) * (2 * np.pi / 24.0)
    dec = dec_deg * DEG
    ant1 = np.tile(a1, n_time).astype(np.uint32)
    ant2 = np.tile(a2, n_time).astype(np.uint32)
    tidx = np.repeat(np.arange(n_time), nbl).astype(np.uint64)
    L = xyz[ant1] - xyz[ant2]
    H = ha[tidx]
    sh, ch, sd, cd = np.sin(H), np.cos(H), np.sin(dec), np.cos(dec)
    uvw = np.stack([-sh * L[:, 0] + ch * L[:, 1],
                    -sd * ch * L[:, 0] - sd * sh * L[:, 1] + cd * L[:, 2],
                    cd * ch * L[:, 0] + cd * sh * L[:, 1] + sd * L[:, 2]], axis=1)
    freqs = np.linspace(f_lo, f_hi, n_chan)
    rows = uvw.shape[0]

    ls, ms = random_directions(n_src, src_radius_deg * DEG, rng)
    fl = rng.uniform(flux[0], flux[1], n_src)
    ns = np.sqrt(1 - ls ** 2 - ms ** 2) - 1
    sources = np.stack([ls, ms, fl], axis=1)

    # predict in float64: phase per (row, source) in metres, then per channel
    model = np.empty((rows, n_chan), dtype=np.complex128)
    k = 2 * np.pi * freqs / C_LIGHT
    if device is not None:
        import torch
        kt = torch.as_tensor(k, device=device)
        lt, mt, nt = (torch.as_tensor(x, device=device) for x in (ls, ms, ns))
        ft = torch.as_tensor(fl, device=device, dtype=torch.complex128)
        step = max(predict_chunk, 1 << 18)
        for r0 in range(0, rows, step):
            r1 = min(rows, r0 + step)
            u = torch.as_tensor(uvw[r0:r1], device=device)
            d = u[:, 0:1] * lt + u[:, 1:2] * mt + u[:, 2:3] * nt
            ph = torch.exp(1j * (d[:, None, :] * kt[None, :, None]))
            model[r0:r1] = (ph @ ft).cpu().numpy()
    else:
        for r0 in range(0, rows, predict_chunk):
            r1 = min(rows, r0 + predict_chunk)
            d = uvw[r0:r1, 0:1] * ls + uvw[r0:r1, 1:2] * ms + uvw[r0:r1, 2:3] * ns
            ph = np.exp(1j * d[:, None, :] * k[None, :, None])
            model[r0:r1] = ph @ fl
    vis = np.zeros((rows, n_chan, n_corr), dtype=np.complex128)
    if n_corr == 1:
        vis[..., 0] = model
    elif n_corr == 2:
        vis[..., 0] = model
        vis[..., 1] = model
    else:
        vis[..., 0] = model
        vis[..., 3] = model
    jones = None
    if jones_sigma is not None and n_corr == 4:
        jones = (np.eye(2)[None, None] +
                 rng.normal(0, jones_sigma, (n_time, n_ant, 2, 2)) +
                 1j * rng.normal(0, jones_sigma, (n_time, n_ant, 2, 2)))
        for r0 in range(0, rows, predict_chunk):
            r1 = min(rows, r0 + predict_chunk)
            Jp = jones[tidx[r0:r1], ant1[r0:r1]][:, None]
            Jq = jones[tidx[r0:r1], ant2[r0:r1]][:, None]
            V = vis[r0:r1].reshape(r1 - r0, n_chan, 2, 2)
            V = Jp @ V @ np.conj(np.swapaxes(Jq, -1, -2))
            vis[r0:r1] = V.reshape(r1 - r0, n_chan, 4)
    if noise > 0:
        s = noise / np.sqrt(2.0)
        vis += rng.normal(0, s, vis.shape) + 1j * rng.normal(0, s, vis.shape)
    flags = rng.uniform(0, 1, vis.shape) < flag_fraction
    return Observation(
        n_ant=n_ant, n_time=n_time, n_chan=n_chan, n_corr=n_corr, ant1=ant1, ant2=ant2,
        time_idx=tidx, uvw=uvw, freqs=freqs, vis=vis.astype(np.complex64),
        weights=np.ones(vis.shape, dtype=np.float32), flags=flags,
        flagged_rows=np.zeros(rows, dtype=bool), field=np.zeros(rows, dtype=np.uint32),
        spw=np.zeros(rows, dtype=np.uint32), phase_centre=(ra_deg * 3600.0, dec_deg * 3600.0),
        sources=sources, jones=jones,
        meta=dict(seed=seed, array=array, array_size_m=array_size_m, hours=hours))


def select_rows(obs, rows):
    """The observation restricted to the given row indexes (order kept)."""
    import dataclasses
    r = np.asarray(rows)
    return dataclasses.replace(
        obs, ant1=obs.ant1[r], ant2=obs.ant2[r], time_idx=obs.time_idx[r], uvw=obs.uvw[r],
        vis=obs.vis[r], weights=obs.weights[r], flags=obs.flags[r],
        flagged_rows=obs.flagged_rows[r], field=obs.field[r], spw=obs.spw[r],
        meta=dict(obs.meta, rows_selected=int(r.size)))


def facet_centres(n_facets, radius_deg, phase_centre, rng):
    """Facet centres (RA, Dec in arcsec) uniform within radius_deg of the phase centre."""
    ra0, dec0 = phase_centre[0] * ARCSEC, phase_centre[1] * ARCSEC
    l, m = random_directions(n_facets, radius_deg * DEG, rng)
    ra, dec = lm_to_radec(l, m, ra0, dec0)
    return np.stack([ra / ARCSEC, dec / ARCSEC], axis=1)


# Workloads (sizes from the imaging configurations).  'small' variants keep the same
# structure at sizes the CPU oracle finishes in seconds.
CONFIGS = {
    "c0": dict(sim=dict(n_ant=64, array="disc", array_size_m=3000.0, n_time=240, n_chan=16,
                        f_lo=1.40e9, f_hi=1.42e9, n_corr=1),
               image=dict(npix=1024, cell=2.0, support=3, oversample=63, facets=0)),
    "c1": dict(sim=dict(n_ant=64, array="disc", array_size_m=3000.0, n_time=240, n_chan=64,
                        f_lo=1.40e9, f_hi=1.42e9, n_corr=4, jones_sigma=0.05),
               image=dict(npix=512, cell=2.0, support=3, oversample=63, facets=25,
                          facet_radius_deg=2.0)),
    "c2": dict(sim=dict(n_ant=128, array="logradial", array_size_m=5000.0, n_time=480,
                        n_chan=32, f_lo=1.0e9, f_hi=1.5e9, n_corr=1),
               image=dict(npix=4096, cell=1.5, support=31, oversample=8, facets=0,
                          wplanes=128)),
}


def batched_facets(obs, n_facets, seed, device=None):
    """The batched uneven-facet workload over an observation of the w-projection workload:
    n_facets facets with sizes log-uniform over 128^2-2048^2 (even) at 1.5", centres within 2 deg
    of the phase centre, each with one of seven w-kernel stacks (16, 24, 32, 48, 64, 96 or 128
    planes, half supports 3 -> 31 linear in |w|, OS x8, complex64) and a quarter of them
    sampling-function facets.  Returns (facets [bullseye_amd.batch.BatchFacet], kernel sets
    [(conv, half_supports)], plane counts, wmax, sizes, psf flags)."""
    from . import kernels
    from .batch import BatchFacet
    im = CONFIGS["c2"]["image"]
    wmax = float(np.float32(np.abs(obs.uvw[:, 2:3] / obs.wavelengths[0][None, :]).max()))
    plane_counts = [16, 24, 32, 48, 64, 96, 128]
    ksets = []
    for P in plane_counts:
        halves = kernels.linear_plane_supports(P, 3, im["support"])
        conv = kernels.wproj_kernels(im["support"], im["oversample"], P, wmax, im["npix"],
                                     im["cell"], half_supports=halves, dtype=np.complex64,
                                     device=device)
        ksets.append((conv, halves))
    rng = np.random.default_rng(seed)
    nf = n_facets
    sizes = (np.round(np.exp(rng.uniform(np.log(128), np.log(2048), nf)) / 2) * 2).astype(int)
    kid = rng.integers(0, len(plane_counts), nf)
    psf = np.zeros(nf, bool)
    psf[rng.choice(nf, nf // 4, replace=False)] = True
    cen = facet_centres(nf, 2.0, obs.phase_centre, rng)
    facets = [BatchFacet(int(sizes[i]), im["cell"], (float(cen[i, 0]), float(cen[i, 1])),
                         int(kid[i]), bool(psf[i])) for i in range(nf)]
    return facets, ksets, plane_counts, wmax, sizes, psf


def differentiable_facet(obs, seed, device=None, n_samples=1_000_000, n_model=500):
    """The differentiable-imaging workload over an observation of the w-projection workload: a
    seeded subset of rows holding n_samples samples (all channels), one 2048^2 facet at 1.5"
    centred 0.2 deg from the phase centre, 32 w-planes (half supports 3 -> 31 linear in |w|, OS
    x8, complex64) and a sparse real model image of n_model pixels ~ U(0, 1).  Returns
    (observation subset, ImagingSpec, half supports, model image [1, 1, 1, 2048, 2048])."""
    from . import kernels
    from .binding import ImagingSpec
    rng = np.random.default_rng(seed)
    keep = np.sort(rng.choice(obs.rows, n_samples // obs.n_chan, replace=False))
    sub = select_rows(obs, keep)
    npix, cell, planes, S, OS = 2048, 1.5, 32, 31, 8
    wmax = float(np.float32(np.abs(sub.uvw[:, 2:3] / sub.wavelengths[0][None, :]).max()))
    halves = kernels.linear_plane_supports(planes, 3, S)
    conv = kernels.wproj_kernels(S, OS, planes, wmax, npix, cell, half_supports=halves,
                                 dtype=np.complex64, device=device)
    ra0, dec0 = sub.phase_centre
    centre = np.array([[ra0 + 0.2 * 3600.0 / np.cos(np.radians(dec0 / 3600.0)), dec0]])
    spec = ImagingSpec(npix=npix, cell=cell, conv=conv, support=S, oversample=OS,
                       facet_centres=centre, wplanes=planes, wmax=wmax)
    img = np.zeros(npix * npix)
    img[rng.choice(img.size, n_model, replace=False)] = rng.uniform(0, 1, n_model)
    return sub, spec, halves, img.reshape(1, 1, 1, npix, npix)
