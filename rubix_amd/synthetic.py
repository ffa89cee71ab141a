"""Seeded synthetic inputs for the particle-to-datacube workloads.

The reference's inputs come from Illustris downloads and the BC03 HDF5
template, neither of which is available offline; these generators produce
inputs of the same shape and physical character (numpy, float32 outputs):

* stars: exponential disk, R ~ Gamma(2, h_R) (surface density ~ exp(-R/h_R)),
  phi ~ U[0, 2pi), height ~ Laplace(0, h_z); flat rotation v_rot along phi_hat
  plus isotropic N(0, sigma^2) per component; one seeded uniform random
  rotation applied to positions and velocities; log-normal masses
  (median 1e5 Msun, sigma_ln 0.5); ages U[0.1, 13.5] Gyr; metallicity
  log-uniform on [1e-4, 0.05].
* SSP grid: layout (n_met, n_age, n_wave) (rubix/spectra/ssp/grid.py:31), log10
  axes geomspace(5e-5, 0.06) and geomspace(0.05, 15) Gyr (slightly wider than
  the star ranges), wavelengths linspace(l0, l1); every template is a
  weighted sum of two Planck curves whose temperatures and weights vary
  smoothly with (age, Z), times 30 seeded Gaussian absorption lines.  All
  values positive.
* geometry: explicit spaxel edges linspace(-L/2, L/2, nb+1), L = kpc/arcsec
  at the redshift (flat LCDM, Planck15 parameters as in
  rubix/cosmology/__init__.py:3) times the field of view; explicit target
  wavelengths; Gaussian PSF (sigma in pixels, rubix/core/psf.py:59-61) and
  25-tap Gaussian LSF (sigma in Angstrom, rubix/telescope/lsf/lsf.py:22-26).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_KMS = 299792.458


def angular_scale_kpc_per_arcsec(z, h=0.6774, om=0.3075):
    """Angular scale of a flat LCDM cosmology (Planck15 parameters).  The
    reference integrates 1/E(z) with a 256-point trapezoid
    (rubix/cosmology/base.py:109-129); here Simpson on 4097 points.  Only
    used to size the synthetic field of view."""
    zz = np.linspace(0.0, z, 4097)
    ez = np.sqrt(om * (1 + zz) ** 3 + (1 - om))
    f = 1.0 / ez
    dz = zz[1] - zz[0]
    integ = dz / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    dc = C_KMS / (100 * h) * integ          # Mpc
    da = dc / (1 + z)
    return da * (math.pi / (180 * 3600)) * 1e3


def _random_rotation(rng):
    """Uniform random rotation matrix (quaternion method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def make_stars(n, seed=0, h_r=3.0, h_z=0.3, v_rot=200.0, sigma_v=50.0,
               mass_median=1e5, mass_sigma_ln=0.5, age_range=(0.1, 13.5),
               met_range=(1e-4, 0.05)):
    rng = np.random.default_rng(seed)
    R = rng.gamma(2.0, h_r, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    zh = rng.laplace(0.0, h_z, n)
    pos = np.stack([R * np.cos(phi), R * np.sin(phi), zh], axis=1)
    vel = np.stack([-v_rot * np.sin(phi), v_rot * np.cos(phi), np.zeros(n)], axis=1)
    vel += rng.normal(0.0, sigma_v, (n, 3))
    rot = _random_rotation(rng)
    pos = pos @ rot.T
    vel = vel @ rot.T
    mass = np.exp(rng.normal(math.log(mass_median), mass_sigma_ln, n))
    age = rng.uniform(age_range[0], age_range[1], n)
    met = 10.0 ** rng.uniform(math.log10(met_range[0]), math.log10(met_range[1]), n)
    f32 = np.float32
    return dict(coords=pos.astype(f32), velocity=vel.astype(f32), mass=mass.astype(f32),
                age=age.astype(f32), metallicity=met.astype(f32))


def make_gas(n, seed=0, h_r=3.0, h_z=0.2, mass_median=2e4, mass_sigma_ln=0.7,
             oxygen_range=(2e-3, 1.5e-2)):
    """Synthetic gas cells: an exponential-ish disk like make_stars, masses in
    solar masses, and metals (n, 10) mass fractions with hydrogen in column 0
    and oxygen in column 4 (the columns the dust step reads)."""
    rng = np.random.default_rng(seed)
    R = rng.gamma(2.0, h_r, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    zh = rng.laplace(0.0, h_z, n)
    pos = np.stack([R * np.cos(phi), R * np.sin(phi), zh], axis=1)
    vel = rng.normal(0.0, 30.0, (n, 3))
    mass = np.exp(rng.normal(math.log(mass_median), mass_sigma_ln, n))
    metals = np.zeros((n, 10))
    metals[:, 0] = rng.uniform(0.70, 0.76, n)
    metals[:, 1] = 0.25
    metals[:, 4] = 10.0 ** rng.uniform(math.log10(oxygen_range[0]), math.log10(oxygen_range[1]), n)
    metals[:, 2:4] = 1e-3
    metals[:, 5:] = 5e-4
    f32 = np.float32
    return dict(coords=pos.astype(f32), velocity=vel.astype(f32), mass=mass.astype(f32),
                metals=metals.astype(f32))


def _planck(lam_A, T):
    lam = lam_A * 1e-10
    hc_k = 1.438776877e-2  # m K
    with np.errstate(over="ignore"):
        return 1.0 / (lam ** 5 * np.expm1(hc_k / (lam * T)))


def make_ssp_grid(n_met, n_age, n_wave, wave_range, seed=1,
                  met_range=(5e-5, 0.06), age_range=(0.05, 15.0)):
    rng = np.random.default_rng(seed)
    met = np.geomspace(met_range[0], met_range[1], n_met)
    age = np.geomspace(age_range[0], age_range[1], n_age)
    wave = np.linspace(wave_range[0], wave_range[1], n_wave)
    lz = (np.log10(met) - np.log10(met[0])) / max(np.log10(met[-1] / met[0]), 1e-30)
    la = (np.log10(age) - np.log10(age[0])) / max(np.log10(age[-1] / age[0]), 1e-30)
    # smooth temperature / weight fields over (Z, age)
    t1_lo, t1_hi = rng.uniform(3000, 4000), rng.uniform(5000, 6000)
    t2_lo, t2_hi = rng.uniform(8000, 11000), rng.uniform(15000, 20000)
    T1 = t1_hi - (t1_hi - t1_lo) * (0.6 * la[None, :] + 0.4 * lz[:, None])
    T2 = t2_hi - (t2_hi - t2_lo) * (0.8 * la[None, :] + 0.2 * lz[:, None])
    w_hot = np.exp(-3.0 * la[None, :]) * (1.0 - 0.3 * lz[:, None]) + 0.02
    centers = rng.uniform(wave[0], wave[-1], 30)
    depth0 = rng.uniform(0.05, 0.5, 30)
    widths = rng.uniform(2.0, 8.0, 30)
    flux = np.empty((n_met, n_age, n_wave))
    b_ref = _planck(np.array([5500.0]), 5800.0)[0]
    for i in range(n_met):
        for j in range(n_age):
            cont = (1 - w_hot[i, j]) * _planck(wave, T1[i, j]) + w_hot[i, j] * _planck(wave, T2[i, j])
            cont = cont / b_ref
            depth = np.clip(depth0 * (0.5 + 0.5 * lz[i] + 0.3 * la[j]), 0.0, 0.9)
            lines = np.ones_like(wave)
            for c, d, s in zip(centers, depth, widths):
                lines *= 1.0 - d * np.exp(-0.5 * ((wave - c) / s) ** 2)
            # luminosity per solar mass fades with age
            flux[i, j] = cont * lines * 1e-4 * (age[j] / age[0]) ** -0.8
    f32 = np.float32
    return dict(ax_met=np.log10(met).astype(f32), ax_age=np.log10(age).astype(f32),
                wave=wave.astype(f32), flux=flux.astype(f32), log_axes=True,
                met=met.astype(f32), age=age.astype(f32))


@dataclass
class Workload:
    name: str
    stars: dict
    grid: dict
    geom: dict
    meta: dict = field(default_factory=dict)

    @property
    def n_stars(self):
        return len(self.stars["mass"])


def make_geometry(nb, arcsec_per_spaxel, z, wave_target, psf_sigma_arcsec, psf_size,
                  lsf_sigma, wave_res):
    fov = nb * arcsec_per_spaxel
    L = angular_scale_kpc_per_arcsec(z) * fov
    edges = np.linspace(-L / 2, L / 2, nb + 1).astype(np.float32)
    sig_px = psf_sigma_arcsec / arcsec_per_spaxel
    return dict(edges=edges, nb=nb, wave_target=np.asarray(wave_target, np.float32), z=float(z),
                psf_sigma=float(sig_px), psf_size=int(psf_size), lsf_sigma=float(lsf_sigma),
                wave_res=float(wave_res), fov_arcsec=float(fov), aperture_kpc=float(L))


def muse_wave(n=3721):
    """4700.15 + 1.25 k, k < 3721, computed in float64 then rounded."""
    return (4700.15 + 1.25 * np.arange(n)).astype(np.float32)


def config1(n_stars=10_000, seed=0):
    """Plain forward at small size: 1e4 stars, grid 6 Z x 20 ages x 1000 l on
    3000-9000 A, 25x25 spaxels at 0.5", 1000 target bins, PSF sigma 0.6" on
    11x11, LSF sigma 1.2 A."""
    stars = make_stars(n_stars, seed=seed)
    grid = make_ssp_grid(6, 20, 1000, (3000.0, 9000.0), seed=1)
    wt = (3400.0 + 5.0 * np.arange(1000)).astype(np.float32)
    geom = make_geometry(25, 0.5, 0.05, wt, 0.6, 11, 1.2, 5.0)
    return Workload("cfg1", stars, grid, geom)


def config2(n_stars=1_000_000, seed=0):
    """Single galaxy at full shape: 1e6 stars, grid 12 Z x 100 ages x 5000 l on
    3000-10000 A, 50x50 spaxels at 0.2" (MUSE), 3721 bins 4700.15 + 1.25 k,
    PSF sigma 0.8" on 15x15, LSF sigma 1.0 A."""
    stars = make_stars(n_stars, seed=seed)
    grid = make_ssp_grid(12, 100, 5000, (3000.0, 10000.0), seed=1)
    geom = make_geometry(50, 0.2, 0.1, muse_wave(), 0.8, 15, 1.0, 1.25)
    return Workload("cfg2", stars, grid, geom)


def config5(n_stars=10_000_000, seed=0):
    """Large galaxy: 1e7 stars into 150x150 spaxels at 0.2" x 3721 bins with
    the small SSP grid (6 Z x 20 ages x 1000 l on 3000-9000 A, z = 0.05) and
    the small workload's PSF (sigma 0.6" on 11x11) and LSF (sigma 1.2 A)."""
    stars = make_stars(n_stars, seed=seed)
    grid = make_ssp_grid(6, 20, 1000, (3000.0, 9000.0), seed=1)
    geom = make_geometry(150, 0.2, 0.05, muse_wave(), 0.6, 11, 1.2, 1.25)
    return Workload("cfg5", stars, grid, geom)


def config4(n_gal=64, seed=3, scale=1.0):
    """Batched ragged workload: n_gal galaxies with star counts log-uniform on
    5e4..2e6 (times scale) and spaxel grids from {25, 50, 100} at 0.2", the
    shared 3721-bin axis and the small SSP grid (z = 0.05), PSF sigma 0.6" on
    11x11, LSF sigma 1.2 A.  Galaxy g's stars use seed 100 + g."""
    ns, nbs = config4_sizes(n_gal, seed)
    grid = make_ssp_grid(6, 20, 1000, (3000.0, 9000.0), seed=1)
    out = []
    for g, (n, nb) in enumerate(zip(ns, nbs)):
        n = max(1, int(round(n * scale)))
        stars = make_stars(n, seed=100 + g)
        geom = make_geometry(int(nb), 0.2, 0.05, muse_wave(), 0.6, 11, 1.2, 1.25)
        out.append(Workload(f"cfg4-{g}", stars, grid, geom))
    return out


def config4_sizes(n_gal=64, seed=3):
    """Star counts log-uniform on 5e4..2e6 and spaxel grids from
    {25, 50, 100} per galaxy (batched ragged workload)."""
    rng = np.random.default_rng(seed)
    ns = np.round(10 ** rng.uniform(math.log10(5e4), math.log10(2e6), n_gal)).astype(np.int64)
    nbs = rng.choice([25, 50, 100], n_gal)
    return ns, nbs


SSP_TEMPLATE_NAMES = {(6, 20, 1000): "synthetic_6x20x1000", (12, 100, 5000): "synthetic_12x100x5000",
                      (6, 20, 3000): "synthetic_6x20x3000_nir",
                      (6, 20, 4000): "synthetic_6x20x4000_wide"}
SSP_TEMPLATE_RANGES = {"synthetic_6x20x1000": (6, 20, 1000, (3000.0, 9000.0)),
                       "synthetic_12x100x5000": (12, 100, 5000, (3000.0, 10000.0)),
                       "synthetic_6x20x3000_nir": (6, 20, 3000, (8000.0, 20000.0)),
                       "synthetic_6x20x4000_wide": (6, 20, 4000, (2500.0, 56000.0))}


def instrument_workload(name, template, n_stars, z=0.05, psf_sigma=1.0, psf_size=5, seed=0,
                        h_r=3.0, telescopes=None):
    """Stars, grid and geometry of a named instrument of the telescope table:
    nb = floor(fov / spatial_res), spaxel edges from the float32 arange of
    the angular aperture at redshift z (the reference's construction,
    rubix/telescope/utils.py:11-37), target wavelengths = the instrument's
    float32 wave_seq, PSF sigma in pixels on psf_size^2 taps, LSF sigma =
    lsf_fwhm / (2 sqrt(2 ln 2)) Angstrom at the instrument's wave_res.  The
    grid is the registered synthetic template `template`."""
    from .telescope import TelescopeFactory, calculate_spatial_bin_edges

    tel = TelescopeFactory(telescopes).create_telescope(name)
    edges, _ = calculate_spatial_bin_edges(tel.fov, tel.sbin, z)
    nm, na, nw, rng = SSP_TEMPLATE_RANGES[template]
    grid = make_ssp_grid(nm, na, nw, rng, seed=1)
    lsf_sigma = float(tel.lsf_fwhm) / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    geom = dict(edges=np.asarray(edges, np.float32), nb=int(tel.sbin),
                wave_target=np.asarray(tel.wave_seq, np.float32), z=float(z),
                psf_sigma=float(psf_sigma), psf_size=int(psf_size), lsf_sigma=lsf_sigma,
                wave_res=float(tel.wave_res), fov_arcsec=float(tel.fov),
                aperture_kpc=float(edges[-1] - edges[0]))
    stars = make_stars(n_stars, seed=seed, h_r=h_r)
    return Workload(name, stars, grid, geom, meta={"template": template})


def instrument_config(w: Workload, device="cuda"):
    """Step-API config naming the instrument itself (no explicit edges or
    wavelengths: the steps derive them from the telescope table as the
    reference does)."""
    g = w.geom
    return {
        "device": device,
        "galaxy": {"dist_z": g["z"], "rotation": {"type": "face-on"}},
        "data": {"args": {"particle_type": ["stars"]}},
        "telescope": {"name": w.name,
                      "psf": {"name": "gaussian", "size": g["psf_size"], "sigma": g["psf_sigma"]},
                      "lsf": {"sigma": g["lsf_sigma"]},
                      "noise": {"signal_to_noise": 1, "noise_distribution": "normal"}},
        "ssp": {"template": {"name": w.meta["template"]}, "method": "linear"},
        "cosmology": {"name": "PLANCK15"},
    }


def step_config(w: Workload, device="cuda"):
    """Step-API config for a workload: a telescope definition with the
    workload's spaxel size and wavelength step, explicit spaxel edges and
    target wavelengths, the Gaussian PSF (sigma in pixels) and LSF, noise at
    S/N 1 (normal), and the registered synthetic SSP template of the same
    grid."""
    g = w.geom
    nb = g["nb"]
    arcsec = g["fov_arcsec"] / nb
    nm, na, nw = w.grid["flux"].shape
    name = SSP_TEMPLATE_NAMES[(nm, na, nw)]
    wt = np.asarray(g["wave_target"], dtype=np.float64)
    tel = {"fov": g["fov_arcsec"], "spatial_res": arcsec,
           "wave_range": [float(wt[0]), float(wt[-1]) + g["wave_res"] / 2], "wave_res": g["wave_res"],
           "lsf_fwhm": None, "signal_to_noise": None, "aperture_type": "square",
           "pixel_type": "square"}
    return {
        "device": device,
        "galaxy": {"dist_z": g["z"], "rotation": {"type": "face-on"}},
        "data": {"args": {"particle_type": ["stars"]}},
        "telescopes": {f"SYNTHETIC_{w.name}": tel},
        "telescope": {"name": f"SYNTHETIC_{w.name}", "spatial_bin_edges": g["edges"],
                      "wave_seq": g["wave_target"],
                      "psf": {"name": "gaussian", "size": g["psf_size"], "sigma": g["psf_sigma"]},
                      "lsf": {"sigma": g["lsf_sigma"]},
                      "noise": {"signal_to_noise": 1, "noise_distribution": "normal"}},
        "ssp": {"template": {"name": name}, "method": "linear",
                "dust": {"extinction_model": "Cardelli89", "Rv": 3.1, "dust_grain_density": 3.5}},
        "cosmology": {"name": "PLANCK15"},
    }
