"""Telescope geometry (setup-time, host side): instrument record, spatial bin
edges and the target wavelength axis.

These are scalars and small arrays computed once per step factory, as the
reference does (rubix/telescope/factory.py:38-100,
rubix/telescope/utils.py:11-53, rubix/cosmology/base.py:109-129,290-311).
Explicit arrays may be given in the step config to avoid the reference's
float-arange length hazards (the spatial edge array can have nb+1 or nb+2
entries, and 4700.15..9351.4 step 1.25 gives 3721 or 3722 wavelengths
depending on the float type).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .config import telescopes_config

C_SPEED = 2.99792458e8  # m/s, rubix/cosmology/base.py:11


@dataclass
class Cosmology:
    """Flat wCDM distances (rubix/cosmology/base.py; PLANCK15 parameters at
    rubix/cosmology/__init__.py:3), evaluated in float32 like the reference."""
    Om0: float = 0.3075
    w0: float = -1.0
    wa: float = 0.0
    h: float = 0.6774

    def _ez(self, z):
        f = np.float32
        z = np.asarray(z, dtype=f)
        a = f(1.0) / (f(1.0) + z)
        de = a ** (f(-3.0) * (f(1.0) + f(self.w0) + f(self.wa))) * np.exp(f(-3.0) * f(self.wa) * (f(1.0) - a))
        t = f(self.Om0) * (f(1.0) + z) ** 3 + (f(1.0) - f(self.Om0)) * de
        return np.sqrt(t).astype(f)

    def comoving_distance_to_z(self, z):
        """256-point trapezoid of 1/E(z) (rubix/cosmology/base.py:109-129 with
        the scan trapezoid of rubix/cosmology/utils.py:40-68), in Mpc."""
        f = np.float32
        zt = np.linspace(0, z, 256).astype(f)
        y = (f(1.0) / self._ez(zt)).astype(f)
        acc = f(0.0)
        for k in range(1, 256):
            acc = f(acc + (zt[k] - zt[k - 1]) * (y[k] + y[k - 1]) / f(2.0))
        return f(acc * f(C_SPEED * 1e-5) / f(self.h))

    def angular_diameter_distance_to_z(self, z):
        return np.float32(self.comoving_distance_to_z(z) / np.float32(1 + z))

    def angular_scale(self, z):
        """kpc / arcsec (rubix/cosmology/base.py:290-311)."""
        da = self.angular_diameter_distance_to_z(z)
        return np.float32(da * np.float32(math.pi / (180 * 60 * 60)) * np.float32(1e3))


PLANCK15 = Cosmology()


def calculate_spatial_bin_edges(fov, spatial_bins, dist_z, cosmology=PLANCK15):
    """rubix/telescope/utils.py:11-37: aperture = angular_scale * fov,
    s = aperture / bins, edges = arange(-aperture/2, aperture/2 + s, s) in
    float32 (may hold bins+1 or bins+2 entries)."""
    f = np.float32
    ang = f(cosmology.angular_scale(dist_z))
    aperture = f(ang * f(fov))
    s = f(aperture / f(spatial_bins))
    return np.arange(-aperture / 2, aperture / 2 + s, s, dtype=f), s


def calculate_wave_seq(wave_range, wave_res):
    """rubix/telescope/utils.py:40-53: arange(start, stop, res) in float32."""
    return np.arange(wave_range[0], wave_range[1], wave_res, dtype=np.float32)


def calculate_wave_edges(wave_seq, wave_res):
    """rubix/telescope/utils.py:56-74: arange(first - res/2, last + res/2,
    res) in float32 (bin edges around the wavelength sequence)."""
    f = np.float32
    start = f(f(wave_seq[0]) - f(wave_res / 2))
    end = f(f(wave_seq[-1]) + f(wave_res / 2))
    return np.arange(start, end, wave_res, dtype=f)


def square_aperture(sbin):
    """rubix/telescope/apertures.py:47-59: every spaxel (flattened sbin x sbin
    ones)."""
    return np.ones(int(sbin) * int(sbin), dtype=np.float32)


def circular_aperture(sbin):
    """rubix/telescope/apertures.py:62-81: spaxel (row r, column c) is in when
    the distance of (c + 1, sbin - r) from (sbin/2 + 0.5, sbin/2 + 0.5) is
    at most sbin/2, evaluated in float32; flattened row-major."""
    f = np.float32
    s = int(sbin)
    centre = f(s / 2 + 0.5)
    xx = np.tile(np.arange(1, s + 1, dtype=f), (s, 1)) - centre
    yy = np.tile(np.arange(s, 0, -1, dtype=f), (s, 1)).T - centre
    rr = np.sqrt(xx * xx + yy * yy)
    return (rr <= f(s / 2)).astype(f).reshape(-1)


def hexagonal_aperture(sbin):
    """rubix/telescope/apertures.py:11-44: pixel (x, y), 1-based, with offsets
    xx = x - c, yy = y - c from the centre c = sbin/2 + 0.5 is in the flat-
    topped hexagon of half-width sbin/2 and half-height sbin*sqrt(3)/4 when
    2 (sbin/4)(sbin sqrt3/4) - (sbin/4)|yy| - (sbin sqrt3/4)|xx| >= 0,
    |xx| < sbin/2 and |yy| < sbin sqrt3/4, all in float32 as the reference
    evaluates them; the mask is indexed [x - 1, y - 1] and flattened."""
    f = np.float32
    s = int(sbin)
    c = s / 2 + 0.5
    off = np.arange(1, s + 1, dtype=np.float64) - c          # exact half-integers
    xx = np.abs(off.astype(f))[:, None]
    yy = np.abs(off.astype(f))[None, :]
    r3 = f(np.sqrt(f(3.0)))
    h = f(f(f(s) * r3) / f(4.0))                             # sbin * sqrt(3) / 4
    q = f(s / 4)
    rr = f(f(2.0) * q * h) - q * yy - h * xx
    inside = (rr >= 0) & (xx < f(s / 2)) & (yy < h)
    return inside.astype(f).reshape(-1)


APERTURES = {"square": square_aperture, "circular": circular_aperture,
             "hexagonal": hexagonal_aperture}


def get_aperture(kind, sbin):
    """rubix/telescope/factory.py:102-110."""
    if kind not in APERTURES:
        raise ValueError(f"Unknown aperture type: {kind}")
    return APERTURES[kind](sbin)


@dataclass
class Telescope:
    """The fields of rubix/telescope/base.py the hot path consumes."""
    name: str
    fov: float
    spatial_res: float
    wave_range: list
    wave_res: float
    lsf_fwhm: Optional[float]
    signal_to_noise: Optional[float]
    sbin: int
    pixel_type: str
    aperture_type: str
    wave_seq: np.ndarray
    aperture_region: Optional[np.ndarray] = None
    wave_edges: Optional[np.ndarray] = None


class TelescopeFactory:
    """rubix/telescope/factory.py:38-100: sbin = floor(fov / spatial_res),
    the aperture mask of the instrument's aperture type, wave_seq =
    arange(wave_range, wave_res) and its bin edges."""

    def __init__(self, telescopes_config_=None):
        self.telescopes_config = telescopes_config_ or telescopes_config

    def create_telescope(self, name: str) -> Telescope:
        if name not in self.telescopes_config:
            raise ValueError(f"Telescope {name} not found in config")
        c = self.telescopes_config[name]
        sbin = int(np.floor(c["fov"] / c["spatial_res"]))
        aperture_type = c.get("aperture_type", "square")
        aperture = get_aperture(aperture_type, sbin)
        wave_seq = calculate_wave_seq(c["wave_range"], c["wave_res"])
        return Telescope(name=name, fov=c["fov"], spatial_res=c["spatial_res"],
                         wave_range=list(c["wave_range"]), wave_res=c["wave_res"],
                         lsf_fwhm=c.get("lsf_fwhm"), signal_to_noise=c.get("signal_to_noise"),
                         sbin=sbin, pixel_type=c.get("pixel_type", "square"),
                         aperture_type=aperture_type, wave_seq=wave_seq,
                         aperture_region=aperture,
                         wave_edges=calculate_wave_edges(wave_seq, c["wave_res"]))
