"""SSP template grid (host arrays) and the template registry.

SSPGrid mirrors rubix/spectra/ssp/grid.py:22-124: fields age, metallicity,
wavelength and flux with flux laid out (n_met, n_age, n_wave).  The reference
loads named templates from HDF5/FITS files through
rubix/spectra/ssp/factory.py; those file formats are out of scope here, so
templates are registered by name (a grid or a zero-argument factory).  The
synthetic grids of the benchmark workloads are pre-registered.

``log_axes``: the age and metallicity axes hold log10 values and lookups
log10 the star values first, i.e. bilinear interpolation in (log Z, log age).
With ``log_axes=False`` the axes are used as given, like the reference.
"""
from __future__ import annotations

from typing import Callable, Dict, Union

import numpy as np


class SSPGrid:
    def __init__(self, age, metallicity, wavelength, flux, log_axes=False):
        self.age = np.asarray(age, dtype=np.float32)
        self.metallicity = np.asarray(metallicity, dtype=np.float32)
        self.wavelength = np.asarray(wavelength, dtype=np.float32)
        self.flux = np.asarray(flux, dtype=np.float32)
        self.log_axes = bool(log_axes)
        if self.flux.shape != (self.metallicity.size, self.age.size, self.wavelength.size):
            raise ValueError("flux must be (n_metallicity, n_age, n_wavelength)")
        self._prepared = {}

    def keys(self):
        return ["age", "metallicity", "wavelength", "flux"]

    def __iter__(self):
        yield from (getattr(self, k) for k in self.keys())

    def prepared(self, z: float, device="cuda"):
        """Device copy + fused-path tables for observed redshift z (cached)."""
        from .ops import PreparedSSP

        key = (float(z), str(device))
        if key not in self._prepared:
            self._prepared[key] = PreparedSSP(self.flux, self.metallicity, self.age,
                                              self.wavelength, z, self.log_axes, device)
        return self._prepared[key]

    def get_lookup_interpolation(self, method: str = "linear", extrap: int = 0, device="cuda"):
        """Bilinear lookup f(metallicity, age) -> (n, n_wave) on the GPU
        (interp2d(method='linear', extrap=0), rubix/spectra/ssp/grid.py:60-124).
        Only the linear method is implemented; extrap must be 0."""
        if method != "linear":
            raise NotImplementedError(
                f"SSP lookup method {method!r} is not implemented on the GPU path; "
                "set config['ssp']['method'] = 'linear' (bilinear interpolation)")
        if extrap != 0:
            raise NotImplementedError("only extrap=0 (zero outside the grid) is implemented")
        from .autograd import SSPLookup

        def lookup(metallicity, age):
            return SSPLookup.apply(metallicity, age, self.prepared(0.0, device))

        return lookup


_REGISTRY: Dict[str, Union[SSPGrid, Callable[[], SSPGrid]]] = {}


def register_template(name: str, grid: Union[SSPGrid, Callable[[], SSPGrid]]):
    """Make a template available to get_ssp(config) under `name`."""
    _REGISTRY[name] = grid


def get_ssp_template(name: str) -> SSPGrid:
    """rubix/spectra/ssp/factory.py:12-76 for registered templates."""
    if name not in _REGISTRY:
        raise ValueError(f"SSP template {name} not found in the supported configuration file.")
    g = _REGISTRY[name]
    if not isinstance(g, SSPGrid):
        g = g()
        _REGISTRY[name] = g
    return g


def _synthetic(n_met, n_age, n_wave, wave_range):
    def make():
        from .synthetic import make_ssp_grid

        d = make_ssp_grid(n_met, n_age, n_wave, wave_range, seed=1)
        return SSPGrid(d["ax_age"], d["ax_met"], d["wave"], d["flux"], log_axes=True)

    return make


register_template("synthetic_6x20x1000", _synthetic(6, 20, 1000, (3000.0, 9000.0)))
register_template("synthetic_12x100x5000", _synthetic(12, 100, 5000, (3000.0, 10000.0)))
# near-infrared (8000-20000 A, 4 A spacing) and ultraviolet-to-5.6 micron
# (2500-56000 A, ~13 A spacing) grids for the instruments beyond the optical
register_template("synthetic_6x20x3000_nir", _synthetic(6, 20, 3000, (8000.0, 20000.0)))
register_template("synthetic_6x20x4000_wide", _synthetic(6, 20, 4000, (2500.0, 56000.0)))
