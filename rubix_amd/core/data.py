"""RubixData containers and the reshape_data step (rubix/core/data.py).

The containers hold torch tensors (device tensors for the GPU steps) under
the reference's field names (rubix/core/data.py:65-355).  reshape_data adds
the leading device axis exactly like rubix/core/data.py:446-487,623-671:
every array attribute is zero-padded to a multiple of the device count and
reshaped to (n_dev, N / n_dev, ...).  Each process drives one GPU, so the
device count is 1 unless a caller overrides `device_count` (the reference
tests monkeypatch jax.device_count the same way).
"""
from __future__ import annotations

import logging
from dataclasses import dataclass, fields
from typing import Callable, Optional

import numpy as np
import torch

from ..logger import get_logger


def device_count() -> int:
    """Devices the leading reshape axis spans (one per process)."""
    return 1


@dataclass
class Galaxy:
    redshift: Optional[object] = None
    center: Optional[object] = None
    halfmassrad_stars: Optional[object] = None


@dataclass
class StarsData:
    coords: Optional[torch.Tensor] = None
    velocity: Optional[torch.Tensor] = None
    mass: Optional[torch.Tensor] = None
    metallicity: Optional[torch.Tensor] = None
    age: Optional[torch.Tensor] = None
    pixel_assignment: Optional[torch.Tensor] = None
    spatial_bin_edges: Optional[torch.Tensor] = None
    mask: Optional[torch.Tensor] = None
    spectra: Optional[torch.Tensor] = None
    datacube: Optional[torch.Tensor] = None

    def __repr__(self):
        parts = ["StarsData:"]
        for f in fields(self):
            v = getattr(self, f.name)
            parts.append(f"{f.name}: None" if v is None else
                         f"{f.name}: shape = {tuple(v.shape)}, dtype = {v.dtype}")
        return "\n\t".join(parts)


@dataclass
class GasData:
    coords: Optional[torch.Tensor] = None
    velocity: Optional[torch.Tensor] = None
    mass: Optional[torch.Tensor] = None
    density: Optional[torch.Tensor] = None
    internal_energy: Optional[torch.Tensor] = None
    metallicity: Optional[torch.Tensor] = None
    metals: Optional[torch.Tensor] = None
    sfr: Optional[torch.Tensor] = None
    electron_abundance: Optional[torch.Tensor] = None
    pixel_assignment: Optional[torch.Tensor] = None
    spatial_bin_edges: Optional[torch.Tensor] = None
    mask: Optional[torch.Tensor] = None
    spectra: Optional[torch.Tensor] = None
    datacube: Optional[torch.Tensor] = None


@dataclass
class RubixData:
    galaxy: Optional[Galaxy] = None
    stars: Optional[StarsData] = None
    gas: Optional[GasData] = None

    def __post_init__(self):
        if self.galaxy is None:
            self.galaxy = Galaxy()
        if self.stars is None:
            self.stars = StarsData()
        if self.gas is None:
            self.gas = GasData()


def reshape_array(arr, n_dev: Optional[int] = None):
    """rubix/core/data.py:446-487 for a tensor or array: pad the leading axis
    with zeros to a multiple of n_dev and reshape to (n_dev, ceil(N/n_dev), ...)."""
    if n_dev is None:
        n_dev = device_count()
    is_np = isinstance(arr, np.ndarray)
    t = torch.as_tensor(arr)
    n = t.shape[0]
    per = (n + n_dev - 1) // n_dev
    total = per * n_dev
    if total > n:
        pad = torch.zeros((total - n,) + tuple(t.shape[1:]), dtype=t.dtype, device=t.device)
        t = torch.cat([t, pad], 0)
    out = t.reshape(n_dev, per, *t.shape[1:])
    return out.numpy() if is_np else out


def process_attributes(obj, logger: logging.Logger) -> None:
    """rubix/core/data.py:623-635: reshape every array attribute."""
    for f in fields(obj):
        v = getattr(obj, f.name)
        if v is None or not isinstance(v, (torch.Tensor, np.ndarray)):
            logger.debug(f"Attribute value of {f.name} is None or not an array")
            continue
        setattr(obj, f.name, reshape_array(v))


def get_reshape_data(config: dict) -> Callable:
    """rubix/core/data.py:638-671."""
    logger = get_logger(config.get("logger", None))

    def reshape_data(rubixdata: RubixData) -> RubixData:
        if rubixdata.stars.coords is not None:
            process_attributes(rubixdata.stars, logger)
        if rubixdata.gas.coords is not None:
            process_attributes(rubixdata.gas, logger)
        return rubixdata

    return reshape_data


def stars_from_arrays(coords, velocity, mass, age, metallicity, device="cuda",
                      requires_grad=False, halfmassrad_stars=None, gas=None) -> RubixData:
    """Convenience: RubixData with the star fields as float32 device tensors
    (and the galaxy's stellar half-mass radius, which rotate_galaxy reads).
    gas: optional dict of gas arrays (coords, velocity, mass, metals, ...)."""
    def t(x):
        v = torch.as_tensor(np.asarray(x, dtype=np.float32), device=device)
        return v.requires_grad_(requires_grad)

    def plain(x):
        return torch.as_tensor(np.asarray(x, np.float32), device=device)

    gas_data = None
    if gas is not None:
        gas_data = GasData(**{k: plain(v) for k, v in gas.items()})
    return RubixData(galaxy=Galaxy(halfmassrad_stars=halfmassrad_stars),
                     stars=StarsData(coords=plain(coords), velocity=t(velocity), mass=t(mass),
                                     age=t(age), metallicity=t(metallicity)),
                     gas=gas_data)
