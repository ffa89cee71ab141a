"""Telescope and spaxel-assignment steps (rubix/core/telescope.py:20-126)."""
from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from .. import ops
from ..autograd import FilterFields
from ..logger import get_logger
from ..telescope import PLANCK15, Telescope, TelescopeFactory, calculate_spatial_bin_edges
from .data import RubixData


def step_device(config: dict) -> torch.device:
    return torch.device(config.get("device", "cuda"))


def get_telescope(config: dict) -> Telescope:
    """rubix/core/telescope.py:20-47.  config["telescopes"] may add or
    override instrument definitions; config["telescope"]["wave_seq"] may give
    the target wavelengths explicitly."""
    table = None
    if "telescopes" in config:
        from ..config import telescopes_config
        table = dict(telescopes_config)
        table.update(config["telescopes"])
    tel = TelescopeFactory(table).create_telescope(config["telescope"]["name"])
    if "wave_seq" in config["telescope"]:
        tel.wave_seq = np.asarray(config["telescope"]["wave_seq"], dtype=np.float32)
    return tel


def get_cosmology(config: dict):
    name = config.get("cosmology", {}).get("name", "PLANCK15")
    if name != "PLANCK15":
        raise ValueError(f"Cosmology {name} not supported")
    return PLANCK15


def get_spatial_bin_edges(config: dict) -> np.ndarray:
    """rubix/core/telescope.py:50-77 (explicit edges in
    config["telescope"]["spatial_bin_edges"] take precedence)."""
    if "spatial_bin_edges" in config["telescope"]:
        return np.asarray(config["telescope"]["spatial_bin_edges"], dtype=np.float32)
    tel = get_telescope(config)
    edges, _ = calculate_spatial_bin_edges(tel.fov, tel.sbin, config["galaxy"]["dist_z"],
                                           get_cosmology(config))
    return edges


def get_spaxel_assignment(config: dict) -> Callable:
    """rubix/core/telescope.py:80-126: square spaxel ids (int32) on the GPU."""
    logger = get_logger(config.get("logger", None))
    tel = get_telescope(config)
    if tel.pixel_type not in ["square"]:
        raise ValueError(f"Pixel type {tel.pixel_type} not supported")
    dev = step_device(config)
    edges = torch.as_tensor(get_spatial_bin_edges(config), device=dev)

    def spaxel_assignment(rubixdata: RubixData) -> RubixData:
        logger.info("Assigning particles to spaxels...")
        for part in (rubixdata.stars, rubixdata.gas):
            if part.coords is not None:
                coords = part.coords.detach().to(dev, torch.float32)
                part.pixel_assignment = ops.spaxel_assign(coords, edges)
                part.spatial_bin_edges = edges
        return rubixdata

    return spaxel_assignment


def get_filter_particles(config: dict) -> Callable:
    """rubix/core/telescope.py:129-201: particles outside the aperture
    (x or y beyond the spatial bin edges' min / max, inclusive) get every
    per-particle field but coords and velocity set to 0 (for gas also not
    metals), and `mask` records which are inside.  One kernel masks the
    fields; gradients w.r.t. the fields are masked the same way."""
    logger = get_logger(config.get("logger", None))
    edges = get_spatial_bin_edges(config)
    lo, hi = float(np.min(edges)), float(np.max(edges))
    types = config.get("data", {}).get("args", {}).get("particle_type", ["stars"])

    def filter_particles(rubixdata: RubixData) -> RubixData:
        logger.info("Filtering particles outside the aperture...")
        for ptype, skip in (("stars", ("coords", "velocity")), ("gas", ("coords", "velocity", "metals"))):
            if ptype not in types:
                continue
            comp = getattr(rubixdata, ptype)
            if comp is None or comp.coords is None:
                continue
            names = [k for k, v in vars(comp).items()
                     if k not in skip and k not in ("mask", "spatial_bin_edges", "datacube")
                     and isinstance(v, torch.Tensor)
                     and v.dim() >= 1 and v.shape[0] == comp.coords.shape[0]]
            coords = comp.coords.detach()
            mask = None
            for i in range(0, len(names), 3):
                group = names[i:i + 3]
                vals = [getattr(comp, k) for k in group]
                flat = [v if v.dim() == 1 else v.reshape(v.shape[0], -1) for v in vals]
                if all(v.dim() == 1 and v.dtype == torch.float32 for v in flat):
                    *outs, mask = FilterFields.apply(coords, lo, hi, *flat)
                else:  # int or multi-column fields: mask computed once, applied per field
                    if mask is None:
                        (mask,) = FilterFields.apply(coords, lo, hi)
                    outs = [torch.where(mask.reshape(-1, *([1] * (v.dim() - 1))), v,
                                        torch.zeros((), dtype=v.dtype, device=v.device))
                            for v in vals]
                for k, o in zip(group, outs):
                    setattr(comp, k, o)
            if mask is None:
                (mask,) = FilterFields.apply(coords, lo, hi)
            comp.mask = mask
        return rubixdata

    return filter_particles
