"""calculate_extinction step of the calc_dusty_ifu chain
(rubix/core/dust.py:16-65 over rubix/spectra/dust/dust_extinction.py:163-358):
per-star dust attenuation of the resampled spectra from the gas column of
the star's spaxel.

Config as the reference's: ssp.dust.{extinction_model, Rv,
dust_grain_density}; the dust-to-gas model and Xco come from the package
config (ssp.dust.dust_to_gas_model, ssp.dust.Xco), as the reference reads
them from its own rubix_config, unless the user's ssp.dust sets them."""
from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from .. import ops
from ..autograd import ApplyExtinction
from ..config import config as pkg_config
from ..logger import get_logger
from ..telescope import calculate_spatial_bin_edges
from .data import RubixData
from .telescope import get_cosmology, get_telescope, step_device

RV_MODELS = list(ops.RV_MODELS)


def spaxel_area(config: dict) -> float:
    """Area of one spaxel in kpc^2 (rubix/core/dust.py:45-55): the spatial bin
    size of calculate_spatial_bin_edges squared, or the spacing of explicit
    config["telescope"]["spatial_bin_edges"] squared."""
    if "spatial_bin_edges" in config["telescope"]:
        e = np.asarray(config["telescope"]["spatial_bin_edges"], dtype=np.float32)
        size = np.float32((e[-1] - e[0]) / np.float32(len(e) - 1))
    else:
        tel = get_telescope(config)
        _, size = calculate_spatial_bin_edges(tel.fov, tel.sbin, config["galaxy"]["dist_z"],
                                              get_cosmology(config))
    return float(np.float32(size) * np.float32(size))


def extinction_inputs(config: dict) -> Callable:
    """Validated dust config -> f(rubixdata) = (A_V per star in input order,
    A(lambda)/A(V) per telescope wavelength), both device tensors."""
    if "dust" not in config["ssp"]:
        raise ValueError("Dust configuration not found in config file.")
    if "extinction_model" not in config["ssp"]["dust"]:
        raise ValueError("Extinction model not found in dust configuration.")
    dust = config["ssp"]["dust"]
    pkg_dust = pkg_config["ssp"]["dust"]
    tel = get_telescope(config)
    n_spaxels = int(tel.sbin) ** 2
    dev = step_device(config)
    wave = torch.as_tensor(np.asarray(tel.wave_seq, dtype=np.float32), device=dev)
    area = spaxel_area(config)

    def inputs(rubixdata: RubixData):
        model = dust["extinction_model"]
        if model not in RV_MODELS:
            raise ValueError(
                f"Extinction model '{model}' is not available. Choose from {RV_MODELS}.")
        gas, stars = rubixdata.gas, rubixdata.stars
        if gas is None or gas.coords is None or gas.pixel_assignment is None:
            raise ValueError("calculate_extinction needs gas particles with spaxel assignments "
                             "(data.args.particle_type must include 'gas')")
        axav = ops.extinction_curve(model, float(dust["Rv"]), wave)
        ext = ops.gas_extinction(gas.mass.reshape(-1).to(dev, torch.float32),
                                 gas.metals.reshape(gas.mass.numel(), -1).to(dev, torch.float32),
                                 dust.get("dust_to_gas_model", pkg_dust["dust_to_gas_model"]),
                                 dust.get("Xco", pkg_dust["Xco"]),
                                 float(dust["dust_grain_density"]), area)
        gc = gas.coords.detach().reshape(-1, 3).to(dev, torch.float32)
        sc = stars.coords.detach().reshape(-1, 3).to(dev, torch.float32)
        edges = stars.spatial_bin_edges
        n_keys = n_spaxels if edges is None else max(n_spaxels, (int(edges.numel()) - 1) ** 2)
        av = ops.column_av(gc, gas.pixel_assignment.reshape(-1), ext, sc,
                           stars.pixel_assignment.reshape(-1), n_spaxels, n_keys)
        return av, axav

    return inputs


def get_extinction(config: dict) -> Callable:
    logger = get_logger(config.get("logger", None))
    inputs = extinction_inputs(config)

    def calculate_extinction(rubixdata: RubixData) -> RubixData:
        logger.info("Applying dust extinction to the spaxel data...")
        av, axav = inputs(rubixdata)
        s = rubixdata.stars.spectra
        out = ApplyExtinction.apply(s.reshape(-1, s.shape[-1]), av, axav)
        rubixdata.stars.spectra = out.reshape(s.shape)
        return rubixdata

    return calculate_extinction


def get_calculate_dusty_datacube_fused(config: dict) -> Callable:
    """calculate_spectra -> scale_spectrum_by_mass -> doppler_shift_and_resampling
    -> calculate_extinction -> calculate_datacube in one fused pass: the
    datacube step of the fused calc_dusty_ifu chain.  Writes stars.datacube
    (nb, nb, n_target); per-star spectra are never materialised."""
    from .ifu import get_calculate_datacube_fused

    logger = get_logger(config.get("logger", None))
    inputs = extinction_inputs(config)
    fused = get_calculate_datacube_fused(config, with_extinction=True)

    def calculate_dusty_datacube_fused(rubixdata: RubixData) -> RubixData:
        logger.info("Calculating Data Cube with dust extinction (fused)...")
        return fused(rubixdata, inputs(rubixdata))

    return calculate_dusty_datacube_fused
