"""convolve_psf step (rubix/core/psf.py:15-75): Gaussian PSF, sigma in
pixels, size x size kernel, per lambda plane, mode "same"."""
from __future__ import annotations

from typing import Callable

import torch

from ..autograd import GaussianPSF
from ..logger import get_logger
from .data import RubixData


def get_convolve_psf(config: dict) -> Callable:
    logger = get_logger(config.get("logger", None))
    if "psf" not in config["telescope"]:
        raise ValueError("PSF configuration not found in telescope configuration")
    if "name" not in config["telescope"]["psf"]:
        raise ValueError("PSF name not found in telescope configuration")
    if config["telescope"]["psf"]["name"] == "gaussian":
        if "size" not in config["telescope"]["psf"]:
            raise ValueError("PSF size not found in telescope configuration")
        if "sigma" not in config["telescope"]["psf"]:
            raise ValueError("PSF sigma not found in telescope configuration")
        size = int(config["telescope"]["psf"]["size"])
        sigma = config["telescope"]["psf"]["sigma"]
    else:
        raise ValueError(f"Unknown PSF kernel name: {config['telescope']['psf']['name']}")

    def convolve_psf(rubixdata: RubixData) -> RubixData:
        """Convolve the input datacube with the PSF kernel."""
        logger.info("Convolving with PSF...")
        cube = rubixdata.stars.datacube
        s = sigma if torch.is_tensor(sigma) else torch.tensor(float(sigma), device=cube.device)
        rubixdata.stars.datacube = GaussianPSF.apply(cube, s, size)
        return rubixdata

    return convolve_psf
