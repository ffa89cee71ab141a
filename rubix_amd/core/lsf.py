"""convolve_lsf step (rubix/core/lsf.py:14-67): 25-tap Gaussian along lambda,
sigma in Angstrom, tap spacing = the telescope's wave_res."""
from __future__ import annotations

from typing import Callable

import torch

from ..autograd import GaussianLSF
from ..logger import get_logger
from .data import RubixData
from .telescope import get_telescope


def get_convolve_lsf(config: dict) -> Callable:
    logger = get_logger(config.get("logger", None))
    if "lsf" not in config["telescope"]:
        raise ValueError("LSF configuration not found in telescope configuration")
    if "sigma" not in config["telescope"]["lsf"]:
        raise ValueError("LSF sigma size not found in telescope configuration")
    sigma = config["telescope"]["lsf"]["sigma"]
    wave_resolution = get_telescope(config).wave_res

    def convolve_lsf(rubixdata: RubixData) -> RubixData:
        """Convolve the input datacube with the LSF."""
        logger.info("Convolving with LSF...")
        cube = rubixdata.stars.datacube
        s = sigma if torch.is_tensor(sigma) else torch.tensor(float(sigma), device=cube.device)
        rubixdata.stars.datacube = GaussianLSF.apply(cube, s, wave_resolution)
        return rubixdata

    return convolve_lsf
