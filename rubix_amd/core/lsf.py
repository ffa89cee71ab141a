"""convolve_lsf step (rubix/core/lsf.py:14-67): 25-tap Gaussian along lambda,
sigma in Angstrom, tap spacing = the telescope's wave_res."""
from __future__ import annotations

from typing import Callable

import torch

from ..autograd import GaussianLSF, lsf_x
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
    n_taps = len(lsf_x(wave_resolution))
    if n_taps % 2 == 0:
        # rubix/telescope/lsf/lsf.py:22-26,61-65,102-105: an even tap count
        # crops n + 1 bins, which the reference cannot reshape to the cube
        logger.warning(f"LSF kernel grid for wave_res {wave_resolution} has {n_taps} taps; the "
                       "reference's convolve_lsf fails to reshape its n + 1 output bins here, "
                       "this step returns the first n (the kernel's x = 0 tap on each bin)")

    def convolve_lsf(rubixdata: RubixData) -> RubixData:
        """Convolve the input datacube with the LSF."""
        logger.info("Convolving with LSF...")
        cube = rubixdata.stars.datacube
        s = sigma if torch.is_tensor(sigma) else torch.tensor(float(sigma), device=cube.device)
        rubixdata.stars.datacube = GaussianLSF.apply(cube, s, wave_resolution)
        return rubixdata

    return convolve_lsf
