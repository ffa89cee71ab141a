"""apply_noise step (rubix/core/noise.py:17-80): adds S/N-scaled noise to
the datacube, with the reference's config keys and error messages.

The reference draws from PRNGKey(0) on every call (rubix/telescope/noise/
noise.py:100); the optional key telescope.noise.seed (default 0) selects
another key.  The noise scale follows calculate_S2N, including its
behaviour when some spaxel has no positive flux: the median is NaN, then 0,
and the noise vanishes everywhere."""
from __future__ import annotations

from typing import Callable

from ..autograd import ApplyNoise
from ..logger import get_logger
from ..ops import NOISE_DISTRIBUTIONS
from .data import RubixData

SUPPORTED_NOISE_DISTRIBUTIONS = list(NOISE_DISTRIBUTIONS)


def get_apply_noise(config: dict) -> Callable:
    if "noise" not in config["telescope"]:
        raise ValueError("Noise information not provided in telescope config")
    if "signal_to_noise" not in config["telescope"]["noise"]:
        raise ValueError("Signal to noise information not provided in noise config")
    if "noise_distribution" not in config["telescope"]["noise"]:
        raise ValueError(
            "Noise distribution not provided in noise config. Currently supported "
            f"distributions are: {SUPPORTED_NOISE_DISTRIBUTIONS}")
    signal_to_noise = config["telescope"]["noise"]["signal_to_noise"]
    noise_distribution = config["telescope"]["noise"]["noise_distribution"]
    seed = int(config["telescope"]["noise"].get("seed", 0))
    logger = get_logger(config.get("logger", None))

    def apply_noise(rubixdata: RubixData) -> RubixData:
        logger.info(
            f"Applying noise to datacube with signal to noise ratio: {signal_to_noise} "
            f"and noise distribution: {noise_distribution}")
        cube = rubixdata.stars.datacube
        rubixdata.stars.datacube = ApplyNoise.apply(cube, signal_to_noise, noise_distribution,
                                                    seed)
        return rubixdata

    return apply_noise
