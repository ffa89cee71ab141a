"""SSP steps (rubix/core/ssp.py:11-98)."""
from __future__ import annotations

from typing import Callable

from ..logger import get_logger
from ..ssp import SSPGrid, get_ssp_template


def get_ssp(config: dict) -> SSPGrid:
    """rubix/core/ssp.py:11-34 (same validation messages)."""
    if "ssp" not in config:
        raise ValueError("Configuration does not contain 'ssp' field")
    if "template" not in config["ssp"]:
        raise ValueError("Configuration does not contain 'template' field")
    if "name" not in config["ssp"]["template"]:
        raise ValueError("Configuration does not contain 'name' field")
    return get_ssp_template(config["ssp"]["template"]["name"])


def lookup_method(config: dict) -> str:
    """rubix/core/ssp.py:56-62: the reference defaults to "cubic"; only
    "linear" (bilinear) runs here, so a missing or other method is an error
    that says how to select the implemented one."""
    method = config["ssp"].get("method", "cubic")
    if method != "linear":
        raise NotImplementedError(
            f"SSP interpolation method {method!r} is not implemented on the GPU path; "
            "set config['ssp']['method'] = 'linear'")
    return method


def get_lookup_interpolation(config: dict) -> Callable:
    """rubix/core/ssp.py:37-65: f(metallicity, age) -> (n, n_wave)."""
    get_logger(config.get("logger", None))
    ssp = get_ssp(config)
    return ssp.get_lookup_interpolation(method=lookup_method(config),
                                        device=config.get("device", "cuda"))


# The reference builds vmap / pmap variants (rubix/core/ssp.py:68-98); the
# GPU lookup is already batched over stars and one process drives one GPU.
get_lookup_interpolation_vmap = get_lookup_interpolation
get_lookup_interpolation_pmap = get_lookup_interpolation
