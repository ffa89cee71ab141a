"""YAML configuration shipped with the package (constants, step chains,
instrument table), loaded with yaml.safe_load."""
from __future__ import annotations

from pathlib import Path

import yaml

HERE = Path(__file__).resolve().parent


def _load(name):
    with open(HERE / name) as f:
        return yaml.safe_load(f)


config = _load("rubix_config.yml")
pipeline_config = _load("pipeline_config.yml")
telescopes_config = _load("telescopes.yaml")
