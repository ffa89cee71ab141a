"""Logging under the reference's logger name "rubix" (rubix/logger.py:10-61)."""
from __future__ import annotations

import logging

_CONFIGURED = False


def get_logger(config=None) -> logging.Logger:
    global _CONFIGURED
    logger = logging.getLogger("rubix")
    if not _CONFIGURED:
        from .config import config as pkg_config

        c = dict(pkg_config.get("logger", {}))
        if config:
            c.update(config)
        logger.setLevel(getattr(logging, str(c.get("log_level", "INFO")).upper(), logging.INFO))
        if not logger.handlers:
            h = logging.StreamHandler()
            h.setFormatter(logging.Formatter(c.get("format", "%(levelname)s %(message)s")))
            logger.addHandler(h)
            if c.get("log_file_path"):
                fh = logging.FileHandler(c["log_file_path"])
                fh.setFormatter(logging.Formatter(c.get("format", "%(levelname)s %(message)s")))
                logger.addHandler(fh)
        _CONFIGURED = True
    return logger
