"""Shared helpers for the test-suite (inputs, tolerances, oracle calls)."""
from __future__ import annotations

import functools

import numpy as np

from oracle import ref_numpy as R
from rubix_amd import synthetic as S


def rel_err(a, b):
    """Norm-wise relative error ||a - b||_2 / ||b||_2 (float64)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def max_rel(a, b):
    """max |a - b| / max |b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b).max()
    return float(np.abs(a - b).max() / (m if m > 0 else 1.0))


def with_kernels(w):
    """Workloads carry PSF/LSF parameters; the oracle builds the kernels."""
    return w


@functools.lru_cache(maxsize=None)
def cfg1(n=10_000):
    return with_kernels(S.config1(n_stars=n))


@functools.lru_cache(maxsize=None)
def cfg2_sample(n=2_000, seed=0):
    """cfg2 grid / cube / wavelengths with a small star sample of the same
    distributions (the oracle is O(N n_lambda) in python loops)."""
    w = S.config2(n_stars=n, seed=seed)
    return with_kernels(w)


@functools.lru_cache(maxsize=None)
def oracle(name, dtype_name, n, ref_knots=True):
    w = cfg1(n) if name == "cfg1" else cfg2_sample(n)
    return R.plain_forward(w.stars, w.grid, w.geom, dtype=np.dtype(dtype_name).type,
                           return_intermediates=True, ref_knots=ref_knots)
