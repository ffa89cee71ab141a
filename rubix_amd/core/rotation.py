"""rotate_galaxy step (rubix/core/rotation.py:10-117, rubix/galaxy/alignment.py:66-265).

The factory validates the rotation config with the reference's messages and
fixes the Euler angles ("face-on" = (0, 0, 0), "edge-on" = (90, 0, 0), or
explicit alpha / beta / gamma in degrees).  The step computes the moment of
inertia tensor of the stars within the galaxy's half-mass radius on the GPU
(fixed-order double reduction), takes its principal axes on the host (a 3x3
symmetric eigenproblem, eigenvectors ordered by ascending eigenvalue as the
reference's argsort does) and rotates coordinates and velocities of every
configured particle type with one kernel each: p' = (p R) E.

Eigenvector signs are whatever the LAPACK symmetric eigensolver returns (the
reference's jnp.linalg.eigh wraps the same routine family on CPU); the sign
choice is not pinned by the reference's tests.  The alignment frame is held
fixed in the adjoint: velocity gradients rotate back through (R E)^T, and no
gradient flows into the masses through R.
"""
from __future__ import annotations

import numpy as np

from ..autograd import RotateRows
from ..logger import get_logger
from .data import RubixData
from .. import ops


def euler_rotation_matrix(alpha: float, beta: float, gamma: float) -> np.ndarray:
    """R = Rz(gamma) Ry(beta) Rx(alpha), degrees (rubix/galaxy/alignment.py:140-205)."""
    a, b, c = np.deg2rad([float(alpha), float(beta), float(gamma)])
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def principal_axes(I) -> np.ndarray:
    """rotation_matrix_from_inertia_tensor (rubix/galaxy/alignment.py:123-137)."""
    vals, vecs = np.linalg.eigh(np.asarray(I, dtype=np.float64))
    return vecs[:, np.argsort(vals)]


def rotate_galaxy(positions, velocities, masses, halfmass_radius, alpha, beta, gamma):
    """rubix/galaxy/alignment.py:240-265 on device tensors."""
    I = ops.inertia_tensor(positions.detach(), masses.detach(), float(halfmass_radius))
    R = principal_axes(I.cpu().numpy())
    E = euler_rotation_matrix(alpha, beta, gamma)
    return RotateRows.apply(positions, R, E), RotateRows.apply(velocities, R, E)


def _scalar(x):
    x = np.asarray(x.detach().cpu() if hasattr(x, "detach") else x, dtype=np.float64).reshape(-1)
    return float(x[0])


def get_galaxy_rotation(config: dict):
    """rubix/core/rotation.py:10-117."""
    if "rotation" not in config["galaxy"]:
        raise ValueError("Rotation information not provided in galaxy config")
    logger = get_logger(config.get("logger", None))
    rot = config["galaxy"]["rotation"]
    if "type" in rot:
        if rot["type"] not in ["face-on", "edge-on"]:
            raise ValueError("Invalid type provided in rotation information")
        alpha, beta, gamma = (0.0, 0.0, 0.0) if rot["type"] == "face-on" else (90.0, 0.0, 0.0)
    else:
        for key in ["alpha", "beta", "gamma"]:
            if key not in rot:
                raise ValueError(f"{key} not provided in rotation information")
        alpha, beta, gamma = rot["alpha"], rot["beta"], rot["gamma"]
    types = config.get("data", {}).get("args", {}).get("particle_type", ["stars"])

    def rotate_galaxy_step(rubixdata: RubixData) -> RubixData:
        logger.info(f"Rotating galaxy with alpha={alpha}, beta={beta}, gamma={gamma}")
        r_half = _scalar(rubixdata.galaxy.halfmassrad_stars)
        for ptype in ["stars", "gas"]:
            if ptype not in types:
                continue
            comp = getattr(rubixdata, ptype)
            assert comp.coords is not None, f"Coordinates not found for {ptype}. "
            assert comp.velocity is not None, f"Velocities not found for {ptype}. "
            assert comp.mass is not None, f"Masses not found for {ptype}. "
            comp.coords, comp.velocity = rotate_galaxy(comp.coords, comp.velocity, comp.mass,
                                                       r_half, alpha, beta, gamma)
        return rubixdata

    rotate_galaxy_step.__name__ = "rotate_galaxy"
    return rotate_galaxy_step
