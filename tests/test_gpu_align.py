"""GPU: aperture filter and galaxy alignment kernels against the oracle.

* filter mask and masked fields: bit-exact (comparisons and selects only);
* inertia tensor: fixed-order double reduction, <= 1e-10 relative to the
  oracle's float64 sums (same float32 radius test);
* rotation (p R) E in float32: <= 1e-6 of the coordinate scale against the
  float64 product; adjoints by the dot-product identity.
"""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops, synthetic as S
from rubix_amd.autograd import FilterFields, RotateRows
from rubix_amd.core.rotation import euler_rotation_matrix, principal_axes

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_filter_bit_exact_and_adjoint():
    st = S.make_stars(100_000, seed=11)
    edges = np.linspace(-6, 6, 31).astype(np.float32)
    coords = torch.as_tensor(st["coords"], device=DEV)
    f = {k: torch.as_tensor(st[k], device=DEV) for k in ("mass", "age", "metallicity")}
    m, a, z, mask = ops.filter_particles(coords, float(edges.min()), float(edges.max()),
                                         f["mass"], f["age"], f["metallicity"])
    ref, rmask = R.filter_particles(st["coords"], edges, {k: st[k] for k in f})
    np.testing.assert_array_equal(mask.cpu().numpy(), rmask)
    for got, k in zip((m, a, z), ("mass", "age", "metallicity")):
        np.testing.assert_array_equal(got.cpu().numpy(), ref[k])
    leaf = f["mass"].clone().requires_grad_(True)
    out, msk = FilterFields.apply(coords, -6.0, 6.0, leaf)
    g = torch.randn_like(out)
    out.backward(g)
    np.testing.assert_array_equal(leaf.grad.cpu().numpy(), np.where(rmask, g.cpu().numpy(), 0))


def test_inertia_tensor_matches_oracle():
    st = S.make_stars(300_000, seed=12)
    coords = torch.as_tensor(st["coords"], device=DEV)
    mass = torch.as_tensor(st["mass"], device=DEV)
    for r_half in (2.0, 5.0, 1e6):
        got = ops.inertia_tensor(coords, mass, r_half).cpu().numpy()
        ref = R.moment_of_inertia_tensor(st["coords"], st["mass"], r_half)
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), r_half
        np.testing.assert_array_equal(got, got.T)
    np.testing.assert_array_equal(
        ops.inertia_tensor(torch.eye(3, device=DEV), torch.ones(3, device=DEV), 2.0).cpu().numpy(),
        2.0 * np.eye(3))


def test_rotation_and_adjoint():
    st = S.make_stars(200_000, seed=13)
    Rm = principal_axes(R.moment_of_inertia_tensor(st["coords"], st["mass"], 5.0))
    E = euler_rotation_matrix(45.0, 30.0, 15.0)
    x = torch.as_tensor(st["velocity"], device=DEV)
    got = ops.rotate_rows(x, Rm, E).cpu().numpy()
    ref = st["velocity"].astype(np.float64) @ Rm.astype(np.float32).astype(np.float64) \
        @ E.astype(np.float32).astype(np.float64)
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()
    xl = x.clone().requires_grad_(True)
    y = RotateRows.apply(xl, Rm, E)
    g = torch.randn_like(y)
    y.backward(g)
    lhs = float((y.double() * g.double()).sum())
    rhs = float((xl.double() * xl.grad.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
