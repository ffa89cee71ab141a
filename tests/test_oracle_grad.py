"""Pins the gradient oracle (torch float64 autograd through the restated
reference op graph, oracle/ref_torch.py) with central finite differences of
the independent numpy float64 forward (oracle/ref_numpy.py), on a small
workload, at 64 random per-star coordinates plus both PSF/LSF sigmas.

Two step sizes: h = 1e-6 |theta| (tight, 1e-5 relative) and the h = 1e-3
|theta| of the differentiation workload.  At h = 1e-3 the chi^2 loss's
curvature gives central differences an O(h^2) truncation error of up to a
few 1e-3 (checked to converge quadratically), so that step is combined with
h/2 by Richardson extrapolation, (4 D(h/2) - D(h)) / 3, which cancels the
h^2 term.  Both sides use exact arithmetic for the shifted knots
(ref_knots=False): with the reference's float32 knots the forward is
piecewise constant at these step sizes; the float32-knot gradient oracle is
the same formula evaluated at the rounded knots.
"""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R, ref_torch as RT
from rubix_amd import synthetic as S


def small_workload(n=60, seed=0):
    stars = S.make_stars(n, seed=seed, h_r=1.0)
    grid = S.make_ssp_grid(4, 6, 300, (3000.0, 9000.0), seed=1)
    wt = (3400.0 + 30.0 * np.arange(150)).astype(np.float32)
    geom = S.make_geometry(8, 0.5, 0.05, wt, 0.55, 5, 40.0, 30.0)
    return S.Workload("small", stars, grid, geom)


def _loss_np(w, stars, psf_sigma, lsf_sigma, tgt, wt):
    geom = dict(w.geom)
    geom["psf_sigma"] = psf_sigma
    geom["lsf_sigma"] = lsf_sigma
    c = R.plain_forward(stars, w.grid, geom, dtype=np.float64, ref_knots=False)
    return float(np.sum(wt * (c - tgt) ** 2))


@pytest.fixture(scope="module")
def setup():
    w = small_workload()
    rng = np.random.default_rng(9)
    base = R.plain_forward(w.stars, w.grid, w.geom, dtype=np.float64, ref_knots=False)
    tgt = base * (1 + 0.05 * rng.normal(size=base.shape))
    wt = 1.0 / np.maximum(np.abs(base), 1e-3 * np.abs(base).max()) ** 2
    st = w.stars
    f64 = torch.float64
    T = {k: torch.tensor(st[k], dtype=f64, requires_grad=True) for k in ("mass", "age", "metallicity")}
    vl = torch.tensor(st["velocity"][:, 2], dtype=f64, requires_grad=True)
    sp = torch.tensor(w.geom["psf_sigma"], dtype=f64, requires_grad=True)
    sl = torch.tensor(w.geom["lsf_sigma"], dtype=f64, requires_grad=True)
    ids = torch.tensor(R.square_spaxel_assignment(st["coords"], w.geom["edges"]).astype(np.int64))
    raw, c1, c2 = RT.full_forward(T["mass"], T["age"], T["metallicity"], vl, ids, sp, sl, w.grid,
                                  w.geom, f64, ref_knots=False)
    loss = torch.sum(torch.tensor(wt) * (c2 - torch.tensor(tgt)) ** 2)
    loss.backward()
    grads = {"mass": T["mass"].grad.numpy(), "age": T["age"].grad.numpy(),
             "metallicity": T["metallicity"].grad.numpy(), "vlos": vl.grad.numpy(),
             "psf_sigma": float(sp.grad), "lsf_sigma": float(sl.grad)}
    return w, tgt, wt, grads, c2.detach().numpy(), base


def test_torch_oracle_forward_matches_numpy(setup):
    w, tgt, wt, grads, c2, base = setup
    np.testing.assert_allclose(c2, base, rtol=1e-10, atol=1e-14 * np.abs(base).max())


def _fd(w, tgt, wt, field, idx, h_rel):
    st = {k: v.astype(np.float64).copy() for k, v in w.stars.items()}
    ps, ls = w.geom["psf_sigma"], w.geom["lsf_sigma"]

    def L(delta):
        s2 = {k: v.copy() for k, v in st.items()}
        p2, l2 = ps, ls
        if field == "psf_sigma":
            p2 = ps + delta
        elif field == "lsf_sigma":
            l2 = ls + delta
        elif field == "vlos":
            s2["velocity"][idx, 2] += delta
        else:
            s2[field][idx] += delta
        return _loss_np(w, s2, p2, l2, tgt, wt)

    if field == "psf_sigma":
        x = ps
    elif field == "lsf_sigma":
        x = ls
    elif field == "vlos":
        x = st["velocity"][idx, 2]
    else:
        x = st[field][idx]
    h = h_rel * max(abs(x), 1e-12)
    lp, lm = L(h), L(-h)
    # round-off of a central difference: ~eps * |L| / h
    return (lp - lm) / (2 * h), 16 * np.finfo(np.float64).eps * max(abs(lp), abs(lm)) / h


def _coords(w, n=64, seed=11):
    rng = np.random.default_rng(seed)
    fields = ["mass", "age", "metallicity", "vlos"]
    return [(fields[rng.integers(0, 4)], int(rng.integers(0, w.n_stars))) for _ in range(n)]


def _fd_rich(w, tgt, wt, field, idx, h_rel):
    if h_rel < 1e-4:
        return _fd(w, tgt, wt, field, idx, h_rel)
    d1, n1 = _fd(w, tgt, wt, field, idx, h_rel)
    d2, n2 = _fd(w, tgt, wt, field, idx, h_rel / 2)
    return (4 * d2 - d1) / 3, (4 * n2 + n1) / 3


@pytest.mark.parametrize("h_rel,tol", [(1e-6, 1e-5), (1e-3, 1e-4)])
def test_gradient_oracle_matches_central_differences(setup, h_rel, tol):
    w, tgt, wt, grads, _, _ = setup
    scale = {f: np.abs(grads[f]).max() for f in ("mass", "age", "metallicity", "vlos")}
    for field, idx in _coords(w):
        fd, noise = _fd_rich(w, tgt, wt, field, idx, h_rel)
        an = grads[field][idx]
        # relative to the coordinate's own size (floor 1e-3 of the field's
        # largest gradient for coordinates whose gradient is ~0), plus the
        # central difference's own round-off
        bound = tol * max(abs(an), 1e-3 * scale[field]) + noise
        assert abs(fd - an) <= bound, (field, idx, fd, an, bound)
    for field in ("psf_sigma", "lsf_sigma"):
        fd, noise = _fd_rich(w, tgt, wt, field, 0, h_rel)
        assert abs(fd - grads[field]) <= tol * abs(grads[field]) + noise, (field, fd, grads[field])
