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


# ------------------------------------------------ the differentiation workload
# The same check at the workload's own size: 10 000 stars of the small
# configuration (25 x 25 x 1000 cube, 6 x 20 x 1000 grid), chi^2 against a
# target rendered from seed 2 with Gaussian noise at S/N = 30, FD at 64
# random per-star coordinates plus both sigmas, h = 1e-3 |theta| with
# Richardson extrapolation.  A per-star coordinate changes only that star's
# spectrum in its own spaxel, so L(theta + h) - L(theta - h) is evaluated
# exactly from the change of that spectrum: the PSF of a one-spaxel cube is
# the kernel placed at that spaxel (the oracle's apply_psf of an impulse
# plane) and the LSF acts on each spectrum alone, so the cube changes by
# psf_impulse (x) lsf(delta spectrum).  The sigmas re-run PSF + LSF on the
# full cube.  The loss is piecewise smooth: its derivative jumps where a
# shifted knot crosses a target bin (v_los; about one crossing per star
# within +-1e-3 |v|, knots and bins being ~5 A apart) or where a star crosses
# a template cell (age, Z).  A central difference over a jump measures
# neither side, so h starts at 1e-3 |theta| and is halved until the
# evaluations at +-h see the same interpolation segments and cells as theta.

def _cfg1_full():
    w = S.config1(n_stars=10_000)
    clean2 = R.plain_forward(S.make_stars(w.n_stars, seed=2), w.grid, w.geom, dtype=np.float64,
                             ref_knots=False)
    rng = np.random.default_rng(2)
    sd = np.abs(clean2) / 30.0
    tgt = clean2 + rng.normal(0.0, 1.0, clean2.shape) * sd
    wt = 1.0 / np.maximum(sd, 1e-6 * np.abs(clean2).max()) ** 2
    return w, tgt, wt


def _star_spectrum(w, mass, age, met, vz, segments=False):
    """One star's resampled spectrum: the oracle's lookup -> mass scale ->
    Doppler -> resample, float64 with exact knots (plain_forward's chunk
    body for one star).  segments=True also returns the template cell and
    the interpolation segment of every target bin."""
    g, geo = w.grid, w.geom
    f64 = np.float64
    spec = R.lookup_linear(np.array([met]), np.array([age]), g["ax_met"], g["ax_age"], g["flux"],
                           log_axes=g.get("log_axes", False), dtype=f64)
    spec = R.scale_spectrum_by_mass(spec, np.array([mass], dtype=f64))
    ssp_wave = R.cosmological_doppler_shift(geo["z"], g["wave"], dtype=f64)
    vel = np.zeros((1, 3))
    vel[0, 2] = vz
    wv = R.velocity_doppler_shift(ssp_wave, vel, "z", dtype=f64)
    tw = np.asarray(geo["wave_target"], f64)
    out = R.resample_spectra(spec, wv, tw)[0]
    if not segments:
        return out
    lz, la = (np.log10(met), np.log10(age)) if g.get("log_axes", False) else (met, age)
    cell = (int(np.searchsorted(np.asarray(g["ax_met"], f64), lz, side="right")),
            int(np.searchsorted(np.asarray(g["ax_age"], f64), la, side="right")))
    seg = np.searchsorted(wv[0], tw, side="right")
    inr = (wv[0] >= tw.min()) & (wv[0] <= tw.max())
    return out, (cell, seg.tobytes(), inr.tobytes())


@pytest.fixture(scope="module")
def cfg1_full():
    w, tgt, wt = _cfg1_full()
    st = w.stars
    f64 = torch.float64
    T = {k: torch.tensor(st[k], dtype=f64, requires_grad=True) for k in ("mass", "age", "metallicity")}
    vl = torch.tensor(st["velocity"][:, 2], dtype=f64, requires_grad=True)
    sp = torch.tensor(w.geom["psf_sigma"], dtype=f64, requires_grad=True)
    sl = torch.tensor(w.geom["lsf_sigma"], dtype=f64, requires_grad=True)
    ids = R.square_spaxel_assignment(st["coords"], w.geom["edges"]).astype(np.int64)
    _, _, c2 = RT.full_forward(T["mass"], T["age"], T["metallicity"], vl, torch.tensor(ids), sp, sl,
                               w.grid, w.geom, f64, ref_knots=False)
    loss = torch.sum(torch.tensor(wt) * (c2 - torch.tensor(tgt)) ** 2)
    loss.backward()
    grads = {"mass": T["mass"].grad.numpy(), "age": T["age"].grad.numpy(),
             "metallicity": T["metallicity"].grad.numpy(), "vlos": vl.grad.numpy(),
             "psf_sigma": float(sp.grad), "lsf_sigma": float(sl.grad)}
    out = R.plain_forward(st, w.grid, w.geom, dtype=np.float64, ref_knots=False,
                          return_intermediates=True)
    return w, tgt, wt, ids, grads, out


def test_star_spectrum_helper_matches_forward():
    """Three stars in one spaxel: the helper's spectra add up to the oracle's
    raw cube row."""
    w = S.config1(n_stars=10_000)
    st = {k: v[:3].copy() for k, v in w.stars.items()}
    st["coords"][:] = st["coords"][0]
    one = S.Workload("three", st, w.grid, w.geom)
    raw = R.plain_forward(st, w.grid, w.geom, dtype=np.float64, ref_knots=False,
                          return_intermediates=True)["cube"]
    p = int(R.square_spaxel_assignment(st["coords"], w.geom["edges"])[0])
    nb = w.geom["nb"]
    tot = sum(_star_spectrum(one, st["mass"][i], st["age"][i], st["metallicity"][i],
                             st["velocity"][i, 2]) for i in range(3))
    np.testing.assert_allclose(raw[p // nb, p % nb], tot, rtol=1e-13, atol=1e-15 * np.abs(tot).max())


def _fd_star(w, tgt, wt, ids, out, field, i, h_rel):
    st = w.stars
    nb = w.geom["nb"]
    p = int(ids[i])
    if p < 0 or p >= nb * nb:
        return 0.0, 0.0
    th = {"mass": float(st["mass"][i]), "age": float(st["age"][i]),
          "metallicity": float(st["metallicity"][i]), "vlos": float(st["velocity"][i, 2])}

    def spec(delta, segments=False):
        t = dict(th)
        t[field] += delta
        return _star_spectrum(w, t["mass"], t["age"], t["metallicity"], t["vlos"], segments)

    K = R.gaussian_kernel_2d(w.geom["psf_size"], w.geom["psf_size"], w.geom["psf_sigma"],
                             dtype=np.float64)
    k = R.lsf_kernel(w.geom["lsf_sigma"], w.geom["wave_res"], dtype=np.float64)
    imp = np.zeros((nb, nb, 1))
    imp[p // nb, p % nb, 0] = 1.0
    foot = R.apply_psf(imp, K)[:, :, 0]
    ys, xs = np.nonzero(foot)
    sl = (slice(ys.min(), ys.max() + 1), slice(xs.min(), xs.max() + 1))
    r = (out["cube_lsf"] - tgt)[sl]
    wts = wt[sl]
    fp = foot[sl][:, :, None]
    s0, seg0 = spec(0.0, True)
    h = h_rel * max(abs(th[field]), 1e-12)
    for _ in range(40):  # the same segments and cell on both sides
        if spec(h, True)[1] == seg0 and spec(-h, True)[1] == seg0:
            break
        h /= 2

    def central(hh):
        dp = R.apply_lsf_kernel((spec(hh) - s0)[None, None, :], k)[0, 0]
        dm = R.apply_lsf_kernel((spec(-hh) - s0)[None, None, :], k)[0, 0]
        Dp, Dm = fp * dp[None, None, :], fp * dm[None, None, :]
        return float(np.sum(wts * (Dp - Dm) * (2 * r + Dp + Dm))) / (2 * hh)

    d1, d2 = central(h), central(h / 2)
    # round-off of the difference form: eps * sum|wt r D| / h
    scale = float(np.sum(np.abs(wts * r))) * np.abs(s0).max() * np.abs(fp).max()
    return (4 * d2 - d1) / 3, 64 * np.finfo(np.float64).eps * scale / h


def _fd_sigma(w, tgt, wt, out, which, h_rel):
    geo = w.geom
    raw = out["cube"]

    def L(ps, ls):
        K = R.gaussian_kernel_2d(geo["psf_size"], geo["psf_size"], ps, dtype=np.float64)
        c = R.apply_lsf_kernel(R.apply_psf(raw, K), R.lsf_kernel(ls, geo["wave_res"], dtype=np.float64))
        return float(np.sum(wt * (c - tgt) ** 2))

    ps, ls = geo["psf_sigma"], geo["lsf_sigma"]
    x = ps if which == "psf_sigma" else ls

    def central(hh):
        if which == "psf_sigma":
            return (L(ps + hh, ls) - L(ps - hh, ls)) / (2 * hh), L(ps, ls)
        return (L(ps, ls + hh) - L(ps, ls - hh)) / (2 * hh), L(ps, ls)

    h = h_rel * x
    (d1, l0), (d2, _) = central(h), central(h / 2)
    return (4 * d2 - d1) / 3, 64 * np.finfo(np.float64).eps * abs(l0) / h


def test_gradient_oracle_fd_on_differentiation_workload(cfg1_full):
    w, tgt, wt, ids, grads, out = cfg1_full
    np.testing.assert_allclose(out["cube_lsf"].sum(), out["cube_lsf"].sum())  # finite
    scale = {f: np.abs(grads[f]).max() for f in ("mass", "age", "metallicity", "vlos")}
    rng = np.random.default_rng(11)
    fields = ["mass", "age", "metallicity", "vlos"]
    worst = 0.0
    for _ in range(64):
        field, i = fields[rng.integers(0, 4)], int(rng.integers(0, w.n_stars))
        fd, noise = _fd_star(w, tgt, wt, ids, out, field, i, 1e-3)
        an = grads[field][i]
        bound = 1e-4 * max(abs(an), 1e-3 * scale[field]) + noise
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * scale[field]))
        assert abs(fd - an) <= bound, (field, i, fd, an, bound)
    for which in ("psf_sigma", "lsf_sigma"):
        fd, noise = _fd_sigma(w, tgt, wt, out, which, 1e-3)
        assert abs(fd - grads[which]) <= 1e-4 * abs(grads[which]) + noise, (which, fd, grads[which])
    print(f"cfg1 (10k stars) gradient oracle vs FD: worst relative {worst:.2e}")
