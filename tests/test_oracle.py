"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against the reference's own known answers
(tests/golden/reference_known_answers.json), against library routines
(numpy / scipy), brute force on tiny inputs, or closed forms chosen so that a
dropped term, a wrong sign/index or a transposed operand fails.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.signal

from oracle import ref_numpy as R

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


# ---------------------------------------------------------------- spaxels

def test_spaxel_known_answer():
    g = GOLD["spaxel_assignment"]
    got = R.square_spaxel_assignment(np.array(g["coords"], np.float32), np.array(g["edges"], np.float32))
    np.testing.assert_array_equal(got, g["expected"])


def _brute_bin(v, edges):
    nb = len(edges) - 1
    if v != v:  # NaN sorts after every edge
        k = len(edges)
    else:
        k = sum(1 for e in edges if e <= v)
    return min(max(k - 1, 0), nb - 1)


def test_spaxel_brute_force():
    rng = np.random.default_rng(0)
    edges = np.sort(rng.uniform(-3, 3, 9)).astype(np.float32)
    c = rng.uniform(-4, 4, (400, 3)).astype(np.float32)
    c[:20, 0] = edges[rng.integers(0, 9, 20)]  # exactly on edges
    c[20:25, 1] = np.nan
    c[25:30, 0] = np.inf
    nb = len(edges) - 1
    exp = [_brute_bin(x, edges) + nb * _brute_bin(y, edges) for x, y in c[:, :2]]
    np.testing.assert_array_equal(R.square_spaxel_assignment(c, edges), exp)


def test_aperture_mask_known_answers():
    g = GOLD["aperture_mask"]
    for case in g["cases"]:
        m = R.mask_particles_outside_aperture(np.array(case["coords"], np.float32), np.array(g["edges"]))
        assert int(m.sum()) == case["inside"]


def test_spatial_bin_edges_known_answer():
    g = GOLD["spatial_bin_edges"]
    edges, size = R.spatial_bin_edges(g["angular_scale"], g["fov"], g["spatial_bins"])
    assert size == pytest.approx(g["expected_bin_size"])
    np.testing.assert_allclose(edges, np.arange(*g["expected_edges_arange"]), atol=1e-6)


def test_sort_matches_brute_force():
    rng = np.random.default_rng(1)
    ids = rng.integers(0, 7, 300).astype(np.int32)
    perm, off = R.sort_by_spaxel(ids, 7)
    exp = sorted(range(len(ids)), key=lambda i: (ids[i], i))
    np.testing.assert_array_equal(perm, exp)
    for s in range(8):
        assert off[s] == sum(1 for v in ids if v < s)


def test_reshape_array_known_answers():
    for case in GOLD["reshape_array"]["cases"]:
        got = R.reshape_array(np.array(case["arr"]), case["n_dev"])
        np.testing.assert_array_equal(got, case["expected"])


# ---------------------------------------------------------------- lookup

def test_lookup_known_answers():
    g = GOLD["lookup_interpolation"]
    for case in g["cases"]:
        got = R.lookup_linear([case["metallicity"]], [case["age"]], g["metallicity"], g["age"],
                              g["flux"], log_axes=False, dtype=np.float64)
        np.testing.assert_allclose(got[0], case["expected"], rtol=1e-12, atol=1e-12)


def test_lookup_brute_force_and_affine():
    rng = np.random.default_rng(2)
    x = np.sort(rng.uniform(0, 1, 4))
    y = np.sort(rng.uniform(0, 1, 5))
    f = rng.uniform(0, 1, (4, 5, 3))
    q = np.stack([rng.uniform(x[0], x[-1], 50), rng.uniform(y[0], y[-1], 50)], 1)
    got = R.lookup_linear(q[:, 0], q[:, 1], x, y, f, dtype=np.float64)
    for n, (xq, yq) in enumerate(q):
        i = max(1, min(len(x) - 1, int(np.sum(x <= xq))))
        j = max(1, min(len(y) - 1, int(np.sum(y <= yq))))
        tx = (xq - x[i - 1]) / (x[i] - x[i - 1])
        ty = (yq - y[j - 1]) / (y[j] - y[j - 1])
        exp = ((1 - tx) * (1 - ty) * f[i - 1, j - 1] + tx * (1 - ty) * f[i, j - 1]
               + (1 - tx) * ty * f[i - 1, j] + tx * ty * f[i, j])
        np.testing.assert_allclose(got[n], exp, rtol=1e-12)
    # an affine field is reproduced exactly (catches swapped axes / weights)
    fa = 0.3 + 2.0 * x[:, None, None] - 1.5 * y[None, :, None] + np.zeros((1, 1, 3))
    got = R.lookup_linear(q[:, 0], q[:, 1], x, y, fa, dtype=np.float64)
    np.testing.assert_allclose(got[:, 0], 0.3 + 2.0 * q[:, 0] - 1.5 * q[:, 1], rtol=1e-12)


def test_lookup_log_axes():
    met = np.array([1e-3, 1e-2, 1e-1])
    age = np.array([0.1, 1.0, 10.0])
    f = np.arange(27, dtype=np.float64).reshape(3, 3, 3)
    got = R.lookup_linear([10 ** -1.5], [10 ** 0.5], np.log10(met), np.log10(age), f, log_axes=True,
                          dtype=np.float64)
    np.testing.assert_allclose(got[0], f[1:, 1:].mean(axis=(0, 1)), rtol=1e-12)
    zero = R.lookup_linear([0.0], [1.0], np.log10(met), np.log10(age), f, log_axes=True,
                           dtype=np.float64)
    assert np.all(zero == 0)


# ---------------------------------------------------------------- doppler / resample

def test_doppler_known_answers():
    g = GOLD["cosmological_doppler_shift"]
    np.testing.assert_allclose(R.cosmological_doppler_shift(g["z"], g["wavelength"], np.float64),
                               g["expected"], rtol=1e-12)
    g = GOLD["velocity_doppler_shift"]
    got = R.velocity_doppler_shift(np.array(g["wavelength"]), np.array(g["velocity"]), g["direction"],
                                   g["speed_of_light"], np.float64)
    np.testing.assert_allclose(got, g["expected"], rtol=1e-9)
    with pytest.raises(ValueError):
        R.velocity_doppler_shift(np.array(g["wavelength"]), np.array(g["velocity"]), "w")


def test_calculate_diff_known_answer():
    g = GOLD["calculate_diff"]
    np.testing.assert_array_equal(R.calculate_diff(np.array(g["vec"])), g["expected"])
    np.testing.assert_array_equal(R.calculate_diff(np.array(g["vec"]), False), g["expected_no_pad"])


def test_interp_matches_numpy():
    rng = np.random.default_rng(3)
    for _ in range(20):
        xp = np.sort(rng.uniform(0, 10, 30))
        fp = rng.normal(size=30)
        x = rng.uniform(-2, 12, 100)
        x[:5] = xp[rng.integers(0, 30, 5)]
        np.testing.assert_allclose(R.interp_jax(x, xp, fp), np.interp(x, xp, fp), rtol=1e-12,
                                   atol=1e-12)


def test_resample_known_answers():
    g = GOLD["resample_spectrum"]
    got = R.resample_spectrum(np.array(g["initial_spectrum"]), np.array(g["initial_wavelength"]),
                              np.array(g["target_wavelength"]))
    np.testing.assert_allclose(got, g["expected"], rtol=1e-12)
    zero = R.resample_spectrum(np.zeros(5), np.array(g["initial_wavelength"]),
                               np.array(g["target_wavelength"]))
    assert np.all(zero == 0) and not np.any(np.isnan(zero))
    g = GOLD["resample_spectrum_vmap"]
    got = R.resample_spectra(np.array(g["initial_spectra"]), np.array(g["initial_wavelengths"]),
                             np.array(g["target_wavelength"]))
    np.testing.assert_allclose(got, g["expected"], rtol=1e-12)


def test_resample_flux_conservation():
    """sum(out * diff(target)) equals the in-range source integral T."""
    rng = np.random.default_rng(4)
    w = np.linspace(3000, 9000, 400) * 1.0003
    S = rng.uniform(0.5, 2.0, 400)
    tw = np.linspace(4000, 8000, 333)
    out = R.resample_spectrum(S, w, tw)
    m = (w >= tw.min()) & (w <= tw.max())
    T = np.sum(S * np.diff(w, prepend=w[0]) * m)
    assert np.sum(out * np.diff(tw, prepend=tw[0])) == pytest.approx(T, rel=1e-12)


def test_calculate_cube_known_answer_and_brute():
    g = GOLD["calculate_cube"]
    got = R.calculate_cube(np.array(g["spectra"], np.float32), np.array(g["spaxel_index"]), g["num_spaxels"])
    np.testing.assert_array_equal(got, g["expected"])
    rng = np.random.default_rng(5)
    sp = rng.normal(size=(50, 4))
    ids = rng.integers(-1, 11, 50)  # -1 and 9, 10 are dropped for nb = 3
    got = R.calculate_cube(sp, ids, 3).reshape(9, 4)
    exp = np.zeros((9, 4))
    for i in range(50):
        if 0 <= ids[i] < 9:
            exp[ids[i]] += sp[i]
    np.testing.assert_allclose(got, exp, rtol=1e-12)


def test_scale_by_mass_known_answer():
    g = GOLD["scale_spectrum_by_mass"]
    got = R.scale_spectrum_by_mass(np.array(g["spectra"], np.float32), np.array(g["mass"], np.float32))
    np.testing.assert_array_equal(got, g["expected"])


# ---------------------------------------------------------------- PSF / LSF

@pytest.mark.parametrize("ky,kx", [(3, 3), (5, 5), (11, 11), (15, 15), (4, 6), (1, 7)])
def test_psf_matches_scipy(ky, kx):
    rng = np.random.default_rng(ky * 10 + kx)
    cube = rng.normal(size=(12, 9, 3))
    K = rng.normal(size=(ky, kx))
    got = R.apply_psf(cube, K)
    for l in range(3):
        np.testing.assert_allclose(got[:, :, l], scipy.signal.convolve2d(cube[:, :, l], K, mode="same"),
                                   rtol=1e-12, atol=1e-12)


def test_psf_reference_cases():
    g = GOLD["psf"]
    K = R.gaussian_kernel_2d(3, 3, 2.0, np.float64)
    assert K.sum() == pytest.approx(g["kernel_3x3_sigma2_sum"])
    K15 = R.gaussian_kernel_2d(15, 15, 4.0, np.float64)
    assert K15[7, 7] == K15.max()
    g1 = np.exp(-0.5 * np.arange(-7, 8) ** 2 / 16.0)
    g1 /= g1.sum()
    np.testing.assert_allclose(K15, np.outer(g1, g1), rtol=1e-12)
    cube = np.zeros(g["delta_cube_shape"])
    cube[5, 5, :] = 1
    out = R.apply_psf(cube, np.array(g["ones_kernel"], float))
    exp = np.zeros_like(cube)
    exp[4:7, 4:7, :] = 1
    np.testing.assert_array_equal(out, exp)


def test_lsf_known_answers():
    g = GOLD["lsf"]
    n = g["wave_bins"]
    spectra = np.zeros((3, n))
    for i, p in enumerate(g["delta_positions"]):
        spectra[i, p] = 1
    out = R.apply_lsf(spectra, g["sigma"], g["wave_resolution"], dtype=np.float64)
    x = np.arange(n)
    for i, p in enumerate(g["delta_positions"]):
        e = np.exp(-0.5 * (x - p) ** 2 / g["sigma"] ** 2)
        np.testing.assert_allclose(out[i], e / e.sum(), atol=1e-5)


def test_lsf_matches_numpy_convolve():
    rng = np.random.default_rng(6)
    cube = rng.normal(size=(4, 3, 60))
    k = R.lsf_kernel(1.3, 1.25, dtype=np.float64)
    assert len(k) == 25
    got = R.apply_lsf_kernel(cube, k)
    for a in range(4):
        for b in range(3):
            full = np.convolve(cube[a, b], k, mode="full")
            np.testing.assert_allclose(got[a, b], full[12:60 + 24 - 12], rtol=1e-12, atol=1e-12)


def test_lsf_preserves_shape_and_integral():
    # tests/test_core_lsf.py:37-79: delta at lambda bin 5 of a 10x10x10 cube
    cube = np.zeros((10, 10, 10))
    cube[:, :, 5] = 1.0
    out = R.apply_lsf(cube, 0.6, 1.25, dtype=np.float64)
    assert out.shape == cube.shape
    assert np.all(out[:, :, 5] != 1)
    np.testing.assert_allclose(out.sum(axis=2), 1.0, rtol=1e-12)


@pytest.mark.parametrize("ny,nx,nl,m,sig", [(25, 25, 7, 11, 1.2), (9, 13, 3, 15, 4.0),
                                             (3, 4, 2, 5, 0.7), (40, 40, 5, 1, 1.0),
                                             (150, 20, 2, 11, 3.0)])
def test_separable_psf_equals_direct(ny, nx, nl, m, sig):
    """The separable PSF restatements (numpy and torch, used for the largest
    cubes) equal the direct convolve2d 'same' sum, and the torch one has the
    same sigma gradient."""
    import torch

    from oracle import ref_torch as RT

    rng = np.random.default_rng(ny * nx + m)
    c = rng.normal(size=(ny, nx, nl))
    d = R.apply_psf(c, R.gaussian_kernel_2d(m, m, sig, dtype=np.float64))
    np.testing.assert_allclose(np.outer(R.gaussian_taps_1d(m, sig), R.gaussian_taps_1d(m, sig)),
                               R.gaussian_kernel_2d(m, m, sig, dtype=np.float64), rtol=1e-14, atol=0)
    s = R.apply_psf_separable(c, R.gaussian_taps_1d(m, sig))
    assert np.abs(s - d).max() <= 1e-14 * np.abs(d).max()
    ct = torch.tensor(c)
    g = rng.normal(size=c.shape)
    sa = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    sb = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    a = RT.apply_psf(ct, RT.gaussian_kernel_2d(m, m, sa))
    b = RT.apply_psf_separable(ct, RT.gaussian_taps_1d(m, sb))
    assert (a - b).abs().max().item() <= 1e-14 * a.abs().max().item()
    (a * torch.tensor(g)).sum().backward()
    (b * torch.tensor(g)).sum().backward()
    assert abs(float(sa.grad) - float(sb.grad)) <= 1e-12 * max(abs(float(sa.grad)), 1e-30)


def test_lsf_even_tap_count():
    """wave_res 1.6 (HECTOR) and 2.7 (CALIFA): the float arange of the kernel
    grid has 26 taps; the full convolution cropped at 12 is n + 1 long (the
    reference's reshape raises) and the oracle keeps its first n bins, the
    x = 0 tap (index 12) on each input bin."""
    for res in (1.6, 2.7):
        k = R.lsf_kernel(1.0, res, dtype=np.float64)
        assert len(k) == 26
        assert abs(float(np.arange(-12 * res, 12 * res + res, res)[12])) < 1e-12
        rng = np.random.default_rng(1)
        spec = rng.normal(size=(3, 40))
        full = np.stack([np.convolve(s, k, mode="full") for s in spec])
        assert full[:, 12:40 + 26 - 1 - 12].shape[1] == 41
        np.testing.assert_allclose(R.apply_lsf_kernel(spec, k), full[:, 12:52], rtol=1e-13,
                                   atol=1e-14)
        # a delta stays centred: the output peaks on the input bin
        d = np.zeros((1, 41))
        d[0, 20] = 1.0
        assert int(np.argmax(R.apply_lsf_kernel(d, R.lsf_kernel(2 * res, res, dtype=np.float64)))) == 20
