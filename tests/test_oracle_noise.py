"""CPU: the apply_noise restatement (oracle/ref_numpy.py noise section)
against published known answers, library routines and the reference's own
noise tests (tests/test_telescope_noise.py, tests/test_core_noise.py)."""
import numpy as np
import pytest
from scipy import special, stats

from oracle import ref_numpy as R


@pytest.mark.parametrize("key,ctr,expected", [
    # Random123 kat_vectors, threefry2x32_20 (also in jax's random_test.py)
    ((0x00000000, 0x00000000), (0x00000000, 0x00000000), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
])
def test_threefry_known_answers(key, ctr, expected):
    y0, y1 = R.threefry2x32(key[0], key[1], np.array([ctr[0]], np.uint32),
                            np.array([ctr[1]], np.uint32))
    assert (int(y0[0]), int(y1[0])) == expected


def test_random_bits_counter_layout():
    """Index i uses the counter (i >> 32, i & 0xffffffff) and the key
    (seed >> 32, seed & 0xffffffff); the output is the XOR of both words."""
    seed = (7 << 32) | 11
    bits = R.random_bits(seed, 5)
    for i in range(5):
        y0, y1 = R.threefry2x32(7, 11, np.array([0], np.uint32), np.array([i], np.uint32))
        assert bits[i] == (y0[0] ^ y1[0])
    assert not np.array_equal(R.random_bits(0, 64), R.random_bits(1, 64))


def test_uniform_draws():
    u = R.sample_noise((1 << 20,), "uniform", seed=0)
    assert u.dtype == np.float32 and u.min() >= 0.0 and u.max() < 1.0
    np.testing.assert_array_equal(u * np.float32(2 ** 23), np.floor(u * np.float32(2 ** 23)))
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 2e-3
    assert stats.kstest(u, "uniform").pvalue > 1e-4


def test_normal_draws_are_the_normal_quantile_of_the_same_uniforms():
    n = 1 << 20
    z = R.sample_noise((n,), "normal", seed=3)
    u = R.sample_noise((n,), "uniform", seed=3).astype(np.float64)
    # jax maps u01 to (-1, 1) as 2 u01 + nextafter(-1, 0); sqrt(2) erfinv(2p - 1) = ndtri(p)
    lo = float(np.nextafter(np.float32(-1), np.float32(0)))
    p = (np.maximum(lo, 2 * u + lo) + 1.0) / 2.0
    np.testing.assert_allclose(z, special.ndtri(p), rtol=2e-6, atol=2e-7)
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.std() - 1.0) < 5e-3
    assert stats.kstest(z, "norm").pvalue > 1e-4
    assert np.isfinite(z).all()


def test_sample_noise_reference_cases():
    # tests/test_telescope_noise.py:61-79
    for t in ("normal", "uniform"):
        x = R.sample_noise((5, 5, 10), t)
        assert x.shape == (5, 5, 10) and not np.isnan(x).any()
    with pytest.raises(ValueError, match="Invalid noise type: invalid"):
        R.sample_noise((5, 5, 10), "invalid")


@pytest.mark.parametrize("n", [1, 2, 25, 24, 1001])
def test_flux_median_is_numpy_median(n):
    rng = np.random.default_rng(n)
    f = rng.uniform(0.1, 10.0, n).astype(np.float32)
    med, lo, hi = R.flux_median(f)
    assert med == np.float32(np.median(f))
    s = np.sort(f)
    assert f[lo] == s[(n - 1) // 2] and f[hi] == s[n // 2]


def test_flux_median_nan_propagation():
    f = np.array([3.0, 0.0, 5.0, 7.0], np.float32)
    assert np.isnan(np.median(np.where(f > 0, f, np.nan)))
    assert R.flux_median(f) == (0.0, -1, -1)
    assert R.flux_median(np.array([1.0, -2.0, 4.0], np.float32))[0] == 0.0


def test_flux_median_ties_take_the_stable_order():
    f = np.array([2.0, 1.0, 2.0, 2.0, 3.0], np.float32)
    med, lo, hi = R.flux_median(f)
    assert med == 2.0 and lo == hi == 2  # ranks 0:1(i1) 1:2(i0) 2:2(i2)


def test_s2n_closed_forms():
    cube = np.full((4, 3, 6), 0.5, np.float32)  # flux 3 everywhere -> median 3
    S = R.calculate_S2N(cube, 0.8)
    np.testing.assert_allclose(S, 1.0 / 0.8, rtol=1e-7)
    cube[1, 2] = 0.0  # one empty spaxel: NaN median -> 0 -> no noise at all
    assert np.all(R.calculate_S2N(cube, 0.8) == 0.0)
    sn = np.arange(1, 13, dtype=np.float32).reshape(4, 3)
    S = R.calculate_S2N(np.full((4, 3, 6), 0.5, np.float32), sn)
    np.testing.assert_allclose(S, 1.0 / sn, rtol=1e-6)


@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_reference_noise_cube_cases(dist):
    # tests/test_telescope_noise.py:15-58 (uniform cube of PRNGKey(0), S2N 0.8)
    cube = R.sample_noise((5, 5, 10), "uniform", seed=0)
    out = R.apply_noise(cube, 0.8, dist)
    noise = out - cube
    assert noise.shape == cube.shape
    assert not np.isnan(noise).any() and not np.isinf(noise).any()
    assert np.all(np.var(noise.reshape(-1, 10), axis=0) > 0)
    zero = np.zeros((5, 5, 10), np.float32)
    assert np.all(R.apply_noise(zero, 0.8, dist) - zero == 0)


def test_apply_noise_scale():
    """noise = cube * z * S with S = sqrt(M)/sn/sqrt(F): for a cube of equal
    rows, (out - cube) / (cube * z) = 1/sn."""
    cube = np.full((6, 5, 40), 0.25, np.float32)
    z = R.sample_noise(cube.shape, "normal", 5)
    out = R.apply_noise(cube, 2.0, "normal", seed=5, dtype=np.float64)
    np.testing.assert_allclose((out - cube) / (cube * z), 0.5, rtol=1e-6)


@pytest.mark.parametrize("shape", [(5, 5, 7), (4, 5, 6)])
@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_vjp_matches_central_differences(shape, dist):
    rng = np.random.default_rng(sum(shape))
    c = rng.uniform(0.5, 2.0, shape)
    g = rng.normal(size=shape)
    sn = rng.uniform(0.5, 3.0, shape[:2])
    dc = R.apply_noise_vjp(c, sn, g, dist, seed=9)

    def L(x):
        return float((g * R.apply_noise(x, sn, dist, seed=9, dtype=np.float64)).sum())

    eps = 1e-6
    for k in range(6):
        v = rng.normal(size=shape)
        fd = (L(c + eps * v) - L(c - eps * v)) / (2 * eps)
        assert abs(fd - (dc * v).sum()) <= 1e-6 * max(1.0, abs(fd)), (k, fd, (dc * v).sum())
    # the median term: perturbing only the median spaxel(s) moves every S
    F = c.sum(-1)
    _, lo, hi = R.flux_median(F)
    v = np.zeros(shape)
    v.reshape(-1, shape[-1])[lo] = 1.0
    fd = (L(c + eps * v) - L(c - eps * v)) / (2 * eps)
    assert abs(fd - (dc * v).sum()) <= 1e-6 * max(1.0, abs(fd))


def test_vjp_with_an_empty_spaxel_is_the_identity():
    c = np.ones((3, 3, 4))
    c[0, 0] = 0.0
    g = np.arange(36.0).reshape(3, 3, 4)
    np.testing.assert_array_equal(R.apply_noise_vjp(c, 1.0, g), g)
