"""GPU: the apply_noise kernels (rbx_noise_sample, rbx_apply_noise,
rbx_apply_noise_bwd) against the noise oracle.  Uniform draws, the flux
median, the median spaxels and the S2N map are compared bit for bit; normal
draws within the erfinv rounding of the two implementations; the adjoint
against the float64 analytic adjoint."""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops
from rubix_amd.autograd import ApplyNoise
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"
NORMAL_RTOL = 1e-6  # device erfinvf vs float64 erfinv rounded once: a few ulp


def _cube(shape, seed, lo=0.1, hi=2.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, shape).astype(np.float32)


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 3])
@pytest.mark.parametrize("seed", [0, 12345, (1 << 40) + 7])
def test_sample_matches_oracle(n, seed):
    u = ops.noise_sample((n,), "uniform", seed).cpu().numpy()
    np.testing.assert_array_equal(u, R.sample_noise((n,), "uniform", seed))
    z = ops.noise_sample((n,), "normal", seed).cpu().numpy()
    np.testing.assert_allclose(z, R.sample_noise((n,), "normal", seed), rtol=NORMAL_RTOL, atol=1e-7)


@pytest.mark.parametrize("shape", [(7, 9, 37), (8, 8, 3721), (1, 1, 5), (50, 50, 200)])
@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_forward_matches_oracle(shape, dist):
    c = _cube(shape, sum(shape))
    sn = np.random.default_rng(1).uniform(0.5, 3.0, shape[:2]).astype(np.float32)
    out, st = ops.apply_noise(torch.from_numpy(c).to(DEV), torch.from_numpy(sn).to(DEV), dist, 0)
    flux = st.flux.cpu().numpy()
    np.testing.assert_allclose(flux, R.flux_image(c), rtol=2e-7)
    med, lo, hi = R.flux_median(flux)
    assert st.median() == med
    assert st.median_spaxels() == (lo, hi)
    s2n = st.s2n.cpu().numpy()
    np.testing.assert_array_equal(s2n, R.calculate_S2N(c, sn, flux=flux))
    ref = R.apply_noise(c, sn, dist, 0, flux=flux)
    got = out.cpu().numpy()
    if dist == "uniform":
        np.testing.assert_array_equal(got, ref)
    else:
        # out = c + c z S: a z error of NORMAL_RTOL |z| moves out by at most
        # NORMAL_RTOL |c z S|, plus one rounding of out
        zs = np.abs(R.sample_noise(shape, dist, 0) * s2n[..., None])
        bound = NORMAL_RTOL * np.abs(c) * zs + 2 * np.spacing(np.abs(ref))
        assert np.all(np.abs(got - ref) <= bound)
    ref64 = R.apply_noise(c.astype(np.float64), sn, dist, 0, dtype=np.float64)
    print(f"{shape} {dist}: rel64={rel_err(got, ref64):.3e}")
    assert rel_err(got, ref64) <= 1e-5


@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_empty_spaxel_means_no_noise(dist):
    c = _cube((6, 7, 50), 3)
    c[2, 3] = 0.0
    cube = torch.from_numpy(c).to(DEV)
    out, st = ops.apply_noise(cube, 1.0, dist, 0)
    assert torch.equal(out, cube)
    assert st.median() == 0.0 and st.median_spaxels() == (-1, -1)
    assert int(st.stats[12:16].view(torch.int32).item()) == 1
    zero = torch.zeros(5, 5, 10, device=DEV)
    assert torch.count_nonzero(ops.apply_noise(zero, 0.8, dist)[0]) == 0


def test_negative_flux_spaxel_means_no_noise():
    c = _cube((4, 4, 8), 4)
    c[0, 1] = -1.0
    cube = torch.from_numpy(c).to(DEV)
    assert torch.equal(ops.apply_noise(cube, 1.0)[0], cube)


@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_reference_noise_cube_properties(dist):
    # tests/test_telescope_noise.py:15-58: uniform cube from PRNGKey(0), S2N 0.8
    cube = ops.noise_sample((5, 5, 10), "uniform", 0)
    out, _ = ops.apply_noise(cube, 0.8, dist)
    noise = (out - cube).cpu().numpy()
    assert noise.shape == (5, 5, 10)
    assert np.isfinite(noise).all()
    assert np.all(np.var(noise.reshape(-1, 10), axis=0) > 0)


def test_tied_fluxes_pick_the_stable_order():
    c = np.ones((3, 3, 4), np.float32)
    c[0, 0] *= 2.0
    c[2, 2] *= 0.5  # fluxes 8, 4 x7, 2: median 4 at the 4th smallest (index 4)
    out, st = ops.apply_noise(torch.from_numpy(c).to(DEV), 1.0)
    med, lo, hi = R.flux_median(R.flux_image(c))
    assert st.median() == med == 4.0
    assert st.median_spaxels() == (lo, hi) == (4, 4)
    c = np.ones((2, 2, 3), np.float32)  # all equal, even count: ranks 1 and 2
    _, st = ops.apply_noise(torch.from_numpy(c).to(DEV), 1.0)
    assert st.median_spaxels() == (1, 2)


@pytest.mark.parametrize("shape", [(5, 5, 7), (4, 6, 300), (16, 16, 1500)])
@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_adjoint_matches_oracle(shape, dist):
    c = _cube(shape, 10 + sum(shape), 0.5, 2.0)
    rng = np.random.default_rng(5)
    g = rng.normal(size=shape).astype(np.float32)
    sn = rng.uniform(0.5, 3.0, shape[:2]).astype(np.float32)
    x = torch.from_numpy(c).to(DEV).requires_grad_(True)
    out = ApplyNoise.apply(x, torch.from_numpy(sn).to(DEV), dist, 0)
    out.backward(torch.from_numpy(g).to(DEV))
    ref = R.apply_noise_vjp(c, sn, g, dist, 0)
    e = rel_err(x.grad.cpu().numpy(), ref)
    print(f"{shape} {dist}: grad rel64={e:.3e}")
    assert e <= 1e-5
    # the flux-path part alone (the adjoint minus its direct term) is small
    # next to g; check it on its own so that an error in it cannot hide
    z = R.sample_noise(shape, dist, 0).astype(np.float64)
    S = R.calculate_S2N(c.astype(np.float64), sn, np.float64)
    direct = g + g * z * S[..., None]
    assert rel_err(x.grad.cpu().numpy() - direct, ref - direct) <= 1e-4


def test_adjoint_with_empty_spaxel_is_the_identity():
    c = _cube((4, 4, 16), 6)
    c[1, 1] = 0.0
    g = np.random.default_rng(2).normal(size=c.shape).astype(np.float32)
    x = torch.from_numpy(c).to(DEV).requires_grad_(True)
    ApplyNoise.apply(x, 1.0, "normal", 0).backward(torch.from_numpy(g).to(DEV))
    np.testing.assert_array_equal(x.grad.cpu().numpy(), g)


def test_seed_changes_the_draws_but_not_the_scale():
    c = torch.from_numpy(_cube((6, 6, 64), 7)).to(DEV)
    a, sa = ops.apply_noise(c, 1.0, "normal", 0)
    b, sb = ops.apply_noise(c, 1.0, "normal", 1)
    assert not torch.equal(a, b)
    assert torch.equal(sa.s2n, sb.s2n)


def test_in_place_call_through_the_c_abi():
    """out may alias cube (the reference's datacube += noise)."""
    from rubix_amd._lib import call, ptr, stream_handle

    c = torch.from_numpy(_cube((5, 6, 33), 8)).to(DEV)
    ref, _ = ops.apply_noise(c, 2.0, "uniform", 3)
    sn = torch.full((5, 6), 2.0, device=DEV)
    flux = torch.empty(5, 6, device=DEV)
    s2n = torch.empty_like(flux)
    stats = torch.empty(16, dtype=torch.uint8, device=DEV)
    call("rbx_apply_noise", ptr(c), 30, 33, ptr(sn), 3, 1, ptr(flux), ptr(s2n), ptr(stats), ptr(c),
         stream_handle(c.device))
    assert torch.equal(c, ref)
