"""PSF and LSF kernels at edge shapes against the float64 oracle.

Forward, cube adjoint and sigma gradient of GaussianPSF / GaussianLSF (the
separable PSF of rubix/telescope/psf/psf.py:33-58 with the kernel of
rubix/telescope/psf/kernels.py:5-31, and the 25-tap LSF of
rubix/telescope/lsf/lsf.py:12-105) against oracle/ref_torch.py in float64
with torch autograd, at shapes the headline cubes do not reach: a kernel
wider than the plane, even and one-tap kernels, one-plane cubes, an axis
too long for the LDS-tiled kernels, spectra shorter than the LSF, one-bin
spectra, and more than one LSF row segment.
The general 2D kernel path (conv2d_planes) is checked on the reference's own
cases (tests/test_telescope_psf.py:22-55: a bright pixel under a 3x3 ones
kernel, which sums exactly), and the LSF on the reference's delta cases
(tests/test_telescope_lsf.py:6-60, atol 1e-5 against the normalised
Gaussian on the bin grid).

Tolerance: norm-wise relative error <= 1e-5 for the float32 GPU results
against float64; sigma gradients to 1e-5 relative, or to float32 rounding of
the sum's scale where the true gradient is zero (one- and two-tap kernels).
"""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R, ref_torch as RT
from rubix_amd import ops
from rubix_amd.autograd import GaussianLSF, GaussianPSF
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-5


def _sigma_close(got, want, gout, cube, sigma, tol=TOL):
    """|got - want| <= tol |want|, or within float32 rounding of the sum's
    scale ||gout|| ||cube|| / sigma (it bounds |dL/dsigma| up to the kernel
    derivative's O(1) norm) where the true gradient is zero or tiny (an
    even two-tap kernel is [1/2, 1/2] for every sigma)."""
    scale = np.linalg.norm(gout) * np.linalg.norm(cube) / sigma
    return abs(got - want) <= max(tol * abs(want), 1e-7 * scale)


PSF_CASES = [  # (ny, nx, nl, size, sigma)
    (1, 1, 3, 5, 1.0),
    (3, 3, 7, 15, 4.0),
    (4, 6, 5, 4, 0.8),
    (7, 5, 1, 1, 1.0),
    (10, 10, 9, 2, 0.6),
    (50, 50, 17, 15, 4.0),
    (21, 13, 1100, 11, 1.2),
    (250, 3, 2, 5, 1.0),  # an axis longer than the LDS-tiled kernels stage
]


@pytest.mark.parametrize("ny,nx,nl,size,sigma", PSF_CASES)
def test_psf_edge_shapes(ny, nx, nl, size, sigma):
    rng = np.random.default_rng(ny * 1000 + nx * 10 + size)
    cube = rng.normal(size=(ny, nx, nl))
    gout = rng.normal(size=(ny, nx, nl))
    print(f"psf case ny={ny} nx={nx} nl={nl} size={size} sigma={sigma}", flush=True)
    c64 = torch.tensor(cube, requires_grad=True)
    s64 = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    want = RT.apply_psf(c64, RT.gaussian_kernel_2d(size, size, s64))
    want.backward(torch.tensor(gout))
    c32 = torch.tensor(cube, dtype=torch.float32, device=DEV, requires_grad=True)
    s32 = torch.tensor(sigma, dtype=torch.float32, device=DEV, requires_grad=True)
    got = GaussianPSF.apply(c32, s32, size)
    got.backward(torch.tensor(gout, dtype=torch.float32, device=DEV))
    torch.cuda.synchronize()
    assert got.shape == (ny, nx, nl)
    errs = (rel_err(got.detach().cpu().numpy(), want.detach().numpy()),
            rel_err(c32.grad.cpu().numpy(), c64.grad.numpy()))
    print("  cube", errs[0], "adjoint", errs[1], "dsigma", float(s32.grad), float(s64.grad))
    assert errs[0] <= TOL and errs[1] <= TOL
    assert _sigma_close(float(s32.grad), float(s64.grad), gout, cube, sigma)


LSF_CASES = [  # (n_spec, nl, sigma, wave_res)
    (5, 1, 1.0, 1.25),
    (3, 10, 1.0, 1.25),
    (4, 24, 2.0, 1.0),
    (4, 25, 1.0, 1.25),
    (4, 26, 0.4, 1.25),
    (6, 100, 2.0, 1.0),
    (2, 5000, 1.0, 1.25),
]


@pytest.mark.parametrize("n_spec,nl,sigma,wave_res", LSF_CASES)
def test_lsf_edge_lengths(n_spec, nl, sigma, wave_res):
    rng = np.random.default_rng(n_spec * 100000 + nl)
    cube = rng.normal(size=(n_spec, nl))
    gout = rng.normal(size=(n_spec, nl))
    print(f"lsf case n_spec={n_spec} nl={nl} sigma={sigma} wave_res={wave_res}", flush=True)
    c64 = torch.tensor(cube, requires_grad=True)
    s64 = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    k = RT.lsf_kernel(s64, wave_res)
    assert k.numel() == 25
    want = RT.apply_lsf(c64, k)
    want.backward(torch.tensor(gout))
    c32 = torch.tensor(cube, dtype=torch.float32, device=DEV, requires_grad=True)
    s32 = torch.tensor(sigma, dtype=torch.float32, device=DEV, requires_grad=True)
    got = GaussianLSF.apply(c32, s32, wave_res)
    got.backward(torch.tensor(gout, dtype=torch.float32, device=DEV))
    torch.cuda.synchronize()
    assert got.shape == (n_spec, nl)
    errs = (rel_err(got.detach().cpu().numpy(), want.detach().numpy()),
            rel_err(c32.grad.cpu().numpy(), c64.grad.numpy()))
    print("  cube", errs[0], "adjoint", errs[1], "dsigma", float(s32.grad), float(s64.grad))
    assert errs[0] <= TOL and errs[1] <= TOL
    assert _sigma_close(float(s32.grad), float(s64.grad), gout, cube, sigma)


@pytest.mark.parametrize("nl", [1, 3])
def test_psf_reference_ones_kernel(nl):
    """tests/test_telescope_psf.py:22-55 on the general 2D kernel path."""
    cube = np.zeros((10, 10, nl), np.float32)
    cube[5, 5, :] = 1
    K = np.ones((3, 3), np.float32)
    got = ops.conv2d_planes(torch.tensor(cube, device=DEV), torch.tensor(K, device=DEV))
    want = np.zeros_like(cube)
    want[4:7, 4:7, :] = 1
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    np.testing.assert_array_equal(got.cpu().numpy(), R.apply_psf(cube.astype(np.float64), K))


@pytest.mark.parametrize("ky,kx", [(4, 6), (1, 7), (15, 15)])
def test_conv2d_planes_matches_oracle(ky, kx):
    rng = np.random.default_rng(ky * 10 + kx)
    cube = rng.normal(size=(12, 9, 33))
    K = rng.normal(size=(ky, kx))
    got = ops.conv2d_planes(torch.tensor(cube, dtype=torch.float32, device=DEV),
                            torch.tensor(K, dtype=torch.float32, device=DEV))
    want = R.apply_psf(cube.astype(np.float32).astype(np.float64),
                       K.astype(np.float32).astype(np.float64))
    assert rel_err(got.cpu().numpy(), want) <= TOL


def test_lsf_reference_delta_cases():
    """tests/test_telescope_lsf.py:6-60: a unit spike at bins 20, 50 and 75 of
    100 becomes the normalised Gaussian around it (sigma 2, resolution 1)."""
    spectra = np.zeros((3, 100), np.float32)
    for i, p in enumerate((20, 50, 75)):
        spectra[i, p] = 1
    got = GaussianLSF.apply(torch.tensor(spectra, device=DEV),
                            torch.tensor(2.0, device=DEV), 1.0).cpu().numpy()
    x = np.arange(100)
    for i, p in enumerate((20, 50, 75)):
        e = np.exp(-0.5 * (x - p) ** 2 / 2.0 ** 2)
        np.testing.assert_allclose(got[i], e / e.sum(), atol=1e-5)


@pytest.mark.parametrize("axis", [0, 1])
def test_batched_psf_pass_equals_per_galaxy(axis):
    """rbx_conv_batch over a ragged batch (nb 1, 7, 101, 3: a one-spaxel
    grid, and a line longer than the batch kernel stages in LDS) against the
    single-galaxy pass on each galaxy's own (nb, nb, n_lambda) cube: the same
    taps in the same order, so bit for bit; the sigma dot (f64, another
    summation order) to 1e-12."""
    nbs, nt, m = [1, 7, 101, 3], 70, 15
    layout = ops.BatchLayout([1] * len(nbs), nbs, DEV)
    rng = np.random.default_rng(40 + axis)
    x = torch.tensor(rng.normal(size=(layout.n_rows, nt)), dtype=torch.float32, device=DEV)
    y = torch.tensor(rng.normal(size=(layout.n_rows, nt)), dtype=torch.float32, device=DEV)
    taps = torch.tensor(rng.uniform(0.0, 1.0, m), dtype=torch.float32, device=DEV)
    dtaps = torch.tensor(rng.normal(size=m), dtype=torch.float64, device=DEV)
    start = (m - 1) // 2
    got = ops.conv_batch(x, layout, axis, taps, start).cpu().numpy()
    dot = float(ops.conv_batch_dot(y, x, layout, axis, dtaps, start))
    want_dot = 0.0
    for g, nb in enumerate(nbs):
        r0 = int(layout.row_begin[g])
        xg = x[r0:r0 + nb * nb].reshape(nb, nb, nt).contiguous()
        yg = y[r0:r0 + nb * nb].reshape(nb, nb, nt).contiguous()
        if axis == 1:
            ref = ops.conv_axis(xg, taps, nb, nb, nt, start)
            want_dot += float(ops.conv_axis_dot(yg, xg, dtaps, nb, nb, nt, start))
        else:
            ref = ops.conv_axis(xg, taps, 1, nb, nb * nt, start)
            want_dot += float(ops.conv_axis_dot(yg, xg, dtaps, 1, nb, nb * nt, start))
        np.testing.assert_array_equal(got[r0:r0 + nb * nb], ref.reshape(nb * nb, nt).cpu().numpy(),
                                      err_msg=f"galaxy {g} (nb {nb}), axis {axis}")
    assert abs(dot - want_dot) <= 1e-12 * max(1.0, abs(want_dot))


def test_batched_psf_table_with_a_gap_row():
    """A galaxy table that is not packed (one row between the two galaxies,
    which belongs to neither) takes the row items: the gap row is zero and
    each galaxy's rows equal its single-galaxy pass bit for bit."""
    from types import SimpleNamespace
    from rubix_amd._lib import RbxGalaxy
    nbs, begins, nt, m = [5, 4], [0, 26], 40, 7
    n_rows = 26 + 16
    arr = (RbxGalaxy * 2)()
    for g in range(2):
        arr[g].star_begin = 0
        arr[g].spaxel_begin = begins[g]
        arr[g].nb = nbs[g]
    table = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(DEV)
    layout = SimpleNamespace(table=table, n_gal=2, n_rows=n_rows)
    rng = np.random.default_rng(77)
    x = torch.tensor(rng.normal(size=(n_rows, nt)), dtype=torch.float32, device=DEV)
    taps = torch.tensor(rng.uniform(0.0, 1.0, m), dtype=torch.float32, device=DEV)
    for axis in (0, 1):
        got = ops.conv_batch(x, layout, axis, taps, 3).cpu().numpy()
        assert np.all(got[25] == 0.0)
        for g, nb in enumerate(nbs):
            xg = x[begins[g]:begins[g] + nb * nb].reshape(nb, nb, nt).contiguous()
            ref = (ops.conv_axis(xg, taps, nb, nb, nt, 3) if axis == 1
                   else ops.conv_axis(xg, taps, 1, nb, nb * nt, 3))
            np.testing.assert_array_equal(got[begins[g]:begins[g] + nb * nb],
                                          ref.reshape(nb * nb, nt).cpu().numpy())
