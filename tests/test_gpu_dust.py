"""GPU: the dust kernels (rbx_extinction_curve, rbx_gas_extinction,
rbx_column_av, rbx_apply_extinction) and the calculate_extinction step
against the dust oracle (float64 curves and gas extinction, the literal
per-spaxel column-density loop)."""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops, synthetic as S
from rubix_amd.autograd import ApplyExtinction
from rubix_amd.core.data import get_reshape_data, stars_from_arrays
from rubix_amd.core.dust import get_extinction, spaxel_area
from rubix_amd.core.ifu import (get_calculate_datacube, get_calculate_spectra,
                                get_doppler_shift_and_resampling, get_scale_spectrum_by_mass)
from rubix_amd.core.pipeline import RubixPipeline
from rubix_amd.core.telescope import get_spaxel_assignment
from tests.helpers import cfg1, rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(x, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=DEV)


@pytest.mark.parametrize("model", ["Cardelli89", "Gordon23"])
@pytest.mark.parametrize("rv", [2.5, 3.1, 4.5])
def test_curves_match_oracle(model, rv):
    # the telescope range and a log grid over every branch (912 A .. 30 um)
    wave = np.concatenate([np.arange(4700.0, 9351.0, 1.25),
                           np.geomspace(912.0, 3.0e5, 4000)]).astype(np.float32)
    got = ops.extinction_curve(model, rv, _t(wave)).cpu().numpy()
    # the reference divides the float32 wavelengths by 1e4 in float32
    w32 = (wave / np.float32(1e4)).astype(np.float32)
    ref = R.extinction_curve(model, w32.astype(np.float64), rv)
    # the curve is a + b k with terms of order 1..3 (G23's infrared b term
    # times 1/Rv - 1/3.1), so near its zero crossings only an absolute bound
    # at the float32 rounding of those terms is meaningful
    err = np.abs(got - ref)
    print(f"{model} Rv={rv}: max abs {err.max():.2e}, "
          f"max rel where |ref| > 0.1: {(err / np.maximum(np.abs(ref), 0.1)).max():.2e}")
    assert np.all(err <= 2e-5 * np.abs(ref) + 2e-6)


@pytest.mark.parametrize("model", ["power law slope free", "broken power law fit"])
@pytest.mark.parametrize("xco", ["MW", "Z"])
def test_gas_extinction_matches_oracle(model, xco):
    g = S.make_gas(50_000, seed=3)
    g["metals"][:100, 4] = 1e-5  # some very low metallicity (other branch of the broken law)
    got = ops.gas_extinction(_t(g["mass"]), _t(g["metals"]), model, xco, 3.5, 0.21).cpu().numpy()
    ref = R.gas_extinction(g["mass"].astype(np.float64), g["metals"].astype(np.float64), model,
                           xco, 3.5, 0.21)
    np.testing.assert_allclose(got, ref, rtol=2e-5)


def _column_case(rng, mode, nsp, G, N):
    nkeys = nsp + int(rng.integers(0, 2))
    gz = rng.normal(size=G).astype(np.float32) * np.float32(3)
    if mode == 1:
        gz = np.abs(gz)
    if mode == 2:
        gz = -np.abs(gz)
    if mode == 3 and G:
        gz[rng.random(G) < 0.3] = 0
    gp = rng.integers(0, nkeys + 1, G).astype(np.int32)
    ext = rng.uniform(0.1, 2, G).astype(np.float32)
    sz = (rng.normal(size=N) * 4).astype(np.float32)
    if mode == 3:
        sz[rng.random(N) < 0.2] = 0
    sp = rng.integers(0, nkeys + 1, N).astype(np.int32)
    return nkeys, gz, gp, ext, sz, sp


def _column_gpu(gz, gp, ext, sz, sp, nsp, nkeys):
    gc = np.zeros((len(gz), 3), np.float32)
    gc[:, 2] = gz
    sc = np.zeros((len(sz), 3), np.float32)
    sc[:, 2] = sz
    if len(gz) == 0:
        gc = np.zeros((0, 3), np.float32)
    return ops.column_av(_t(gc), _t(gp, torch.int32), _t(ext), _t(sc), _t(sp, torch.int32), nsp,
                         nkeys).cpu().numpy()


def test_column_av_small_cases_match_literal_oracle():
    """300 random small cases: other spaxels' gas above / below / both, z = 0
    ties, empty segments, ids outside the grid, 0 .. 2 gas cells."""
    rng = np.random.default_rng(0)
    for trial in range(300):
        mode = trial % 4
        nsp = int(rng.integers(1, 6))
        G = int(rng.integers(0, 30))
        N = int(rng.integers(1, 40))
        nkeys, gz, gp, ext, sz, sp = _column_case(rng, mode, nsp, G, N)
        ref = R.column_av(gz, gp, ext, sz, sp, nsp)
        got = _column_gpu(gz, gp, ext, sz, sp, nsp, nkeys)
        assert np.allclose(got, ref, rtol=2e-6, atol=1e-6, equal_nan=True), (trial, got, ref)


def test_column_av_galaxy_matches_literal_oracle():
    nb = 12
    g = S.make_gas(6000, seed=5)
    st = S.make_stars(3000, seed=6)
    edges = np.linspace(-12, 12, nb + 1).astype(np.float32)
    gp = R.square_spaxel_assignment(g["coords"], edges).astype(np.int32)
    sp = R.square_spaxel_assignment(st["coords"], edges).astype(np.int32)
    ext = np.random.default_rng(1).uniform(0.0, 0.05, len(gp)).astype(np.float32)
    ref = R.column_av(g["coords"][:, 2], gp, ext, st["coords"][:, 2], sp, nb * nb)
    got = ops.column_av(_t(g["coords"]), _t(gp, torch.int32), _t(ext), _t(st["coords"]),
                        _t(sp, torch.int32), nb * nb, nb * nb).cpu().numpy()
    # the oracle's cumulative sums are float32 (as the reference's), the GPU's double
    np.testing.assert_allclose(got, ref, rtol=3e-5, atol=1e-6)


def test_non_finite_gas_extinction_poisons_every_star():
    ext = np.array([1.0, np.inf, 1.0], np.float32)
    got = _column_gpu(np.array([1, 2, 3], np.float32), np.zeros(3, np.int32), ext,
                      np.array([1.5, 2.5], np.float32), np.zeros(2, np.int32), 1, 1)
    assert np.isnan(got).all()


def test_apply_extinction_and_adjoint():
    rng = np.random.default_rng(2)
    spec = rng.uniform(0, 1, (300, 517)).astype(np.float32)
    av = rng.uniform(0, 3, 300).astype(np.float32)
    wave = np.linspace(4700, 9350, 517).astype(np.float32)
    axav = ops.extinction_curve("Cardelli89", 3.1, _t(wave))
    x = _t(spec).requires_grad_(True)
    out = ApplyExtinction.apply(x, _t(av), axav)
    ref = R.apply_spaxel_extinction(spec.astype(np.float64), wave, av, "Cardelli89", 3.1)
    assert rel_err(out.detach().cpu().numpy(), ref) <= 1e-5
    g = rng.normal(size=spec.shape).astype(np.float32)
    out.backward(_t(g))
    fac = ref / spec.astype(np.float64)
    assert rel_err(x.grad.cpu().numpy(), g * fac) <= 1e-5


def test_reference_mock_case():
    # tests/test_dust_extinction.py:37-98 through the C ABI
    gc = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], np.float32)
    gp = np.array([0, 1], np.int32)
    metals = np.array([[0.01, 0.02, 0.03, 0.04, 0.05], [0.06, 0.07, 0.08, 0.09, 0.1]],
                      np.float32)
    spec = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    wave = np.array([5000.0, 6000.0], np.float32)
    for model in ("Cardelli89", "Gordon23"):
        ext = ops.gas_extinction(_t([1.0, 2.0]), _t(metals), "power law slope free", "MW", 3.0,
                                 1.0)
        av = ops.column_av(_t(gc), _t(gp, torch.int32), ext, _t(gc), _t(gp, torch.int32), 2, 2)
        out = ops.apply_extinction(_t(spec), av, ops.extinction_curve(model, 3.1, _t(wave)))
        ext_r = R.gas_extinction(np.array([1.0, 2.0]), metals.astype(np.float64),
                                 "power law slope free", "MW", 3.0, 1.0)
        av_r = R.column_av(gc[:, 2], gp, ext_r, gc[:, 2], gp, 2)
        ref = R.apply_spaxel_extinction(spec, wave, av_r, model, 3.1)
        got = out.cpu().numpy()
        assert got.shape == (2, 2) and np.all(got >= 0)
        np.testing.assert_allclose(got, ref, rtol=1e-5)


def _dusty_data(w, n_gas, seed):
    st = w.stars
    return stars_from_arrays(st["coords"], st["velocity"], st["mass"], st["age"],
                             st["metallicity"], gas=S.make_gas(n_gas, seed=seed))


def test_extinction_step_matches_oracle():
    """spaxel_assignment -> ... -> doppler_shift_and_resampling ->
    calculate_extinction -> calculate_datacube against the oracle's steps."""
    w = cfg1(3000)
    cfg = S.step_config(w)
    cfg["data"]["args"]["particle_type"] = ["stars", "gas"]
    rd = _dusty_data(w, 4000, 11)
    gas_np = S.make_gas(4000, seed=11)
    for f in (get_spaxel_assignment, get_reshape_data, get_calculate_spectra,
              get_scale_spectrum_by_mass, get_doppler_shift_and_resampling, get_extinction,
              get_calculate_datacube):
        rd = f(cfg)(rd)
    cube = rd.stars.datacube.detach().cpu().numpy()

    st, geom, grid = w.stars, w.geom, w.grid
    edges = geom["edges"]
    nb = len(edges) - 1
    ids = R.square_spaxel_assignment(st["coords"], edges)
    gids = R.square_spaxel_assignment(gas_np["coords"], edges)
    spec = R.lookup_linear(st["metallicity"], st["age"], grid["ax_met"], grid["ax_age"],
                           grid["flux"], log_axes=grid.get("log_axes", False), dtype=np.float64)
    spec = R.scale_spectrum_by_mass(spec, st["mass"].astype(np.float64))
    ssp_wave = R.cosmological_doppler_shift(geom["z"], grid["wave"], dtype=np.float32)
    wk = R.velocity_doppler_shift(ssp_wave, st["velocity"], "z", dtype=np.float32)
    res = R.resample_spectra(spec, wk.astype(np.float64), geom["wave_target"].astype(np.float64))
    area = spaxel_area(cfg)
    ext = R.gas_extinction(gas_np["mass"].astype(np.float64), gas_np["metals"].astype(np.float64),
                           "broken power law fit", "Z", 3.5, area)
    av = R.column_av(gas_np["coords"][:, 2], gids, ext.astype(np.float32), st["coords"][:, 2],
                     ids, nb * nb)
    print(f"A_V: median {np.median(av):.3f}, max {av.max():.3f}")
    assert av.max() > 0.05  # the test exercises a visible attenuation
    res = R.apply_spaxel_extinction(res, geom["wave_target"], av.astype(np.float64), "Cardelli89",
                                    3.1)
    ref = R.calculate_cube(res, ids, nb)
    e = rel_err(cube, ref)
    print(f"dusty cube rel64={e:.3e}")
    assert e <= 1e-5


def test_calc_dusty_ifu_chain_runs_and_differentiates():
    w = S.config1(n_stars=2000, seed=7)
    cfg = S.step_config(w)
    cfg["data"]["args"]["particle_type"] = ["stars", "gas"]
    st = w.stars
    rd = stars_from_arrays(st["coords"], st["velocity"], st["mass"], st["age"], st["metallicity"],
                           requires_grad=True, halfmassrad_stars=4.0,
                           gas=S.make_gas(3000, seed=8))
    p = RubixPipeline(cfg, "calc_dusty_ifu")
    leaves = (rd.stars.mass, rd.stars.age)
    p.gradient(rd, lambda c: (c * c).sum())
    for x in leaves:
        assert x.grad is not None and torch.isfinite(x.grad).all()
    assert float(leaves[0].grad.abs().sum()) > 0


@pytest.mark.parametrize("gas_scale", [1.0, 400.0])
def test_fused_dusty_chain_equals_per_step_chain(gas_scale):
    """calc_dusty_ifu_fused (extinction inside the fused kernels) against
    calc_dusty_ifu (spectra materialised, then attenuated): cube and every
    gradient.  The fused kernels form 10^y as 2^(y log2 10) with the
    hardware exp2, the per-step kernel with exp10f; gas masses scaled by 400
    raise A_V 400-fold (A_V is linear in the gas column), so the exponents
    reach the size where the rounding of y log2 10 is largest."""
    w = S.config1(n_stars=3000, seed=9)
    cfg = S.step_config(w)
    cfg["data"]["args"]["particle_type"] = ["stars", "gas"]
    st = w.stars
    gas = S.make_gas(5000, seed=10)
    gas["mass"] = gas["mass"] * np.float32(gas_scale)
    cubes, grads = [], []
    for chain in ("calc_dusty_ifu", "calc_dusty_ifu_fused"):
        rd = stars_from_arrays(st["coords"], st["velocity"], st["mass"], st["age"],
                               st["metallicity"], requires_grad=True, halfmassrad_stars=4.0,
                               gas=gas)
        leaves = (rd.stars.mass, rd.stars.age, rd.stars.metallicity, rd.stars.velocity)
        out = RubixPipeline(cfg, chain).run(rd)
        cube = out.stars.datacube
        ((cube * cube).sum()).backward()
        cubes.append(cube.detach().cpu().numpy())
        grads.append([x.grad.cpu().numpy() for x in leaves])
    e = rel_err(cubes[1], cubes[0])
    print(f"gas x{gas_scale}: fused vs per-step dusty cube: {e:.3e}, "
          f"total flux {float(np.sum(cubes[0], dtype=np.float64)):.4g}")
    assert e <= 1e-5
    for name, a, b in zip(("mass", "age", "metallicity", "velocity"), grads[1], grads[0]):
        eg = rel_err(a, b)
        print(f"  d/d{name}: {eg:.3e}")
        assert eg <= 1e-5, name


def test_fused_extinction_with_zero_av_is_the_plain_cube():
    w = cfg1(2000)
    st = w.stars
    ssp = ops.PreparedSSP(w.grid["flux"], w.grid["ax_met"], w.grid["ax_age"], w.grid["wave"],
                          w.geom["z"], w.grid.get("log_axes", False), DEV)
    tg = ops.TargetGrid(w.geom["wave_target"], DEV)
    nb = len(w.geom["edges"]) - 1
    ids = ops.spaxel_assign(_t(st["coords"]), _t(w.geom["edges"]))
    args = (_t(st["mass"]), _t(st["age"]), _t(st["metallicity"]), _t(st["velocity"]), ids,
            nb * nb, nb * nb)
    plain, _ = ops.cube_forward(ssp, tg, *args)
    axav = ops.extinction_curve("Gordon23", 3.1, _t(w.geom["wave_target"]))
    zero, _ = ops.cube_forward(ssp, tg, *args,
                               extinction=(torch.zeros(len(st["mass"]), device=DEV), axav))
    assert torch.equal(plain, zero)
