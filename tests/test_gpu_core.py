"""GPU: the drop-in step API end to end.  The per-step chain and the fused
chain run through RubixPipeline on RubixData and are checked against the
float64 oracle: from spaxel assignment (calc_ifu*_aligned) and from raw
particles through rotate_galaxy and filter_particles (calc_ifu,
calc_ifu_fused).  Every step writes the reference's fields with the
reference's shapes."""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import synthetic as S
from rubix_amd.core.data import get_reshape_data, stars_from_arrays
from rubix_amd.core.ifu import (get_calculate_datacube, get_calculate_spectra,
                                get_doppler_shift_and_resampling, get_scale_spectrum_by_mass)
from rubix_amd.core.pipeline import RubixPipeline
from rubix_amd.core.telescope import get_spaxel_assignment
from tests.helpers import cfg1, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _data(w, requires_grad=False):
    st = w.stars
    return stars_from_arrays(st["coords"], st["velocity"], st["mass"], st["age"], st["metallicity"],
                             requires_grad=requires_grad)


@pytest.mark.parametrize("chain", ["calc_ifu_aligned", "calc_ifu_fused_aligned"])
def test_pipeline_matches_oracle(chain):
    w = cfg1(4000)
    cfg = S.step_config(w)
    out = RubixPipeline(cfg, chain).run(_data(w))
    ref = R.plain_forward(w.stars, w.grid, w.geom, dtype=np.float64)
    cube = out.stars.datacube
    assert tuple(cube.shape) == (25, 25, 1000)
    assert rel_err(cube.detach().cpu().numpy(), ref) <= TOL
    if chain == "calc_ifu_fused_aligned":
        assert out.stars.spectra is None


@pytest.mark.parametrize("chain,rotation", [("calc_ifu", {"type": "face-on"}),
                                            ("calc_ifu_fused", {"type": "edge-on"}),
                                            ("calc_ifu_fused", {"alpha": 45.0, "beta": 30.0,
                                                                "gamma": 15.0})])
def test_full_chain_from_raw_particles(chain, rotation):
    """rotate_galaxy -> filter_particles -> ... -> convolve_lsf -> apply_noise
    against the oracle's alignment, aperture filter, plain forward and noise
    (float64)."""
    w = S.config1(n_stars=4000, seed=7)
    cfg = S.step_config(w)
    cfg["galaxy"]["rotation"] = rotation
    r_half = 4.0
    st = w.stars
    rd = stars_from_arrays(st["coords"], st["velocity"], st["mass"], st["age"], st["metallicity"],
                           halfmassrad_stars=r_half)
    out = RubixPipeline(cfg, chain).run(rd)
    if "type" in rotation:
        ang = (0.0, 0.0, 0.0) if rotation["type"] == "face-on" else (90.0, 0.0, 0.0)
    else:
        ang = (rotation["alpha"], rotation["beta"], rotation["gamma"])
    p, v = R.rotate_galaxy(st["coords"], st["velocity"], st["mass"], r_half, *ang)
    fields, mask = R.filter_particles(p.astype(np.float32), w.geom["edges"],
                                      {k: st[k] for k in ("mass", "age", "metallicity")})
    ref_stars = dict(coords=p.astype(np.float32), velocity=v.astype(np.float32), **fields)
    ref = R.plain_forward(ref_stars, w.grid, w.geom, dtype=np.float64)
    ref = R.apply_noise(ref, 1.0, "normal", 0, dtype=np.float64)
    got_p = out.stars.coords.detach().cpu().numpy().reshape(-1, 3)
    assert np.abs(got_p - p).max() <= 1e-5 * np.abs(p).max()
    np.testing.assert_array_equal(out.stars.mask.cpu().numpy().reshape(-1), mask)
    e = rel_err(out.stars.datacube.detach().cpu().numpy(), ref)
    print(f"{chain} {rotation}: rel64={e:.3e}, {int(mask.sum())} of {len(mask)} inside")
    assert e <= TOL


def test_steps_fields_and_shapes():
    w = cfg1(1000)
    cfg = S.step_config(w)
    rd = _data(w)
    rd = get_spaxel_assignment(cfg)(rd)
    assert rd.stars.pixel_assignment.dtype == torch.int32 and rd.stars.pixel_assignment.shape == (1000,)
    np.testing.assert_array_equal(rd.stars.pixel_assignment.cpu().numpy(),
                                  R.square_spaxel_assignment(w.stars["coords"], w.geom["edges"]))
    rd = get_reshape_data(cfg)(rd)
    assert rd.stars.mass.shape == (1, 1000) and rd.stars.velocity.shape == (1, 1000, 3)
    rd = get_calculate_spectra(cfg)(rd)
    assert rd.stars.spectra.shape == (1, 1000, 1000)
    spec = R.lookup_linear(w.stars["metallicity"], w.stars["age"], w.grid["ax_met"], w.grid["ax_age"],
                           w.grid["flux"], True, np.float64)
    assert rel_err(rd.stars.spectra[0].cpu().numpy(), spec) <= TOL
    rd = get_scale_spectrum_by_mass(cfg)(rd)
    rd = get_doppler_shift_and_resampling(cfg)(rd)
    assert rd.stars.spectra.shape == (1, 1000, 1000)
    rd = get_calculate_datacube(cfg)(rd)
    assert rd.stars.datacube.shape == (25, 25, 1000)


def test_pipeline_gradient_fused_equals_plain():
    w = cfg1(500)
    cfg = S.step_config(w)
    grads = []
    for chain in ("calc_ifu_aligned", "calc_ifu_fused_aligned"):
        rd = _data(w, requires_grad=True)
        leaves = (rd.stars.mass, rd.stars.age, rd.stars.metallicity, rd.stars.velocity)
        p = RubixPipeline(cfg, chain)
        p.gradient(rd, lambda c: (c * c).sum())
        grads.append([x.grad.cpu().numpy() for x in leaves])
    for a, b in zip(*grads):
        assert rel_err(a, b) <= TOL


@pytest.mark.parametrize("chain", ["calc_ifu", "calc_ifu_fused"])
def test_chain_from_rubix_galaxy_file(tmp_path, chain):
    """RubixPipeline(config).run() with no data reads
    config["output_path"]/rubix_galaxy.h5, as the reference's pipeline does
    (get_rubix_data -> prepare_input: centring on the galaxy center, median
    velocity within 10 kpc), then rotate_galaxy (face-on, the file's
    half-mass radius) ... apply_noise.  The file is the h5py-written fixture
    tests/golden/rubix_galaxy_small.h5 (300 stars); the oracle starts from
    its arrays restated in numpy."""
    import os
    import shutil
    golden = os.path.join(os.path.dirname(__file__), "golden")
    shutil.copy(os.path.join(golden, "rubix_galaxy_small.h5"), tmp_path / "rubix_galaxy.h5")
    exp = np.load(os.path.join(golden, "hdf5_expected.npz"))
    w = S.config1(n_stars=10, seed=7)  # geometry and SSP grid only
    cfg = S.step_config(w)
    cfg["output_path"] = str(tmp_path)
    out = RubixPipeline(cfg, chain).run()
    center = exp["galaxy__center"].astype(np.float32)
    pos, vel = exp["stars__coords"], exp["stars__velocity"]
    vel = vel - np.median(vel[np.linalg.norm(pos - center, axis=1) < 10], axis=0)
    pos = pos - center
    st = {k: exp[f"stars__{k}"] for k in ("mass", "age", "metallicity")}
    p, v = R.rotate_galaxy(pos, vel, st["mass"], 3.75, 0.0, 0.0, 0.0)
    fields, mask = R.filter_particles(p.astype(np.float32), w.geom["edges"], st)
    ref_stars = dict(coords=p.astype(np.float32), velocity=v.astype(np.float32), **fields)
    ref = R.plain_forward(ref_stars, w.grid, w.geom, dtype=np.float64)
    ref = R.apply_noise(ref, 1.0, "normal", 0, dtype=np.float64)
    np.testing.assert_array_equal(out.stars.mask.cpu().numpy().reshape(-1), mask)
    e = rel_err(out.stars.datacube.detach().cpu().numpy(), ref)
    print(f"{chain} from rubix_galaxy.h5: rel64={e:.3e}, {int(mask.sum())} of {len(mask)} inside")
    assert e <= TOL
