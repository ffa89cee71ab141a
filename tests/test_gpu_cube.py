"""GPU parity: fused datacube, PSF and LSF against the oracle.

Tolerance (float32 path): norm-wise relative error <= 1e-5 against both the
float64 oracle and the float32 restatement of the reference.  Elementwise
the float32 reference itself differs from exact arithmetic by up to ~1e-5 of
the cube maximum (knot rounding w = fl(wave * a) moves interpolation
weights), so the elementwise check is max|diff| / max|ref| <= 5e-5.
"""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops
from rubix_amd.autograd import render
from tests.helpers import cfg1, cfg2_sample, max_rel, oracle, rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-5


def _device_inputs(w):
    t = {k: torch.as_tensor(v, device=DEV) for k, v in w.stars.items()}
    g = w.grid
    ssp = ops.PreparedSSP(g["flux"], g["ax_met"], g["ax_age"], g["wave"], w.geom["z"],
                          g["log_axes"], DEV)
    tgt = ops.TargetGrid(w.geom["wave_target"], DEV)
    edges = torch.as_tensor(w.geom["edges"], device=DEV)
    return t, ssp, tgt, edges


def _render(w, t=None):
    if t is None:
        t, ssp, tgt, edges = _device_inputs(w)
    g = w.geom
    with torch.no_grad():
        ids, raw, c1, c2 = render(t["coords"], t["velocity"], t["mass"], t["age"], t["metallicity"],
                                  edges, ssp, tgt, g["nb"], torch.tensor(g["psf_sigma"]),
                                  g["psf_size"], torch.tensor(g["lsf_sigma"]), g["wave_res"],
                                  return_all=True)
    return ids.cpu().numpy(), raw.cpu().numpy(), c1.cpu().numpy(), c2.cpu().numpy()


@pytest.mark.parametrize("name,n", [("cfg1", 10_000), ("cfg2", 2_000)])
def test_fused_cube_matches_oracle(name, n):
    w = cfg1(n) if name == "cfg1" else cfg2_sample(n)
    ids, raw, c1, c2 = _render(w)
    o64 = oracle(name, "float64", n)
    o32 = oracle(name, "float32", n)
    ox = oracle(name, "float64", n, ref_knots=False)
    np.testing.assert_array_equal(ids, o64["ids"])
    report = {}
    for key, got in (("cube", raw), ("cube_psf", c1), ("cube_lsf", c2)):
        e64 = rel_err(got, o64[key])
        e32 = rel_err(got, o32[key])
        m64 = max_rel(got, o64[key])
        report[key] = (e64, e32, m64)
        print(f"{name} {key}: rel64={e64:.3e} rel32={e32:.3e} maxrel64={m64:.3e} "
              f"exact-arithmetic rel={rel_err(got, ox[key]):.3e} "
              f"(oracle32 vs 64: {rel_err(o32[key], o64[key]):.3e})")
    for key, (e64, e32, m64) in report.items():
        assert e64 <= TOL, (key, e64)
        assert e32 <= TOL, (key, e32)
        assert m64 <= 5e-5, (key, m64)


@pytest.mark.parametrize("name", ["cfg1", "cfg2"])
def test_deterministic_bitwise(name):
    """Two runs give identical bits.  cfg2's 24 MB template grid sorts the
    stars into 8 age bands (chunks per band and spaxel); cfg1's fits one."""
    w = cfg1(10_000) if name == "cfg1" else cfg2_sample(3_000)
    a = _render(w)
    b = _render(w)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("chunk", [16, 65, 1024])
def test_chunk_lengths_match_oracle(chunk):
    """cfg1 in full (10k stars, up to 211 in one spaxel) with chunks of 16, 65
    and 1024 stars: one-star remainders, odd star counts for the fused
    kernel's pairs of stars, and chunks spanning several 64-star batches of
    staged records.  Raw cube and the per-star gradients of <G, cube>
    against the float64 oracle."""
    w = cfg1(10_000)
    t, ssp, tgt, edges = _device_inputs(w)
    nb = w.geom["nb"]
    ids = ops.spaxel_assign(t["coords"], edges)
    layout = ops.BatchLayout([w.n_stars], [nb], DEV)
    tune = ops.Tuning(chunk=chunk)
    cube, ctx = ops.batch_forward(ssp, tgt, layout, t["mass"], t["age"], t["metallicity"],
                                  t["velocity"], ids, tune=tune)
    o = oracle("cfg1", "float64", 10_000)
    e = rel_err(cube.cpu().numpy().reshape(o["cube"].shape), o["cube"])
    print(f"cfg1 chunk {chunk}: raw cube rel64={e:.3e}")
    assert e <= TOL
    ref = ops.batch_forward(ssp, tgt, layout, t["mass"], t["age"], t["metallicity"],
                            t["velocity"], ids)[0]
    gen = torch.Generator(device=DEV).manual_seed(2)
    G = torch.randn(cube.shape, generator=gen, device=DEV)
    got = ops.batch_backward(ssp, tgt, layout, ctx, G, tune=tune)
    _, ctx0 = ops.batch_forward(ssp, tgt, layout, t["mass"], t["age"], t["metallicity"],
                                t["velocity"], ids)
    want = ops.batch_backward(ssp, tgt, layout, ctx0, G)
    # the chunking changes only the order of the cube's f64 partial sums; the
    # per-star gradients do not depend on it
    for a, b in zip(got, want):
        assert torch.equal(a, b)
    assert rel_err(cube.cpu().numpy(), ref.cpu().numpy()) <= 1e-6


def test_edge_stars():
    """Out-of-grid stars (age 0 -> log10 = -inf, like filtered particles),
    zero mass, stars outside the field of view (clipped into the border
    spaxels) and an empty star set."""
    w = cfg1(2_000)
    st = {k: v.copy() for k, v in w.stars.items()}
    st["age"][:50] = 0.0
    st["metallicity"][50:100] = 0.0
    st["mass"][100:150] = 0.0
    st["metallicity"][150:200] = 1.0   # above the grid
    st["coords"][200:260, :2] = 1e3   # far outside the FOV
    import copy
    w2 = copy.copy(w)
    w2.stars = st
    ids, raw, c1, c2 = _render(w2)
    o = R.plain_forward(st, w.grid, w.geom, dtype=np.float64, return_intermediates=True)
    assert np.isfinite(raw).all()
    assert rel_err(raw, o["cube"]) <= TOL
    assert rel_err(c2, o["cube_lsf"]) <= TOL


def test_empty_star_set():
    w = cfg1(1000)
    t, ssp, tgt, edges = _device_inputs(w)
    t = {k: v[:0] for k, v in t.items()}
    g = w.geom
    with torch.no_grad():
        out = render(t["coords"], t["velocity"], t["mass"], t["age"], t["metallicity"], edges, ssp,
                     tgt, g["nb"], torch.tensor(g["psf_sigma"]), g["psf_size"],
                     torch.tensor(g["lsf_sigma"]), g["wave_res"])
    assert out.shape == (g["nb"], g["nb"], tgt.n)
    assert float(out.abs().max()) == 0.0


def test_dropped_ids_nb_plus_two():
    """Edges with nb+2 entries: ids are strided by len(edges)-1 and ids >= nb^2
    are dropped by the reduction (rubix/spectra/ifu.py:286)."""
    w = cfg1(3000)
    t, ssp, tgt, _ = _device_inputs(w)
    L = float(w.geom["edges"][-1] - w.geom["edges"][0])
    nb = w.geom["nb"]
    edges = np.linspace(-L / 2, L / 2 + L / nb, nb + 2).astype(np.float32)
    ids = ops.spaxel_assign(t["coords"], torch.as_tensor(edges, device=DEV))
    nbins = len(edges) - 1
    with torch.no_grad():
        cube, _ = ops.cube_forward(ssp, tgt, t["mass"], t["age"], t["metallicity"], t["velocity"],
                                   ids, nb * nb, nbins * nbins)
    o_ids = R.square_spaxel_assignment(w.stars["coords"], edges)
    np.testing.assert_array_equal(ids.cpu().numpy(), o_ids)
    geom = dict(w.geom)
    geom["edges"] = w.geom["edges"]
    # oracle: same chain with the quirky ids
    from oracle.ref_numpy import (calculate_cube, cosmological_doppler_shift, lookup_linear,
                                  resample_spectra, scale_spectrum_by_mass, velocity_doppler_shift)
    spec = lookup_linear(w.stars["metallicity"], w.stars["age"], w.grid["ax_met"], w.grid["ax_age"],
                         w.grid["flux"], True, np.float64)
    spec = scale_spectrum_by_mass(spec, w.stars["mass"].astype(np.float64))
    sw = cosmological_doppler_shift(w.geom["z"], w.grid["wave"], np.float64)
    ww = velocity_doppler_shift(sw, w.stars["velocity"], "z", dtype=np.float64)
    res = resample_spectra(spec, ww, w.geom["wave_target"].astype(np.float64))
    ref = calculate_cube(res, o_ids, nb).reshape(nb * nb, -1)
    assert rel_err(cube.cpu().numpy(), ref) <= TOL


def test_descending_targets_use_gather_path():
    """The kernels stage each (star, slab) knot window in LDS only for
    ascending targets; a descending target axis takes the per-bin gather
    path.  (The reference's calculate_diff then gives negative steps, so Nn
    and the cube change sign: the oracle with the same reversed axis is the
    reference.)  Cube and adjoint against the float64 oracle."""
    import dataclasses

    from oracle import ref_torch as RT
    w0 = cfg1(2000)
    geom = dict(w0.geom, wave_target=w0.geom["wave_target"][::-1].copy())
    w = dataclasses.replace(w0, geom=geom)
    t, ssp, tgt, edges = _device_inputs(w)
    assert not tgt.nondecreasing
    ids = ops.spaxel_assign(t["coords"], edges)
    nb = geom["nb"]
    cube, ctx = ops.cube_forward(ssp, tgt, t["mass"], t["age"], t["metallicity"], t["velocity"],
                                 ids, nb * nb, nb * nb)
    st = w.stars
    f64 = torch.float64
    leaves = [torch.tensor(st[k], dtype=f64, requires_grad=True) for k in ("mass", "age", "metallicity")]
    vl = torch.tensor(st["velocity"][:, 2], dtype=f64, requires_grad=True)
    ids64 = torch.tensor(R.square_spaxel_assignment(st["coords"], geom["edges"]).astype(np.int64))
    ref = RT.star_cube(*leaves, vl, ids64, w.grid, geom, f64, True)
    e = rel_err(cube.cpu().numpy().reshape(ref.shape), ref.detach().numpy())
    print(f"descending targets: cube rel64={e:.3e}")
    assert e <= TOL
    gen = torch.Generator(device=DEV).manual_seed(5)
    dc = torch.randn(cube.shape, generator=gen, device=DEV)
    got = ops.cube_backward(ssp, tgt, ctx, dc)
    ref.backward(dc.cpu().double().reshape(ref.shape))
    for name, a, b in zip(("mass", "age", "metallicity", "v_los"), got, leaves + [vl]):
        e = rel_err(a.cpu().numpy(), b.grad.numpy())
        print(f"descending targets: d/d{name} rel64={e:.3e}")
        assert e <= TOL, name


def test_coarse_locator_runs_the_correction_paths(monkeypatch):
    """A knot locator with cells wider than the knot spacing (forced here by a
    small cell budget; a grid needing more than 2^23 cells takes it) holds
    ~4 knots per cell, so the guesses are off by up to a few knots.  The
    forward then takes its exact window searches, and the adjoint's short
    path corrects most segments inside its 64-knot windows (reading the
    segment's row values again after the correction) and sends the groups
    whose segments leave the window to the general path.  Cube and all four
    gradients against the float64 oracle, on the cfg2 grid (5000 knots) and
    target axis (3721 ascending bins)."""
    from oracle import ref_torch as RT
    w = cfg2_sample(2000)
    monkeypatch.setattr(ops.PreparedSSP, "MAX_LOCATE", 1200)
    t, ssp, tgt, edges = _device_inputs(w)
    assert ssp.locate_multi and ssp.n_locate < w.grid["wave"].size / 3
    ids = ops.spaxel_assign(t["coords"], edges)
    geom = w.geom
    nb = geom["nb"]
    cube, ctx = ops.cube_forward(ssp, tgt, t["mass"], t["age"], t["metallicity"], t["velocity"],
                                 ids, nb * nb, nb * nb)
    st = w.stars
    f64 = torch.float64
    leaves = [torch.tensor(st[k], dtype=f64, requires_grad=True) for k in ("mass", "age", "metallicity")]
    vl = torch.tensor(st["velocity"][:, 2], dtype=f64, requires_grad=True)
    ids64 = torch.tensor(R.square_spaxel_assignment(st["coords"], geom["edges"]).astype(np.int64))
    ref = RT.star_cube(*leaves, vl, ids64, w.grid, geom, f64, True)
    e = rel_err(cube.cpu().numpy().reshape(ref.shape), ref.detach().numpy())
    print(f"coarse locator: cube rel64={e:.3e}")
    assert e <= TOL
    gen = torch.Generator(device=DEV).manual_seed(11)
    dc = torch.randn(cube.shape, generator=gen, device=DEV)
    got = ops.cube_backward(ssp, tgt, ctx, dc)
    ref.backward(dc.cpu().double().reshape(ref.shape))
    for name, a, b in zip(("mass", "age", "metallicity", "v_los"), got, leaves + [vl]):
        e = rel_err(a.cpu().numpy(), b.grad.numpy())
        print(f"coarse locator: d/d{name} rel64={e:.3e}")
        assert e <= TOL, name


def test_dense_template_grid_uses_gather_path():
    """20000 template knots (0.3 A apart): every slab window exceeds the LDS
    window, so the forward gathers per bin; checked against the oracle."""
    from rubix_amd import synthetic as S
    w = cfg1(1500)
    grid = S.make_ssp_grid(6, 20, 20000, (3000.0, 9000.0), seed=1)
    t, _, tgt, edges = _device_inputs(w)
    ssp = ops.PreparedSSP(grid["flux"], grid["ax_met"], grid["ax_age"], grid["wave"], w.geom["z"],
                          grid["log_axes"], DEV)
    ids = ops.spaxel_assign(t["coords"], edges)
    nb = w.geom["nb"]
    cube, _ = ops.cube_forward(ssp, tgt, t["mass"], t["age"], t["metallicity"], t["velocity"], ids,
                               nb * nb, nb * nb)
    ref = R.plain_forward(w.stars, grid, w.geom, dtype=np.float64, return_intermediates=True)["cube"]
    e = rel_err(cube.cpu().numpy().reshape(ref.shape), ref)
    print(f"dense grid cube rel64={e:.3e}")
    assert e <= TOL


def test_age_bands_change_only_the_summation_order():
    """cfg2's 24 MB template grid sorts the stars by (age band, spaxel) and
    reduces each spaxel's chunks band by band when the key count allows
    (n_keys <= 8192); with n_keys = 8193 the same stars take one band.
    The per-star work is the same, so the gradients are bitwise equal, and
    the cubes can differ only by the order of the f64 chunk additions.  The
    terms are float32 values, whose f64 sums are exact while they span less
    than about 2^29 in magnitude, so at 100 000 stars (~40 per spaxel) the
    cubes agree bit for bit as well (the check allows rounding)."""
    w = cfg2_sample(100_000)
    t, ssp, tgt, edges = _device_inputs(w)
    ids = ops.spaxel_assign(t["coords"], edges)
    nb = w.geom["nb"]
    out = {}
    for n_keys in (nb * nb, 8193):
        cube, ctx = ops.cube_forward(ssp, tgt, t["mass"], t["age"], t["metallicity"],
                                     t["velocity"], ids, nb * nb, n_keys)
        gen = torch.Generator(device=DEV).manual_seed(3)
        dc = torch.randn(cube.shape, generator=gen, device=DEV)
        out[n_keys] = (cube.cpu().numpy(), [g.cpu().numpy() for g in ops.cube_backward(ssp, tgt, ctx, dc)])
    (c8, g8), (c1, g1) = out[nb * nb], out[8193]
    e = rel_err(c8, c1)
    print(f"8 bands vs 1 band: cube rel={e:.3e}, max|diff|/max={max_rel(c8, c1):.3e}")
    assert e <= 1e-7 and max_rel(c8, c1) <= 1e-6
    for a, b in zip(g8, g1):
        np.testing.assert_array_equal(a, b)
