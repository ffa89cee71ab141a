"""GPU: batched ragged mode (many galaxies with uneven star counts and spaxel
grids in one fused pass, optionally replayed from a HIP graph).

* The batch cube rows of each galaxy are bitwise the single-galaxy fused
  cube, and the per-star gradients bitwise the single-galaxy adjoint (the
  per-star work does not see the galaxy; chunks never cross a spaxel).
* After PSF + LSF each galaxy matches the float64 oracle (<= 1e-5 norm-wise).
* Launch tunings change nothing but the grouping of the cube's chunk sums.
* Every replay of the captured forward + adjoint equals the eager step.
"""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops, synthetic as S
from rubix_amd.autograd import render
from rubix_amd.batch import GalaxyBatch, GraphedStep
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-5
WT = (3400.0 + 5.0 * np.arange(1000)).astype(np.float32)


def _batch(counts=(3000, 0, 5000, 1, 800), nbs=(25, 10, 40, 25, 7)):
    grid = S.make_ssp_grid(6, 20, 1000, (3000.0, 9000.0), seed=1)
    stars, geoms = [], []
    for g, (n, nb) in enumerate(zip(counts, nbs)):
        stars.append(S.make_stars(n, seed=50 + g))
        geoms.append(S.make_geometry(nb, 0.5, 0.05, WT, 0.6, 11, 1.2, 5.0))
    return grid, stars, geoms


def _prepared(grid, geom):
    ssp = ops.PreparedSSP(grid["flux"], grid["ax_met"], grid["ax_age"], grid["wave"], geom["z"],
                          grid["log_axes"], DEV)
    return ssp, ops.TargetGrid(geom["wave_target"], DEV)


def _gb(stars, geoms, requires_grad=True):
    return GalaxyBatch(stars, [g["nb"] for g in geoms], [g["edges"] for g in geoms], DEV,
                       requires_grad=requires_grad)


def test_batch_cube_and_adjoint_bitwise_equal_single_galaxy():
    grid, stars, geoms = _batch()
    ssp, tg = _prepared(grid, geoms[0])
    gb = _gb(stars, geoms, requires_grad=False)
    L, t = gb.layout, gb.t
    ids = torch.cat([ops.spaxel_assign(t["coords"][L.stars_of(g)], gb.edges[g])
                     for g in range(L.n_gal)])
    cube, ctx = ops.batch_forward(ssp, tg, L, t["mass"], t["age"], t["metallicity"],
                                  t["velocity"], ids)
    gen = torch.Generator(device=DEV).manual_seed(0)
    dcube = torch.randn(cube.shape, generator=gen, device=DEV)
    grads = ops.batch_backward(ssp, tg, L, ctx, dcube)
    for g in range(L.n_gal):
        s, r = L.stars_of(g), L.rows_of(g)
        nb = L.nb[g]
        c1, ctx1 = ops.cube_forward(ssp, tg, t["mass"][s], t["age"][s], t["metallicity"][s],
                                    t["velocity"][s], ids[s], nb * nb, nb * nb)
        assert torch.equal(cube[r], c1), g
        if L.n_stars[g]:
            g1 = ops.cube_backward(ssp, tg, ctx1, dcube[r].contiguous())
            for a, b in zip(grads, g1):
                assert torch.equal(a[s], b), g


def test_render_batched_matches_oracle_per_galaxy():
    grid, stars, geoms = _batch()
    ssp, tg = _prepared(grid, geoms[0])
    gb = _gb(stars, geoms, requires_grad=False)
    geo = geoms[0]
    sp = torch.tensor(geo["psf_sigma"], device=DEV)
    sl = torch.tensor(geo["lsf_sigma"], device=DEV)
    out = gb.render(ssp, tg, sp, geo["psf_size"], sl, geo["wave_res"]).cpu().numpy()
    for g, (st, geom) in enumerate(zip(stars, geoms)):
        nb = geom["nb"]
        got = out[gb.layout.rows_of(g)].reshape(nb, nb, -1)
        if len(st["mass"]) == 0:
            assert not got.any()
            continue
        ref = R.plain_forward(st, grid, geom, dtype=np.float64)
        e = rel_err(got, ref)
        print(f"galaxy {g} nb={nb} n={len(st['mass'])}: rel64={e:.3e}")
        assert e <= TOL


def test_batched_gradients_equal_per_galaxy_render():
    grid, stars, geoms = _batch()
    ssp, tg = _prepared(grid, geoms[0])
    geo = geoms[0]
    gb = _gb(stars, geoms)
    sp = torch.tensor(geo["psf_sigma"], device=DEV, requires_grad=True)
    sl = torch.tensor(geo["lsf_sigma"], device=DEV, requires_grad=True)
    out = gb.render(ssp, tg, sp, geo["psf_size"], sl, geo["wave_res"])
    gen = torch.Generator(device=DEV).manual_seed(1)
    cot = torch.randn(out.shape, generator=gen, device=DEV)
    out.backward(cot)
    L = gb.layout
    dsp = dsl = 0.0
    for g in range(L.n_gal):
        s = L.stars_of(g)
        if L.n_stars[g] == 0:
            continue
        leaf = {k: gb.t[k][s].detach().clone().requires_grad_(True)
                for k in ("mass", "age", "metallicity", "velocity")}
        sp1 = sp.detach().clone().requires_grad_(True)
        sl1 = sl.detach().clone().requires_grad_(True)
        nb = L.nb[g]
        c = render(gb.t["coords"][s], leaf["velocity"], leaf["mass"], leaf["age"],
                   leaf["metallicity"], gb.edges[g], ssp, tg, nb, sp1, geo["psf_size"], sl1,
                   geo["wave_res"])
        assert torch.equal(c.reshape(nb * nb, -1), out.detach()[L.rows_of(g)]), g
        c.backward(cot[L.rows_of(g)].reshape(c.shape))
        for k in leaf:
            assert torch.equal(leaf[k].grad, gb.t[k].grad[s]), (g, k)
        dsp += float(sp1.grad)
        dsl += float(sl1.grad)
    assert abs(float(sp.grad) - dsp) <= 1e-6 * abs(dsp)
    assert abs(float(sl.grad) - dsl) <= 1e-6 * abs(dsl)


@pytest.mark.parametrize("tune", [ops.Tuning(256, 1, 0), ops.Tuning(256, 2, 3),
                                  ops.Tuning(256, 8, 1), ops.Tuning(16, 4, 0),
                                  ops.Tuning(1024, 4, 2)])
def test_tunings_agree(tune):
    grid, stars, geoms = _batch()
    ssp, tg = _prepared(grid, geoms[0])
    gb = _gb(stars, geoms, requires_grad=False)
    L, t = gb.layout, gb.t
    ids = torch.cat([ops.spaxel_assign(t["coords"][L.stars_of(g)], gb.edges[g])
                     for g in range(L.n_gal)])
    args = (t["mass"], t["age"], t["metallicity"], t["velocity"], ids)
    c0, x0 = ops.batch_forward(ssp, tg, L, *args)
    c1, x1 = ops.batch_forward(ssp, tg, L, *args, tune=tune)
    if tune.chunk == 256:
        assert torch.equal(c0, c1)
    else:
        assert rel_err(c1.cpu().numpy(), c0.cpu().numpy()) <= 1e-7
    dcube = torch.ones_like(c0)
    g0 = ops.batch_backward(ssp, tg, L, x0, dcube)
    g1 = ops.batch_backward(ssp, tg, L, x1, dcube, tune=tune)
    for a, b in zip(g0, g1):
        assert torch.equal(a, b)


def test_graph_replays_equal_eager():
    grid, stars, geoms = _batch(counts=(2000, 700, 3000), nbs=(25, 10, 40))
    ssp, tg = _prepared(grid, geoms[0])
    geo = geoms[0]
    gb = _gb(stars, geoms)
    sp = torch.tensor(geo["psf_sigma"], device=DEV, requires_grad=True)
    sl = torch.tensor(geo["lsf_sigma"], device=DEV, requires_grad=True)

    def rend():
        return gb.render(ssp, tg, sp, geo["psf_size"], sl, geo["wave_res"])

    with torch.no_grad():
        tgt = rend() * 1.05
        w = 1.0 / torch.clamp(tgt.abs() / 30.0, min=1e-6 * float(tgt.abs().max())) ** 2
    step = GraphedStep(rend, gb.leaves + [sp, sl], tgt, w)
    cube_e, grads_e = step.eager()
    cube_e = cube_e.clone()
    grads_e = [g.clone() for g in grads_e]
    for rep in range(3):
        cube_g, grads_g = step.replay()
        torch.cuda.synchronize()
        assert torch.equal(cube_g, cube_e), rep
        for a, b in zip(grads_g, grads_e):
            assert torch.equal(a, b), rep
        step.eager()  # eager work between replays
        torch.cuda.synchronize()
