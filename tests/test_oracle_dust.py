"""CPU: the dust restatement (oracle/ref_numpy.py dust section) against the
reference's own known answers (tests/test_dust_classes.py,
tests/test_dust_extinction.py), published curve values, closed forms and the
literal per-spaxel column-density algorithm on hand-checkable cases.

CCM89 is pinned by published values of the Cardelli, Clayton & Mathis (1989)
curve at R(V) = 3.1 in every branch (FUV, far-NUV, NUV, optical, IR): the
table the dust_extinction package (K. Gordon et al.; the implementation the
reference follows, notebooks/dust_extinction.ipynb) checks its CCM89 model
against.  Every coefficient of the restatement, perturbed by 0.1 %, fails
that check.  G23 has no published table of curve values here: its
coefficient tables are checked against the numbers of the reference's source
(tests/golden/dust_coefficients.json, extracted from the file's syntax tree
by tests/golden/make_dust_coefficients.py), and its structure against the
normalisation it is built on (A(0.55 um)/A(V) = 1, E(0.44 - 0.55)/A(V) =
1/R(V)), the agreement of neighbouring branches across their overlaps and
the CCM89 curve.
"""
import copy
import json
import os

import numpy as np
import pytest

from oracle import ref_numpy as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# x [1/micron] and A(x)/A(V) of CCM89 at R(V) = 3.1 (published, 8 decimals)
CCM89_X = np.array([10.0, 9.0, 8.0, 7.0, 6.0, 5.0, 4.6, 4.0, 3.0, 2.78, 2.27, 1.82, 1.43, 1.11,
                    0.8, 0.63, 0.46])
CCM89_AXAV = np.array([5.23835484, 4.13406452, 3.33685933, 2.77962453, 2.52195399, 2.84252644,
                       3.18598916, 2.31531711, 1.64254927, 1.56880904, 1.32257836, 1.0,
                       0.75125994, 0.4780346, 0.28206957, 0.19200814, 0.11572348])


def test_ccm89_published_values():
    got = R.cardelli89(CCM89_X, 3.1)
    np.testing.assert_allclose(got, CCM89_AXAV, rtol=0, atol=6e-9)


_CCM89_TABLES = ["_CCM89_IR", "_CCM89_OPT_A", "_CCM89_OPT_B", "_CCM89_NUV_A", "_CCM89_NUV_B",
                 "_CCM89_FNUV_A", "_CCM89_FNUV_B", "_CCM89_FUV_A", "_CCM89_FUV_B"]


@pytest.mark.parametrize("table", _CCM89_TABLES)
def test_ccm89_every_coefficient_is_pinned(table, monkeypatch):
    """A 0.1 % change of any coefficient (the optical constants 1 and 0
    excepted: those are the normalisation a(1.82) = 1, b(1.82) = 0 that the
    x = 1.82 value pins) moves some published value by more than 1e-6."""
    base = list(getattr(R, table))
    for i, c in enumerate(base):
        if c in (0.0, 1.0) and table.startswith("_CCM89_OPT"):
            continue
        bad = copy.copy(base)
        bad[i] = c * 1.001
        monkeypatch.setattr(R, table, type(getattr(R, table))(bad))
        dev = np.abs(R.cardelli89(CCM89_X, 3.1) - CCM89_AXAV).max()
        assert dev > 1e-6, (table, i, dev)
    monkeypatch.setattr(R, table, type(getattr(R, table))(base))
    assert np.abs(R.cardelli89(CCM89_X, 3.1) - CCM89_AXAV).max() < 6e-9


def test_g23_tables_match_reference_source():
    with open(os.path.join(GOLDEN, "dust_coefficients.json")) as f:
        ref = json.load(f)
    assert R._G23_IR_A == ref["ir_a"]
    assert R._G23_OPT_A == ref["opt_a"]
    assert R._G23_OPT_B == ref["opt_b"]
    assert [list(R._G23_UV_A), list(R._G23_UV_B)] == ref["fm90_calls"]
    assert ref["powerlaw_calls"] == [{"amplitude": R._G23_IR_B[0], "x_0": R._G23_IR_B[1],
                                      "alpha": R._G23_IR_B[2]}]
    assert list(R._G23_BRANCHES["ir"]) == ref["ir_mask"]
    assert list(R._G23_BRANCHES["opt"]) == ref["opt_mask"]
    assert list(R._G23_BRANCHES["uv"]) == ref["uv_mask"]
    assert list(R._G23_BRANCHES["optir"]) == ref["optir_waves"]
    assert list(R._G23_BRANCHES["uvopt"]) == ref["uvopt_waves"]


def _g23_ab(w):
    a = R.gordon23(np.asarray(w, dtype=np.float64), 3.1)
    b = (R.gordon23(np.asarray(w, dtype=np.float64), 2.5) - a) / (1 / 2.5 - 1 / 3.1)
    return a, b


def test_g23_normalisation():
    """G23 is fit as A(lambda)/A(V) = a + b (1/R(V) - 1/3.1) with monochromatic
    V = 0.55 um and B = 0.44 um: a(V) = 1, b(V) = 0 and A(B)/A(V) - 1 =
    E(B - V)/A(V) = 1/R(V) for every R(V), i.e. a(B) = 1 + 1/3.1, b(B) = 1,
    to the accuracy of the fit."""
    (av,), (bv,) = _g23_ab([0.55])
    (ab,), (bb,) = _g23_ab([0.44])
    assert abs(av - 1) < 0.02 and abs(bv) < 0.01
    assert abs(ab - (1 + 1 / 3.1)) < 0.03 and abs(bb - 1) < 0.06
    for rv in (2.5, 3.1, 4.0, 5.5):
        ebv = R.gordon23(np.array([0.44]), rv)[0] - R.gordon23(np.array([0.55]), rv)[0]
        assert abs(ebv * rv - 1) < 0.06, (rv, ebv * rv)


def test_g23_branches_agree_across_overlaps():
    """Where two branches are blended (0.9-1.1 um and 0.3-0.33 um) each one
    alone is within a few per cent of the other: the blend joins two fits of
    the same curve."""
    w = np.linspace(0.9, 1.1, 21)
    op = R._g23_optical(1 / w, R._G23_OPT_A)
    ir = R._g23_nirmir(w, R._G23_IR_A)
    assert np.abs(op / ir - 1).max() < 0.05
    w = np.linspace(0.3, 0.33, 7)
    uv = R.fm90(1 / w, *R._G23_UV_A)
    op = R._g23_optical(1 / w, R._G23_OPT_A)
    assert np.abs(uv / op - 1).max() < 0.04


def test_g23_close_to_ccm89():
    """Both are Milky Way average curves: between 0.17 and 3.2 um they agree
    to 6 % in the UV/optical and drift to 16 % below CCM89 in the NIR."""
    w = np.linspace(0.17, 3.2, 200)
    r = R.gordon23(w, 3.1) / R.cardelli89(1 / w, 3.1)
    assert r.min() > 0.83 and r.max() < 1.05
    vis = (w > 0.3) & (w < 1.0)
    assert np.abs(r[vis] - 1).max() < 0.06


def test_generic_model_known_answers():
    # tests/test_dust_classes.py:14-80
    x = np.array([1.0, 2.0, 3.0])
    np.testing.assert_allclose(R.powerlaw1d(x, 2.0, 1.0, 1.0), [2.0, 1.0, 0.66666667], rtol=1e-6)
    np.testing.assert_allclose(R.polynomial1d(x, [1.0, 2.0, 3.0]), [6.0, 17.0, 34.0])
    np.testing.assert_allclose(R.polynomial1d(x, [1.0]), [1.0, 1.0, 1.0])
    np.testing.assert_allclose(R.drude1d(x, 1.0, 1.0, 1.0), [1.0, 0.30769232, 0.12328766],
                               rtol=1e-6)
    with pytest.raises(ValueError, match="0 is not an allowed value for x_0"):
        R.drude1d(x, 1.0, 0.0, 1.0)
    np.testing.assert_allclose(R.modified_drude(x, 1.0, 1.0, 1.0, 0.0),
                               [1.0, 0.30769232, 0.12328766], rtol=1e-6)
    np.testing.assert_allclose(R.fm90(np.array([4.0, 5.0, 6.0]), 0.10, 0.70, 3.23, 0.41, 4.59,
                                      0.95), [4.1879544, 5.723751, 4.7574277], rtol=1e-6)


@pytest.mark.parametrize("kw,msg", [
    (dict(C1=6.0, C2=0.5, C3=3.0, C4=0.5, xo=4.5, gamma=0.5), "C1 is out of bounds: 6.0"),
    (dict(C1=0.0, C2=-1.5, C3=3.0, C4=0.5, xo=4.5, gamma=0.5), "C2 is out of bounds: -1.5"),
    (dict(C1=0.0, C2=0.5, C3=7.0, C4=0.5, xo=4.5, gamma=0.5), "C3 is out of bounds: 7.0"),
    (dict(C1=0.0, C2=0.5, C3=3.0, C4=-1.5, xo=4.5, gamma=0.5), "C4 is out of bounds: -1.5"),
    (dict(C1=0.0, C2=0.5, C3=3.0, C4=0.5, xo=5.5, gamma=0.5), "xo is out of bounds: 5.5"),
    (dict(C1=0.0, C2=0.5, C3=3.0, C4=0.5, xo=4.5, gamma=0.1), "gamma is out of bounds: 0.1"),
])
def test_fm90_bounds(kw, msg):
    # tests/test_dust_classes.py:83-107
    with pytest.raises(ValueError, match=msg):
        R.fm90(np.array([4.0, 5.0, 6.0]), **kw)


def test_curves_reference_ranges_and_closed_forms():
    # tests/test_dust_classes.py:110-170: values in [0, 10]
    w = np.array([0.5, 1.0, 2.0, 3.0, 5.0, 8.0, 10.0])
    c = R.cardelli89(w, 3.1)
    assert c.shape == w.shape and np.all(c >= 0) and np.all(c <= 10)
    g = R.gordon23(np.array([0.1, 0.3, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0, 30.0]), 3.1)
    assert np.all(g >= 0) and np.all(g <= 10)
    # CCM89 at y = x - 1.82 = 0: a = 1, b = 0 -> A/A(V) = 1 for every Rv
    for rv in (2.0, 3.1, 5.0):
        assert abs(R.cardelli89(np.array([1.82]), rv)[0] - 1.0) < 1e-12
    # G23 is normalised to A(V): ~1 at 0.55 micron for every Rv
    for rv in (2.5, 3.1, 5.0):
        assert abs(R.gordon23(np.array([0.55]), rv)[0] - 1.0) < 0.02
    # IR branch of CCM89 at 1 micron^-1: a + b / Rv = (0.574 - 0.527 / Rv)
    assert abs(R.cardelli89(np.array([1.0]), 3.1)[0] - (0.574 - 0.527 / 3.1)) < 1e-12
    # extinguish with Ebv: Av = Rv * Ebv (tests/test_dust_classes.py:127-143)
    ax = R.cardelli89(w, 3.1)
    np.testing.assert_allclose(R.extinguish(ax, 3.1 * 1.0), np.power(10.0, -0.4 * ax * 3.1))


def test_extinction_model_choice():
    with pytest.raises(ValueError, match="Extinction model 'InvalidModel' is not available"):
        R.extinction_curve("InvalidModel", np.array([0.5]), 3.1)


@pytest.mark.parametrize("model", ["power law slope free", "broken power law fit"])
@pytest.mark.parametrize("xco", ["MW", "Z"])
def test_dust_to_gas(model, xco):
    # tests/test_dust_extinction.py:178-212: shape and sign
    r = R.dust_to_gas_ratio(np.array([8.0, 8.5]), model, xco)
    assert r.shape == (2,) and np.all(r >= 0)
    # closed form at solar abundance (x_sol = 8.69, above either transition)
    assert abs(R.dust_to_gas_ratio(np.array([8.69]), model, xco)[0] - 10 ** -2.21) < 1e-15


def test_dust_to_gas_broken_law_below_transition():
    r = R.dust_to_gas_ratio(np.array([7.5]), "broken power law fit", "Z")
    assert abs(r[0] - 1 / 10 ** (0.96 + 3.10 * (8.69 - 7.5))) < 1e-18
    with pytest.raises(NotImplementedError):
        R.dust_to_gas_ratio(np.array([8.0]), "power law slope fixed", "MW")


def test_gas_extinction_closed_form():
    metals = np.zeros((1, 10))
    metals[0, 0], metals[0, 4] = 0.7, 16 * 0.7 * 10 ** (8.69 - 12)  # log(O/H) + 12 = 8.69
    ext = R.gas_extinction(np.array([1.0e4]), metals, "power law slope free", "MW", 3.5, 0.25)
    conv = 1.989e33 / 3.08568e21 ** 2
    expect = 3 * np.pi * 1.0e4 * 10 ** -2.21 * conv / (0.4 * np.log(10) * 5448e-8 * 3.5) / 0.25
    assert abs(ext[0] / expect - 1) < 1e-12


def test_column_av_single_spaxel():
    """All gas in spaxel 0: cumulative [1, 2, 3] at z = [1, 2, 3]; interpolation
    inside, linear extrapolation below, clamp above."""
    gz = np.array([3.0, 1.0, 2.0], np.float32)
    gp = np.zeros(3, np.int32)
    ext = np.ones(3, np.float32)
    sz = np.array([2.5, 1.0, 0.0, -1.0, 5.0], np.float32)
    av = R.column_av(gz, gp, ext, sz, np.zeros(5, np.int32), 1)
    np.testing.assert_allclose(av, [2.5, 1.0, 0.0, -1.0, 3.0], atol=1e-6)


def test_column_av_other_spaxels_bound_the_range():
    """A gas cell of another spaxel at negative z becomes the key -1e30 |z|
    with value 0: a star below its own spaxel's gas then takes the first
    cumulative value instead of extrapolating; one at positive z does the
    same above the last."""
    gz = np.array([1.0, 2.0, -0.5, 0.7], np.float32)
    gp = np.array([0, 0, 1, 1], np.int32)
    ext = np.array([1.0, 2.0, 5.0, 5.0], np.float32)
    sz = np.array([0.0, 1.5, 9.0], np.float32)
    av = R.column_av(gz, gp, ext, sz, np.zeros(3, np.int32), 2)
    np.testing.assert_allclose(av, [1.0, 2.0, 3.0], rtol=1e-6)
    # without cells elsewhere the same star below extrapolates: 1 + (0-1)*(3-1)/(2-1) = -1
    av = R.column_av(gz[:2], gp[:2], ext[:2], sz[:1], np.zeros(1, np.int32), 2)
    np.testing.assert_allclose(av, [-1.0], rtol=1e-6)


def test_column_av_ids_outside_the_grid_get_zero():
    gz = np.array([1.0, 2.0], np.float32)
    av = R.column_av(gz, np.zeros(2, np.int32), np.ones(2, np.float32),
                     np.array([1.5, 1.5], np.float32), np.array([0, 4], np.int32), 1)
    np.testing.assert_allclose(av, [1.5, 0.0])


def test_reference_mock_case():
    # tests/test_dust_extinction.py:37-98: two gas cells, two stars, two bins
    gz = np.array([0.3, 0.6], np.float32)
    gp = np.array([0, 1], np.int32)
    metals = np.array([[0.01, 0.02, 0.03, 0.04, 0.05], [0.06, 0.07, 0.08, 0.09, 0.1]])
    ext = R.gas_extinction(np.array([1.0, 2.0]), metals, "power law slope free", "MW", 3.0, 1.0)
    av = R.column_av(gz, gp, ext, gz, gp, 2)
    spec = np.array([[1.0, 2.0], [3.0, 4.0]])
    for model in ("Cardelli89", "Gordon23"):
        out = R.apply_spaxel_extinction(spec, np.array([5000.0, 6000.0]), av, model, 3.1)
        assert out.shape == (2, 2) and np.all(out >= 0)
