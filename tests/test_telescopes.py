"""The instrument table and telescope factory (rubix/telescope/factory.py:38-110,
rubix/telescope/telescopes.yaml:1-150, rubix/telescope/apertures.py:11-81).

The aperture masks are pinned by exact integer geometry: offsets from the
centre are half-integers, so "inside the circle" is 4 (xx^2 + yy^2) <= s^2 in
integers, and the hexagon's sides are sqrt(3)(s/2 - |xx|) >= |yy| and
|yy| < s sqrt(3)/4, decided exactly by squaring.  The float32 evaluation the
reference does gives the same mask for every instrument size: no pixel centre
lies within float32 rounding of an edge.
"""
import numpy as np
import pytest
import yaml

from rubix_amd.config import telescopes_config
from rubix_amd.core.telescope import get_telescope
from rubix_amd.telescope import (TelescopeFactory, calculate_wave_edges, circular_aperture,
                                 hexagonal_aperture, square_aperture)

# sbin = floor(fov / spatial_res) and the float32 arange length of the
# wavelength axis of every instrument of the reference's table
EXPECTED = {
    "MUSE": (25, 3721), "NIRSpec_PRISM_CLEAR": (30, 411), "NIRSpec_G140M_F070LP": (30, 878),
    "NIRSpec_G140M_F100LP": (30, 1411), "NIRSpec_G235M_F170LP": (30, 1357),
    "NIRSpec_G395M_F290LP": (30, 1291), "NIRSpec_G140H_F070LP": (30, 2324),
    "NIRSpec_G140H_F100LP": (30, 3756), "NIRSpec_G235H_F170LP": (30, 3644),
    "NIRSpec_G395H_F290LP": (30, 3468), "MaNGA_12": (24, 6474), "MaNGA_32": (64, 6474),
    "SAMI": (30, 1924), "HECTOR": (300, 1369), "CALIFA": (74, 389),
}


def test_table_has_every_instrument():
    assert set(telescopes_config) == set(EXPECTED)
    for name, c in telescopes_config.items():
        assert set(c) == {"fov", "spatial_res", "wave_range", "wave_res", "lsf_fwhm",
                          "signal_to_noise", "aperture_type", "pixel_type"}, name
        assert c["pixel_type"] == "square"


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_create_every_instrument(name):
    t = TelescopeFactory().create_telescope(name)
    sbin, nt = EXPECTED[name]
    assert t.sbin == sbin
    assert t.wave_seq.dtype == np.float32 and len(t.wave_seq) == nt
    lo, hi = telescopes_config[name]["wave_range"]
    assert t.wave_seq[0] == np.float32(lo)
    # float32 steps of wave_res (the arange may end within rounding of hi)
    d = np.diff(t.wave_seq.astype(np.float64))
    assert np.all(np.abs(d - t.wave_res) <= 1e-6 * hi + 1e-3)
    assert t.wave_seq[-1] <= hi + t.wave_res
    assert t.aperture_region.shape == (sbin * sbin,)
    assert set(np.unique(t.aperture_region)) <= {0.0, 1.0}
    # the edges bracket the sequence by half a step on each side
    assert abs(float(t.wave_edges[0]) - (float(t.wave_seq[0]) - t.wave_res / 2)) < 1e-2
    assert len(t.wave_edges) in (nt, nt + 1, nt + 2)


def test_step_config_names_an_instrument():
    cfg = {"telescope": {"name": "HECTOR"}}
    t = get_telescope(cfg)
    assert t.name == "HECTOR" and t.sbin == 300 and t.aperture_type == "hexagonal"


def test_unknown_names_and_apertures():
    with pytest.raises(ValueError, match="Telescope NOPE not found in config"):
        TelescopeFactory().create_telescope("NOPE")
    bad = {"T": dict(telescopes_config["MUSE"], aperture_type="unknown")}
    with pytest.raises(ValueError, match="Unknown aperture type: unknown"):
        TelescopeFactory(bad).create_telescope("T")


def test_factory_from_yaml_file(tmp_path):
    """rubix's test_telescope_factory_with_file_path case: a 100" field at 10"
    per spaxel, square aperture of 10 x 10 ones."""
    cfg = {"telescope1": {"fov": 100.0, "spatial_res": 10, "aperture_type": "square",
                          "wave_range": [400.0, 700.0], "wave_res": 0.5, "lsf_fwhm": 0.1,
                          "signal_to_noise": 100.0, "pixel_type": "square"}}
    p = tmp_path / "t.yaml"
    p.write_text(yaml.safe_dump(cfg))
    table = yaml.safe_load(p.read_text())
    t = TelescopeFactory(table).create_telescope("telescope1")
    assert (t.fov, t.spatial_res, t.wave_res, t.lsf_fwhm, t.signal_to_noise) == (100.0, 10, 0.5, 0.1, 100.0)
    np.testing.assert_array_equal(t.aperture_region, np.ones(100, np.float32))


def _exact_circle(s):
    # offsets doubled: 2 xx = 2 c - (2 s + 1 - 2 k) are odd / even integers
    idx = np.arange(1, s + 1)
    xx2 = 2 * idx - (s + 1)                   # 2 (x - (s/2 + 0.5))
    yy2 = 2 * (s - np.arange(s)) - (s + 1)    # rows run y = s .. 1
    X = np.tile(xx2, (s, 1))
    Y = np.tile(yy2, (s, 1)).T
    return (X * X + Y * Y <= s * s).astype(np.float32).reshape(-1)


def _exact_hexagon(s):
    idx = np.arange(1, s + 1)
    a2 = np.abs(2 * idx - (s + 1))            # 2 |offset|, integers
    X = a2[:, None]                           # mask[x - 1, y - 1]
    Y = a2[None, :]
    # |xx| < s/2  <=>  X < s
    # |yy| < s sqrt(3)/4  <=>  4 |yy| < s sqrt 3  <=>  (2 Y)^2 < 3 s^2
    # sqrt(3) (s/2 - |xx|) >= |yy|  <=>  s - X >= 0 and 3 (s - X)^2 >= Y^2
    side = (s - X >= 0) & (3 * (s - X) ** 2 >= Y ** 2)
    return ((X < s) & (4 * Y * Y < 3 * s * s) & side).astype(np.float32).reshape(-1)


@pytest.mark.parametrize("s", [1, 2, 3, 10, 24, 25, 30, 64, 74, 300])
def test_apertures_match_exact_geometry(s):
    np.testing.assert_array_equal(square_aperture(s), np.ones(s * s, np.float32))
    np.testing.assert_array_equal(circular_aperture(s), _exact_circle(s))
    np.testing.assert_array_equal(hexagonal_aperture(s), _exact_hexagon(s))


def test_aperture_areas():
    """Large masks approach the shapes' areas: pi s^2 / 4 and the hexagon of
    circumradius s/2, (3 sqrt 3 / 8) s^2."""
    s = 300
    assert abs(circular_aperture(s).sum() / (np.pi * s * s / 4) - 1) < 0.01
    assert abs(hexagonal_aperture(s).sum() / (3 * np.sqrt(3) / 8 * s * s) - 1) < 0.01
    h = hexagonal_aperture(s).reshape(s, s)
    # symmetric under x -> -x and y -> -y; vertices on the x axis
    np.testing.assert_array_equal(h, h[::-1, :])
    np.testing.assert_array_equal(h, h[:, ::-1])
    assert h[:, s // 2].sum() > h[s // 2, :].sum()


def test_wave_edges():
    seq = np.arange(4700.15, 4710.0, 1.25, dtype=np.float32)
    e = calculate_wave_edges(seq, 1.25)
    assert e.dtype == np.float32
    np.testing.assert_allclose(e[:len(seq)], seq - 0.625, atol=1e-3)
