"""Host-side behaviour of the drop-in step API (no GPU needed): config
validation messages, reshape_data, telescope setup, pipeline assembly."""
import json
import os

import numpy as np
import pytest
import torch

from rubix_amd.core import data as D
from rubix_amd.core.lsf import get_convolve_lsf
from rubix_amd.core.noise import get_apply_noise
from rubix_amd.core.pipeline import RubixPipeline
from rubix_amd.core.psf import get_convolve_psf
from rubix_amd.core.ssp import get_ssp, lookup_method
from rubix_amd.pipeline import LinearTransformerPipeline
from rubix_amd.telescope import PLANCK15, TelescopeFactory, calculate_spatial_bin_edges

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


@pytest.mark.parametrize("cfg,msg", GOLD["config_errors"]["psf"])
def test_psf_config_errors(cfg, msg):
    with pytest.raises(ValueError) as e:
        get_convolve_psf(cfg)
    assert str(e.value) == msg


@pytest.mark.parametrize("cfg,msg", GOLD["config_errors"]["lsf"])
def test_lsf_config_errors(cfg, msg):
    with pytest.raises(ValueError) as e:
        get_convolve_lsf(cfg)
    assert msg in str(e.value)


@pytest.mark.parametrize("cfg,msg", GOLD["config_errors"]["noise"])
def test_noise_config_errors(cfg, msg):
    with pytest.raises(ValueError) as e:
        get_apply_noise(cfg)
    assert str(e.value) == msg


def test_noise_step_accepts_the_reference_config():
    cfg = {"telescope": {"noise": {"signal_to_noise": 1, "noise_distribution": "normal"}}}
    assert callable(get_apply_noise(cfg))
    assert "apply_noise" in __import__("rubix_amd.core.pipeline", fromlist=["x"]).FACTORIES


def test_ssp_config_errors():
    # rubix/core/ssp.py:23-31 messages
    with pytest.raises(ValueError, match="does not contain 'ssp' field"):
        get_ssp({})
    with pytest.raises(ValueError, match="does not contain 'template' field"):
        get_ssp({"ssp": {}})
    with pytest.raises(ValueError, match="does not contain 'name' field"):
        get_ssp({"ssp": {"template": {}}})
    with pytest.raises(ValueError, match="not found in the supported configuration file"):
        get_ssp({"ssp": {"template": {"name": "BruzualCharlot2003"}}})
    with pytest.raises(NotImplementedError, match="method 'cubic'"):
        lookup_method({"ssp": {"template": {"name": "x"}}})
    assert lookup_method({"ssp": {"method": "linear"}}) == "linear"


@pytest.mark.parametrize("case", GOLD["reshape_array"]["cases"])
def test_reshape_array_known_answers(case):
    got = D.reshape_array(torch.tensor(case["arr"]), case["n_dev"])
    np.testing.assert_array_equal(got.numpy(), case["expected"])


def test_reshape_data_step_uses_device_count(monkeypatch):
    monkeypatch.setattr(D, "device_count", lambda: 2)
    rd = D.RubixData(stars=D.StarsData(coords=torch.zeros(3, 3), mass=torch.arange(3.0)))
    out = D.get_reshape_data({})(rd)
    assert out.stars.coords.shape == (2, 2, 3)
    np.testing.assert_array_equal(out.stars.mass.numpy(), [[0, 1], [2, 0]])


def test_telescope_factory_muse():
    tel = TelescopeFactory().create_telescope("MUSE")
    assert tel.sbin == 25
    assert tel.pixel_type == "square"
    np.testing.assert_array_equal(tel.wave_seq, np.arange(4700.15, 9351.4, 1.25, dtype=np.float32))
    with pytest.raises(ValueError, match="not found in config"):
        TelescopeFactory().create_telescope("HUBBLE")


def test_angular_scale_and_edges():
    # flat LCDM, Planck15 parameters: compare with an independent quadrature
    from scipy.integrate import quad

    z = 0.1
    om, h = 0.3075, 0.6774
    dc = 299792.458 / (100 * h) * quad(lambda x: 1 / np.sqrt(om * (1 + x) ** 3 + 1 - om), 0, z)[0]
    expect = dc / (1 + z) * np.pi / (180 * 3600) * 1e3
    assert PLANCK15.angular_scale(z) == pytest.approx(expect, rel=1e-4)
    edges, s = calculate_spatial_bin_edges(5.0, 25, z)
    assert len(edges) in (26, 27)
    assert s == pytest.approx(expect * 5.0 / 25, rel=1e-4)


def _fn(name):
    def f(x):
        return x + [name]
    f.__name__ = name
    return f


def test_linear_pipeline_order_and_errors():
    cfg = {"Transformers": {"b": {"name": "b", "depends_on": "a"},
                            "a": {"name": "a", "depends_on": None},
                            "c": {"name": "c", "depends_on": "b"}}}
    p = LinearTransformerPipeline(cfg, [_fn("c"), _fn("a"), _fn("b")])
    assert p.apply([]) == ["a", "b", "c"]
    with pytest.raises(ValueError, match="already present"):
        LinearTransformerPipeline(cfg, [_fn("a"), _fn("a")])
    bad = {"Transformers": {"a": {"name": "a", "depends_on": None},
                            "b": {"name": "b", "depends_on": "a"},
                            "c": {"name": "c", "depends_on": "a"}}}
    with pytest.raises(ValueError, match="not linear"):
        LinearTransformerPipeline(bad, [_fn("a"), _fn("b"), _fn("c")])
    with pytest.raises(ValueError, match="not registered"):
        LinearTransformerPipeline(cfg, [_fn("a"), _fn("b")])


def test_pipeline_unknown_name():
    with pytest.raises(ValueError, match="not found"):
        RubixPipeline({"pipeline": {"name": "no_such_chain"}})
