"""rubix_galaxy.h5 loading without h5py (rubix_amd.hdf5, rubix_amd.utils,
rubix_amd.core.data.prepare_input).

The fixtures under tests/golden/ were written by the HDF5 library itself
(h5py 3.3.0 on HDF5 1.10.6, tests/golden/make_hdf5_fixtures.py): the
rubix_galaxy layout of create_rubix_galaxy
(rubix/galaxy/input_handler/base.py:16-75) with h5py's defaults and with
libver="latest", and a file of the other storage forms the reader accepts.
Every array must come back bit for bit with its dtype, every string and
attribute as written.  prepare_input is checked on the CPU against a numpy
restatement of rubix/core/data.py:491-604 and rubix/galaxy/alignment.py:10-62
(np.median, np.random.choice with the reference's seed)."""
import json
import os
import shutil

import numpy as np
import pytest
import torch

from rubix_amd.hdf5 import H5File, HDF5Error
from rubix_amd.utils import load_galaxy_data
from rubix_amd.core.data import get_rubix_data, prepare_input

G = os.path.join(os.path.dirname(__file__), "golden")
EXP = np.load(os.path.join(G, "hdf5_expected.npz"))
META = json.load(open(os.path.join(G, "hdf5_expected.json")))
GALAXY_FILES = ("rubix_galaxy_small.h5", "rubix_galaxy_latest.h5")


def expected(path):
    return EXP[path.replace("/", "__")]


@pytest.mark.parametrize("name", GALAXY_FILES)
def test_galaxy_file_reads_exactly(name):
    with H5File(os.path.join(G, name)) as f:
        assert f.keys() == ["galaxy", "meta", "particles"]
        assert f["particles"].keys() == ["gas", "stars"]
        for ptype in ("stars", "gas"):
            assert f[f"particles/{ptype}"].keys() == sorted(META["units"][ptype])
            for field, unit in META["units"][ptype].items():
                d = f[f"particles/{ptype}/{field}"]
                got, want = d[()], expected(f"{ptype}/{field}")
                assert got.dtype == want.dtype and got.shape == want.shape
                np.testing.assert_array_equal(got, want)
                assert d.attrs == {"unit": unit}
        for field, unit in META["units"]["galaxy"].items():
            np.testing.assert_array_equal(f[f"galaxy/{field}"][()], expected(f"galaxy/{field}"))
            assert f[f"galaxy/{field}"].attrs["unit"] == unit
        meta = {k: f[f"meta/{k}"][()] for k in f["meta"].keys()}
        assert meta == {"redshift": 0.1, "simulation": b"TNG50-1", "snapshot": 99,
                        "subhalo_id": 14}


def test_storage_forms():
    with H5File(os.path.join(G, "hdf5_formats.h5")) as f:
        for k in ("chunked_gzip", "chunked_shuffle_gzip", "big_endian", "compact"):
            got = f[k][()]
            want = EXP[k]
            assert got.dtype == want.dtype.newbyteorder("=") and got.shape == want.shape, k
            np.testing.assert_array_equal(got, want)
        assert f["chunked_gzip"][3:5, 2].tolist() == EXP["chunked_gzip"][3:5, 2].tolist()
        s = f["scalar_int"][()]
        assert s == META["scalar_int"] and s.dtype == np.int64
        assert [x.decode() for x in f["vlen_strings"][()]] == META["vlen_strings"]
        assert f["big_endian"].attrs == {"fixed": META["fixed_attr"]}
        assert f["attrs30"].attrs == META["attrs30"]
        assert f["many"].keys() == META["many"]
        with pytest.raises(HDF5Error, match="committed"):
            f["uses_committed_type"][()]
        for k in META["many"]:
            np.testing.assert_array_equal(f[f"many/{k}"][()], EXP[f"many__{k}"])


def test_errors(tmp_path):
    bad = tmp_path / "not.h5"
    bad.write_bytes(b"plain text, no signature" * 40)
    with pytest.raises(HDF5Error, match="not an HDF5 file"):
        H5File(str(bad))
    cut = tmp_path / "cut.h5"
    cut.write_bytes(open(os.path.join(G, "rubix_galaxy_small.h5"), "rb").read()[:3000])
    with pytest.raises(HDF5Error):
        with H5File(str(cut)) as f:
            f["particles/stars/coords"][()]
    with H5File(os.path.join(G, "rubix_galaxy_small.h5")) as f:
        with pytest.raises(KeyError):
            f["particles/dark_matter"]
    with pytest.raises(FileNotFoundError, match="does not exist"):
        load_galaxy_data(str(tmp_path / "missing.h5"))


@pytest.mark.parametrize("name", GALAXY_FILES)
def test_load_galaxy_data(name):
    """rubix/utils.py:153-197: the same dictionaries h5py would give."""
    data, units = load_galaxy_data(os.path.join(G, name))
    np.testing.assert_array_equal(data["subhalo_center"], expected("galaxy/center"))
    assert data["subhalo_halfmassrad_stars"] == 3.75 and data["redshift"] == 0.1
    assert units == {"galaxy": META["units"]["galaxy"], "stars": META["units"]["stars"],
                     "gas": META["units"]["gas"]}
    for ptype in ("stars", "gas"):
        assert sorted(data["particle_data"][ptype]) == sorted(META["units"][ptype])
        for field in META["units"][ptype]:
            np.testing.assert_array_equal(data["particle_data"][ptype][field],
                                          expected(f"{ptype}/{field}"))


def _numpy_prepare(types, subset=None):
    """rubix/core/data.py:491-604 + rubix/galaxy/alignment.py:10-62 in numpy
    (float32 as jnp with 64-bit types off).  The gas subset draws from the
    star count and, as jnp indexing does, clamps indices past the gas arrays
    (subset 150 > 120 gas particles)."""
    center = expected("galaxy/center").astype(np.float32)
    out = {}
    for ptype in types:
        fields = {k: expected(f"{ptype}/{k}") for k in META["units"][ptype]}
        fields = {k: v.astype(np.float32) if v.dtype == np.float64 else v for k, v in fields.items()}
        pos, vel = fields["coords"], fields["velocity"]
        mask = np.linalg.norm(pos - center, axis=1) < 10
        fields["coords"] = pos - center
        fields["velocity"] = vel - np.median(vel[mask], axis=0)
        if subset:
            np.random.seed(42)
            n = len(out["stars"]["coords"]) if "stars" in out else len(fields["coords"])
            idx = np.random.choice(np.arange(n), size=subset, replace=False)
            fields = {k: v[np.minimum(idx, len(v) - 1)] for k, v in fields.items()}
        out[ptype] = fields
    return out


@pytest.mark.parametrize("subset", [None, 100, 150])
def test_prepare_input_matches_numpy_restatement(tmp_path, subset):
    shutil.copy(os.path.join(G, "rubix_galaxy_small.h5"), tmp_path / "rubix_galaxy.h5")
    cfg = {"device": "cpu", "output_path": str(tmp_path),
           "data": {"args": {"particle_type": ["stars", "gas"]},
                    "subset": {"use_subset": subset is not None, "subset_size": subset}}}
    rd = prepare_input(cfg)
    want = _numpy_prepare(["stars", "gas"], subset)
    assert rd.galaxy.redshift == 0.1 and rd.galaxy.halfmassrad_stars == 3.75
    assert rd.galaxy.center_unit == "kpc" and rd.galaxy.redshift_unit == "dimensionless"
    for ptype in ("stars", "gas"):
        holder = getattr(rd, ptype)
        for field, v in want[ptype].items():
            got = getattr(holder, field)
            assert isinstance(got, torch.Tensor) and got.dtype == torch.float32, (ptype, field)
            np.testing.assert_array_equal(got.numpy(), v, err_msg=f"{ptype}.{field}")
            assert getattr(holder, field + "_unit") == META["units"][ptype][field]
    # the same through get_rubix_data (the file exists, so nothing is converted)
    rd2 = get_rubix_data(cfg)
    np.testing.assert_array_equal(rd2.stars.coords.numpy(), rd.stars.coords.numpy())


def test_prepare_input_errors(tmp_path):
    cfg = {"device": "cpu", "output_path": str(tmp_path),
           "data": {"args": {"particle_type": ["stars"]}}}
    with pytest.raises(FileNotFoundError, match="not part of rubix_amd"):
        get_rubix_data(cfg)
    shutil.copy(os.path.join(G, "rubix_galaxy_small.h5"), tmp_path / "rubix_galaxy.h5")
    rd = prepare_input(cfg)
    assert rd.gas.coords is None
    # a center outside the particle bounds is rejected as in the reference
    from rubix_amd.core.data import center_particles
    rd.galaxy.center = np.array([1e3, 0.0, 0.0])
    with pytest.raises(ValueError, match="Center is not within the bounds"):
        center_particles(rd, "stars")
