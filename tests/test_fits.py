"""CPU: the FITS writer / reader (rubix_amd/core/fits.py) against the FITS
Standard 4.0 layout, byte for byte where the standard fixes it, and a round
trip of a datacube through store_fits / load_fits."""
import numpy as np
import torch

from rubix_amd import synthetic as S
from rubix_amd.core.data import RubixData, StarsData
from rubix_amd.core.fits import load_fits, store_fits, write_cube_fits


def test_layout_follows_the_standard(tmp_path):
    cube = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    p = tmp_path / "c.fits"
    write_cube_fits(str(p), cube, [("PIPELINE", "calc_ifu"), ("DIST_Z", 0.1)], [("EXTNAME", "DATA")])
    raw = p.read_bytes()
    assert len(raw) % 2880 == 0
    # mandatory first cards, fixed format: logical / integer values end in column 30
    assert raw[:80].decode() == "SIMPLE  =                    T / conforms to FITS standard".ljust(80)
    assert raw[80:160].decode() == "BITPIX  =                    8".ljust(80)
    assert raw[160:240].decode() == "NAXIS   =                    0".ljust(80)
    ext = raw.index(b"XTENSION")
    assert ext == 2880
    assert raw[ext:ext + 80].decode() == "XTENSION= 'IMAGE   '           / Image extension".ljust(80)
    assert raw[ext + 80:ext + 160].decode() == "BITPIX  =                  -32".ljust(80)
    assert raw[ext + 240:ext + 320].decode() == "NAXIS1  =                    2".ljust(80)  # cube.T
    assert raw[ext + 400:ext + 480].decode() == "NAXIS3  =                    4".ljust(80)
    # data: the second header is one block; big-endian cube.T follows
    data = np.frombuffer(raw[5760:5760 + 96], dtype=">f4")
    np.testing.assert_array_equal(data, cube.T.reshape(-1))
    assert raw[5760 + 96:] == b"\0" * (2880 - 96)


def test_store_and_load_round_trip(tmp_path):
    w = S.config1(n_stars=10)
    cfg = S.step_config(w, device="cpu")
    cfg.update({"pipeline": {"name": "calc_ifu"}, "simulation": {"name": "IllustrisTNG"},
                "cosmology": {"name": "PLANCK15"}})
    cfg["data"].update({"load_galaxy_args": {"id": 12}, "subset": {"use_subset": True}})
    cfg["data"]["args"]["snapshot"] = 99
    cube = np.random.default_rng(0).normal(size=(5, 5, 7)).astype(np.float32)
    rd = RubixData(stars=StarsData(datacube=torch.from_numpy(cube)))
    name = store_fits(cfg, rd, str(tmp_path) + "/out/")
    assert name.endswith("IllustrisTNG_id12_snap99_stars_subsetTrue.fits")
    prim, hdr, back = load_fits(name)
    np.testing.assert_array_equal(back, cube)
    assert prim["PIPELINE"] == "calc_ifu" and prim["GALAXYID"] == 12 and prim["SUBSET"] is True
    assert prim["N_DISTR"] == "normal" and prim["DIST_Z"] == cfg["galaxy"]["dist_z"]
    assert hdr["EXTNAME"] == "DATA" and hdr["CTYPE3"] == "AWAV" and hdr["CRPIX1"] == 2.0
    assert hdr["NAXIS1"] == 5 and hdr["NAXIS3"] == 7
