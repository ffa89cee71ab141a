"""Writes tests/golden/reference_cube.fits with astropy, card for card as the
reference's store_fits builds it (rubix/core/fits.py:13-100: a header of run
keys on an empty primary HDU, an ImageHDU of datacube.T named DATA with the
world-coordinate keys, HDUList.writeto).

Run with an interpreter that has astropy (this container's
/opt/conda/bin/python3.9: astropy 4.3.1; the test suite's own interpreter
has none):

    /opt/conda/bin/python3.9 tests/golden/make_fits_fixture.py

The run keys are those of CONFIG below and the telescope values those of
MUSE in rubix_amd's telescope table (spatial_res 0.2 arcsec, wave_res
1.25 A, wave_range[0] 4700.15 A); tests/test_fits.py rebuilds the same
config, writes the same cube with rubix_amd.core.fits.store_fits and
compares the files byte for byte.
"""
import os

import numpy as np
from astropy.io import fits

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG = {
    "pipeline": {"name": "calc_ifu"},
    "galaxy": {"dist_z": 0.1, "rotation": {"type": "face-on"}},
    "simulation": {"name": "IllustrisTNG"},
    "data": {"args": {"snapshot": 99}, "load_galaxy_args": {"id": 12},
             "subset": {"use_subset": True}},
    "ssp": {"template": {"name": "BruzualCharlot2003"}},
    "telescope": {"name": "MUSE", "psf": {"name": "gaussian", "size": 5, "sigma": 0.6},
                  "lsf": {"sigma": 0.5},
                  "noise": {"signal_to_noise": 100, "noise_distribution": "normal"}},
    "cosmology": {"name": "PLANCK15"},
}
SPATIAL_RES, WAVE_RES, WAVE0 = 0.2, 1.25, 4700.15

cube = np.random.default_rng(0).normal(size=(5, 5, 7)).astype(np.float32)
config = CONFIG
hdr = fits.Header()
hdr["SIMPLE"] = "T /conforms to FITS standard"
hdr["PIPELINE"] = config["pipeline"]["name"]
hdr["DIST_z"] = config["galaxy"]["dist_z"]
hdr["ROTATION"] = config["galaxy"]["rotation"]["type"]
hdr["SIM"] = config["simulation"]["name"]
galaxy_id = config["data"]["load_galaxy_args"]["id"]
snapshot = config["data"]["args"]["snapshot"]
hdr["GALAXYID"] = galaxy_id
object_name = f"{config['simulation']['name']} {galaxy_id}"
hdr["SNAPSHOT"] = snapshot
hdr["SUBSET"] = config["data"]["subset"]["use_subset"]
hdr["SSP"] = config["ssp"]["template"]["name"]
hdr["INSTR"] = config["telescope"]["name"]
hdr["PSF"] = config["telescope"]["psf"]["name"]
hdr["PSF_SIZE"] = config["telescope"]["psf"]["size"]
hdr["PSFSIGMA"] = config["telescope"]["psf"]["sigma"]
hdr["LSF"] = config["telescope"]["lsf"]["sigma"]
hdr["S_TO_N"] = config["telescope"]["noise"]["signal_to_noise"]
hdr["N_DISTR"] = config["telescope"]["noise"]["noise_distribution"]
hdr["COSMO"] = config["cosmology"]["name"]

hdr1 = fits.Header()
hdr1["EXTNAME"] = "DATA"
hdr1["OBJECT"] = object_name
hdr1["BUNIT"] = "erg/(s*cm^2*A)"
hdr1["CRPIX1"] = (cube.shape[0] - 1) / 2
hdr1["CRPIX2"] = (cube.shape[1] - 1) / 2
hdr1["CD1_1"] = SPATIAL_RES / 3600
hdr1["CD1_2"] = 0
hdr1["CD2_1"] = 0
hdr1["CD2_2"] = SPATIAL_RES / 3600
hdr1["CUNIT1"] = "deg"
hdr1["CUNIT2"] = "deg"
hdr1["CTYPE1"] = "RA---TAN"
hdr1["CTYPE2"] = "DEC--TAN"
hdr1["CTYPE3"] = "AWAV"
hdr1["CUNIT3"] = "Angstrom"
hdr1["CD3_3"] = WAVE_RES
hdr1["CRPIX3"] = 1
hdr1["CRVAL3"] = WAVE0
hdr1["CD1_3"] = 0
hdr1["CD2_3"] = 0
hdr1["CD3_1"] = 0
hdr1["CD3_2"] = 0

hdul = fits.HDUList([fits.PrimaryHDU(header=hdr), fits.ImageHDU(cube.T, header=hdr1)])
hdul.writeto(os.path.join(HERE, "reference_cube.fits"), overwrite=True)
print("wrote reference_cube.fits")
