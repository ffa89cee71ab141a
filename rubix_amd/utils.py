"""read_yaml and load_galaxy_data (rubix/utils.py:47-65, 153-197).

load_galaxy_data reads a rubix_galaxy.h5 file (the layout
create_rubix_galaxy writes, rubix/galaxy/input_handler/base.py:16-75)
through rubix_amd.hdf5, a numpy HDF5 reader, in place of h5py."""
from __future__ import annotations

import os

import yaml

from .hdf5 import H5File


def read_yaml(path_to_file: str) -> dict:
    """rubix/utils.py:47-65."""
    try:
        with open(path_to_file, "r") as cfgfile:
            cfg = yaml.safe_load(cfgfile)
    except Exception as e:
        raise RuntimeError(
            f"Something went wrong while reading yaml file {str(path_to_file)}") from e
    return cfg


def load_galaxy_data(path_to_file: str):
    """rubix/utils.py:153-197: (galaxy_data, units) of a rubix_galaxy.h5 file.

    galaxy_data: subhalo_center, subhalo_halfmassrad_stars, redshift (numpy
    values as stored) and particle_data[type][field] (numpy arrays);
    units: units["galaxy"][field] and units[type][field] from each dataset's
    "unit" attribute."""
    galaxy_data = {}
    units = {}
    if not os.path.exists(path_to_file):
        raise FileNotFoundError(f"File {str(path_to_file)} does not exist")
    with H5File(path_to_file) as f:
        galaxy_data["subhalo_center"] = f["galaxy/center"][()]
        galaxy_data["subhalo_halfmassrad_stars"] = f["galaxy/halfmassrad_stars"][()]
        galaxy_data["redshift"] = f["galaxy/redshift"][()]
        units["galaxy"] = {}
        for key in f["galaxy"].keys():
            units["galaxy"][key] = f["galaxy"][key].attrs["unit"]
        galaxy_data["particle_data"] = {}
        for key in f["particles"].keys():
            galaxy_data["particle_data"][key] = {}
            units[key] = {}
            for field in f["particles"][key].keys():
                galaxy_data["particle_data"][key][field] = f[f"particles/{key}/{field}"][()]
                units[key][field] = f[f"particles/{key}/{field}"].attrs["unit"]
    return galaxy_data, units
