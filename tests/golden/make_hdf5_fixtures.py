"""Writes the HDF5 fixtures of tests/test_hdf5.py with the real HDF5 library.

Run with an interpreter that has h5py (this container's
/opt/conda/bin/python3.9: h5py 3.3.0 on HDF5 1.10.6; the test suite's own
interpreter has no h5py):

    /opt/conda/bin/python3.9 tests/golden/make_hdf5_fixtures.py

Files (small, committed):
  rubix_galaxy_small.h5   the layout create_rubix_galaxy writes
                          (rubix/galaxy/input_handler/base.py:16-75): groups
                          meta / galaxy / particles/<type>, one dataset per
                          field with a "unit" string attribute; h5py defaults
                          (superblock 0, version 1 object headers, symbol-table
                          groups, contiguous layout, variable-length strings).
  rubix_galaxy_latest.h5  the same content with libver="latest" (superblock 3,
                          version 2 object headers, link messages, version 3
                          attribute messages, layout message version 4).
  hdf5_formats.h5         chunked datasets with deflate and shuffle, a
                          big-endian integer array, a compact dataset, a
                          fixed-length string attribute, a variable-length
                          string dataset, a scalar, 40 members in one group
                          (several symbol-table nodes), 30 attributes on
                          one dataset (object header continuation) and a
                          dataset of a committed datatype (rejected).
  hdf5_expected.npz       every array above (allow_pickle=False)
  hdf5_expected.json      every string / attribute / member list above
"""
import json
import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(2024)

N_STARS, N_GAS = 300, 120
stars = {
    "coords": rng.normal(0.0, 5.0, (N_STARS, 3)).astype(np.float32),
    "velocity": rng.normal(0.0, 100.0, (N_STARS, 3)).astype(np.float32),
    "mass": np.exp(rng.normal(np.log(1e5), 0.5, N_STARS)).astype(np.float32),
    "metallicity": (10 ** rng.uniform(-4, np.log10(0.05), N_STARS)).astype(np.float32),
    "age": rng.uniform(0.1, 13.5, N_STARS).astype(np.float32),
}
gas = {
    "coords": rng.normal(0.0, 6.0, (N_GAS, 3)),
    "velocity": rng.normal(0.0, 80.0, (N_GAS, 3)),
    "mass": np.exp(rng.normal(np.log(3e4), 0.3, N_GAS)),
    "density": rng.uniform(1e-3, 1.0, N_GAS),
    "internal_energy": rng.uniform(1e2, 1e4, N_GAS),
    "metallicity": rng.uniform(0.0, 0.04, N_GAS),
    "metals": rng.uniform(0.0, 0.7, (N_GAS, 10)),
    "sfr": rng.uniform(0.0, 1.0, N_GAS),
    "electron_abundance": rng.uniform(0.0, 1.2, N_GAS),
}
galaxy = {"center": np.array([0.25, -0.5, 0.125]), "halfmassrad_stars": np.float64(3.75),
          "redshift": np.float64(0.1)}
meta = {"simulation": "TNG50-1", "snapshot": 99, "subhalo_id": 14, "redshift": 0.1}
units = {
    "galaxy": {"center": "kpc", "halfmassrad_stars": "kpc", "redshift": "dimensionless"},
    "stars": {"coords": "kpc", "velocity": "km/s", "mass": "Msun", "metallicity": "",
              "age": "Gyr"},
    "gas": {"coords": "kpc", "velocity": "km/s", "mass": "Msun", "density": "Msun/kpc^3",
            "internal_energy": "km^2/s^2", "metallicity": "", "metals": "",
            "sfr": "Msun/yr", "electron_abundance": ""},
}


def write_galaxy(path, **kw):
    with h5py.File(path, "w", **kw) as f:
        m = f.create_group("meta")
        g = f.create_group("galaxy")
        p = f.create_group("particles")
        for k, v in meta.items():
            m.create_dataset(k, data=v)
        for k, v in galaxy.items():
            g.create_dataset(k, data=v)
            g[k].attrs["unit"] = units["galaxy"][k]
        for ptype, fields in (("stars", stars), ("gas", gas)):
            p.create_group(ptype)
            for field, v in fields.items():
                p[ptype].create_dataset(field, data=v)
                p[ptype][field].attrs["unit"] = units[ptype][field]


write_galaxy(os.path.join(HERE, "rubix_galaxy_small.h5"))
write_galaxy(os.path.join(HERE, "rubix_galaxy_latest.h5"), libver="latest")

arrays = {f"stars/{k}": v for k, v in stars.items()}
arrays.update({f"gas/{k}": v for k, v in gas.items()})
arrays.update({f"galaxy/{k}": np.asarray(v) for k, v in galaxy.items()})

chunked = rng.normal(size=(37, 11)).astype(np.float32)
shuffled = rng.integers(-1000, 1000, (50, 7)).astype(np.int32)
big_endian = np.arange(24, dtype=">u2").reshape(4, 6)
compact = np.linspace(0.0, 1.0, 9)
many = {f"item{i:02d}": np.full(3, i, dtype=np.int64) for i in range(40)}
with h5py.File(os.path.join(HERE, "hdf5_formats.h5"), "w") as f:
    f.create_dataset("chunked_gzip", data=chunked, chunks=(8, 4), compression="gzip",
                     compression_opts=4)
    f.create_dataset("chunked_shuffle_gzip", data=shuffled, chunks=(16, 7), shuffle=True,
                     compression="gzip")
    f.create_dataset("big_endian", data=big_endian)
    dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
    dcpl.set_layout(h5py.h5d.COMPACT)
    space = h5py.h5s.create_simple(compact.shape)
    dsid = h5py.h5d.create(f.id, b"compact", h5py.h5t.IEEE_F64LE, space, dcpl=dcpl)
    dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, compact)
    f.create_dataset("scalar_int", data=np.int64(-42))
    f.create_dataset("vlen_strings", data=["alpha", "beta", "", "a longer string value"],
                     dtype=h5py.string_dtype())
    f["big_endian"].attrs["fixed"] = np.bytes_("fixed-len")
    grp = f.create_group("many")
    for k, v in many.items():
        grp.create_dataset(k, data=v)
    d = f.create_dataset("attrs30", data=np.arange(5.0))
    for i in range(30):
        d.attrs[f"a{i:02d}"] = i * 1.5
    f["committed_f4"] = np.dtype("<f4")  # a named (committed) datatype
    f.create_dataset("uses_committed_type", data=np.ones(3, "<f4"), dtype=f["committed_f4"])
arrays.update({"chunked_gzip": chunked, "chunked_shuffle_gzip": shuffled,
               "big_endian": big_endian.astype(np.uint16), "compact": compact})
arrays.update({f"many/{k}": v for k, v in many.items()})
np.savez(os.path.join(HERE, "hdf5_expected.npz"), **{k.replace("/", "__"): v for k, v in arrays.items()})
with open(os.path.join(HERE, "hdf5_expected.json"), "w") as fh:
    json.dump({"units": units, "meta": meta, "scalar_int": -42,
               "vlen_strings": ["alpha", "beta", "", "a longer string value"],
               "fixed_attr": "fixed-len", "attrs30": {f"a{i:02d}": i * 1.5 for i in range(30)},
               "many": sorted(many)}, fh, indent=1, sort_keys=True)
print("wrote", sorted(os.listdir(HERE)))
