"""store_fits / load_fits (rubix/core/fits.py:13-115) without astropy: a
FITS writer and reader for the datacube, written to the FITS Standard 4.0
(2880-byte blocks of 80-character header cards, big-endian data).

store_fits writes what the reference writes: an empty primary HDU whose
header holds the run's configuration keys, and an IMAGE extension "DATA"
holding datacube.T (NAXIS1 = the cube's first axis, NAXIS3 = wavelength) as
IEEE float32 with the cube's world coordinate keys.  load_fits returns
(primary header, data header, cube in the (nb, nb, n_wave) layout); the
reference returns an mpdaf Cube, which is not available here.  The bytes
are those astropy.io.fits writes for the reference's HDU list (checked
against an astropy-written file in tests/test_fits.py)."""
from __future__ import annotations

import os
from typing import Dict, Tuple

import numpy as np

from ..logger import get_logger
from .telescope import get_telescope

BLOCK = 2880
CARD = 80


def _float(v: float) -> str:
    """A real value as astropy.io.fits (which the reference writes with)
    renders it in a fixed-format card: 16 significant digits ('G'), a '.0'
    appended to integral values, a two-digit signed exponent, and a mantissa
    cut short (not rounded) so that the field fits its 20 columns."""
    s = f"{float(v):.16G}"
    if "E" in s:
        mant, exp = s.split("E")
        sign = exp[0] if exp[0] in "+-" else ""
        s = f"{mant}E{sign}{int(exp.lstrip('+-')):02d}"
    elif "." not in s and "INF" not in s and "NAN" not in s:
        s += ".0"
    if len(s) > 20:
        e = s.find("E")
        s = s[:20] if e < 0 else s[:20 - (len(s) - e)] + s[e:]
    return s


def _value(v) -> str:
    """Fixed-format value field (columns 11-80 of a card)."""
    if isinstance(v, (bool, np.bool_)):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20}"
    if isinstance(v, (float, np.floating)):
        return f"{_float(v):>20}"
    quoted = "'" + f"{str(v).replace(chr(39), chr(39) * 2):<8}" + "'"
    return f"{quoted:<20}"


def _card(key: str, value=None, comment: str = "") -> bytes:
    key = key.upper()
    if len(key) > 8:
        raise ValueError(f"FITS keyword longer than 8 characters: {key}")
    if key == "END":
        text = "END"
    else:
        text = f"{key:<8}= {_value(value)}"
        if comment:
            text += f" / {comment}"
    if len(text) > CARD:
        raise ValueError(f"FITS card longer than 80 characters: {text}")
    return text.ljust(CARD).encode("ascii")


def _header(cards) -> bytes:
    body = b"".join(_card(*c) for c in cards) + _card("END")
    pad = (-len(body)) % BLOCK
    return body + b" " * pad


def write_cube_fits(path: str, cube: np.ndarray, primary: list, data_keys: list) -> None:
    """Primary HDU (no data) + IMAGE extension of cube.T as float32."""
    arr = np.ascontiguousarray(np.asarray(cube, dtype=np.float32).T)
    shape_fits = arr.shape[::-1]  # NAXIS1 is the fastest (last numpy) axis
    # mandatory cards with the comments astropy.io.fits gives them
    prim = [("SIMPLE", True, "conforms to FITS standard"), ("BITPIX", 8, "array data type"),
            ("NAXIS", 0, "number of array dimensions"), ("EXTEND", True)] + list(primary)
    ext = [("XTENSION", "IMAGE", "Image extension"), ("BITPIX", -32, "array data type"),
           ("NAXIS", arr.ndim, "number of array dimensions")]
    ext += [(f"NAXIS{i + 1}", int(n)) for i, n in enumerate(shape_fits)]
    ext += [("PCOUNT", 0, "number of parameters"), ("GCOUNT", 1, "number of groups")]
    ext += list(data_keys)
    data = arr.astype(">f4").tobytes()
    data += b"\0" * ((-len(data)) % BLOCK)
    with open(path, "wb") as f:
        f.write(_header(prim))
        f.write(_header(ext))
        f.write(data)


def _parse_value(s: str):
    s = s.strip()
    if s.startswith("'"):
        end = s.find("'", 1)
        while end != -1 and end + 1 < len(s) and s[end + 1] == "'":
            end = s.find("'", end + 2)
        return s[1:end].replace("''", "'").rstrip()
    v = s.split("/", 1)[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v)


def _read_header(f) -> Dict[str, object]:
    hdr: Dict[str, object] = {}
    while True:
        block = f.read(BLOCK)
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        for i in range(0, BLOCK, CARD):
            card = block[i:i + CARD].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                return hdr
            if card[8:10] == "= ":
                hdr[key] = _parse_value(card[10:])


def read_cube_fits(path: str) -> Tuple[dict, dict, np.ndarray]:
    with open(path, "rb") as f:
        prim = _read_header(f)
        ext = _read_header(f)
        dims = [int(ext[f"NAXIS{i + 1}"]) for i in range(int(ext["NAXIS"]))]
        if int(ext["BITPIX"]) != -32:
            raise ValueError(f"unsupported BITPIX {ext['BITPIX']}")
        n = int(np.prod(dims))
        data = np.frombuffer(f.read(4 * n), dtype=">f4").reshape(dims[::-1])
    return prim, ext, np.ascontiguousarray(data.T).astype(np.float32)


def store_fits(config: dict, data, filepath: str) -> str:
    """rubix/core/fits.py:13-100; returns the file name written."""
    logger = get_logger(config.get("logger", None))
    args = config["data"]["args"]
    if "cube_type" not in args or args["cube_type"] == "stars":
        datacube, parttype = data.stars.datacube, "stars"
    elif args["cube_type"] == "gas":
        datacube, parttype = data.gas.datacube, "gas"
    else:
        raise ValueError(f"cube_type {args['cube_type']!r} is not 'stars' or 'gas'")
    cube = datacube.detach().cpu().numpy() if hasattr(datacube, "detach") else np.asarray(datacube)
    tel = get_telescope(config)
    galaxy_id = config["data"]["load_galaxy_args"]["id"]
    snapshot = args["snapshot"]
    tele = config["telescope"]
    primary = [
        ("PIPELINE", config["pipeline"]["name"]),
        ("DIST_Z", float(config["galaxy"]["dist_z"])),
        ("ROTATION", config["galaxy"]["rotation"]["type"]),
        ("SIM", config["simulation"]["name"]),
        ("GALAXYID", galaxy_id),
        ("SNAPSHOT", snapshot),
        ("SUBSET", config["data"]["subset"]["use_subset"]),
        ("SSP", config["ssp"]["template"]["name"]),
        ("INSTR", tele["name"]),
        ("PSF", tele["psf"]["name"]),
        ("PSF_SIZE", tele["psf"]["size"]),
        ("PSFSIGMA", float(tele["psf"]["sigma"])),
        ("LSF", float(tele["lsf"]["sigma"])),
        ("S_TO_N", tele["noise"]["signal_to_noise"]),
        ("N_DISTR", tele["noise"]["noise_distribution"]),
        ("COSMO", config["cosmology"]["name"]),
    ]
    data_keys = [
        ("EXTNAME", "DATA"),
        ("OBJECT", f"{config['simulation']['name']} {galaxy_id}"),
        ("BUNIT", "erg/(s*cm^2*A)"),
        ("CRPIX1", (cube.shape[0] - 1) / 2),
        ("CRPIX2", (cube.shape[1] - 1) / 2),
        ("CD1_1", tel.spatial_res / 3600),
        ("CD1_2", 0),
        ("CD2_1", 0),
        ("CD2_2", tel.spatial_res / 3600),
        ("CUNIT1", "deg"),
        ("CUNIT2", "deg"),
        ("CTYPE1", "RA---TAN"),
        ("CTYPE2", "DEC--TAN"),
        ("CTYPE3", "AWAV"),
        ("CUNIT3", "Angstrom"),
        ("CD3_3", float(tel.wave_res)),
        ("CRPIX3", 1),
        ("CRVAL3", float(tel.wave_range[0])),
        ("CD1_3", 0),
        ("CD2_3", 0),
        ("CD3_1", 0),
        ("CD3_2", 0),
    ]
    name = (f"{filepath}{config['simulation']['name']}_id{galaxy_id}_snap{snapshot}_"
            f"{parttype}_subset{config['data']['subset']['use_subset']}.fits")
    os.makedirs(os.path.dirname(name) or ".", exist_ok=True)
    write_cube_fits(name, cube, primary, data_keys)
    logger.info(f"Datacube saved to {name}")
    return name


def load_fits(filepath: str):
    """rubix/core/fits.py:103-115: (primary header, data header, cube)."""
    return read_cube_fits(filepath)
