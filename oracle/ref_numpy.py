"""numpy restatement of rubix's particle -> datacube path (forward).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU baseline as the checker.  Never imported by ``rubix_amd``.

Every function follows the reference op for op and cites the file:line it
restates (paths relative to the AstroAI-Lab/rubix checkout).  ``dtype``
selects the arithmetic: ``np.float32`` mirrors the reference (which never
enables x64; the SSP grid is cast to float32 at rubix/spectra/ssp/grid.py:147),
``np.float64`` is the high-precision oracle that errors are reported against.

Pinning (what checks each function, in tests/test_oracle.py):
* square_spaxel_assignment - reference known answer [4, 14]
  (tests/test_telescope_utils.py:20-35) + brute-force loop + NaN/edge cases.
* sort_by_spaxel           - np.argsort(kind="stable") identity + invariants.
* lookup_linear            - reference lookup known answers
  (tests/test_ssp_grid.py:653-698: node exactness, 0 outside, range bound),
  brute-force bilinear on tiny grids, reproduction of affine functions.
* velocity/cosmological doppler, calculate_diff - reference known answers
  (tests/test_spectra_ifu.py:29-50,150-172).
* interp_jax               - np.interp (library routine) incl. clamping.
* resample_spectrum        - reference 5-point recipe and zero spectrum
  (tests/test_spectra_ifu.py:175-228), flux conservation identity.
* calculate_cube           - reference known answer
  (tests/test_spectra_ifu.py:231-257) + brute force.
* apply_psf                - scipy.signal.convolve2d(mode="same") on random
  planes and kernels (odd sizes) + the reference delta/ones-kernel cases.
* apply_lsf                - np.convolve(mode="full") + crop, and the
  reference delta -> normalised Gaussian known answer
  (tests/test_telescope_lsf.py:6-60).
* gaussian_kernel_2d       - sums to 1, centre max, equals outer(g, g).
* mask_particles_outside_aperture, filter_particles - reference boundary
  cases (tests/test_telescope_utils.py:38-79) + brute force.
* moment_of_inertia_tensor - reference known answer diag(2, 2, 2)
  (tests/test_galaxy_alignment.py:76-90) + brute-force double loop.
* rotation_matrix_from_inertia_tensor - diagonalises I (R^T I R diagonal,
  eigenvalues ascending), identity for diag(1, 2, 3)
  (tests/test_galaxy_alignment.py:93-108).
* euler_rotation_matrix    - reference known answer for (90, 0, 0)
  (tests/test_galaxy_alignment.py:132-146), scipy Rotation.from_euler
  composition, orthonormal with det 1.
* threefry2x32             - the published Threefry-2x32-20 known-answer
  vectors (Random123 kat_vectors; also asserted by jax's random tests).
* sample_noise             - uniform: the 2^-23 grid, moments, KS test;
  normal: sqrt(2) erfinv(2p - 1) = ndtri(p) (scipy) for the same p, moments,
  KS test; the reference's shape / no-NaN / invalid-type cases
  (tests/test_telescope_noise.py:61-79).  The mapping from key and index to
  words follows jax's published construction, but jax is absent, so the
  draws themselves are "parity unpinned" against the reference.
* flux_median, calculate_S2N - np.median (library) on odd and even counts,
  NaN propagation, constant-flux closed form S = 1/sn, zero cube.
* apply_noise              - the reference's cases (zeros -> zero noise, no
  NaN / inf, variation across spaxels; tests/test_telescope_noise.py:15-58).
* apply_noise_vjp          - central finite differences of the float64
  forward (odd and even spaxel counts, so both midpoint weights count).
"""
from __future__ import annotations

import numpy as np

# rubix/config/rubix_config.yml:8
SPEED_OF_LIGHT = 299792.458


# ---------------------------------------------------------------- geometry

def spatial_bin_edges(ang_scale, fov, spatial_bins, dtype=np.float32):
    """rubix/telescope/utils.py:11-37 with the angular scale given.

    aperture = ang_scale * fov; s = aperture / bins;
    edges = arange(-aperture/2, aperture/2 + s, s).  The float arange can
    yield bins+1 or bins+2 edges depending on rounding (a reference quirk
    that the synthetic configs avoid by passing explicit edges).
    """
    a = dtype(ang_scale) * dtype(fov)
    s = a / dtype(spatial_bins)
    return np.arange(-a / 2, a / 2 + s, s, dtype=dtype), s


def square_spaxel_assignment(coords, edges):
    """rubix/telescope/utils.py:77-152.

    ix = digitize(x, edges) - 1 (== searchsorted(edges, x, 'right') - 1),
    clipped to [0, nbins-1] with nbins = len(edges) - 1; same for y;
    id = ix + nbins * iy as int32.
    """
    coords = np.asarray(coords)
    edges = np.asarray(edges)
    nb = len(edges) - 1
    ix = np.digitize(coords[:, 0], edges) - 1
    iy = np.digitize(coords[:, 1], edges) - 1
    ix = np.clip(ix, 0, nb - 1)
    iy = np.clip(iy, 0, nb - 1)
    return (ix + nb * iy).astype(np.int32)


def mask_particles_outside_aperture(coords, edges):
    """rubix/telescope/utils.py:155-178 (boundary inclusive)."""
    lo, hi = np.min(edges), np.max(edges)
    m = (coords[:, 0] >= lo) & (coords[:, 0] <= hi)
    m &= (coords[:, 1] >= lo) & (coords[:, 1] <= hi)
    return m


def filter_particles(coords, edges, fields):
    """rubix/core/telescope.py:129-201: every per-particle field except coords
    and velocity becomes jnp.where(mask, value, 0); returns (fields, mask)."""
    mask = mask_particles_outside_aperture(coords, edges)
    out = {k: np.where(mask, np.asarray(v), np.zeros((), dtype=np.asarray(v).dtype))
           for k, v in fields.items()}
    return out, mask


def moment_of_inertia_tensor(positions, masses, halfmass_radius):
    """rubix/galaxy/alignment.py:66-120: I_ii = sum m (|r|^2 - r_i^2),
    I_ij = -sum m r_i r_j over particles with |r| <= halfmass_radius (the
    radius test in float32 as the reference evaluates it; sums in float64)."""
    p32 = np.asarray(positions, dtype=np.float32)
    d = np.sqrt((p32[:, 0] * p32[:, 0] + p32[:, 1] * p32[:, 1]) + p32[:, 2] * p32[:, 2])
    keep = d <= np.float32(halfmass_radius)
    p = p32[keep].astype(np.float64)
    m = np.asarray(masses, dtype=np.float64)[keep]
    q = np.sum(p * p, axis=1)
    I = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            if i == j:
                I[i, j] = np.sum(m * q - m * p[:, i] ** 2)
            else:
                I[i, j] = -np.sum(m * p[:, i] * p[:, j])
    return I


def rotation_matrix_from_inertia_tensor(I):
    """rubix/galaxy/alignment.py:123-137: eigenvectors of I (eigh) as columns,
    ordered by ascending eigenvalue."""
    vals, vecs = np.linalg.eigh(np.asarray(I, dtype=np.float64))
    return vecs[:, np.argsort(vals)]


def euler_rotation_matrix(alpha, beta, gamma):
    """rubix/galaxy/alignment.py:140-205: R = Rz(gamma) Ry(beta) Rx(alpha),
    angles in degrees, elementary rotations about the fixed axes."""
    a, b, c = np.deg2rad([alpha, beta, gamma])
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def rotate_galaxy(positions, velocities, masses, halfmass_radius, alpha, beta, gamma):
    """rubix/galaxy/alignment.py:240-265: p' = (p R) E with R the principal
    axes of the inertia tensor and E the Euler matrix (row vectors)."""
    R = rotation_matrix_from_inertia_tensor(
        moment_of_inertia_tensor(positions, masses, halfmass_radius))
    E = euler_rotation_matrix(alpha, beta, gamma)
    p = np.asarray(positions, dtype=np.float64) @ R @ E
    v = np.asarray(velocities, dtype=np.float64) @ R @ E
    return p, v


def sort_by_spaxel(ids, n_seg):
    """Stable sort-by-spaxel index (new in the build; analogue of the
    lexsort + searchsorted(side='left') at
    rubix/spectra/dust/dust_extinction.py:237-267).

    Returns perm (stable argsort of ids) and offsets[s] = first sorted
    position with id >= s, for s in [0, n_seg]; ids >= n_seg sort last
    (they are the stars segment_sum drops).
    """
    ids = np.asarray(ids)
    perm = np.argsort(ids, kind="stable").astype(np.int32)
    srt = ids[perm]
    offsets = np.searchsorted(srt, np.arange(n_seg + 1), side="left").astype(np.int32)
    return perm, offsets


def reshape_array(arr, n_dev=1):
    """rubix/core/data.py:446-487: zero-pad the leading axis to a multiple
    of n_dev and reshape to (n_dev, ceil(N/n_dev), ...)."""
    arr = np.asarray(arr)
    n = arr.shape[0]
    per = (n + n_dev - 1) // n_dev
    tot = per * n_dev
    if tot > n:
        pad = [(0, tot - n)] + [(0, 0)] * (arr.ndim - 1)
        arr = np.pad(arr, pad, mode="constant")
    return arr.reshape(n_dev, per, *arr.shape[1:])


# ---------------------------------------------------------------- spectra

def lookup_linear(metallicity, age, ax_met, ax_age, flux, log_axes=False,
                  dtype=np.float32):
    """Bilinear SSP lookup, ``interp2d(method="linear", extrap=0)`` as bound
    at rubix/spectra/ssp/grid.py:113-120 (x = metallicity, y = age,
    f = flux with layout (n_met, n_age, n_wave), grid.py:31).

    Cell: i = clip(searchsorted(x, xq, 'right'), 1, nx-1) (same for y);
    value = [(x2-xq)(y2-yq) f[i-1,j-1] + (xq-x1)(y2-yq) f[i,j-1]
             + (x2-xq)(yq-y1) f[i-1,j] + (xq-x1)(yq-y1) f[i,j]] / (dx dy);
    exactly 0 where xq or yq lies strictly outside its axis range.
    ``log_axes``: the axes hold log10 values and the queries are log10'd
    first (the bilinear-in-(log Z, log age) parity mode).
    Returns (n_q, n_wave).
    """
    Z = np.asarray(metallicity, dtype=dtype)
    t = np.asarray(age, dtype=dtype)
    x = np.asarray(ax_met, dtype=dtype)
    y = np.asarray(ax_age, dtype=dtype)
    f = np.asarray(flux, dtype=dtype)
    with np.errstate(divide="ignore", invalid="ignore"):
        return _lookup_linear(Z, t, x, y, f, log_axes, dtype)


def _lookup_linear(Z, t, x, y, f, log_axes, dtype):
    xq = np.log10(Z) if log_axes else Z
    yq = np.log10(t) if log_axes else t
    i = np.clip(np.searchsorted(x, xq, side="right"), 1, len(x) - 1)
    j = np.clip(np.searchsorted(y, yq, side="right"), 1, len(y) - 1)
    x1, x2 = x[i - 1], x[i]
    y1, y2 = y[j - 1], y[j]
    ax2 = (x2 - xq)[:, None]
    ax1 = (xq - x1)[:, None]
    ay2 = (y2 - yq)[:, None]
    ay1 = (yq - y1)[:, None]
    num = (ax2 * ay2 * f[i - 1, j - 1] + ax1 * ay2 * f[i, j - 1]
           + ax2 * ay1 * f[i - 1, j] + ax1 * ay1 * f[i, j])
    den = ((x2 - x1) * (y2 - y1))[:, None]
    out = num / den
    outside = (xq < x[0]) | (xq > x[-1]) | (yq < y[0]) | (yq > y[-1])
    out = np.where(outside[:, None], dtype(0), out).astype(dtype)
    return out


def scale_spectrum_by_mass(spectra, mass):
    """rubix/core/ifu.py:149-157: spectra * mass[..., None]."""
    return spectra * np.expand_dims(mass, -1).astype(spectra.dtype)


def cosmological_doppler_shift(z, wavelength, dtype=np.float32):
    """rubix/spectra/ifu.py:66-80: (1 + z) * wavelength (python-float
    factor, evaluated in the array's float type)."""
    return (dtype(1 + z) * np.asarray(wavelength, dtype=dtype)).astype(dtype)


def velocity_doppler_shift(wavelength, velocity, direction="z",
                           c=SPEED_OF_LIGHT, dtype=np.float32):
    """rubix/spectra/ifu.py:168-220: w[i, k] = wavelength[k] * exp(v_i / c)
    for the chosen velocity component (rubix_config.yml:245 uses "z")."""
    comp = {"x": 0, "y": 1, "z": 2}
    if direction not in comp:
        raise ValueError(
            f"{direction} is not a valid direction. Supported directions are 'x', 'y', or 'z'.")
    v = np.asarray(velocity, dtype=dtype)[:, comp[direction]]
    # v / c in the array's float type; exp correctly rounded to that type
    # (evaluated in float64 and rounded once)
    a = np.exp((v / dtype(c)).astype(np.float64)).astype(dtype)
    return (np.asarray(wavelength, dtype=dtype)[None, :] * a[:, None]).astype(dtype)


def calculate_diff(vec, pad_with_zero=True):
    """rubix/spectra/ifu.py:83-102: diff with the first element prepended
    (so the first difference is 0)."""
    vec = np.asarray(vec)
    if pad_with_zero:
        return np.diff(vec, prepend=vec[..., :1], axis=-1)
    return np.diff(vec, axis=-1)


def interp_jax(x, xp, fp, left=None):
    """``jnp.interp`` semantics (used at rubix/spectra/ifu.py:250):
    i = clip(searchsorted(xp, x, 'right'), 1, n-1);
    f = fp[i-1] + (x - xp[i-1]) / (xp[i] - xp[i-1]) * (fp[i] - fp[i-1]);
    f = fp[0] where x < xp[0] (unless left == "extrapolate");
    f = fp[-1] where x > xp[-1]."""
    x = np.asarray(x)
    xp = np.asarray(xp)
    fp = np.asarray(fp)
    n = len(xp)
    i = np.clip(np.searchsorted(xp, x, side="right"), 1, n - 1)
    df = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = x - xp[i - 1]
    eps = np.spacing(np.finfo(xp.dtype).eps)
    dx0 = np.abs(dx) <= eps
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where(dx0, fp[i - 1], fp[i - 1] + (delta / np.where(dx0, 1, dx)) * df)
    if left != "extrapolate":
        f = np.where(x < xp[0], fp[0], f)
    f = np.where(x > xp[-1], fp[-1], f)
    return f.astype(fp.dtype)


def nan_to_num(x):
    """``jnp.nan_to_num(x, nan=0.0)``: nan -> 0, +-inf -> +-finfo.max."""
    return np.nan_to_num(x, nan=0.0)


def resample_spectrum(initial_spectrum, initial_wavelength, target_wavelength):
    """rubix/spectra/ifu.py:223-266 for one spectrum.

    mask = (w >= min(target)) & (w <= max(target));
    T = sum(S * diff(w) * mask); P = interp(target, w, S);
    Nn = sum(P * diff(target)); out = P * nan_to_num(T / Nn)."""
    S = np.asarray(initial_spectrum)
    w = np.asarray(initial_wavelength)
    tw = np.asarray(target_wavelength)
    m = (w >= np.min(tw)) & (w <= np.max(tw))
    iwd = calculate_diff(w) * m
    total = np.sum(S * iwd, dtype=S.dtype)
    P = interp_jax(tw, w, S)
    new_total = np.sum(P * calculate_diff(tw), dtype=S.dtype)
    with np.errstate(divide="ignore", invalid="ignore"):
        sf = S.dtype.type(total) / S.dtype.type(new_total)
    sf = nan_to_num(sf)
    return (P * sf).astype(S.dtype)


def resample_spectra(spectra, wavelengths, target_wavelength):
    """vmap of resample_spectrum over stars (rubix/core/ifu.py:163-199)."""
    out = np.empty((spectra.shape[0], len(target_wavelength)), dtype=spectra.dtype)
    for i in range(spectra.shape[0]):
        out[i] = resample_spectrum(spectra[i], wavelengths[i], target_wavelength)
    return out


def calculate_cube(spectra, spaxel_index, num_spaxels):
    """rubix/spectra/ifu.py:269-288: segment_sum over spaxel ids into
    num_spaxels**2 segments (ids outside [0, nb^2) dropped), reshaped to
    (nb, nb, n_wave)."""
    spectra = np.asarray(spectra)
    ids = np.asarray(spaxel_index)
    ns = num_spaxels * num_spaxels
    out = np.zeros((ns, spectra.shape[-1]), dtype=spectra.dtype)
    ok = (ids >= 0) & (ids < ns)
    np.add.at(out, ids[ok], spectra[ok])
    return out.reshape(num_spaxels, num_spaxels, spectra.shape[-1])


# ---------------------------------------------------------------- PSF / LSF

def gaussian_kernel_2d(m, n, sigma, dtype=np.float32):
    """rubix/telescope/psf/kernels.py:5-31 (sigma in pixels)."""
    x = np.arange(-((m - 1) / 2), ((m - 1) / 2) + 1).astype(dtype)
    y = np.arange(-((n - 1) / 2), ((n - 1) / 2) + 1).astype(dtype)
    X, Y = np.meshgrid(x, y, indexing="ij")
    s = dtype(sigma)
    v = np.exp(-(X ** 2 + Y ** 2) / (dtype(2) * s ** 2)).astype(dtype)
    return (v / np.sum(v, dtype=dtype)).astype(dtype)


def convolve2d_same(plane, kernel):
    """Zero-padded true 2D convolution cropped to the input size, centred as
    ``convolve2d(mode="same")`` (rubix/telescope/psf/psf.py:9-11): the crop
    starts at (k-1)//2 of the full output on each axis."""
    plane = np.asarray(plane)
    kernel = np.asarray(kernel)
    ny, nx = plane.shape[:2]
    ky, kx = kernel.shape
    sy, sx = (ky - 1) // 2, (kx - 1) // 2
    out = np.zeros_like(plane)
    for a in range(ky):
        for b in range(kx):
            dy = sy - a
            dx = sx - b
            # out[y, x] += K[a, b] * in[y + dy, x + dx]
            y0, y1 = max(0, -dy), min(ny, ny - dy)
            x0, x1 = max(0, -dx), min(nx, nx - dx)
            if y0 >= y1 or x0 >= x1:
                continue
            out[y0:y1, x0:x1] += kernel[a, b] * plane[y0 + dy:y1 + dy, x0 + dx:x1 + dx]
    return out


def apply_psf(datacube, psf_kernel):
    """rubix/telescope/psf/psf.py:33-58: convolve every lambda plane of the
    (ny, nx, n_wave) cube with the kernel, mode "same".  The plane loop is
    vectorised over lambda (the cube's trailing axis)."""
    return convolve2d_same(np.asarray(datacube), np.asarray(psf_kernel))


def gaussian_taps_1d(m, sigma, dtype=np.float64):
    """The 1D factor g of gaussian_kernel_2d: exp(-x^2 / 2 sigma^2) over
    x = arange(-(m-1)/2, (m-1)/2 + 1), normalised to sum 1.  In exact
    arithmetic gaussian_kernel_2d(m, m, sigma) = outer(g, g), since
    exp(-(x^2 + y^2)/2s^2) = exp(-x^2/2s^2) exp(-y^2/2s^2) and the sum of the
    outer product is the product of the sums."""
    x = np.arange(-((m - 1) / 2), ((m - 1) / 2) + 1).astype(dtype)
    v = np.exp(-(x ** 2) / (dtype(2) * dtype(sigma) ** 2))
    return (v / v.sum()).astype(dtype)


def apply_psf_separable(datacube, g):
    """apply_psf with the kernel outer(g, g): the 'same' zero-padded
    convolution along the y axis, then along the x axis (a separable
    restatement of convolve2d(mode="same") for outer-product kernels, for
    cubes where the direct sum over m^2 taps is too slow)."""
    cube = np.asarray(datacube)
    g = np.asarray(g)
    m = len(g)
    s = (m - 1) // 2

    def along(a, axis):
        n = a.shape[axis]
        out = np.zeros_like(a)
        for t in range(m):
            d = s - t  # out[i] += g[t] * a[i + d]
            lo, hi = max(0, -d), min(n, n - d)
            if lo >= hi:
                continue
            dst = [slice(None)] * a.ndim
            src = [slice(None)] * a.ndim
            dst[axis] = slice(lo, hi)
            src[axis] = slice(lo + d, hi + d)
            out[tuple(dst)] += g[t] * a[tuple(src)]
        return out

    return along(along(cube, 0), 1)


def gaussian1d(x, sigma):
    """rubix/telescope/lsf/lsf.py:12-15."""
    x = np.asarray(x)
    r = np.exp(-0.5 * (x ** 2) / x.dtype.type(sigma) ** 2)
    return r / np.sum(r, dtype=x.dtype)


def lsf_kernel(sigma, wave_res, factor=12, dtype=np.float32):
    """rubix/telescope/lsf/lsf.py:22-26: x = arange(-f*dl, f*dl + dl, dl)
    (25 taps for factor 12 when the float arange is exact), Gaussian,
    normalised to sum 1.  sigma in Angstrom."""
    x = np.arange(-factor * wave_res, factor * wave_res + wave_res, wave_res).astype(dtype)
    return gaussian1d(x, sigma).astype(dtype)


def apply_lsf(datacube, lsf_sigma, wave_resolution, extend_factor=12, dtype=None):
    """rubix/telescope/lsf/lsf.py:29-105: flatten to (n_spec, n_wave), full
    convolution with the LSF kernel along lambda, crop
    [factor : n + len(k) - 1 - factor], reshape back."""
    cube = np.asarray(datacube)
    dt = dtype or cube.dtype.type
    k = lsf_kernel(lsf_sigma, wave_resolution, extend_factor, dtype=dt)
    return apply_lsf_kernel(cube, k, extend_factor)


def apply_lsf_kernel(datacube, k, extend_factor=12):
    """Full convolution of every spectrum with k, cropped to
    [extend_factor, n + L - 1 - extend_factor) (rubix/telescope/lsf/lsf.py:61-65).
    With an even tap count L that crop holds n + 1 bins and the reference's
    reshape to the cube raises (lsf.py:102-105); the first n bins are returned
    here (the kernel's x = 0 tap on the input bin, as for odd L)."""
    cube = np.asarray(datacube)
    shape = cube.shape
    spec = cube.reshape(-1, shape[-1])
    n = spec.shape[1]
    L = len(k)
    full = np.zeros((spec.shape[0], n + L - 1), dtype=spec.dtype)
    for t in range(L):
        full[:, t:t + n] += k[t] * spec
    end = n + L - 1 - extend_factor
    crop = full[:, extend_factor:end]
    return crop[:, :n].reshape(shape)


# ---------------------------------------------------------------- noise

SUPPORTED_NOISE_DISTRIBUTIONS = ["normal", "uniform"]  # rubix/telescope/noise/noise.py:5
_U32 = np.uint32


def _rotl32(x, r):
    return (x << _U32(r)) | (x >> _U32(32 - r))


def threefry2x32(k0, k1, c0, c1):
    """Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror, Shaw, "Parallel
    random numbers: as easy as 1, 2, 3", SC 2011), vectorised over the
    counter words c0, c1 (uint32 arrays): the block function behind
    jax.random's default PRNG."""
    rot = (13, 15, 26, 6, 17, 29, 16, 24)
    ks = (_U32(k0), _U32(k1), _U32(k0) ^ _U32(k1) ^ _U32(0x1BD11BDA))
    x0 = np.asarray(c0, dtype=_U32) + ks[0]
    x1 = np.asarray(c1, dtype=_U32) + ks[1]
    for r in range(20):
        x0 = x0 + x1
        x1 = _rotl32(x1, rot[r % 8])
        x1 = x1 ^ x0
        if r % 4 == 3:
            i = r // 4 + 1
            x0 = x0 + ks[i % 3]
            x1 = x1 + ks[(i + 1) % 3] + _U32(i)
    return x0, x1


def random_bits(seed, n):
    """32-bit words for flat indices 0..n-1 under key PRNGKey(seed) =
    (seed >> 32, seed & 0xffffffff): Threefry of the counter (i >> 32,
    i & 0xffffffff), the two output words XOR-ed (jax.random's counter layout
    with jax_threefry_partitionable, the default from jax 0.5)."""
    i = np.arange(n, dtype=np.uint64)
    hi = (i >> np.uint64(32)).astype(_U32)
    lo = (i & np.uint64(0xFFFFFFFF)).astype(_U32)
    seed = int(seed)
    y0, y1 = threefry2x32(seed >> 32, seed & 0xFFFFFFFF, hi, lo)
    return y0 ^ y1


def sample_noise(shape, type="normal", seed=0):
    """rubix/telescope/noise/noise.py:8-32 (key = PRNGKey(seed); the reference
    passes seed 0).  Uniform: the top 23 bits as the mantissa of a float in
    [1, 2), minus 1.  Normal: sqrt(2) erfinv(u), u uniform on
    [nextafter(-1, 0), 1) (jax.random.uniform / normal), float32 throughout
    with erfinv evaluated in float64 and rounded."""
    from scipy.special import erfinv

    if type not in SUPPORTED_NOISE_DISTRIBUTIONS:
        raise ValueError(f"Invalid noise type: {type}. Supported types: "
                         f"{SUPPORTED_NOISE_DISTRIBUTIONS}")
    n = int(np.prod(shape))
    bits = random_bits(seed, n)
    f = ((bits >> _U32(9)) | _U32(0x3F800000)).view(np.float32) - np.float32(1.0)
    if type == "uniform":
        return f.reshape(shape)
    lo = np.nextafter(np.float32(-1.0), np.float32(0.0))
    u = np.maximum(lo, f * np.float32(2.0) + lo)
    e = erfinv(u.astype(np.float64)).astype(np.float32)
    return (np.float32(np.sqrt(2.0)) * e).reshape(shape)


def flux_median(flux):
    """jnp.median of where(flux > 0, flux, NaN) then nan_to_num
    (rubix/telescope/noise/noise.py:62-69): NaN (hence 0) as soon as one
    spaxel has no positive flux, else the midpoint (lo + hi) * 0.5 of the
    ranks floor((n-1)/2) and ceil((n-1)/2).  Also returns those two spaxels
    in a stable sort's order (-1, -1 for a zero median)."""
    f = np.asarray(flux).reshape(-1)
    if f.size == 0 or not np.all(f > 0):
        return f.dtype.type(0), -1, -1
    order = np.argsort(f, kind="stable")
    n = f.size
    lo, hi = order[(n - 1) // 2], order[n // 2]
    return (f[lo] + f[hi]) * f.dtype.type(0.5), int(lo), int(hi)


def flux_image(datacube, dtype=np.float32):
    """rubix/telescope/noise/noise.py:62: sum over the wavelength axis (in
    float64, rounded once to dtype)."""
    return np.asarray(datacube, dtype=np.float64).sum(axis=-1).astype(dtype)


def calculate_S2N(datacube, signal_to_noise, dtype=np.float32, flux=None):
    """rubix/telescope/noise/noise.py:35-78 (signal_to_noise a number or the
    per-spaxel array the apply_noise step passes):
    S = where(flux > 0, (sqrt(median) / sn) / sqrt(flux), 0), each operation
    in dtype."""
    F = flux_image(datacube, dtype) if flux is None else np.asarray(flux, dtype=dtype)
    med, _, _ = flux_median(F)
    sn = np.broadcast_to(np.asarray(signal_to_noise, dtype=dtype), F.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        S = (np.sqrt(dtype(med)) / sn) / np.sqrt(F)
    return np.where(F > 0, S, dtype(0)).astype(dtype)


def apply_noise(datacube, signal_to_noise, distribution="normal", seed=0, dtype=np.float32,
                flux=None):
    """The apply_noise step (rubix/core/noise.py:67-78 over calculate_noise_cube,
    rubix/telescope/noise/noise.py:81-115): cube + cube * (z * S[..., None])."""
    c = np.asarray(datacube, dtype=dtype)
    z = sample_noise(c.shape, distribution, seed).astype(dtype)
    S = calculate_S2N(c, signal_to_noise, dtype, flux)
    return c + c * (z * S[..., None])


def apply_noise_vjp(datacube, signal_to_noise, g, distribution="normal", seed=0):
    """Adjoint of apply_noise with respect to the cube, float64 analytic:
    d out / d c = 1 + z S (direct) plus the flux path, dS/dF = -S/(2F) and
    dS/dM = S/(2M) with M moving with the flux of the median spaxels (weight
    1/2 for each midpoint rank).  M = 0 (a spaxel without positive flux)
    leaves S = 0 and the adjoint g."""
    c = np.asarray(datacube, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    z = sample_noise(c.shape, distribution, seed).astype(np.float64)
    F = c.sum(axis=-1)
    med, lo, hi = flux_median(F)
    if lo < 0:
        return g.copy()
    S = calculate_S2N(c, signal_to_noise, np.float64, F)
    A = (g * c * z).sum(axis=-1)
    dM = (A * S).sum() / (2.0 * med)
    dF = (-A * S / (2.0 * F)).reshape(-1)
    dF[lo] += 0.5 * dM
    dF[hi] += 0.5 * dM
    return g + g * z * S[..., None] + dF.reshape(F.shape)[..., None]


# ---------------------------------------------------------------- dust

# rubix/config/rubix_config.yml:3-7, 145-150
MSUN_TO_GRAMS = 1.989e33
KPC_TO_CM = 3.08568e21
RV_MODELS = ["Cardelli89", "Gordon23"]  # rubix/spectra/dust/extinction_models.py:10


def powerlaw1d(x, amplitude, x_0, alpha):
    """rubix/spectra/dust/generic_models.py:18-48: A (x / x0)^-alpha."""
    return amplitude * (x / x_0) ** (-alpha)


def polynomial1d(x, coeffs):
    """generic_models.py:51-97 with the default domain = window = (-1, 1)
    (the identity map): Horner from the highest coefficient."""
    coeffs = list(coeffs)
    if len(coeffs) == 1:
        return coeffs[-1] * np.ones_like(x)
    c0 = coeffs[-1]
    for i in range(2, len(coeffs) + 1):
        c0 = coeffs[-i] + c0 * x
    return c0


def drude1d(x, amplitude=1.0, x_0=1.0, fwhm=1.0):
    """generic_models.py:100-155."""
    if x_0 == 0:
        raise ValueError("0 is not an allowed value for x_0")
    return amplitude * ((fwhm / x_0) ** 2) / ((x / x_0 - x_0 / x) ** 2 + (fwhm / x_0) ** 2)


def modified_drude(x, scale, x_o, gamma_o, asym):
    """generic_models.py:158-194."""
    gamma = 2.0 * gamma_o / (1.0 + np.exp(asym * (x - x_o)))
    return scale * ((gamma / x_o) ** 2) / ((x / x_o - x_o / x) ** 2 + (gamma / x_o) ** 2)


def fm90(x, C1=0.10, C2=0.70, C3=3.23, C4=0.41, xo=4.59, gamma=0.95):
    """generic_models.py:197-335 (Fitzpatrick & Massa 1990), with the
    reference's parameter bounds and messages."""
    bounds = {"C1": (-10.0, 5.0), "C2": (-0.1, 5.0), "C3": (-1.0, 6.0), "C4": (-0.5, 1.5),
              "xo": (4.5, 4.9), "gamma": (0.6, 1.7)}
    for k, v in (("C1", C1), ("C2", C2), ("C3", C3), ("C4", C4), ("xo", xo), ("gamma", gamma)):
        if not (bounds[k][0] <= v <= bounds[k][1]):
            raise ValueError(f"{k} is out of bounds: {v}")
    e = C1 + C2 * x
    x2 = x ** 2
    e = e + C3 * (x2 / ((x2 - xo ** 2) ** 2 + x2 * (gamma ** 2)))
    m = x >= 5.9
    y = np.where(m, x - 5.9, 0.0)
    return np.where(m, e + C4 * (0.5392 * (y ** 2) + 0.05644 * (y ** 3)), e)


def smoothstep1(x, x_min, x_max):
    """helpers.py:68-84 with N = 1: t = clip((x - x_min)/(x_max - x_min), 0, 1),
    (comb(1,0) comb(3,1) - comb(2,1) comb(3,0) t) t^2 = (3 - 2t) t^2."""
    t = np.clip((x - x_min) / (x_max - x_min), 0, 1)
    return (3.0 - 2.0 * t) * t ** 2


# CCM89 coefficients (rubix/spectra/dust/extinction_models.py:124-177):
# IR a, b = coeff * x^1.61; optical 7th-order polynomials in y = x - 1.82
# (lowest order first, constant 1 / 0); NUV (c0, c1, c2, x_c, w) of
# c0 + c1 x + c2 / ((x - x_c)^2 + w); far-NUV and FUV polynomial terms.
_CCM89_IR = (0.574, -0.527, 1.61)
_CCM89_OPT_A = [1.0, 0.17699, -0.50447, -0.02427, 0.72085, 0.01979, -0.77530, 0.32999]
_CCM89_OPT_B = [0.0, 1.41338, 2.28305, 1.07233, -5.38434, -0.62251, 5.30260, -2.09002]
_CCM89_NUV_A = (1.752, -0.316, -0.104, 4.67, 0.341)
_CCM89_NUV_B = (-3.09, 1.825, 1.206, 4.62, 0.263)
_CCM89_FNUV_A = (-0.04473, -0.009779)   # (y^2, y^3), y = x - 5.9
_CCM89_FNUV_B = (0.2130, 0.1207)
_CCM89_FUV_A = [-1.073, -0.628, 0.137, -0.070]   # y = x - 8, lowest order first
_CCM89_FUV_B = [13.670, 4.257, -0.420, 0.374]


def _poly(y, c):
    return sum(ci * y ** i for i, ci in enumerate(c))


def cardelli89(wave, Rv):
    """A(x)/A(V) of rubix/spectra/dust/extinction_models.py:104-178 (CCM89).
    The branches test the argument as a wavenumber x in 1/micron; the
    reference's dust step passes wavelength / 1e4, i.e. microns (see
    apply_spaxel_extinction)."""
    w = np.asarray(wave, dtype=np.float64)
    a = np.zeros_like(w)
    b = np.zeros_like(w)
    ir = (0.3 <= w) & (w < 1.1)
    opt = (1.1 <= w) & (w < 3.3)
    nuv = (3.3 <= w) & (w <= 8.0)
    fnuv = (5.9 <= w) & (w <= 8)
    fuv = (8 < w) & (w <= 10)
    ia, ib, ie = _CCM89_IR
    a = np.where(ir, ia * w ** ie, a)
    b = np.where(ir, ib * w ** ie, b)
    y = w - 1.82
    a = np.where(opt, _poly(y, _CCM89_OPT_A), a)
    b = np.where(opt, _poly(y, _CCM89_OPT_B), b)
    with np.errstate(divide="ignore", invalid="ignore"):
        for c, arr in ((_CCM89_NUV_A, "a"), (_CCM89_NUV_B, "b")):
            v = c[0] + c[1] * w + c[2] / ((w - c[3]) ** 2 + c[4])
            if arr == "a":
                a = np.where(nuv, v, a)
            else:
                b = np.where(nuv, v, b)
    y = w - 5.9
    a = np.where(fnuv, a + (_CCM89_FNUV_A[0] * y ** 2 + _CCM89_FNUV_A[1] * y ** 3), a)
    b = np.where(fnuv, b + (_CCM89_FNUV_B[0] * y ** 2 + _CCM89_FNUV_B[1] * y ** 3), b)
    y = w - 8.0
    a = np.where(fuv, _poly(y, _CCM89_FUV_A), a)
    b = np.where(fuv, _poly(y, _CCM89_FUV_B), b)
    return a + b / Rv


_G23_IR_A = [0.38526, 1.68467, 0.78791, 4.30578, 4.78338,
             0.06652, 9.8434, 2.21205, -0.24703, 0.0267, 19.58294, 17., -0.27]
_G23_OPT_A = [-0.35848, 0.7122, 0.08746, -0.05403, 0.00674,
              0.03893, 2.288, 0.243, 0.02965, 2.054, 0.179, 0.01747, 1.587, 0.243]
_G23_OPT_B = [0.12354, -2.68335, 2.01901, -0.39299, 0.03355,
              0.18453, 2.288, 0.243, 0.19728, 2.054, 0.179, 0.1713, 1.587, 0.243]
_G23_IR_B = (-1.01251, 1.0, -1.06099)        # power law (amplitude, x_0, alpha)
_G23_UV_A = (0.81297, 0.2775, 1.06295, 0.11303, 4.60, 0.99)    # FM90 C1..C4, xo, gamma
_G23_UV_B = (-2.97868, 1.89808, 3.10334, 0.65484, 4.60, 0.99)
# wavelength bounds (micron): branches and the two smoothstep overlaps
_G23_BRANCHES = {"ir": (1.0, 35.0), "opt": (0.3, 1.1), "uv": (0.09, 0.3),
                 "optir": (0.9, 1.1), "uvopt": (0.3, 0.33)}


def _g23_nirmir(w, p):
    """extinction_models.py:401-445: broken power law + two modified Drudes."""
    scale, alpha, alpha2, swave, swidth = p[:5]
    pow1 = scale * w ** (-1.0 * alpha)
    ratio = swave ** (-1.0 * alpha) / swave ** (-1.0 * alpha2)
    pow2 = scale * ratio * w ** (-1.0 * alpha2)
    wt = smoothstep1(w, swave - swidth / 2, swave + swidth / 2)
    ax = pow1 * (1.0 - wt) + pow2 * wt
    ax = ax + modified_drude(w, *p[5:9])
    return ax + modified_drude(w, *p[9:13])


def _g23_optical(x, p):
    """extinction_models.py:315-351: polynomial (5 coefficients) + 3 Drudes."""
    return (polynomial1d(x, p[:5]) + drude1d(x, p[5], p[6], p[7]) + drude1d(x, p[8], p[9], p[10])
            + drude1d(x, p[11], p[12], p[13]))


def gordon23(wave, Rv):
    """A(lambda)/A(V) of extinction_models.py:255-398 (G23), wave in microns."""
    w = np.asarray(wave, dtype=np.float64)
    a = np.zeros_like(w)
    b = np.zeros_like(w)
    B = _G23_BRANCHES
    ir = (B["ir"][0] <= w) & (w < B["ir"][1])
    opt = (B["opt"][0] <= w) & (w < B["opt"][1])
    uv = (B["uv"][0] <= w) & (w <= B["uv"][1])
    optir = (w >= B["optir"][0]) & (w <= B["optir"][1])
    uvopt = (w >= B["uvopt"][0]) & (w <= B["uvopt"][1])
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ir_a = _g23_nirmir(w, _G23_IR_A)
        ir_b = powerlaw1d(w, *_G23_IR_B)
        op_a = _g23_optical(1 / w, _G23_OPT_A)
        op_b = _g23_optical(1 / w, _G23_OPT_B)
        uv_a = fm90(1 / w, *_G23_UV_A)
        uv_b = fm90(1 / w, *_G23_UV_B)
    a = np.where(ir, ir_a, a)
    b = np.where(ir, ir_b, b)
    a = np.where(opt, op_a, a)
    b = np.where(opt, op_b, b)
    wt = smoothstep1(w, *B["optir"])
    a = np.where(optir, (1.0 - wt) * op_a + wt * ir_a, a)
    b = np.where(optir, (1.0 - wt) * op_b + wt * ir_b, b)
    a = np.where(uv, uv_a, a)
    b = np.where(uv, uv_b, b)
    wt = smoothstep1(w, *B["uvopt"])
    a = np.where(uvopt, (1.0 - wt) * uv_a + wt * op_a, a)
    b = np.where(uvopt, (1.0 - wt) * uv_b + wt * op_b, b)
    return a + b * (1 / Rv - 1 / 3.1)


def extinction_curve(model, wave_um, Rv):
    if model not in RV_MODELS:
        raise ValueError(f"Extinction model '{model}' is not available. Choose from {RV_MODELS}.")
    return cardelli89(wave_um, Rv) if model == "Cardelli89" else gordon23(wave_um, Rv)


def extinguish(axav, Av):
    """dust_baseclasses.py:125-164: 10^(-0.4 A(x)/A(V) Av)."""
    return np.power(10.0, -0.4 * axav * Av)


def dust_to_gas_ratio(log_OH, model, Xco):
    """rubix/spectra/dust/dust_extinction.py:13-90 (Remy-Ruyer et al. 2014,
    table 1)."""
    x_sol = 8.69
    if model == "power law slope fixed":
        raise NotImplementedError("power law slope fixed not implemented yet.")
    table = {("MW", "power law slope free"): (2.21, 1.62),
             ("Z", "power law slope free"): (2.21, 2.02)}
    if (Xco, model) in table:
        a, alpha = table[(Xco, model)]
        return 1 / 10 ** (a + alpha * (x_sol - log_OH))
    broken = {"MW": (2.21, 1.00, 0.68, 3.08, 7.96), "Z": (2.21, 1.00, 0.96, 3.10, 8.10)}
    if model == "broken power law fit" and Xco in broken:
        a, ah, b, al, xt = broken[Xco]
        return 1 / np.where(log_OH > xt, 10 ** (a + ah * (x_sol - log_OH)),
                            10 ** (b + al * (x_sol - log_OH)))
    raise ValueError(f"unknown dust-to-gas model {model!r} / Xco {Xco!r}")


def gas_extinction(gas_mass, gas_metals, model, Xco, grain_density, spaxel_area):
    """Per gas cell: log(O/H) + 12 from metals[:, 4] / (16 metals[:, 0]),
    dust mass = mass * D/G, and calculate_extinction
    (dust_extinction.py:93-160, 3 pi Sigma C / (0.4 ln 10 lambda_V rho)) over
    the spaxel area (dust_extinction.py:269-287)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        log_OH = 12 + np.log10(gas_metals[:, 4] / (16 * gas_metals[:, 0]))
        dg = dust_to_gas_ratio(log_OH, model, Xco)
        dust = gas_mass * dg
        conv = MSUN_TO_GRAMS / KPC_TO_CM ** 2
        ext = 3 * np.pi * dust * conv / (0.4 * np.log(10) * (5448 * 1e-8) * grain_density)
    return ext / spaxel_area


def column_av(gas_z, gas_pix, gas_ext, star_z, star_pix, n_spaxels):
    """Av per star, the scan of dust_extinction.py:237-340 restated literally:
    gas in lexsort (pixel, z) order; for each spaxel s the keys are
    z * where(own, 1, 1e30) (float32), the values the masked cumulative sum
    of the sorted extinction, stably sorted by key, and jnp.interp with
    left = "extrapolate" at each star's z; a star keeps the value of its
    own spaxel.  gas_ext is per gas cell in input order.  float32."""
    f = np.float32
    gz = np.asarray(gas_z, f)
    order = np.lexsort((gz, np.asarray(gas_pix)))
    gz_s = gz[order]
    gp_s = np.asarray(gas_pix)[order]
    ext_s = np.asarray(gas_ext, f)[order]
    sz = np.asarray(star_z, f)
    sp = np.asarray(star_pix)
    av = np.zeros(len(sz), f)
    big = f(1e30)
    for s in range(n_spaxels):
        own = gp_s == s
        with np.errstate(invalid="ignore", over="ignore"):
            cum = np.cumsum(ext_s * own.astype(f), dtype=f) * own.astype(f)
            keys = gz_s * np.where(own, f(1), big)
        o = np.argsort(keys, kind="stable")
        xp, fp = keys[o], cum[o]
        stars = sp == s
        if not stars.any() or len(xp) == 0:
            continue
        if len(xp) == 1:
            av[stars] = fp[0]
            continue
        with np.errstate(invalid="ignore", divide="ignore"):
            av[stars] = interp_jax(sz[stars], xp, fp, left="extrapolate")
    return av


def apply_spaxel_extinction(spectra, wave_target, av, model, Rv):
    """spectra (N, n_wave) times 10^(-0.4 A(lambda)/A(V) Av) with the curve
    evaluated at wave_target / 1e4 (dust_extinction.py:330-358)."""
    ax = extinction_curve(model, np.asarray(wave_target, np.float64) / 1e4, Rv)
    return np.asarray(spectra) * extinguish(ax[None, :], np.asarray(av)[:, None])


# ---------------------------------------------------------------- path

def plain_forward(stars, grid, geom, dtype=np.float32, chunk=4096,
                  return_intermediates=False, ref_knots=True, psf="direct", raw_only=False):
    """The calc_ifu step chain (rubix/config/pipeline_config.yml:1-60) from
    spaxel_assignment to convolve_lsf, at D = 1:

    spaxel_assignment -> calculate_spectra -> scale_spectrum_by_mass ->
    doppler_shift_and_resampling -> calculate_datacube -> convolve_psf ->
    convolve_lsf.

    ``stars``: dict coords (N,3), velocity (N,3), mass, age, metallicity.
    ``grid``: dict ax_met, ax_age, wave, flux, log_axes.
    ``geom``: dict edges, wave_target, z, psf_size, psf_sigma (pixels),
    lsf_sigma (Angstrom), wave_res; the kernels are built in ``dtype``.
    Stars are processed in chunks; the cube accumulates across chunks.

    ``ref_knots`` (float64 only): the shifted source knots are the float32
    values the reference computes, fl32(fl32((1+z) wave) * fl32(exp(v/c))),
    and everything after them is float64.  The knots are inputs of the
    interpolation whose rounding moves interpolation weights by ~1e-4, so this
    is the float64 evaluation of the reference's own float32 algorithm;
    ``ref_knots=False`` gives exact arithmetic throughout.
    ``psf="separable"`` applies the PSF as two 1D passes (apply_psf_separable),
    equal to the direct 2D sum in exact arithmetic, for large cubes.
    ``raw_only`` returns the datacube before the PSF (calculate_datacube's
    output).
    """
    edges = np.asarray(geom["edges"], dtype=np.float32)
    nb = len(edges) - 1
    ids = square_spaxel_assignment(np.asarray(stars["coords"], dtype=np.float32), edges)
    wave_t = np.asarray(geom["wave_target"], dtype=dtype)
    kd = np.float32 if (ref_knots or dtype == np.float32) else dtype
    ssp_wave = cosmological_doppler_shift(geom["z"], grid["wave"], dtype=kd)
    n = len(ids)
    cube = np.zeros((nb * nb, len(wave_t)), dtype=dtype)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        spec = lookup_linear(stars["metallicity"][s:e], stars["age"][s:e],
                             grid["ax_met"], grid["ax_age"], grid["flux"],
                             log_axes=grid.get("log_axes", False), dtype=dtype)
        spec = scale_spectrum_by_mass(spec, np.asarray(stars["mass"][s:e], dtype=dtype))
        w = velocity_doppler_shift(ssp_wave, stars["velocity"][s:e], "z", dtype=kd).astype(dtype)
        res = resample_spectra(spec, w, wave_t)
        part = calculate_cube(res, ids[s:e], nb).reshape(nb * nb, -1)
        cube += part
    cube = cube.reshape(nb, nb, -1)
    if raw_only:
        return cube
    m = geom["psf_size"]
    if psf == "separable":
        c_psf = apply_psf_separable(cube, gaussian_taps_1d(m, geom["psf_sigma"], dtype=dtype))
    else:
        K = gaussian_kernel_2d(m, m, geom["psf_sigma"], dtype=dtype)
        c_psf = apply_psf(cube, K)
    c_lsf = apply_lsf_kernel(c_psf, lsf_kernel(geom["lsf_sigma"], geom["wave_res"], dtype=dtype))
    if return_intermediates:
        return dict(ids=ids, cube=cube, cube_psf=c_psf, cube_lsf=c_lsf)
    return c_lsf
