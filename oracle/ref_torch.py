"""torch (CPU) restatement of the plain per-step path, for autograd.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference has no adjoint (rubix/core/pipeline.py:203-208), and jax is not
available here, so ``jax.grad`` of the reference is replaced by torch
autograd through the same op graph: the same searchsorted/clip cell choice,
the same ``where`` branches for interpolation clamping and the out-of-grid
zero, the in-range mask without gradient and ``nan_to_num`` whose gradient is
zero at non-finite values.  In float64 this is the gradient oracle; in
float32 on all host cores it is the timed CPU baseline of the plain path.

Functions cite the reference lines they mirror; ``ref_numpy`` holds the
forward definitions they are checked against.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

SPEED_OF_LIGHT = 299792.458  # rubix/config/rubix_config.yml:8


def lookup_linear(Z, t, ax_met, ax_age, flux, log_axes=False):
    """interp2d(method='linear', extrap=0) (rubix/spectra/ssp/grid.py:113-120);
    flux (n_met, n_age, n_wave); returns (n, n_wave)."""
    xq = torch.log10(Z) if log_axes else Z
    yq = torch.log10(t) if log_axes else t
    nx, ny = ax_met.numel(), ax_age.numel()
    i = torch.searchsorted(ax_met, xq.detach().contiguous(), right=True).clamp(1, nx - 1)
    j = torch.searchsorted(ax_age, yq.detach().contiguous(), right=True).clamp(1, ny - 1)
    x1, x2 = ax_met[i - 1], ax_met[i]
    y1, y2 = ax_age[j - 1], ax_age[j]
    ax2 = (x2 - xq)[:, None]
    ax1 = (xq - x1)[:, None]
    ay2 = (y2 - yq)[:, None]
    ay1 = (yq - y1)[:, None]
    num = (ax2 * ay2 * flux[i - 1, j - 1] + ax1 * ay2 * flux[i, j - 1]
           + ax2 * ay1 * flux[i - 1, j] + ax1 * ay1 * flux[i, j])
    out = num / ((x2 - x1) * (y2 - y1))[:, None]
    outside = (xq < ax_met[0]) | (xq > ax_met[-1]) | (yq < ax_age[0]) | (yq > ax_age[-1])
    # where() with a zero branch: gradient flows only inside the grid.
    # Queries that are non-finite (log10(0)) would give nan in the unused
    # branch; mask them before the arithmetic so no nan reaches the grad.
    return torch.where(outside[:, None], torch.zeros_like(out), out)


def safe_queries(Z, t, log_axes, ax_met, ax_age):
    """Replace queries that are outside the grid by an inside value (the
    result is zeroed by the out-of-grid ``where`` anyway) so that the unused
    branch of ``where`` carries no inf/nan into the backward pass."""
    xq = torch.log10(Z) if log_axes else Z
    yq = torch.log10(t) if log_axes else t
    outside = (xq < ax_met[0]) | (xq > ax_met[-1]) | (yq < ax_age[0]) | (yq > ax_age[-1])
    outside = outside | ~torch.isfinite(xq) | ~torch.isfinite(yq)
    return outside


def interp(x, xp, fp):
    """Batched ``jnp.interp``: x (m,), xp/fp (B, n) -> (B, m).
    i = clip(searchsorted(xp, x, right), 1, n-1); lerp; clamp to fp[0] /
    fp[-1] outside [xp[0], xp[-1]] (rubix/spectra/ifu.py:250)."""
    B, n = xp.shape
    xx = x[None, :].expand(B, -1).contiguous()
    i = torch.searchsorted(xp.detach().contiguous(), xx, right=True).clamp(1, n - 1)
    xp0 = torch.gather(xp, 1, i - 1)
    xp1 = torch.gather(xp, 1, i)
    fp0 = torch.gather(fp, 1, i - 1)
    fp1 = torch.gather(fp, 1, i)
    f = fp0 + ((xx - xp0) / (xp1 - xp0)) * (fp1 - fp0)
    f = torch.where(xx < xp[:, :1], fp[:, :1].expand(-1, x.numel()), f)
    f = torch.where(xx > xp[:, -1:], fp[:, -1:].expand(-1, x.numel()), f)
    return f


def calculate_diff(v):
    """rubix/spectra/ifu.py:83-102 (prepend first element)."""
    return torch.diff(v, prepend=v[..., :1], dim=-1)


def resample(S, w, tw):
    """Batched rubix/spectra/ifu.py:223-266."""
    mask = ((w >= tw.min()) & (w <= tw.max())).to(S.dtype).detach()
    iwd = calculate_diff(w) * mask
    total = torch.sum(S * iwd, dim=-1)
    P = interp(tw, w, S)
    new_total = torch.sum(P * calculate_diff(tw)[None, :], dim=-1)
    sf = torch.nan_to_num(total / new_total, nan=0.0)
    return P * sf[:, None]


def calculate_cube(spec, ids, nb):
    """rubix/spectra/ifu.py:269-288 (segment_sum; out-of-range ids drop)."""
    ns = nb * nb
    ok = (ids >= 0) & (ids < ns)
    out = torch.zeros(ns, spec.shape[-1], dtype=spec.dtype)
    out = out.index_add(0, ids[ok].long(), spec[ok])
    return out.reshape(nb, nb, -1)


def gaussian_kernel_2d(m, n, sigma):
    """rubix/telescope/psf/kernels.py:5-31 with a tensor sigma (pixels)."""
    dt = sigma.dtype
    x = torch.arange(-((m - 1) / 2), ((m - 1) / 2) + 1, dtype=dt)
    y = torch.arange(-((n - 1) / 2), ((n - 1) / 2) + 1, dtype=dt)
    X, Y = torch.meshgrid(x, y, indexing="ij")
    v = torch.exp(-(X ** 2 + Y ** 2) / (2 * sigma ** 2))
    return v / v.sum()


def apply_psf(cube, K):
    """rubix/telescope/psf/psf.py:33-58: per lambda plane, convolve2d 'same'
    (crop of the full convolution starting at (k-1)//2), as a sum of shifted
    copies of the cube weighted by the kernel taps."""
    ny, nx, nl = cube.shape
    ky, kx = K.shape
    sy, sx = (ky - 1) // 2, (kx - 1) // 2
    out = torch.zeros_like(cube)
    for a in range(ky):
        for b in range(kx):
            dy, dx = sy - a, sx - b
            y0, y1 = max(0, -dy), min(ny, ny - dy)
            x0, x1 = max(0, -dx), min(nx, nx - dx)
            if y0 >= y1 or x0 >= x1:
                continue
            pad = F.pad(K[a, b] * cube[y0 + dy:y1 + dy, x0 + dx:x1 + dx],
                        (0, 0, x0, nx - x1, y0, ny - y1))
            out = out + pad
    return out


def gaussian_taps_1d(m, sigma):
    """1D factor g of gaussian_kernel_2d (outer(g, g) in exact arithmetic)."""
    x = torch.arange(-((m - 1) / 2), ((m - 1) / 2) + 1, dtype=sigma.dtype)
    v = torch.exp(-(x ** 2) / (2 * sigma ** 2))
    return v / v.sum()


def apply_psf_separable(cube, g):
    """apply_psf with the kernel outer(g, g) as two 1D 'same' zero-padded
    convolutions (y axis, then x axis): conv1d with the flipped taps over
    each line, padding (m - 1)/2 on both sides (odd m)."""
    ny, nx, nl = cube.shape
    m = g.numel()
    k = torch.flip(g, dims=(0,))[None, None, :]
    # along y: lines are (x, lambda) columns
    a = cube.permute(1, 2, 0).reshape(-1, 1, ny)
    a = F.conv1d(a, k, padding=(m - 1) // 2).reshape(nx, nl, ny).permute(2, 0, 1)
    b = a.permute(0, 2, 1).reshape(-1, 1, nx)
    b = F.conv1d(b, k, padding=(m - 1) // 2).reshape(ny, nl, nx).permute(0, 2, 1)
    return b


def lsf_kernel(sigma, wave_res, factor=12):
    """rubix/telescope/lsf/lsf.py:12-26 with a tensor sigma (Angstrom)."""
    import numpy as np
    x = np.arange(-factor * wave_res, factor * wave_res + wave_res, wave_res)
    x = torch.tensor(x, dtype=sigma.dtype)
    r = torch.exp(-0.5 * x ** 2 / sigma ** 2)
    return r / r.sum()


def apply_lsf(cube, k, factor=12):
    """rubix/telescope/lsf/lsf.py:29-105 (full convolution + crop
    [factor, n + L - 1 - factor)).  An even tap count L (the float arange of
    the kernel grid gives 26 taps for wave_res 1.6 or 2.7) makes that crop
    n + 1 long, which the reference cannot reshape to the cube (it raises);
    here the first n bins of the crop are returned, the LSF centred on the
    x = 0 tap like the odd case."""
    shape = cube.shape
    spec = cube.reshape(-1, shape[-1])
    n = spec.shape[1]
    L = k.numel()
    full = F.conv1d(spec[:, None, :], torch.flip(k, dims=(0,))[None, None, :], padding=L - 1)
    return full[:, 0, factor:factor + n].reshape(shape)


def reference_knots(ssp_wave32, vel_los, a):
    """Shifted knots with the reference's float32 values and the derivative of
    the exact product: value fl32(fl32((1+z) wave) * fl32(exp(v/c))), gradient
    d(wave * a)/da (straight-through the rounding, as autodiff of the float32
    reference treats it)."""
    import numpy as np
    v32 = vel_los.detach().to(torch.float32)
    a32 = torch.exp((v32 / np.float32(SPEED_OF_LIGHT)).double()).to(torch.float32)
    w32 = (ssp_wave32[None, :] * a32[:, None]).to(a.dtype)
    w = ssp_wave32.to(a.dtype)[None, :] * a[:, None]
    return w + (w32 - w).detach()


def star_cube(mass, age, met, vel_los, ids, grid, geom, dtype, ref_knots=True, extinction=None):
    """Raw datacube (before PSF/LSF) from per-star tensors, the
    calculate_spectra -> scale_spectrum_by_mass ->
    doppler_shift_and_resampling -> calculate_datacube chain
    (rubix/core/ifu.py:63-121,149-157,269-293,327-339).  ``ref_knots``: see
    reference_knots (float64 evaluation of the reference's float32 knots).
    ``extinction``: optional (av per star, A(lambda)/A(V) per target bin)
    tensors; the resampled spectra are multiplied by 10^(-0.4 av axav) before
    the cube (calc_dusty_ifu's calculate_extinction step,
    rubix/spectra/dust/dust_extinction.py:330-358)."""
    axm = torch.as_tensor(grid["ax_met"], dtype=dtype)
    axa = torch.as_tensor(grid["ax_age"], dtype=dtype)
    flux = torch.as_tensor(grid["flux"], dtype=dtype)
    log_axes = grid.get("log_axes", False)
    outside = safe_queries(met.detach(), age.detach(), log_axes, axm, axa)
    # keep the where-selected branch finite (see safe_queries)
    one_m = torch.where(outside, torch.full_like(met, float(10 ** axm[0]) if log_axes else float(axm[0])), met)
    one_a = torch.where(outside, torch.full_like(age, float(10 ** axa[0]) if log_axes else float(axa[0])), age)
    spec = lookup_linear(one_m, one_a, axm, axa, flux, log_axes)
    spec = torch.where(outside[:, None], torch.zeros_like(spec), spec)
    spec = spec * mass[:, None]
    wave32 = torch.as_tensor(grid["wave"], dtype=torch.float32)
    ssp_wave32 = torch.tensor(float(1 + geom["z"]), dtype=torch.float32) * wave32
    a = torch.exp(vel_los / SPEED_OF_LIGHT)
    if ref_knots or dtype == torch.float32:
        w = reference_knots(ssp_wave32, vel_los, a)
    else:
        ssp_wave = (1 + geom["z"]) * torch.as_tensor(grid["wave"], dtype=dtype)
        w = ssp_wave[None, :] * a[:, None]
    tw = torch.as_tensor(geom["wave_target"], dtype=dtype)
    res = resample(spec, w, tw)
    if extinction is not None:
        av, axav = extinction
        res = res * torch.pow(torch.tensor(10.0, dtype=res.dtype),
                              -0.4 * av.to(res.dtype)[:, None] * axav.to(res.dtype)[None, :])
    nb = len(geom["edges"]) - 1
    return calculate_cube(res, ids, nb)


def full_forward(mass, age, met, vel_los, ids, sigma_psf, sigma_lsf, grid, geom,
                 dtype=torch.float64, ref_knots=True, psf="direct", extinction=None):
    """Whole path with tensors for every differentiable input; returns
    (raw cube, cube after PSF, cube after LSF).  psf="separable": the PSF as
    two 1D passes (equal in exact arithmetic), for large cubes; extinction:
    see star_cube."""
    cube = star_cube(mass, age, met, vel_los, ids, grid, geom, dtype, ref_knots, extinction)
    m = geom["psf_size"]
    if psf == "separable":
        c1 = apply_psf_separable(cube, gaussian_taps_1d(m, sigma_psf))
    else:
        K = gaussian_kernel_2d(m, m, sigma_psf)
        c1 = apply_psf(cube, K)
    k = lsf_kernel(sigma_lsf, geom["wave_res"])
    c2 = apply_lsf(c1, k)
    return cube, c1, c2
