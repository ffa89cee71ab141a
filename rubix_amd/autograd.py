"""torch.autograd.Function wrappers: every forward kernel with its hand-written
adjoint.

* FusedDatacube   - fused lookup + mass scaling + Doppler + resampling +
                    segment reduction; adjoint w.r.t. mass, age, metallicity
                    and the line-of-sight velocity component.
* GaussianPSF     - separable Gaussian PSF (sigma in pixels, as
                    rubix/core/psf.py:59-61 passes it); adjoint w.r.t. the cube
                    and sigma.
* GaussianLSF     - 25-tap Gaussian LSF along lambda (sigma in Angstrom,
                    rubix/telescope/lsf/lsf.py:22-26); adjoint w.r.t. the cube
                    and sigma.
* BatchedDatacube, BatchedPSF - the same for a ragged batch of galaxies in
                    one pass (render_batched).

The reference has no adjoint (rubix/core/pipeline.py:203-208) and keeps the
PSF/LSF sigmas as python floats captured by the step closures
(rubix/core/psf.py:59-61, rubix/core/lsf.py:50); here sigma is a tensor so it
can be differentiated.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops


# ---------------------------------------------------------------- taps

def psf_x(size: int):
    """Tap positions of gaussian_kernel_2d (rubix/telescope/psf/kernels.py:26-30):
    arange(-(m-1)/2, (m-1)/2 + 1)."""
    return np.arange(-((size - 1) / 2), ((size - 1) / 2) + 1, dtype=np.float64)


def lsf_x(wave_res: float, factor: int = 12):
    """Tap positions of the LSF kernel (rubix/telescope/lsf/lsf.py:22-26):
    arange(-f dl, f dl + dl, dl)."""
    return np.arange(-factor * wave_res, factor * wave_res + wave_res, wave_res, dtype=np.float64)


def gaussian_taps_2d_factor(size: int, sigma: float):
    """1D factor g of gaussian_kernel_2d(size, size, sigma)
    (rubix/telescope/psf/kernels.py:5-31): K = outer(g, g) with
    g = exp(-x^2 / 2 sigma^2) / sum.  Returns (g, dg/dsigma) in float64 numpy."""
    return _norm_gauss(psf_x(size), float(sigma))


def lsf_taps(sigma: float, wave_res: float, factor: int = 12):
    """LSF kernel of rubix/telescope/lsf/lsf.py:22-26 and its sigma-derivative:
    g = exp(-x^2 / 2 sigma^2) / sum."""
    return _norm_gauss(lsf_x(wave_res, factor), float(sigma))


def _norm_gauss(x, sigma):
    e = np.exp(-0.5 * x * x / (sigma * sigma))
    g = e / e.sum()
    x2 = x * x / sigma ** 3
    dg = g * (x2 - np.sum(g * x2))
    return g, dg


_XGRID = {}


def _xgrid(key, make, device):
    """Device copy of a tap-position grid, made once per (grid, device), so a
    step that is replayed from a HIP graph copies nothing from the host."""
    k = (key, str(device))
    t = _XGRID.get(k)
    if t is None:
        t = torch.as_tensor(make(), dtype=torch.float64, device=device)
        _XGRID[k] = t
    return t


def _taps(key, make, sigma, device):
    """(float32 taps, float64 d taps/d sigma) of the normalised Gaussian on the
    grid `make()`.  A device-tensor sigma is evaluated on the device (no host
    synchronisation); a python number on the host.  Both evaluate
    g = exp(-x^2 / 2 sigma^2) / sum in float64 and round g once to float32."""
    if isinstance(sigma, torch.Tensor) and sigma.is_cuda:
        x = _xgrid(key, make, device)
        s = sigma.detach().to(device=device, dtype=torch.float64).reshape(())
        e = torch.exp(-0.5 * x * x / (s * s))
        g = e / e.sum()
        x2 = x * x / (s * s * s)
        dg = g * (x2 - (g * x2).sum())
        return g.to(torch.float32), dg
    g, dg = _norm_gauss(make(), float(sigma))
    return _dev(g, device), _dev64(dg, device)


def _dev(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device=device)


def _dev64(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)


# ---------------------------------------------------------------- cube

class FusedDatacube(torch.autograd.Function):
    """(mass, age, metallicity, velocity) -> raw datacube (nb*nb, n_target)."""

    @staticmethod
    def forward(ctx, mass, age, met, velocity, ids, ssp, target, n_spaxel, n_keys, vel_comp, c,
                extinction=None):
        cube, cctx = ops.cube_forward(ssp, target, mass.detach(), age.detach(), met.detach(),
                                      velocity.detach(), ids, n_spaxel, n_keys, vel_comp, c,
                                      extinction)
        ctx.cctx = cctx
        ctx.ssp = ssp
        ctx.target = target
        ctx.vel_comp = vel_comp
        ctx.vshape = velocity.shape
        return cube

    @staticmethod
    def backward(ctx, dcube):
        need = ctx.needs_input_grad
        dm, dt, dz, dv = ops.cube_backward(ctx.ssp, ctx.target, ctx.cctx, dcube.contiguous(),
                                           need=(need[0], need[1], need[2], need[3]))
        dvel = None
        if need[3]:
            dvel = torch.zeros(ctx.vshape, dtype=torch.float32, device=dcube.device)
            dvel[:, ctx.vel_comp] = dv
        return dm, dt, dz, dvel, None, None, None, None, None, None, None, None


# ---------------------------------------------------------------- PSF

class GaussianPSF(torch.autograd.Function):
    """Separable Gaussian PSF over the (ny, nx, n_lambda) cube."""

    @staticmethod
    def forward(ctx, cube, sigma, size):
        ny, nx, nl = cube.shape
        gt, dgt = _taps(("psf", size), lambda: psf_x(size), sigma, cube.device)
        start = (size - 1) // 2
        cube = cube.contiguous()
        tmp = ops.conv_axis(cube, gt, ny, nx, nl, start)             # along x
        out = ops.conv_axis(tmp, gt, 1, ny, nx * nl, start)          # along y
        ctx.save_for_backward(cube, tmp)
        ctx.taps = (gt, dgt)
        ctx.size = size
        return out

    @staticmethod
    def backward(ctx, gout):
        cube, tmp = ctx.saved_tensors
        gt, dgt = ctx.taps
        m = ctx.size
        ny, nx, nl = cube.shape
        start = (m - 1) // 2
        rstart = m - 1 - start
        grev = torch.flip(gt, (0,)).contiguous()
        gout = gout.contiguous()
        v = ops.conv_axis(gout, grev, 1, ny, nx * nl, rstart)        # Y^T
        gin = ops.conv_axis(v, grev, ny, nx, nl, rstart) if ctx.needs_input_grad[0] else None
        dsig = None
        if ctx.needs_input_grad[1]:
            d1 = ops.conv_axis_dot(gout, tmp, dgt, 1, ny, nx * nl, start)   # <G, Y' X C>
            d2 = ops.conv_axis_dot(v, cube, dgt, ny, nx, nl, start)         # <Y^T G, X' C>
            dsig = (d1 + d2).to(torch.float32)
        return gin, dsig, None


# ---------------------------------------------------------------- LSF

class GaussianLSF(torch.autograd.Function):
    """Gaussian LSF along lambda (full convolution cropped at the extend
    factor) over any (..., n_lambda) cube."""

    @staticmethod
    def forward(ctx, cube, sigma, wave_res, factor=12):
        shape = cube.shape
        nl = shape[-1]
        ns = cube.numel() // nl
        kt, dkt = _taps(("lsf", float(wave_res), factor), lambda: lsf_x(wave_res, factor), sigma,
                        cube.device)
        cube = cube.contiguous()
        out = ops.conv_axis(cube, kt, ns, nl, 1, factor)
        ctx.save_for_backward(cube)
        ctx.taps = (kt, dkt)
        ctx.factor = factor
        return out

    @staticmethod
    def backward(ctx, gout):
        (cube,) = ctx.saved_tensors
        kt, dkt = ctx.taps
        L = kt.numel()
        nl = cube.shape[-1]
        ns = cube.numel() // nl
        gout = gout.contiguous()
        gin = None
        if ctx.needs_input_grad[0]:
            krev = torch.flip(kt, (0,)).contiguous()
            gin = ops.conv_axis(gout, krev, ns, nl, 1, L - 1 - ctx.factor)
        dsig = None
        if ctx.needs_input_grad[1]:
            dsig = ops.conv_axis_dot(gout, cube, dkt, ns, nl, 1, ctx.factor).to(torch.float32)
        return gin, dsig, None, None


# ---------------------------------------------------------------- alignment

class FilterFields(torch.autograd.Function):
    """filter_particles on up to three per-particle fields; the adjoint masks
    the incoming gradients the same way."""

    @staticmethod
    def forward(ctx, coords, lo, hi, *fields):
        *outs, mask = ops.filter_particles(coords.detach(), lo, hi, *[f.detach() for f in fields])
        ctx.save_for_backward(mask)
        ctx.mark_non_differentiable(mask)
        return (*outs, mask)

    @staticmethod
    def backward(ctx, *grads):
        (mask,) = ctx.saved_tensors
        outs = [None if g is None else ops.apply_mask(mask, g.contiguous()) for g in grads[:-1]]
        return (None, None, None, *outs)


class RotateRows(torch.autograd.Function):
    """Row vectors x -> (x first) second; adjoint g -> (g second^T) first^T.
    The matrices are constants (the alignment frame is not differentiated)."""

    @staticmethod
    def forward(ctx, x, first, second):
        ctx.mats = (np.asarray(first, np.float64), np.asarray(second, np.float64))
        return ops.rotate_rows(x.detach(), first, second)

    @staticmethod
    def backward(ctx, g):
        first, second = ctx.mats
        return ops.rotate_rows(g.contiguous(), second.T, first.T), None, None


# ---------------------------------------------------------------- noise

class ApplyNoise(torch.autograd.Function):
    """apply_noise: cube -> cube + cube * z * S2N(cube).  The adjoint
    differentiates through the S2N map (flux image and median); the draws z
    are recomputed from the seed."""

    @staticmethod
    def forward(ctx, cube, signal_to_noise, distribution, seed):
        out, state = ops.apply_noise(cube.detach(), signal_to_noise, distribution, seed)
        ctx.save_for_backward(cube)
        ctx.state = state
        return out

    @staticmethod
    def backward(ctx, g):
        (cube,) = ctx.saved_tensors
        return ops.apply_noise_bwd(cube, g.contiguous(), ctx.state), None, None, None


# ---------------------------------------------------------------- dust

class ApplyExtinction(torch.autograd.Function):
    """Per-star spectra (N, n_wave) times 10^(-0.4 A(lambda)/A(V) A_V); the
    adjoint multiplies by the same factor.  A_V and the curve are constants
    (they depend on coordinates and gas, not on the differentiated star
    fields)."""

    @staticmethod
    def forward(ctx, spectra, av, axav):
        ctx.save_for_backward(av, axav)
        return ops.apply_extinction(spectra.detach(), av, axav)

    @staticmethod
    def backward(ctx, g):
        av, axav = ctx.saved_tensors
        return ops.apply_extinction(g.contiguous(), av, axav), None, None


# ---------------------------------------------------------------- batches

class BatchedDatacube(torch.autograd.Function):
    """Every galaxy of a ragged batch in one fused pass: concatenated
    (mass, age, metallicity, velocity) and per-galaxy spaxel ids -> batch
    cube (n_rows, n_target)."""

    @staticmethod
    def forward(ctx, mass, age, met, velocity, ids, ssp, target, layout, tune, vel_comp, c):
        cube, cctx = ops.batch_forward(ssp, target, layout, mass.detach(), age.detach(),
                                       met.detach(), velocity.detach(), ids, tune, vel_comp, c)
        ctx.cctx = cctx
        ctx.args = (ssp, target, layout, tune, vel_comp)
        ctx.vshape = velocity.shape
        return cube

    @staticmethod
    def backward(ctx, dcube):
        ssp, target, layout, tune, vel_comp = ctx.args
        need = ctx.needs_input_grad
        dm, dt, dz, dv = ops.batch_backward(ssp, target, layout, ctx.cctx, dcube.contiguous(),
                                            need=(need[0], need[1], need[2], need[3]), tune=tune)
        dvel = None
        if need[3]:
            dvel = torch.zeros(ctx.vshape, dtype=torch.float32, device=dcube.device)
            dvel[:, vel_comp] = dv
        return dm, dt, dz, dvel, None, None, None, None, None, None, None


class BatchedPSF(torch.autograd.Function):
    """Separable Gaussian PSF over every galaxy of a batch cube (n_rows,
    n_target), each on its own (nb, nb) grid."""

    @staticmethod
    def forward(ctx, cube, sigma, size, layout):
        gt, dgt = _taps(("psf", size), lambda: psf_x(size), sigma, cube.device)
        start = (size - 1) // 2
        cube = cube.contiguous()
        tmp = ops.conv_batch(cube, layout, 1, gt, start)   # along x
        out = ops.conv_batch(tmp, layout, 0, gt, start)    # along y
        ctx.save_for_backward(cube, tmp)
        ctx.taps = (gt, dgt)
        ctx.size = size
        ctx.layout = layout
        return out

    @staticmethod
    def backward(ctx, gout):
        cube, tmp = ctx.saved_tensors
        gt, dgt = ctx.taps
        m = ctx.size
        layout = ctx.layout
        start = (m - 1) // 2
        rstart = m - 1 - start
        grev = torch.flip(gt, (0,)).contiguous()
        gout = gout.contiguous()
        v = ops.conv_batch(gout, layout, 0, grev, rstart)
        gin = ops.conv_batch(v, layout, 1, grev, rstart) if ctx.needs_input_grad[0] else None
        dsig = None
        if ctx.needs_input_grad[1]:
            d1 = ops.conv_batch_dot(gout, tmp, layout, 0, dgt, start)
            d2 = ops.conv_batch_dot(v, cube, layout, 1, dgt, start)
            dsig = (d1 + d2).to(torch.float32)
        return gin, dsig, None, None


# ---------------------------------------------------------------- per step

class SSPLookup(torch.autograd.Function):
    """calculate_spectra: (metallicity, age) -> spectra (N, n_wave)."""

    @staticmethod
    def forward(ctx, met, age, ssp):
        met_d, age_d = met.detach().contiguous(), age.detach().contiguous()
        ctx.save_for_backward(met_d, age_d)
        ctx.ssp = ssp
        return ops.ssp_lookup(ssp, met_d, age_d)

    @staticmethod
    def backward(ctx, dspec):
        met, age = ctx.saved_tensors
        dm, da = ops.ssp_lookup_bwd(ctx.ssp, met, age, dspec.contiguous())
        return dm.view_as(met), da.view_as(age), None


class ScaleByMass(torch.autograd.Function):
    """scale_spectrum_by_mass: spectra[i, :] * mass[i]."""

    @staticmethod
    def forward(ctx, spectra, mass):
        spectra_d, mass_d = spectra.detach().contiguous(), mass.detach().contiguous()
        ctx.save_for_backward(spectra_d, mass_d)
        return ops.scale_rows(spectra_d, mass_d)

    @staticmethod
    def backward(ctx, dout):
        spectra, mass = ctx.saved_tensors
        dout = dout.contiguous()
        ds = ops.scale_rows(dout, mass) if ctx.needs_input_grad[0] else None
        dm = ops.row_dot(dout, spectra).view_as(mass) if ctx.needs_input_grad[1] else None
        return ds, dm


class Resample(torch.autograd.Function):
    """doppler_shift_and_resampling: spectra (N, n_wave) + velocity (N, 3) ->
    (N, n_target)."""

    @staticmethod
    def forward(ctx, spectra, velocity, ssp, target, vel_comp, c):
        spectra_d, vel_d = spectra.detach().contiguous(), velocity.detach().contiguous()
        ctx.save_for_backward(spectra_d, vel_d)
        ctx.args = (ssp, target, vel_comp, c)
        return ops.resample(ssp, target, spectra_d, vel_d, vel_comp, c)

    @staticmethod
    def backward(ctx, dout):
        spectra, vel = ctx.saved_tensors
        ssp, target, vel_comp, c = ctx.args
        need_s, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        ds, dv = ops.resample_bwd(ssp, target, spectra, vel, dout.contiguous(), vel_comp, c,
                                  need_spectra=need_s, need_velocity=need_v)
        dvel = None
        if need_v:
            dvel = torch.zeros_like(vel)
            dvel[:, vel_comp] = dv
        return ds, dvel, None, None, None, None


class SegmentCube(torch.autograd.Function):
    """calculate_datacube: segment sum of (N, n_target) rows by spaxel id."""

    @staticmethod
    def forward(ctx, spectra, ids, n_spaxel, n_keys):
        ctx.save_for_backward(ids)
        ctx.n_spaxel = n_spaxel
        return ops.segment_cube(spectra.detach(), ids, n_spaxel, n_keys)

    @staticmethod
    def backward(ctx, dcube):
        (ids,) = ctx.saved_tensors
        return ops.segment_cube_bwd(dcube.contiguous(), ids, ctx.n_spaxel), None, None, None


def render_plain(coords, velocity, mass, age, met, edges, ssp, target, nb, psf_sigma, psf_size,
                 lsf_sigma, wave_res, vel_comp=2, c=ops.SPEED_OF_LIGHT):
    """The plain per-step chain (one kernel per reference step, spectra
    materialised), differentiable like render()."""
    ids = ops.spaxel_assign(coords, edges)
    nbins = edges.numel() - 1
    spec = SSPLookup.apply(met, age, ssp)
    spec = ScaleByMass.apply(spec, mass)
    spec = Resample.apply(spec, velocity, ssp, target, vel_comp, c)
    raw = SegmentCube.apply(spec, ids, nb * nb, nbins * nbins).view(nb, nb, target.n)
    c1 = GaussianPSF.apply(raw, psf_sigma, psf_size)
    return GaussianLSF.apply(c1, lsf_sigma, wave_res)


# ---------------------------------------------------------------- composition

def render(coords, velocity, mass, age, met, edges, ssp, target, nb, psf_sigma, psf_size,
           lsf_sigma, wave_res, vel_comp=2, c=ops.SPEED_OF_LIGHT, return_all=False):
    """spaxel assignment -> fused datacube -> PSF -> LSF, differentiable in
    mass, age, metallicity, velocity, psf_sigma and lsf_sigma (tensors).
    Returns the (nb, nb, n_target) cube after the LSF."""
    ids = ops.spaxel_assign(coords, edges)
    nbins = edges.numel() - 1
    raw = FusedDatacube.apply(mass, age, met, velocity, ids, ssp, target, nb * nb, nbins * nbins,
                              vel_comp, c)
    raw = raw.view(nb, nb, target.n)
    c1 = GaussianPSF.apply(raw, psf_sigma, psf_size)
    c2 = GaussianLSF.apply(c1, lsf_sigma, wave_res)
    if return_all:
        return ids, raw, c1, c2
    return c2


def render_batched(coords, velocity, mass, age, met, edges, layout, ssp, target, psf_sigma,
                   psf_size, lsf_sigma, wave_res, tune=None, vel_comp=2, c=ops.SPEED_OF_LIGHT):
    """spaxel assignment (per galaxy, its own edges) -> one fused pass over
    the whole batch -> PSF per galaxy -> LSF; differentiable like render().
    Star arrays are the galaxies' concatenated arrays (layout.star_begin);
    returns the batch cube (n_rows, n_target)."""
    ids = torch.cat([ops.spaxel_assign(coords[layout.stars_of(g)], edges[g])
                     for g in range(layout.n_gal)])
    raw = BatchedDatacube.apply(mass, age, met, velocity, ids, ssp, target, layout, tune,
                                vel_comp, c)
    c1 = BatchedPSF.apply(raw, psf_sigma, psf_size, layout)
    return GaussianLSF.apply(c1, lsf_sigma, wave_res)
