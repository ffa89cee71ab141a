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

def gaussian_taps_2d_factor(size: int, sigma: float):
    """1D factor g of gaussian_kernel_2d(size, size, sigma)
    (rubix/telescope/psf/kernels.py:5-31): K = outer(g, g) with
    g = exp(-x^2 / 2 sigma^2) / sum, x = arange(-(m-1)/2, (m-1)/2 + 1).
    Returns (g, dg/dsigma) in float64 numpy."""
    x = np.arange(-((size - 1) / 2), ((size - 1) / 2) + 1, dtype=np.float64)
    return _norm_gauss(x, float(sigma))


def lsf_taps(sigma: float, wave_res: float, factor: int = 12):
    """LSF kernel of rubix/telescope/lsf/lsf.py:22-26 and its sigma-derivative:
    x = arange(-f dl, f dl + dl, dl), g = exp(-x^2 / 2 sigma^2) / sum."""
    x = np.arange(-factor * wave_res, factor * wave_res + wave_res, wave_res, dtype=np.float64)
    return _norm_gauss(x, float(sigma))


def _norm_gauss(x, sigma):
    e = np.exp(-0.5 * x * x / (sigma * sigma))
    g = e / e.sum()
    x2 = x * x / sigma ** 3
    dg = g * (x2 - np.sum(g * x2))
    return g, dg


def _dev(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device=device)


# ---------------------------------------------------------------- cube

class FusedDatacube(torch.autograd.Function):
    """(mass, age, metallicity, velocity) -> raw datacube (nb*nb, n_target)."""

    @staticmethod
    def forward(ctx, mass, age, met, velocity, ids, ssp, target, n_spaxel, n_keys, vel_comp, c):
        cube, cctx = ops.cube_forward(ssp, target, mass.detach(), age.detach(), met.detach(),
                                      velocity.detach(), ids, n_spaxel, n_keys, vel_comp, c)
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
        return dm, dt, dz, dvel, None, None, None, None, None, None, None


# ---------------------------------------------------------------- PSF

class GaussianPSF(torch.autograd.Function):
    """Separable Gaussian PSF over the (ny, nx, n_lambda) cube."""

    @staticmethod
    def forward(ctx, cube, sigma, size):
        ny, nx, nl = cube.shape
        g, dg = gaussian_taps_2d_factor(size, float(sigma))
        dev = cube.device
        gt, dgt = _dev(g, dev), _dev(dg, dev)
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
        k, dk = lsf_taps(float(sigma), float(wave_res), factor)
        dev = cube.device
        kt, dkt = _dev(k, dev), _dev(dk, dev)
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
