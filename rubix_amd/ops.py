"""Tensor-level wrappers of the HIP kernels (device tensors in, device tensors out).

Every function here runs a hand-written HIP kernel through the C ABI on
torch's current stream.  There is no CPU or torch fallback: inputs must be
ROCm device tensors and the library must be built, otherwise the call raises.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import (RbxExtinction, RbxGalaxy, RbxRotation, RbxSsp, RbxStars, RbxTarget, RbxTuning,
                   call, ptr, stream_handle)

SPEED_OF_LIGHT = 299792.458  # km/s, rubix/config/rubix_config.yml:8
VELOCITY_COMPONENT = {"x": 0, "y": 1, "z": 2}


def _need_gpu(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise RuntimeError(
                "rubix_amd kernels need ROCm device tensors (got "
                f"{type(t).__name__} on {getattr(t, 'device', '?')}); there is no CPU path")


def _f32(t, name):
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    return t.contiguous()


# ---------------------------------------------------------------- spaxels

def spaxel_assign(coords: torch.Tensor, edges: torch.Tensor) -> torch.Tensor:
    """square_spaxel_assignment (rubix/telescope/utils.py:77-152) on the GPU.
    coords (N, >=2) float32, edges (n_edges,) float32 -> int32 ids (N,)."""
    _need_gpu(coords, edges)
    coords = _f32(coords, "coords")
    edges = _f32(edges, "edges")
    if coords.dim() != 2 or coords.shape[1] < 2:
        raise ValueError(f"coords must be (N, >=2), got {tuple(coords.shape)}")
    n = coords.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=coords.device)
    call("rbx_spaxel_assign", ptr(coords), n, coords.shape[1], ptr(edges), edges.numel(),
         ptr(out), stream_handle(coords.device))
    return out


def spaxel_sort(ids: torch.Tensor, n_keys: int):
    """Stable sort-by-spaxel index: (perm, sorted_ids, offsets[n_keys + 1])."""
    _need_gpu(ids)
    if ids.dtype != torch.int32:
        raise TypeError("ids must be int32")
    ids = ids.contiguous()
    n = ids.numel()
    dev = ids.device
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    srt = torch.empty(n, dtype=torch.int32, device=dev)
    offsets = torch.empty(n_keys + 1, dtype=torch.int32, device=dev)
    L = _lib.load()
    wsb = L.rbx_spaxel_sort_workspace_size(n, n_keys)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    call("rbx_spaxel_sort", ptr(ids), n, n_keys, ptr(perm), ptr(srt), ptr(offsets), ptr(ws), wsb,
         stream_handle(dev))
    return perm, srt, offsets


# ---------------------------------------------------------------- alignment

def filter_particles(coords, lo: float, hi: float, *fields):
    """filter_particles (rubix/core/telescope.py:129-201): mask = coords x, y
    within [lo, hi]; each field (N,) float32 becomes field where mask else 0.
    Returns (masked fields..., mask (N,) bool)."""
    _need_gpu(coords, *fields)
    coords = _f32(coords, "coords")
    if len(fields) > 3:
        raise ValueError("at most three fields per call")
    n = coords.shape[0]
    ins = [_f32(f, "field") for f in fields]
    for f in ins:
        if f.numel() != n:
            raise ValueError("fields must have one value per particle")
    outs = [torch.empty_like(f) for f in ins]
    mask = torch.empty(n, dtype=torch.uint8, device=coords.device)
    pad_in = ins + [None] * (3 - len(ins))
    pad_out = outs + [None] * (3 - len(outs))
    call("rbx_filter_particles", ptr(coords), n, coords.shape[1], lo, hi, *[ptr(t) for t in pad_in],
         *[ptr(t) for t in pad_out], ptr(mask), stream_handle(coords.device))
    return (*outs, mask.bool())


def apply_mask(mask, x):
    """x where mask else 0, row-wise for (N, ...) x (the filter's adjoint)."""
    _need_gpu(mask, x)
    x = _f32(x, "x")
    m = mask.to(torch.uint8).contiguous()
    n = m.numel()
    out = torch.empty_like(x)
    call("rbx_apply_mask", ptr(m), n, x.numel() // max(n, 1), ptr(x), ptr(out),
         stream_handle(x.device))
    return out


def inertia_tensor(coords, mass, halfmass_radius: float):
    """moment_of_inertia_tensor (rubix/galaxy/alignment.py:66-120) as a
    (3, 3) float64 device tensor (fixed-order double reduction)."""
    _need_gpu(coords, mass)
    coords = _f32(coords, "coords")
    mass = _f32(mass, "mass")
    n = coords.shape[0]
    L = _lib.load()
    wsb = L.rbx_inertia_workspace_size(n)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=coords.device)
    out = torch.empty(3, 3, dtype=torch.float64, device=coords.device)
    call("rbx_inertia_tensor", ptr(coords), ptr(mass), n, coords.shape[1], float(halfmass_radius),
         ptr(out), ptr(ws), wsb, stream_handle(coords.device))
    return out


def rotate_rows(x, first, second):
    """(x first) second for row vectors x (N, >=3) with 3x3 matrices
    (apply_init_rotation then apply_rotation, rubix/galaxy/alignment.py)."""
    _need_gpu(x)
    x = _f32(x, "x")
    rot = RbxRotation()
    a = np.asarray(first, dtype=np.float32).reshape(9)
    b = np.asarray(second, dtype=np.float32).reshape(9)
    for k in range(9):
        rot.first[k] = float(a[k])
        rot.second[k] = float(b[k])
    out = torch.empty_like(x)
    call("rbx_rotate_rows", ptr(x), x.shape[0], x.shape[1], rot, ptr(out), stream_handle(x.device))
    return out


# ---------------------------------------------------------------- SSP tables

class PreparedSSP:
    """Device-resident SSP grid for one redshift, with the tables the fused
    kernels read: the (1+z)-shifted wavelengths (float32, as
    rubix/core/ifu.py:261 computes them), per-template prefix sums of
    flux * diff(wave_obs) (float64) and the knot-locator table."""

    MAX_LOCATE = 1 << 23

    def __init__(self, flux, ax_met, ax_age, wavelength, z, log_axes, device):
        dev = torch.device(device)
        self.device = dev
        self.flux = torch.as_tensor(flux, dtype=torch.float32).to(dev).contiguous()
        self.ax_met = torch.as_tensor(ax_met, dtype=torch.float32).to(dev).contiguous()
        self.ax_age = torch.as_tensor(ax_age, dtype=torch.float32).to(dev).contiguous()
        wave = torch.as_tensor(wavelength, dtype=torch.float32).to(dev)
        _need_gpu(self.flux)
        nm, na, nw = self.flux.shape
        if self.ax_met.numel() != nm or self.ax_age.numel() != na or wave.numel() != nw:
            raise ValueError("flux must be (n_met, n_age, n_wave) matching the axes")
        if nm < 2 or na < 2 or nw < 2:
            raise ValueError("the SSP grid needs >= 2 metallicities, ages and wavelengths")
        self.z = float(z)
        self.log_axes = bool(log_axes)
        # (1 + z) * wavelength in float32 (python-float factor, array dtype)
        self.wave_obs = (torch.tensor(1.0 + self.z, dtype=torch.float32, device=dev) * wave).contiguous()
        w = self.wave_obs.double().cpu().numpy()
        dw = np.diff(w)
        if not np.all(dw > 0):
            raise ValueError("SSP wavelengths must be strictly increasing")
        for ax, nm_ in ((self.ax_met, "metallicity"), (self.ax_age, "age")):
            if not bool(torch.all(ax[1:] > ax[:-1])):
                raise ValueError(f"SSP {nm_} axis must be strictly increasing")
        # knot locator: cells of width h <= the smallest knot spacing hold at
        # most one knot each (one probe per query); a grid that would need more
        # than MAX_LOCATE cells gets wider cells and the kernels walk.
        h = float(dw.min())
        span = float(w[-1] - w[0])
        nloc = int(math.floor(span / h)) + 3
        self.locate_multi = False
        if nloc > self.MAX_LOCATE:
            nloc = self.MAX_LOCATE
            h = span / (nloc - 3)
            self.locate_multi = True
        self.n_locate = nloc
        self.locate_x0 = float(np.float32(w[0]))
        inv_h = np.float32(1.0 / h)
        # the float32 1/h must not widen the cells past the smallest spacing
        while not self.locate_multi and 1.0 / float(inv_h) > float(dw.min()):
            inv_h = np.nextafter(inv_h, np.float32(np.inf))
        self.locate_inv_h = float(inv_h)
        if span * self.locate_inv_h + 2 > nloc - 1:
            self.n_locate = nloc = int(math.floor(span * self.locate_inv_h)) + 3
        self.cum = torch.empty(nm * na * nw, dtype=torch.float64, device=dev)
        self.locate = torch.empty(5 * nloc, dtype=torch.int32, device=dev)
        self.shape = (nm, na, nw)
        self._struct = self._make_struct()
        call("rbx_ssp_prepare", self._struct, ptr(self.cum), ptr(self.locate), stream_handle(dev))

    def _make_struct(self):
        nm, na, nw = self.shape
        s = RbxSsp()
        s.flux = self.flux.data_ptr()
        s.ax_met = self.ax_met.data_ptr()
        s.ax_age = self.ax_age.data_ptr()
        s.wave_obs = self.wave_obs.data_ptr()
        s.cum = self.cum.data_ptr()
        s.locate = self.locate.data_ptr()
        s.n_met, s.n_age, s.n_wave = nm, na, nw
        s.n_locate = self.n_locate
        s.locate_x0 = self.locate_x0
        s.locate_inv_h = self.locate_inv_h
        s.log_axes = int(self.log_axes)
        s.locate_multi = int(self.locate_multi)
        return s

    @property
    def struct(self):
        return self._struct


class TargetGrid:
    """Telescope wavelength axis on the device (rubix/core/ifu.py:254-255)
    with calculate_diff (rubix/spectra/ifu.py:83-102) and its min/max."""

    def __init__(self, wave, device):
        w = torch.as_tensor(wave, dtype=torch.float32).to(device).contiguous()
        _need_gpu(w)
        if w.numel() < 1:
            raise ValueError("empty target wavelength grid")
        self.wave = w
        d = torch.zeros_like(w)
        if w.numel() > 1:
            d[1:] = w[1:] - w[:-1]
        self.dwave = d.contiguous()
        self.pairs = torch.stack([w, self.dwave], dim=1).contiguous()
        self.wmin = float(w.min())
        self.wmax = float(w.max())
        self.ascending = bool(torch.all(w[1:] > w[:-1])) if w.numel() > 1 else True
        self.nondecreasing = bool(torch.all(w[1:] >= w[:-1])) if w.numel() > 1 else True
        s = RbxTarget()
        s.wave = w.data_ptr()
        s.dwave = self.dwave.data_ptr()
        s.n = w.numel()
        s.wmin = self.wmin
        s.wmax = self.wmax
        s.ascending = int(self.nondecreasing)
        s.pairs = self.pairs.data_ptr()
        self._struct = s

    @property
    def n(self):
        return self.wave.numel()

    @property
    def struct(self):
        return self._struct


# ---------------------------------------------------------------- fused cube

@dataclass
class CubeContext:
    """What the fused forward leaves for the adjoint (workspace + inputs)."""
    stars: RbxStars
    workspace: torch.Tensor
    keep: tuple
    ext: object = None  # RbxExtinction or None


def _stars_struct(mass, age, met, velocity, ids, n_spaxel, n_keys, vel_comp, c):
    n = mass.numel()
    for t, nm in ((age, "age"), (met, "metallicity"), (ids, "ids")):
        if t.numel() != n:
            raise ValueError(f"{nm} has {t.numel()} entries, mass has {n}")
    if velocity.dim() != 2 or velocity.shape[0] != n:
        raise ValueError(f"velocity must be (N, 3), got {tuple(velocity.shape)}")
    s = RbxStars()
    s.n = n
    s.mass = mass.data_ptr()
    s.age = age.data_ptr()
    s.metallicity = met.data_ptr()
    s.velocity = velocity.data_ptr()
    s.vel_stride = velocity.shape[1]
    s.vel_comp = vel_comp
    s.n_spaxel = n_spaxel
    s.n_keys = n_keys
    s.ids = ids.data_ptr()
    s.speed_of_light = c
    return s


def cube_forward(ssp: PreparedSSP, target: TargetGrid, mass, age, met, velocity, ids,
                 n_spaxel: int, n_keys: int, vel_comp: int = 2, c: float = SPEED_OF_LIGHT,
                 extinction=None):
    """Fused calculate_spectra -> scale_spectrum_by_mass ->
    doppler_shift_and_resampling [-> calculate_extinction] -> calculate_datacube.
    Returns the raw cube (n_spaxel, n_target) and the context the adjoint
    needs.  extinction: optional (av per star, axav per target bin) device
    tensors (rubix/spectra/dust/dust_extinction.py:330-358)."""
    _need_gpu(mass, age, met, velocity, ids)
    mass = _f32(mass, "mass")
    age = _f32(age, "age")
    met = _f32(met, "metallicity")
    velocity = _f32(velocity, "velocity")
    if ids.dtype != torch.int32:
        raise TypeError("ids must be int32")
    ids = ids.contiguous()
    st = _stars_struct(mass, age, met, velocity, ids, n_spaxel, n_keys, vel_comp, c)
    L = _lib.load()
    wsb = L.rbx_cube_workspace_size(st.n, n_keys, target.n)
    dev = mass.device
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    cube = torch.empty(n_spaxel, target.n, dtype=torch.float32, device=dev)
    if extinction is None:
        call("rbx_cube_fwd", ssp.struct, target.struct, st, ptr(cube), ptr(ws), wsb,
             stream_handle(dev))
        return cube, CubeContext(st, ws, (mass, age, met, velocity, ids))
    av = _f32(extinction[0], "av").reshape(-1)
    axav = _f32(extinction[1], "axav").reshape(-1)
    if av.numel() != st.n or axav.numel() != target.n:
        raise ValueError(f"extinction: av has {av.numel()} entries for {st.n} stars, axav "
                         f"{axav.numel()} for {target.n} bins")
    ext = RbxExtinction(ptr(av), ptr(axav))
    call("rbx_cube_fwd_ext", ssp.struct, target.struct, st, ext, ptr(cube), ptr(ws), wsb,
         stream_handle(dev))
    return cube, CubeContext(st, ws, (mass, age, met, velocity, ids, av, axav), ext)


def cube_backward(ssp: PreparedSSP, target: TargetGrid, ctx: CubeContext, dcube,
                  need=(True, True, True, True)):
    """Adjoint of cube_forward: (dmass, dage, dmet, dv_los) for dL/dcube."""
    _need_gpu(dcube)
    dcube = _f32(dcube, "dcube")
    n = ctx.stars.n
    dev = dcube.device
    outs = [torch.empty(n, dtype=torch.float32, device=dev) if k else None for k in need]
    if ctx.ext is None:
        call("rbx_cube_bwd", ssp.struct, target.struct, ctx.stars, ptr(dcube),
             *[ptr(o) for o in outs], ptr(ctx.workspace), ctx.workspace.numel(),
             stream_handle(dev))
    else:
        call("rbx_cube_bwd_ext", ssp.struct, target.struct, ctx.stars, ctx.ext, ptr(dcube),
             *[ptr(o) for o in outs], ptr(ctx.workspace), ctx.workspace.numel(),
             stream_handle(dev))
    return tuple(outs)


# ---------------------------------------------------------------- convolutions

def conv_axis(x, taps, n_outer, n_axis, n_inner, start):
    """1D convolution along the middle axis of an (n_outer, n_axis, n_inner)
    view: out[o,x,i] = sum_t taps[t] * in[o, x + start - t, i] (zero pad)."""
    _need_gpu(x, taps)
    x = _f32(x, "input")
    taps = _f32(taps, "taps")
    out = torch.empty_like(x)
    call("rbx_conv_axis", ptr(x), ptr(out), n_outer, n_axis, n_inner, ptr(taps), taps.numel(),
         start, stream_handle(x.device))
    return out


def conv2d_planes(cube, kernel, sy=None, sx=None):
    """Per-lambda-plane 2D convolution, mode "same" (general kernel)."""
    _need_gpu(cube, kernel)
    cube = _f32(cube, "cube")
    kernel = _f32(kernel, "kernel")
    ny, nx, nl = cube.shape
    ky, kx = kernel.shape
    sy = (ky - 1) // 2 if sy is None else sy
    sx = (kx - 1) // 2 if sx is None else sx
    out = torch.empty_like(cube)
    call("rbx_conv2d_planes", ptr(cube), ptr(out), ny, nx, nl, ptr(kernel), ky, kx, sy, sx,
         stream_handle(cube.device))
    return out


def conv_axis_dot(a, b, taps, n_outer, n_axis, n_inner, start):
    """sum(a * conv_axis(b)) as a device float64 scalar (fixed order); taps
    float64."""
    _need_gpu(a, b, taps)
    a = _f32(a, "a")
    b = _f32(b, "b")
    if taps.dtype != torch.float64:
        raise TypeError("conv_axis_dot taps must be float64")
    taps = taps.contiguous()
    L = _lib.load()
    total = n_outer * n_axis * n_inner
    wsb = L.rbx_dot_workspace_size(total)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=a.device)
    res = torch.empty((), dtype=torch.float64, device=a.device)
    call("rbx_conv_axis_dot", ptr(a), ptr(b), n_outer, n_axis, n_inner, ptr(taps), taps.numel(),
         start, ptr(res), ptr(ws), wsb, stream_handle(a.device))
    return res


# ---------------------------------------------------------------- batches

TUNE_TWO_PASS = 1  # rbx_tuning.flags: star_norm + per-tile accumulation forward


@dataclass(frozen=True)
class Tuning:
    """Launch tuning of the fused kernels (rbx_tuning): stars per reduction
    chunk, wavefronts per workgroup of the per-star kernels, persistent
    workgroups per CU (0 = one workgroup per work item), flags
    (TUNE_TWO_PASS)."""
    chunk: int = 256
    waves: int = 4
    persist: int = 0
    flags: int = 0

    @property
    def struct(self):
        return RbxTuning(self.chunk, self.waves, self.persist, self.flags)


class BatchLayout:
    """Galaxy table of a ragged batch: galaxy g owns stars
    [star_begin[g], star_begin[g] + n_stars[g]) of the concatenated star
    arrays and rows [row_begin[g], row_begin[g] + nb[g]^2) of the batch cube
    (n_rows, n_target)."""

    def __init__(self, n_stars, nbs, device):
        n_stars = [int(v) for v in n_stars]
        nbs = [int(v) for v in nbs]
        if len(n_stars) != len(nbs) or not n_stars:
            raise ValueError("need one star count and one nb per galaxy")
        if min(n_stars) < 0 or min(nbs) < 1:
            raise ValueError("star counts must be >= 0 and nb >= 1")
        self.n_stars = n_stars
        self.nb = nbs
        self.star_begin = np.concatenate([[0], np.cumsum(n_stars)[:-1]]).astype(np.int64)
        self.row_begin = np.concatenate([[0], np.cumsum([b * b for b in nbs])[:-1]]).astype(np.int64)
        self.n_gal = len(nbs)
        self.n_total = int(sum(n_stars))
        self.n_rows = int(sum(b * b for b in nbs))
        if self.n_rows > 2 ** 31 - 3:
            raise ValueError("batch cube has too many rows")
        arr = (RbxGalaxy * self.n_gal)()
        for g in range(self.n_gal):
            arr[g].star_begin = int(self.star_begin[g])
            arr[g].spaxel_begin = int(self.row_begin[g])
            arr[g].nb = nbs[g]
        raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        self.table = torch.from_numpy(raw).to(device)

    def stars_of(self, g):
        b = int(self.star_begin[g])
        return slice(b, b + self.n_stars[g])

    def rows_of(self, g):
        b = int(self.row_begin[g])
        return slice(b, b + self.nb[g] ** 2)


def batch_forward(ssp: PreparedSSP, target: TargetGrid, layout: BatchLayout, mass, age, met,
                  velocity, ids, tune: Tuning = None, vel_comp: int = 2, c: float = SPEED_OF_LIGHT):
    """Fused forward of every galaxy of a batch in one pass: ids are each
    galaxy's own spaxel ids; returns the batch cube (n_rows, n_target) and the
    adjoint's context."""
    _need_gpu(mass, age, met, velocity, ids)
    mass, age, met, velocity = (_f32(mass, "mass"), _f32(age, "age"), _f32(met, "metallicity"),
                                _f32(velocity, "velocity"))
    if ids.dtype != torch.int32:
        raise TypeError("ids must be int32")
    if mass.numel() != layout.n_total:
        raise ValueError(f"batch has {layout.n_total} stars, got {mass.numel()}")
    ids = ids.contiguous()
    st = _stars_struct(mass, age, met, velocity, ids, layout.n_rows, layout.n_rows + 1, vel_comp, c)
    tune = tune or Tuning()
    L = _lib.load()
    wsb = L.rbx_batch_workspace_size(st.n, layout.n_rows, target.n, tune.struct)
    if wsb == 0:
        raise ValueError(f"invalid batch shape or tuning {tune}")
    dev = mass.device
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    cube = torch.empty(layout.n_rows, target.n, dtype=torch.float32, device=dev)
    call("rbx_batch_fwd", ssp.struct, target.struct, st, ptr(layout.table), layout.n_gal,
         tune.struct, ptr(cube), ptr(ws), wsb, stream_handle(dev))
    return cube, CubeContext(st, ws, (mass, age, met, velocity, ids))


def batch_backward(ssp: PreparedSSP, target: TargetGrid, layout: BatchLayout, ctx: CubeContext,
                   dcube, need=(True, True, True, True), tune: Tuning = None):
    """Adjoint of batch_forward: (dmass, dage, dmet, dv_los) for dL/dcube."""
    _need_gpu(dcube)
    dcube = _f32(dcube, "dcube")
    n = ctx.stars.n
    dev = dcube.device
    tune = tune or Tuning()
    outs = [torch.empty(n, dtype=torch.float32, device=dev) if k else None for k in need]
    call("rbx_batch_bwd", ssp.struct, target.struct, ctx.stars, ptr(layout.table), layout.n_gal,
         tune.struct, ptr(dcube), *[ptr(o) for o in outs], ptr(ctx.workspace),
         ctx.workspace.numel(), stream_handle(dev))
    return tuple(outs)


def conv_batch(x, layout: BatchLayout, axis: int, taps, start: int):
    """One separable-PSF pass over every galaxy of a batch cube (n_rows,
    n_target): axis 1 along x, 0 along y, zero padded inside each galaxy."""
    _need_gpu(x, taps)
    x = _f32(x, "cube")
    taps = _f32(taps, "taps")
    out = torch.empty_like(x)
    call("rbx_conv_batch", ptr(x), ptr(out), ptr(layout.table), layout.n_gal, layout.n_rows,
         x.shape[-1], axis, ptr(taps), taps.numel(), start, stream_handle(x.device))
    return out


def conv_batch_dot(a, b, layout: BatchLayout, axis: int, taps, start: int):
    """sum(a * conv_batch(b)) as a device float64 scalar; taps float64."""
    _need_gpu(a, b, taps)
    a = _f32(a, "a")
    b = _f32(b, "b")
    if taps.dtype != torch.float64:
        raise TypeError("conv_batch_dot taps must be float64")
    taps = taps.contiguous()
    L = _lib.load()
    wsb = L.rbx_dot_workspace_size(a.numel())
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=a.device)
    res = torch.empty((), dtype=torch.float64, device=a.device)
    call("rbx_conv_batch_dot", ptr(a), ptr(b), ptr(layout.table), layout.n_gal, layout.n_rows,
         a.shape[-1], axis, ptr(taps), taps.numel(), start, ptr(res), ptr(ws), wsb,
         stream_handle(a.device))
    return res


# ---------------------------------------------------------------- per step

def _rows(x, name):
    x = _f32(x, name)
    if x.dim() != 2:
        raise ValueError(f"{name} must be 2-D (stars, wavelengths), got {tuple(x.shape)}")
    return x


def ssp_lookup(ssp: PreparedSSP, met, age):
    """calculate_spectra: interp2d linear lookup -> (N, n_wave)."""
    _need_gpu(met, age)
    met = _f32(met, "metallicity").reshape(-1)
    age = _f32(age, "age").reshape(-1)
    n = met.numel()
    out = torch.empty(n, ssp.shape[2], dtype=torch.float32, device=met.device)
    call("rbx_ssp_lookup", ssp.struct, ptr(met), ptr(age), n, ptr(out), stream_handle(met.device))
    return out


def ssp_lookup_bwd(ssp: PreparedSSP, met, age, dspec):
    _need_gpu(met, age, dspec)
    met = _f32(met, "metallicity").reshape(-1)
    age = _f32(age, "age").reshape(-1)
    dspec = _rows(dspec, "dspectra")
    n = met.numel()
    dm = torch.empty(n, dtype=torch.float32, device=met.device)
    da = torch.empty(n, dtype=torch.float32, device=met.device)
    call("rbx_ssp_lookup_bwd", ssp.struct, ptr(met), ptr(age), n, ptr(dspec), ptr(dm), ptr(da),
         stream_handle(met.device))
    return dm, da


def scale_rows(x, scale):
    """scale_spectrum_by_mass: x[i, :] * scale[i]."""
    _need_gpu(x, scale)
    x = _rows(x, "spectra")
    scale = _f32(scale, "mass").reshape(-1)
    out = torch.empty_like(x)
    call("rbx_scale_rows", ptr(x), ptr(scale), x.shape[0], x.shape[1], ptr(out),
         stream_handle(x.device))
    return out


def row_dot(a, b):
    _need_gpu(a, b)
    a = _rows(a, "a")
    b = _rows(b, "b")
    out = torch.empty(a.shape[0], dtype=torch.float32, device=a.device)
    call("rbx_row_dot", ptr(a), ptr(b), a.shape[0], a.shape[1], ptr(out), stream_handle(a.device))
    return out


def resample(ssp: PreparedSSP, target: TargetGrid, spectra, velocity, vel_comp=2,
             c=SPEED_OF_LIGHT):
    """doppler_shift_and_resampling on materialised spectra (N, n_wave)."""
    _need_gpu(spectra, velocity)
    spectra = _rows(spectra, "spectra")
    velocity = _f32(velocity, "velocity")
    n = spectra.shape[0]
    if spectra.shape[1] != ssp.shape[2] or velocity.shape[0] != n:
        raise ValueError("spectra must be (N, n_wave) and velocity (N, 3)")
    out = torch.empty(n, target.n, dtype=torch.float32, device=spectra.device)
    call("rbx_resample", ssp.struct, target.struct, ptr(spectra), ptr(velocity), velocity.shape[1],
         vel_comp, c, n, ptr(out), stream_handle(spectra.device))
    return out


def resample_bwd(ssp: PreparedSSP, target: TargetGrid, spectra, velocity, dout, vel_comp=2,
                 c=SPEED_OF_LIGHT, need_spectra=True, need_velocity=True):
    _need_gpu(spectra, velocity, dout)
    if not target.ascending:
        raise ValueError("the resampling adjoint needs increasing target wavelengths")
    spectra = _rows(spectra, "spectra")
    velocity = _f32(velocity, "velocity")
    dout = _rows(dout, "dout")
    n = spectra.shape[0]
    dev = spectra.device
    ds = torch.empty_like(spectra) if need_spectra else None
    dv = torch.empty(n, dtype=torch.float32, device=dev) if need_velocity else None
    call("rbx_resample_bwd", ssp.struct, target.struct, ptr(spectra), ptr(velocity),
         velocity.shape[1], vel_comp, c, n, ptr(dout), ptr(ds), ptr(dv), stream_handle(dev))
    return ds, dv


def segment_cube(spectra, ids, n_spaxel, n_keys):
    """calculate_datacube: deterministic segment sum -> (n_spaxel, n_wave)."""
    _need_gpu(spectra, ids)
    spectra = _rows(spectra, "spectra")
    if ids.dtype != torch.int32:
        raise TypeError("ids must be int32")
    ids = ids.contiguous().reshape(-1)
    n, nw = spectra.shape
    L = _lib.load()
    wsb = L.rbx_segment_cube_workspace_size(n, n_keys, nw)
    ws = torch.empty(wsb, dtype=torch.uint8, device=spectra.device)
    cube = torch.empty(n_spaxel, nw, dtype=torch.float32, device=spectra.device)
    call("rbx_segment_cube", ptr(spectra), ptr(ids), n, n_spaxel, n_keys, nw, ptr(cube), ptr(ws),
         wsb, stream_handle(spectra.device))
    return cube


def segment_cube_bwd(dcube, ids, n_spaxel):
    _need_gpu(dcube, ids)
    dcube = _rows(dcube, "dcube")
    ids = ids.contiguous().reshape(-1)
    n = ids.numel()
    out = torch.empty(n, dcube.shape[1], dtype=torch.float32, device=dcube.device)
    call("rbx_segment_cube_bwd", ptr(dcube), ptr(ids), n, n_spaxel, dcube.shape[1], ptr(out),
         stream_handle(dcube.device))
    return out


# ---------------------------------------------------------------- noise

NOISE_DISTRIBUTIONS = {"normal": 0, "uniform": 1}  # rubix/telescope/noise/noise.py:5


def _noise_code(distribution: str) -> int:
    if distribution not in NOISE_DISTRIBUTIONS:
        raise ValueError(f"Invalid noise type: {distribution}. Supported types: "
                         f"{list(NOISE_DISTRIBUTIONS)}")
    return NOISE_DISTRIBUTIONS[distribution]


def _seed(seed: int) -> int:
    seed = int(seed)
    if not 0 <= seed < 2 ** 64:
        raise ValueError(f"seed must be in [0, 2**64), got {seed}")
    return seed


def noise_sample(shape, distribution: str = "normal", seed: int = 0, device=None):
    """sample_noise (rubix/telescope/noise/noise.py:8-32) for PRNGKey(seed):
    standard normal or uniform [0, 1) draws at every flat index of shape."""
    code = _noise_code(distribution)
    device = torch.device(device if device is not None else "cuda")
    out = torch.empty(tuple(shape), dtype=torch.float32, device=device)
    _need_gpu(out)
    call("rbx_noise_sample", out.numel(), _seed(seed), code, ptr(out), stream_handle(device))
    return out


@dataclass
class NoiseState:
    """What the adjoint of apply_noise needs besides the cube: the flux image,
    the S2N map and the median statistics (device struct)."""
    flux: torch.Tensor
    s2n: torch.Tensor
    stats: torch.Tensor  # 16 bytes: rbx_noise_stats
    seed: int
    code: int

    def median(self) -> float:
        return float(self.stats[:4].view(torch.float32).item())

    def median_spaxels(self):
        st = self.stats[4:12].view(torch.int32).cpu()
        return int(st[0]), int(st[1])


def _cube_rows(cube):
    if cube.dim() < 2:
        raise ValueError(f"datacube must have a wavelength axis, got shape {tuple(cube.shape)}")
    cube = _f32(cube, "datacube")
    nt = cube.shape[-1]
    return cube, cube.numel() // nt if nt else 0, nt


def apply_noise(cube, signal_to_noise, distribution: str = "normal", seed: int = 0):
    """calculate_noise_cube + the datacube update of the apply_noise step
    (rubix/telescope/noise/noise.py:35-115, rubix/core/noise.py:67-78):
    returns (cube + noise, NoiseState).  signal_to_noise: a number or a
    per-spaxel array of cube.shape[:-1]."""
    _need_gpu(cube)
    code = _noise_code(distribution)
    cube, n_spaxel, nt = _cube_rows(cube)
    sn = torch.as_tensor(signal_to_noise, dtype=torch.float32, device=cube.device)
    sn = torch.broadcast_to(sn, cube.shape[:-1]).contiguous()
    dev = cube.device
    flux = torch.empty(cube.shape[:-1], dtype=torch.float32, device=dev)
    s2n = torch.empty_like(flux)
    stats = torch.empty(16, dtype=torch.uint8, device=dev)
    out = torch.empty_like(cube)
    call("rbx_apply_noise", ptr(cube), n_spaxel, nt, ptr(sn), _seed(seed), code, ptr(flux),
         ptr(s2n), ptr(stats), ptr(out), stream_handle(dev))
    return out, NoiseState(flux, s2n, stats, _seed(seed), code)


def apply_noise_bwd(cube, dout, state: NoiseState):
    """Adjoint of apply_noise with respect to the cube (the draws are
    recomputed from the seed)."""
    _need_gpu(cube, dout)
    cube, n_spaxel, nt = _cube_rows(cube)
    dout = _f32(dout, "dout")
    L = _lib.load()
    wsb = L.rbx_noise_workspace_size(n_spaxel)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=cube.device)
    dcube = torch.empty_like(cube)
    call("rbx_apply_noise_bwd", ptr(cube), ptr(dout), n_spaxel, nt, ptr(state.flux),
         ptr(state.s2n), ptr(state.stats), state.seed, state.code, ptr(ws), wsb, ptr(dcube),
         stream_handle(cube.device))
    return dcube


# ---------------------------------------------------------------- dust

RV_MODELS = {"Cardelli89": 0, "Gordon23": 1}  # rubix/spectra/dust/extinction_models.py:10
DUST_TO_GAS_MODELS = {"power law slope free": 0, "broken power law fit": 1}
XCO_MODELS = {"MW": 0, "Z": 1}


def extinction_curve(model: str, rv: float, wave):
    """A(x)/A(V) of the Cardelli89 or Gordon23 model at wave / 1e4 (wave in
    Angstrom, as rubix/spectra/dust/dust_extinction.py:351 passes it)."""
    if model not in RV_MODELS:
        raise ValueError(f"Extinction model '{model}' is not available. Choose from "
                         f"{list(RV_MODELS)}.")
    _need_gpu(wave)
    wave = _f32(wave, "wave").reshape(-1)
    out = torch.empty_like(wave)
    call("rbx_extinction_curve", RV_MODELS[model], float(rv), ptr(wave), wave.numel(), ptr(out),
         stream_handle(wave.device))
    return out


def gas_extinction(mass, metals, dust_to_gas_model: str, xco: str, grain_density: float,
                   spaxel_area: float):
    """Per gas cell V-band extinction over the spaxel area
    (rubix/spectra/dust/dust_extinction.py:13-160, 269-287)."""
    if dust_to_gas_model == "power law slope fixed":
        raise NotImplementedError("power law slope fixed not implemented yet.")
    if dust_to_gas_model not in DUST_TO_GAS_MODELS:
        raise ValueError(f"unknown dust-to-gas model {dust_to_gas_model!r}; supported: "
                         f"{list(DUST_TO_GAS_MODELS)}")
    if xco not in XCO_MODELS:
        raise ValueError(f"unknown Xco {xco!r}; supported: {list(XCO_MODELS)}")
    _need_gpu(mass, metals)
    mass = _f32(mass, "gas mass").reshape(-1)
    metals = _f32(metals, "gas metals")
    metals = metals.reshape(mass.numel(), -1)
    if metals.shape[1] < 5:
        raise ValueError(f"gas metals need at least 5 columns (H ... O), got {metals.shape[1]}")
    out = torch.empty_like(mass)
    call("rbx_gas_extinction", ptr(mass), ptr(metals), mass.numel(), metals.shape[1],
         DUST_TO_GAS_MODELS[dust_to_gas_model], XCO_MODELS[xco], float(grain_density),
         float(spaxel_area), ptr(out), stream_handle(mass.device))
    return out


def column_av(gas_coords, gas_pix, gas_ext, star_coords, star_pix, n_spaxels: int, n_keys: int):
    """A_V per star: the cumulative gas extinction of its spaxel along z at
    the star's z (rubix/spectra/dust/dust_extinction.py:237-327)."""
    _need_gpu(gas_coords, gas_pix, gas_ext, star_coords, star_pix)
    gc = _rows(gas_coords.reshape(-1, gas_coords.shape[-1]), "gas coords")
    sc = _rows(star_coords.reshape(-1, star_coords.shape[-1]), "star coords")
    gp = gas_pix.reshape(-1).to(torch.int32).contiguous()
    sp = star_pix.reshape(-1).to(torch.int32).contiguous()
    ge = _f32(gas_ext, "gas extinction").reshape(-1)
    L = _lib.load()
    wsb = L.rbx_column_av_workspace_size(gc.shape[0], n_keys)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=sc.device)
    av = torch.empty(sc.shape[0], dtype=torch.float32, device=sc.device)
    call("rbx_column_av", ptr(gc), gc.shape[1], ptr(gp), ptr(ge), gc.shape[0], ptr(sc), sc.shape[1],
         ptr(sp), sc.shape[0], int(n_spaxels), int(n_keys), ptr(ws), wsb, ptr(av),
         stream_handle(sc.device))
    return av


def apply_extinction(spectra, av, axav):
    """spectra (N, n_wave) * 10^(-0.4 axav * av) (dust_extinction.py:330-358)."""
    _need_gpu(spectra, av, axav)
    x = _rows(spectra, "spectra")
    av = _f32(av, "av").reshape(-1)
    axav = _f32(axav, "axav").reshape(-1)
    if av.numel() != x.shape[0] or axav.numel() != x.shape[1]:
        raise ValueError(f"apply_extinction: spectra {tuple(x.shape)}, av {av.numel()}, "
                         f"axav {axav.numel()}")
    out = torch.empty_like(x)
    call("rbx_apply_extinction", ptr(x), ptr(av), ptr(axav), x.shape[0], x.shape[1], ptr(out),
         stream_handle(x.device))
    return out
