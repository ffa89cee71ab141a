"""Batched ragged mode: many galaxies with uneven star counts and spaxel grids
rendered (and differentiated) in one pass of the fused kernels, with the
whole forward + adjoint step captured once in a HIP graph and replayed.

GalaxyBatch    - concatenated device star arrays + the galaxy table (ops.BatchLayout)
                 + each galaxy's spaxel edges.
GraphedStep    - forward + adjoint of a chi^2 loss captured in a torch.cuda.CUDAGraph
                 (HIP graph): the cotangent 2 (C - C_target) w is formed
                 elementwise (no reduction inside the graph) and the gradients
                 come from torch.autograd.grad, so a replay launches the same
                 kernels on the same buffers with no host work.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from .autograd import render_batched


class GalaxyBatch:
    """Device-resident inputs of a batch.  `stars` is a list of per-galaxy
    dicts (coords (n,3), velocity (n,3), mass, age, metallicity), `nbs` the
    spaxel grid side of each galaxy and `edges` each galaxy's spaxel edges."""

    FIELDS = ("coords", "velocity", "mass", "age", "metallicity")

    def __init__(self, stars, nbs, edges, device, requires_grad=True):
        self.layout = ops.BatchLayout([len(s["mass"]) for s in stars], nbs, device)
        self.t = {}
        for k in self.FIELDS:
            arr = np.concatenate([np.asarray(s[k], dtype=np.float32) for s in stars])
            self.t[k] = torch.as_tensor(arr, device=device)
        if requires_grad:
            for k in ("mass", "age", "metallicity", "velocity"):
                self.t[k].requires_grad_(True)
        self.edges = [torch.as_tensor(np.asarray(e, dtype=np.float32), device=device) for e in edges]

    @property
    def leaves(self):
        return [self.t[k] for k in ("mass", "age", "metallicity", "velocity")]

    def render(self, ssp, target, psf_sigma, psf_size, lsf_sigma, wave_res, tune=None, **kw):
        t = self.t
        kw = {**{k: t[k] for k in ("mass", "age", "metallicity", "velocity")}, **kw}
        return render_batched(t["coords"], kw["velocity"], kw["mass"], kw["age"],
                              kw["metallicity"], self.edges, self.layout, ssp, target,
                              psf_sigma, psf_size, lsf_sigma, wave_res, tune)


class GraphedStep:
    """Forward + adjoint of chi^2 = sum w (C - C_target)^2 for fixed input
    tensors, captured once and replayed.  `render()` must read only `inputs`
    (and constant device state) and return the cube.  After replay(),
    `cube` and `grads` (one per input, in order) hold the step's results."""

    def __init__(self, render, inputs, c_target, weight, warmup=2):
        self.render = render
        self.inputs = list(inputs)
        self.c_target = c_target
        self.weight = weight
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self._step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.cube, self.grads = self._step()

    def _step(self):
        cube = self.render()
        cot = 2.0 * (cube.detach() - self.c_target) * self.weight
        grads = torch.autograd.grad(cube, self.inputs, grad_outputs=cot, allow_unused=True)
        return cube.detach(), grads

    def eager(self):
        """One step without the graph (same arithmetic)."""
        return self._step()

    def replay(self):
        self.graph.replay()
        return self.cube, self.grads
