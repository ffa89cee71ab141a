"""Benchmark of the particle -> datacube path, forward + adjoint, on one GPU.

Workload (one "step"): 1,000,000 synthetic stars -> 50 x 50 x 3721 datacube
with a 12 Z x 100 age x 5000 lambda SSP grid, PSF sigma 4 px on 15 x 15,
LSF sigma 1.0 A (25 taps); spaxel assignment + fused datacube + PSF + LSF,
a chi^2 loss against a fixed target cube, and the full adjoint (gradients
w.r.t. every star's mass, age, metallicity and v_los and both sigmas).
Inputs are resident in HBM before timing starts.

Prints one JSON line.  Metric: star*lambda-bins/s = N * n_target / t(fwd+adj).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "star·λ-bins/s fwd+adjoint, 1M stars→50²×3721 cube; % HBM roofline vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n, n_src, n_tgt, n_cube):
    """Plain per-step model: every intermediate written once, read once, FP32.
    forward per star 16 n_src + 8 n_tgt; adjoint per star 24 n_src + 8 n_tgt;
    PSF + LSF 16 B per cube element forward and 16 adjoint."""
    fwd = n * (16 * n_src + 8 * n_tgt) + 16 * n_cube
    adj = n * (24 * n_src + 8 * n_tgt) + 16 * n_cube
    return fwd, adj


def cpu_baseline(w, sample, threads):
    """The oracle's torch-CPU restatement of the plain per-step path (float32,
    autograd adjoint) on `sample` stars of the same workload, all host cores."""
    import numpy as np
    import torch

    from oracle import ref_numpy as R, ref_torch as RT

    torch.set_num_threads(threads)
    st = {k: v[:sample] for k, v in w.stars.items()}
    f32 = torch.float32
    ids = torch.tensor(R.square_spaxel_assignment(st["coords"], w.geom["edges"]).astype(np.int64))
    mass = torch.tensor(st["mass"], dtype=f32, requires_grad=True)
    age = torch.tensor(st["age"], dtype=f32, requires_grad=True)
    met = torch.tensor(st["metallicity"], dtype=f32, requires_grad=True)
    vl = torch.tensor(st["velocity"][:, 2], dtype=f32, requires_grad=True)
    sp = torch.tensor(w.geom["psf_sigma"], dtype=f32, requires_grad=True)
    sl = torch.tensor(w.geom["lsf_sigma"], dtype=f32, requires_grad=True)
    t0 = time.perf_counter()
    _, _, c2 = RT.full_forward(mass, age, met, vl, ids, sp, sl, w.grid, w.geom, f32)
    loss = (c2 * c2).sum()
    loss.backward()
    dt = time.perf_counter() - t0
    nt = len(w.geom["wave_target"])
    return sample * nt / dt, dt


def _model_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibrate_copy(n_bytes=1 << 30):
    """A plain copy of known size: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch read n_bytes each when the counters are exact."""
    import torch

    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n_bytes, "copies": 2}), flush=True)


def timed_leg(step, warmup, steps):
    """W untimed steps, then exactly K timed steps between two synchronisations.
    step(ev) records three events: before the forward, after the loss, after
    the backward.  Returns (elapsed_s, fwd_ms list, adj_ms list)."""
    import torch

    for _ in range(warmup):
        step(None)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        step(evs[i])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fwd = [a.elapsed_time(b) for a, b, _ in evs]
    adj = [b.elapsed_time(c) for _, b, c in evs]
    return elapsed, fwd, adj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-stars", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="timed steps of the plain per-step GPU path (0 = skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (twice): the reference "
                         "workload for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a "
              "design (sharding by stars + a lambda-slab reduce-scatter) and is not built.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from rubix_amd import _lib, ops, synthetic as S
    from rubix_amd.autograd import FusedDatacube, GaussianLSF, GaussianPSF, render_plain

    dev = torch.device("cuda", 0)
    t_gen = time.perf_counter()
    w = S.config2(n_stars=args.n_stars)
    log(f"workload generated in {time.perf_counter() - t_gen:.1f}s: N={w.n_stars}")
    g, geo = w.grid, w.geom
    nb, nt = geo["nb"], len(geo["wave_target"])
    n_src = g["flux"].shape[-1]
    t = {k: torch.as_tensor(v, device=dev) for k, v in w.stars.items()}
    for k in ("mass", "age", "metallicity", "velocity"):
        t[k].requires_grad_(True)
    ssp = ops.PreparedSSP(g["flux"], g["ax_met"], g["ax_age"], g["wave"], geo["z"], g["log_axes"], dev)
    tg = ops.TargetGrid(geo["wave_target"], dev)
    edges = torch.as_tensor(geo["edges"], device=dev)
    sp = torch.tensor(geo["psf_sigma"], device=dev, requires_grad=True)
    sl = torch.tensor(geo["lsf_sigma"], device=dev, requires_grad=True)
    nbins = edges.numel() - 1

    def fused(mass, age, met, vel):
        ids = ops.spaxel_assign(t["coords"], edges)
        raw = FusedDatacube.apply(mass, age, met, vel, ids, ssp, tg, nb * nb, nbins * nbins, 2,
                                  ops.SPEED_OF_LIGHT)
        c1 = GaussianPSF.apply(raw.view(nb, nb, nt), sp, geo["psf_size"])
        return GaussianLSF.apply(c1, sl, geo["wave_res"])

    def plain(mass, age, met, vel):
        return render_plain(t["coords"], vel, mass, age, met, edges, ssp, tg, nb, sp,
                            geo["psf_size"], sl, geo["wave_res"])

    with torch.no_grad():
        # fixed target: a second galaxy (seed 2) of the same size, S/N = 30 weights
        st2 = S.make_stars(args.n_stars, seed=2)
        t2 = {k: torch.as_tensor(v, device=dev) for k, v in st2.items()}
        tgt = fused(t2["mass"], t2["age"], t2["metallicity"], t2["velocity"]).clone()
        inv_var = 1.0 / torch.clamp((tgt.abs() / 30.0) ** 2, min=(1e-6 * float(tgt.abs().max())) ** 2)
        del t2

    def make_step(render):
        def step(ev):
            for k in ("mass", "age", "metallicity", "velocity"):
                t[k].grad = None
            sp.grad = None
            sl.grad = None
            if ev:
                ev[0].record()
            out = render(t["mass"], t["age"], t["metallicity"], t["velocity"])
            loss = ((out - tgt) ** 2 * inv_var).sum()
            if ev:
                ev[1].record()
            loss.backward()
            if ev:
                ev[2].record()
        return step

    _lib.TIMER.start()
    elapsed, fwd_ms, adj_ms = timed_leg(make_step(fused), args.warmup, args.steps)
    calls = _lib.TIMER.stop()
    ms = elapsed / args.steps * 1e3
    n = w.n_stars
    value = n * nt / (elapsed / args.steps)
    # the timer also saw the warmup steps' calls: keep the last K of each
    per_call = {k: float(np.median(v[-args.steps:])) for k, v in calls.items()}
    log("per-call ms (median):", json.dumps({k: round(v, 4) for k, v in per_call.items()}))

    fwd_b, adj_b = algorithmic_bytes(n, n_src, nt, nb * nb * nt)
    step_frac = (fwd_b + adj_b) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
    # dominant launch: the per-star adjoint rbx_cube_bwd (star_bwd_kernel, then
    # the small per-star star_grad_kernel)
    bwd_ms = per_call.get("rbx_cube_bwd", float("nan"))
    bwd_bytes = n * (24 * n_src + 8 * nt)
    achieved = bwd_bytes / (bwd_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "rbx_cube_bwd (star_bwd_kernel + star_grad_kernel)",
                "kernel_ms": round(bwd_ms, 4), "algorithmic_bytes_per_launch": bwd_bytes,
                "step_effective_frac": round(step_frac, 4)}
    split = {"fwd_ms_median": round(float(np.median(fwd_ms)), 4),
             "adj_ms_median": round(float(np.median(adj_ms)), 4),
             "step_ms_median": round(float(np.median(np.add(fwd_ms, adj_ms))), 4)}

    plain_res = None
    if args.plain_steps > 0:
        # the plain per-step chain: one kernel per reference step, per-star
        # spectra materialised (N x n_source twice, N x n_target once)
        torch.cuda.reset_peak_memory_stats(dev)
        p_el, p_fwd, p_adj = timed_leg(make_step(plain), 1, args.plain_steps)
        p_ms = p_el / args.plain_steps * 1e3
        plain_res = {"ms_per_step": round(p_ms, 3), "value": n * nt / (p_el / args.plain_steps),
                     "fwd_ms_median": round(float(np.median(p_fwd)), 3),
                     "adj_ms_median": round(float(np.median(p_adj)), 3),
                     "hbm_frac": round((fwd_b + adj_b) / (p_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "peak_hbm_gb": round(torch.cuda.max_memory_allocated(dev) / 1e9, 2),
                     "fused_speedup": round(p_ms / ms, 3), "steps": args.plain_steps}
        log("plain:", json.dumps(plain_res))

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        rate, dt = cpu_baseline(w, args.cpu_sample, threads)
        cpu = {"value": rate, "unit": "star·λ-bins/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_sample} stars of the same workload (full SSP grid, 50x50x3721 cube, "
                         f"PSF + LSF), float32 torch-CPU restatement of the plain per-step path, "
                         f"forward + autograd adjoint, one pass {dt:.1f}s on {_model_name()}"}
    res = {
        "metric": METRIC, "value": value, "unit": "star·λ-bins/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"single galaxy, {n} stars -> {nb}x{nb}x{nt} cube, SSP 12x100x{n_src}, "
                               "PSF 15x15 sigma 4 px, LSF sigma 1.0 A; fused forward + adjoint of chi^2",
                   "n_stars": n, "n_spaxel": nb * nb, "n_target": nt, "n_source": n_src},
        "roofline": roofline, "cpu_baseline": cpu, "split": split, "plain": plain_res,
        "per_call_ms": {k: round(v, 4) for k, v in per_call.items()},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
