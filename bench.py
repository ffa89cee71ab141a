"""Benchmark of the particle -> datacube path, forward + adjoint, on one GPU.

Workload (one "step"): 1,000,000 synthetic stars -> 50 x 50 x 3721 datacube
with a 12 Z x 100 age x 5000 lambda SSP grid, PSF sigma 4 px on 15 x 15,
LSF sigma 1.0 A (25 taps); spaxel assignment + fused datacube + PSF + LSF,
a chi^2 loss against a fixed target cube, and the full adjoint (gradients
w.r.t. every star's mass, age, metallicity and v_los and both sigmas).
Inputs are resident in HBM before timing starts.

Prints one JSON line.  Metric: star*lambda-bins/s = N * n_target / t(fwd+adj).

Roofline of the dominant launch: the per-star kernels are bound by the VALU
issue slots (their HBM traffic is a few per cent of the 8 TB/s peak), so
`bound` is "valu".  A SIMD issues a wave64 VALU instruction every 4 cycles, or
two in one 4-cycle slot when the hardware pairs them (rocprofv3
SQ_ACTIVE_INST_VALU2 counts those slots).  achieved = the launch's VALU issue
slots (SQ_INSTS_VALU - SQ_ACTIVE_INST_VALU2, totals over its waves, committed
under profiles/) / the launch's duration measured here with HIP events;
peak = 4 SIMDs x CUs x clock / 4 cycles, so frac <= 1 and frac = 1 is every
SIMD issuing in every slot.  `valu_wave_instructions` is the instruction count.
`hbm_frac_physical` is the launch's measured HBM bytes (FETCH_SIZE and
WRITE_SIZE, calibrated against a device copy) over its duration and 8 TB/s;
`effective_frac` the plain per-step byte model (every intermediate written
and read once in float32) over the duration, which a fused kernel that
recomputes instead of storing can exceed.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "star·λ-bins/s fwd+adjoint, 1M stars→50²×3721 cube; % HBM roofline vs plain"
HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4          # MI355X peak engine clock
SIMDS_PER_CU = 4
CYCLES_PER_VALU = 4      # one VALU issue slot: a wave64 instruction on a 16-lane SIMD


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n, n_src, n_tgt, n_cube):
    """Plain per-step model: every intermediate written once, read once, FP32.
    forward per star 16 n_src + 8 n_tgt; adjoint per star 24 n_src + 8 n_tgt;
    PSF + LSF 16 B per cube element forward and 16 adjoint."""
    fwd = n * (16 * n_src + 8 * n_tgt) + 16 * n_cube
    adj = n * (24 * n_src + 8 * n_tgt) + 16 * n_cube
    return fwd, adj


def cpu_baseline(w, sample, threads, reps=5, warmup=1, extinction=None, n_full=None):
    """The oracle's torch-CPU restatement of the plain per-step path (float32,
    autograd adjoint) on the first `sample` stars of workload w (its full
    grid, cube, PSF and LSF), all `threads` host cores: `warmup` untimed
    passes, then the median of `reps` timed passes.  The per-star stages
    (lookup, mass scaling, Doppler resampling, cube) and the cube stages
    (PSF, LSF) are timed apart, forward and adjoint each, so the rate can be
    stated for the workload's own star count n_full: the per-star time
    scales with the stars, the cube stages do not.
    Returns a dict of median stage times (s) and the rates at `sample` and
    at n_full stars (star*lambda-bins/s of forward + adjoint)."""
    import numpy as np
    import torch

    from oracle import ref_numpy as R, ref_torch as RT

    torch.set_num_threads(threads)
    st = {k: v[:sample] for k, v in w.stars.items()}
    f32 = torch.float32
    geo = w.geom
    ids = torch.tensor(R.square_spaxel_assignment(st["coords"], geo["edges"]).astype(np.int64))
    ext = None
    if extinction is not None:
        ext = (torch.tensor(extinction[0][:sample], dtype=f32), torch.tensor(extinction[1], dtype=f32))
    times = {"fwd_star": [], "fwd_cube": [], "adj_cube": [], "adj_star": []}
    m = geo["psf_size"]
    for it in range(warmup + reps):
        mass = torch.tensor(st["mass"], dtype=f32, requires_grad=True)
        age = torch.tensor(st["age"], dtype=f32, requires_grad=True)
        met = torch.tensor(st["metallicity"], dtype=f32, requires_grad=True)
        vl = torch.tensor(st["velocity"][:, 2], dtype=f32, requires_grad=True)
        sp = torch.tensor(geo["psf_sigma"], dtype=f32, requires_grad=True)
        sl = torch.tensor(geo["lsf_sigma"], dtype=f32, requires_grad=True)
        t0 = time.perf_counter()
        cube = RT.star_cube(mass, age, met, vl, ids, w.grid, geo, f32, extinction=ext)
        t1 = time.perf_counter()
        cd = cube.detach().requires_grad_(True)
        c1 = RT.apply_psf(cd, RT.gaussian_kernel_2d(m, m, sp))
        c2 = RT.apply_lsf(c1, RT.lsf_kernel(sl, geo["wave_res"]))
        loss = (c2 * c2).sum()
        t2 = time.perf_counter()
        loss.backward()
        t3 = time.perf_counter()
        cube.backward(cd.grad)
        t4 = time.perf_counter()
        if it >= warmup:
            for k, v in zip(times, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                times[k].append(v)
    med = {k: float(np.median(v)) for k, v in times.items()}
    nt = len(geo["wave_target"])
    star_s = med["fwd_star"] + med["adj_star"]
    cube_s = med["fwd_cube"] + med["adj_cube"]
    n_full = n_full or sample
    return {"sample_rate": sample * nt / (star_s + cube_s),
            "rate": n_full * nt / (star_s * n_full / sample + cube_s),
            "stage_s": {k: round(v, 4) for k, v in med.items()}}


def cpu_leg(w, sample, threads, label, reps=5, warmup=1, n_full=None, **kw):
    """One cpu_baseline measurement as a JSON-ready dict."""
    r = cpu_baseline(w, sample, threads, reps, warmup, n_full=n_full, **kw)
    return {"workload": label, "stars": sample, "stars_rated": n_full or sample,
            "value": r["rate"], "value_at_sample": r["sample_rate"], "stage_s": r["stage_s"],
            "reps": reps, "warmup": warmup}


def load_counters(name):
    """The JSON block of profiles/<name>.md: per C-ABI launch, its kernels'
    waves, VALU instructions per wave and HBM KiB fetched / written per
    dispatch (rocprofv3 counters, with the commands that made them), and
    the copy calibration of FETCH_SIZE / WRITE_SIZE."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    return dict(json.loads(m.group(1)), source=f"profiles/{name}.md") if m else None


def valu_peak_per_s():
    import torch

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return SIMDS_PER_CU * cus * CLOCK_GHZ * 1e9 / CYCLES_PER_VALU, cus


def roofline_of(counters, call, call_ms, algo_bytes):
    """Roofline object of one C-ABI launch from its committed counters and
    its duration measured in this run (None when no counters match)."""
    if not counters or call not in counters.get("launches", {}):
        return None
    L = counters["launches"][call]
    cal = counters["copy_calibration"]
    valu = sum(k["waves"] * k["valu_per_wave"] for k in L["kernels"].values())
    # 4-cycle slots that issued two VALU instructions (0 where not counted)
    dual = sum(k["waves"] * k.get("valu2_per_wave", 0.0) for k in L["kernels"].values())
    slots = valu - dual
    traffic = sum(k["fetch_kib"] * 1024 * cal["fetch_scale"] + k["write_kib"] * 1024 * cal["write_scale"]
                  for k in L["kernels"].values())
    peak, cus = valu_peak_per_s()
    t = call_ms * 1e-3
    achieved = slots / t
    return {"bound": "valu", "achieved": round(achieved / 1e9, 2), "peak": round(peak / 1e9, 2),
            "unit": "G VALU issue slots/s", "frac": round(achieved / peak, 4),
            "traffic": round(traffic), "kernel": call, "kernels": sorted(L["kernels"]),
            "kernel_ms": round(call_ms, 4), "valu_wave_instructions": round(valu),
            "valu_dual_issue_slots": round(dual), "valu_issue_slots": round(slots),
            "hbm_frac_physical": round(traffic / t / (HBM_PEAK_GBS * 1e9), 4),
            "effective_frac": round(algo_bytes / t / (HBM_PEAK_GBS * 1e9), 4),
            "algorithmic_bytes_per_launch": algo_bytes, "cus": cus, "clock_ghz": CLOCK_GHZ,
            "counters": counters["source"]}


# Kernels of each C-ABI launch (name prefixes after the namespace).
LAUNCH_KERNELS = {
    "fwd": ("band_key_kernel", "batch_keys_kernel", "fill_i32_kernel", "sort_walk_kernel",
            "scan_reduce_kernel", "scan_top_kernel", "scan_down_kernel", "sort_offsets_kernel",
            "star_prep_kernel", "chunk_count_kernel", "chunk_fill_kernel", "cube_fused_kernel",
            "star_norm_kernel", "star_sf_kernel", "cube_accum_kernel", "cube_combine_kernel"),
    "bwd": ("star_bwd_reg_kernel", "star_bwd_kernel", "star_grad_kernel"),
    "noise_fwd": ("flux_rows_kernel", "noise_s2n_kernel", "add_noise_kernel"),
    "noise_bwd": ("noise_bwd_rows_kernel", "noise_coef_kernel", "add_rows_kernel"),
}


def _kernel_key(name):
    n = name.split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1] if "<" not in n else n[n.find("::") + 2:] if "::" in n else n


def _pmc_rows(d):
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        rows += list(csv.DictReader(open(f)))
    return rows


def summarize_counters(args):
    """Markdown summary (with a JSON block bench.py reads) of rocprofv3 PMC
    runs of one workload: per C-ABI launch (--launch NAME=COUNT:fwd|bwd, COUNT
    launches of it in the run) the total SQ_WAVES, SQ_INSTS_VALU, FETCH_SIZE and
    WRITE_SIZE of its kernels per launch (and SQ_ACTIVE_INST_VALU2,
    SQ_INSTS_VALU_INT32 where the VALU pass counted them), and the copy
    calibration."""
    per = {}
    for kind, d in (("valu", args.valu_dir), ("fetch", args.fetch_dir), ("write", args.write_dir)):
        for r in _pmc_rows(d):
            k = _kernel_key(r["Kernel_Name"])
            c = r["Counter_Name"]
            e = per.setdefault(k, {})
            e[c] = e.get(c, 0.0) + float(r["Counter_Value"])
            if kind == "valu" and c == "SQ_WAVES":
                e["dispatches"] = e.get("dispatches", 0) + 1

    def copy_kib(d, counter):
        vals = [float(r["Counter_Value"]) for r in _pmc_rows(d)
                if "copyBuffer" in r["Kernel_Name"] and r["Counter_Name"] == counter]
        return max(vals)

    cal_bytes = 1 << 30
    fetch_scale = cal_bytes / 1024 / copy_kib(args.copy_fetch_dir, "FETCH_SIZE")
    write_scale = cal_bytes / 1024 / copy_kib(args.copy_write_dir, "WRITE_SIZE")
    launches = {}
    for spec in args.launch:
        name, rest = spec.split("=")
        count, kind = rest.split(":")
        count = int(count)
        ks = {}
        for k, e in per.items():
            base = k.split("<")[0]
            if base in LAUNCH_KERNELS[kind] and "SQ_WAVES" in e:
                ks[k] = {"waves": e["SQ_WAVES"] / count,
                         "valu_per_wave": e["SQ_INSTS_VALU"] / e["SQ_WAVES"],
                         "fetch_kib": e.get("FETCH_SIZE", 0.0) / count,
                         "write_kib": e.get("WRITE_SIZE", 0.0) / count,
                         "dispatches_per_launch": e["dispatches"] / count}
                if "SQ_ACTIVE_INST_VALU2" in e:
                    ks[k]["valu2_per_wave"] = e["SQ_ACTIVE_INST_VALU2"] / e["SQ_WAVES"]
                if "SQ_INSTS_VALU_INT32" in e:
                    ks[k]["int32_per_wave"] = e["SQ_INSTS_VALU_INT32"] / e["SQ_WAVES"]
        launches[name] = {"kernels": ks}
    label = args.workload_label or args.workload
    data = {"workload": label, "launches": launches,
            "copy_calibration": {"bytes": cal_bytes, "fetch_scale": round(fetch_scale, 5),
                                 "write_scale": round(write_scale, 5)}}
    out = [f"# Counters per launch: {label}", "",
           "Commands (from the repository root with TMPDIR=/tmp, after `python -m rubix_amd.build`):", "", "```"]
    out += [c for c in args.command]
    out += ["```", "",
            f"Copy calibration: a 1 GiB device copy reads FETCH_SIZE x {fetch_scale:.4f} and "
            f"WRITE_SIZE x {write_scale:.4f} = 1 GiB per dispatch; the bytes below are scaled by "
            "these factors.", "",
            "VALU issue slots = VALU instructions - dual-issue slots (SQ_ACTIVE_INST_VALU2: "
            "4-cycle slots in which a SIMD issued two VALU instructions).", "",
            "| launch | kernel | waves | VALU / wave | dual-issue slots / wave | VALU issue slots | "
            "HBM read B | HBM written B |",
            "|---|---|---|---|---|---|---|---|"]
    for ln, L in launches.items():
        for k, e in sorted(L["kernels"].items(), key=lambda kv: -kv[1]["waves"] * kv[1]["valu_per_wave"]):
            v2 = e.get("valu2_per_wave", 0.0)
            out.append(f"| {ln} | {k} | {e['waves']:.4g} | {e['valu_per_wave']:.1f} | {v2:.1f} | "
                       f"{e['waves'] * (e['valu_per_wave'] - v2):.4g} | "
                       f"{e['fetch_kib'] * 1024 * fetch_scale:.4g} | "
                       f"{e['write_kib'] * 1024 * write_scale:.4g} |")
    out += ["", "```json", json.dumps(data, indent=1), "```"]
    print("\n".join(out))


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
    # no Python garbage collection inside the timed steps (a full collection
    # of the autograd objects can stall the host for tens of ms)
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for i in range(steps):
            step(evs[i])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    finally:
        gc.enable()
    fwd = [a.elapsed_time(b) for a, b, _ in evs]
    adj = [b.elapsed_time(c) for _, b, c in evs]
    # start-to-start spans of the steps on the GPU (the last: start to end);
    # their sum against the wall time shows host-side stalls
    timed_leg.spans = [evs[i][0].elapsed_time(evs[i + 1][0]) for i in range(steps - 1)] + \
        [evs[-1][0].elapsed_time(evs[-1][2])]
    return elapsed, fwd, adj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("single", "large", "batched", "noise", "dust"),
                    default="single",
                    help="single: 1M stars -> 50x50x3721 (the headline); large: 10M stars -> "
                         "150x150x3721; batched: 64 ragged galaxies in one pass; noise: the "
                         "apply_noise step on a 50x50x3721 cube; dust: the single workload "
                         "with 1M gas cells and dust extinction (fused calc_dusty_ifu)")
    ap.add_argument("--n-stars", type=int, default=None)
    ap.add_argument("--workload-label", dest="workload_label", default=None)
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="stars of the workload timed on the host (default per workload)")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="timed host passes after one warm-up (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="timed steps of the plain per-step GPU path (0 = skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched workload: default launch tuning, no autotuning")
    ap.add_argument("--persist", type=int, default=None,
                    help="batched workload: k persistent workgroups per CU for the per-star "
                         "kernels (default tuning otherwise, no autotuning)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (twice): the reference "
                         "workload for scaling the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--counters-run", action="store_true",
                    help="one forward (the target) and one forward + adjoint step of the "
                         "workload and nothing else: the run rocprofv3 PMC passes profile")
    ap.add_argument("--summarize-counters", action="store_true",
                    help="print the counters summary of PMC output directories (no GPU)")
    ap.add_argument("--valu-dir")
    ap.add_argument("--fetch-dir")
    ap.add_argument("--write-dir")
    ap.add_argument("--copy-fetch-dir")
    ap.add_argument("--copy-write-dir")
    ap.add_argument("--launch", action="append", default=[],
                    help="NAME=COUNT:fwd|bwd for --summarize-counters")
    ap.add_argument("--command", action="append", default=[],
                    help="profiling command line recorded in the summary")
    args = ap.parse_args()
    if args.summarize_counters:
        summarize_counters(args)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.counters_run:
        args.steps, args.warmup, args.no_cpu, args.plain_steps = 1, 0, True, 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a "
              "design (sharding by stars + a lambda-slab reduce-scatter) and is not built.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "large":
        return run_large(args)
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "noise":
        return run_noise(args)
    if args.workload == "dust":
        args.plain_steps = 0
        return run_single(args, dust=True)
    return run_single(args)


def run_single(args, dust=False):
    import numpy as np
    import torch

    from rubix_amd import _lib, ops, synthetic as S
    from rubix_amd.autograd import FusedDatacube, GaussianLSF, GaussianPSF, render_plain

    dev = torch.device("cuda", 0)
    t_gen = time.perf_counter()
    w = S.config2(n_stars=args.n_stars or 1_000_000)
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

    ext_of = None
    if dust:
        # gas cells of the same disk; A_V from the column of each star's
        # spaxel, recomputed inside every step (Cardelli89, Rv 3.1)
        gas = {k: torch.as_tensor(v, device=dev) for k, v in S.make_gas(1_000_000, seed=3).items()}
        axav = ops.extinction_curve("Cardelli89", 3.1, torch.as_tensor(geo["wave_target"],
                                                                       device=dev))
        area = float(((geo["edges"][-1] - geo["edges"][0]) / nb) ** 2)

        def ext_of():
            gids = ops.spaxel_assign(gas["coords"], edges)
            sids = ops.spaxel_assign(t["coords"], edges)
            gext = ops.gas_extinction(gas["mass"], gas["metals"], "broken power law fit", "Z", 3.5,
                                      area)
            av = ops.column_av(gas["coords"], gids, gext, t["coords"], sids, nb * nb,
                               nbins * nbins)
            return av, axav

    def fused(mass, age, met, vel):
        ids = ops.spaxel_assign(t["coords"], edges)
        raw = FusedDatacube.apply(mass, age, met, vel, ids, ssp, tg, nb * nb, nbins * nbins, 2,
                                  ops.SPEED_OF_LIGHT, ext_of() if ext_of else None)
        c1 = GaussianPSF.apply(raw.view(nb, nb, nt), sp, geo["psf_size"])
        return GaussianLSF.apply(c1, sl, geo["wave_res"])

    def plain(mass, age, met, vel):
        return render_plain(t["coords"], vel, mass, age, met, edges, ssp, tg, nb, sp,
                            geo["psf_size"], sl, geo["wave_res"])

    with torch.no_grad():
        # fixed target: a second galaxy (seed 2) of the same size, S/N = 30 weights
        st2 = S.make_stars(w.n_stars, seed=2)
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
    # dominant launch: the per-star adjoint rbx_cube_bwd (star_bwd_reg_kernel,
    # then the small per-star star_grad_kernel); the forward launch second
    bwd_name = "rbx_cube_bwd_ext" if dust else "rbx_cube_bwd"
    fwd_name = "rbx_cube_fwd_ext" if dust else "rbx_cube_fwd"
    counters = None
    if (n, nb, nt, n_src) == (1_000_000, 50, 3721, 5000):
        counters = load_counters("counters_dust_1M_gas1M_50x50x3721" if dust
                                 else "counters_1M_50x50x3721")
    roofline = roofline_of(counters, bwd_name, per_call.get(bwd_name, float("nan")),
                           n * (24 * n_src + 8 * nt))
    if roofline is not None:
        roofline["forward_launch"] = roofline_of(counters, fwd_name,
                                                 per_call.get(fwd_name, float("nan")),
                                                 n * (16 * n_src + 8 * nt))
        roofline["step_effective_frac"] = round(step_frac, 4)
    split = {"fwd_ms_median": round(float(np.median(fwd_ms)), 4),
             "adj_ms_median": round(float(np.median(adj_ms)), 4),
             "step_ms_median": round(float(np.median(np.add(fwd_ms, adj_ms))), 4),
             "step_spans_ms": [round(x, 3) for x in timed_leg.spans]}

    plain_res = None
    if args.plain_steps > 0:
        # the plain per-step chain: one kernel per reference step, per-star
        # spectra materialised (N x n_source twice, N x n_target once)
        torch.cuda.reset_peak_memory_stats(dev)
        p_el, p_fwd, p_adj = timed_leg(make_step(plain), 1, args.plain_steps)
        p_ms = p_el / args.plain_steps * 1e3
        plain_roof_ms = (fwd_b + adj_b) / (HBM_PEAK_GBS * 1e9) * 1e3
        plain_res = {"ms_per_step": round(p_ms, 3), "value": n * nt / (p_el / args.plain_steps),
                     "fwd_ms_median": round(float(np.median(p_fwd)), 3),
                     "adj_ms_median": round(float(np.median(p_adj)), 3),
                     "hbm_frac": round((fwd_b + adj_b) / (p_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "peak_hbm_gb": round(torch.cuda.max_memory_allocated(dev) / 1e9, 2),
                     "fused_speedup": round(p_ms / ms, 3),
                     "plain_roofline_ms": round(plain_roof_ms, 3),
                     "speedup_vs_plain_roofline": round(plain_roof_ms / ms, 3),
                     "steps": args.plain_steps}
        log("plain:", json.dumps(plain_res))

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        ext_host = None
        if dust:
            av, axav = ext_of()
            ext_host = (av.cpu().numpy(), axav.cpu().numpy())
        sample = min(args.cpu_sample or 25_000, n)
        legs = [cpu_leg(w, sample, threads,
                        f"{sample} stars of this workload (full grid, cube, PSF, LSF)",
                        reps=args.cpu_reps, extinction=ext_host, n_full=n)]
        if not dust:
            w1 = S.config1()
            legs.append(cpu_leg(w1, w1.n_stars, threads,
                                "the whole small workload: 10k stars -> 25x25x1000, SSP 6x20x1000",
                                reps=args.cpu_reps))
        cpu = {"value": legs[0]["value"], "unit": "star·λ-bins/s", "cores": threads, "kind": "port",
               "sample": (f"{sample} stars of the same workload: float32 torch-CPU restatement "
                          f"of the plain per-step path"
                          f"{' with dust extinction (A_V per star given)' if dust else ''}, "
                          f"forward + autograd adjoint, 1 warm-up + median of {args.cpu_reps} "
                          f"passes on {threads} threads of {_model_name()}; per-star stages "
                          f"scaled to the {n} stars, the cube stages (PSF, LSF) once"),
               "legs": legs, "gpu_over_cpu": round(value / legs[0]["value"], 1)}
    workload = (f"single galaxy, {n} stars -> {nb}x{nb}x{nt} cube, SSP 12x100x{n_src}, "
                "PSF 15x15 sigma 4 px, LSF sigma 1.0 A; fused forward + adjoint of chi^2")
    if dust:
        workload += ("; dust extinction from 1000000 gas cells (Cardelli89, Rv 3.1, broken power "
                     "law D/G, A_V recomputed each step)")
    res = {
        "metric": METRIC + (" (with dust extinction)" if dust else ""), "value": value,
        "unit": "star·λ-bins/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload,
                   "n_stars": n, "n_spaxel": nb * nb, "n_target": nt, "n_source": n_src},
        "roofline": roofline, "cpu_baseline": cpu, "split": split, "plain": plain_res,
        "per_call_ms": {k: round(v, 4) for k, v in per_call.items()},
    }
    print(json.dumps(res), flush=True)



def _chi2_weights(tgt):
    import torch

    return 1.0 / torch.clamp((tgt.abs() / 30.0) ** 2, min=(1e-6 * float(tgt.abs().max())) ** 2)


def run_large(args):
    """10M stars -> 150 x 150 x 3721 with the small SSP grid: fused forward +
    adjoint of chi^2 with per-star spectra recomputed in the adjoint (149 GB
    if stored).  Reports peak HBM, and the recompute overhead
    t(fused adjoint) / t(stored-spectra adjoint of the plain path) at 1M
    stars of the same geometry, where storing fits."""
    import numpy as np
    import torch

    from rubix_amd import _lib, ops, synthetic as S
    from rubix_amd.autograd import render, render_plain

    dev = torch.device("cuda", 0)
    n = args.n_stars or 10_000_000
    t_gen = time.perf_counter()
    w = S.config5(n_stars=n)
    log(f"workload generated in {time.perf_counter() - t_gen:.1f}s: N={n}")
    g, geo = w.grid, w.geom
    nb, nt = geo["nb"], len(geo["wave_target"])
    n_src = g["flux"].shape[-1]
    ssp = ops.PreparedSSP(g["flux"], g["ax_met"], g["ax_age"], g["wave"], geo["z"], g["log_axes"], dev)
    tg = ops.TargetGrid(geo["wave_target"], dev)
    edges = torch.as_tensor(geo["edges"], device=dev)

    def leaves(m):
        t = {k: torch.as_tensor(v[:m], device=dev) for k, v in w.stars.items()}
        for k in ("mass", "age", "metallicity", "velocity"):
            t[k].requires_grad_(True)
        sp = torch.tensor(geo["psf_sigma"], device=dev, requires_grad=True)
        sl = torch.tensor(geo["lsf_sigma"], device=dev, requires_grad=True)
        return t, sp, sl

    def make_step(t, sp, sl, fn, tgt, wt):
        def fwd():
            return fn(t["coords"], t["velocity"], t["mass"], t["age"], t["metallicity"], edges,
                      ssp, tg, nb, sp, geo["psf_size"], sl, geo["wave_res"])

        def step(ev):
            for k in ("mass", "age", "metallicity", "velocity"):
                t[k].grad = None
            sp.grad = None
            sl.grad = None
            if ev:
                ev[0].record()
            out = fwd()
            loss = ((out - tgt) ** 2 * wt).sum()
            if ev:
                ev[1].record()
            loss.backward()
            if ev:
                ev[2].record()
        return fwd, step

    t, sp, sl = leaves(n)
    with torch.no_grad():
        fwd0, _ = make_step(t, sp, sl, render, None, None)
        tgt = fwd0() * 1.05
        wt = _chi2_weights(tgt)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    _, step = make_step(t, sp, sl, render, tgt, wt)
    _lib.TIMER.start()
    elapsed, fwd_ms, adj_ms = timed_leg(step, args.warmup, args.steps)
    calls = _lib.TIMER.stop()
    if args.counters_run:
        return
    per_call = {k: float(np.median(v[-args.steps:])) for k, v in calls.items()}
    ms = elapsed / args.steps * 1e3
    peak = torch.cuda.max_memory_allocated(dev) / 1e9
    fwd_b, adj_b = algorithmic_bytes(n, n_src, nt, nb * nb * nt)
    counters = load_counters("counters_10M_150x150x3721") if (n, nb, nt) == (10_000_000, 150, 3721) \
        else None
    roofline = roofline_of(counters, "rbx_cube_bwd", per_call.get("rbx_cube_bwd", float("nan")),
                           n * (24 * n_src + 8 * nt))
    if roofline is not None:
        roofline["forward_launch"] = roofline_of(counters, "rbx_cube_fwd",
                                                 per_call.get("rbx_cube_fwd", float("nan")),
                                                 n * (16 * n_src + 8 * nt))
    del t, sp, sl, tgt, wt
    torch.cuda.empty_cache()

    # recompute overhead at 1M stars of the same geometry
    m = min(n, 1_000_000)
    overhead = {}
    for name, fn in (("recompute", render), ("stored", render_plain)):
        t, sp, sl = leaves(m)
        with torch.no_grad():
            fwd0, _ = make_step(t, sp, sl, render, None, None)
            tgt = fwd0() * 1.05
            wt = _chi2_weights(tgt)
        _, step = make_step(t, sp, sl, fn, tgt, wt)
        _, f, a = timed_leg(step, 1, 3)
        overhead[name] = {"fwd_ms_median": round(float(np.median(f)), 3),
                          "adj_ms_median": round(float(np.median(a)), 3)}
        del t, sp, sl, tgt, wt
        torch.cuda.empty_cache()
    ratio = overhead["recompute"]["adj_ms_median"] / overhead["stored"]["adj_ms_median"]
    value = n * nt / (elapsed / args.steps)
    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        sample = min(args.cpu_sample or 10_000, n)
        leg = cpu_leg(w, sample, threads, f"{sample} stars of this workload (full grid, cube, PSF, "
                      "LSF)", reps=2, n_full=n)
        cpu = {"value": leg["value"], "unit": "star·λ-bins/s", "cores": threads, "kind": "port",
               "sample": (f"{sample} stars of the same workload: float32 torch-CPU restatement of "
                          f"the plain per-step path, forward + autograd adjoint, 1 warm-up + median "
                          f"of 2 passes on {threads} threads of {_model_name()}; per-star stages "
                          f"scaled to the {n} stars, the cube stages (PSF, LSF) once"),
               "legs": [leg], "gpu_over_cpu": round(value / leg["value"], 1)}
    res = {
        "metric": "star·λ-bins/s fwd+adjoint, 10M stars→150²×3721 cube (recompute)",
        "value": value, "unit": "star·λ-bins/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "roofline": roofline, "cpu_baseline": cpu,
        "per_call_ms": {k: round(v, 4) for k, v in per_call.items()},
        "config": {"workload": f"single galaxy, {n} stars -> {nb}x{nb}x{nt} cube, SSP 6x20x{n_src}, "
                               "PSF 11x11 sigma 3 px, LSF sigma 1.2 A; fused forward + adjoint of chi^2",
                   "n_stars": n, "n_spaxel": nb * nb, "n_target": nt, "n_source": n_src},
        "split": {"fwd_ms_median": round(float(np.median(fwd_ms)), 3),
                  "adj_ms_median": round(float(np.median(adj_ms)), 3)},
        "peak_hbm_gb": round(peak, 2),
        "per_star_spectra_if_stored_gb": round(n * nt * 4 / 1e9, 1),
        "step_effective_frac": round((fwd_b + adj_b) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        "recompute_overhead": {"n_stars": m, "adj_ratio_recompute_over_stored": round(ratio, 4),
                               **overhead},
    }
    print(json.dumps(res), flush=True)


def run_batched(args):
    """64 galaxies (star counts log-uniform on 5e4..2e6, spaxel grids from
    {25, 50, 100}, shared 3721-bin axis and small SSP grid): chi^2 forward +
    adjoint of the whole batch in one fused pass per stage, launch tuning
    from the autotuner, the step captured in a HIP graph and replayed.
    Compared with looping the plain per-galaxy path and the fused
    single-galaxy path over the same galaxies."""
    import numpy as np
    import torch

    from rubix_amd import _lib, autotune, ops, synthetic as S
    from rubix_amd.autograd import render, render_plain
    from rubix_amd.batch import GalaxyBatch, GraphedStep

    dev = torch.device("cuda", 0)
    t_gen = time.perf_counter()
    ws = S.config4()
    if args.n_stars:
        ws = S.config4(scale=args.n_stars / sum(w.n_stars for w in ws))
    log(f"batch generated in {time.perf_counter() - t_gen:.1f}s: "
        f"{len(ws)} galaxies, {sum(w.n_stars for w in ws)} stars")
    grid, geo = ws[0].grid, ws[0].geom
    nt = len(geo["wave_target"])
    n_src = grid["flux"].shape[-1]
    ssp = ops.PreparedSSP(grid["flux"], grid["ax_met"], grid["ax_age"], grid["wave"], geo["z"],
                          grid["log_axes"], dev)
    tg = ops.TargetGrid(geo["wave_target"], dev)
    gb = GalaxyBatch([w.stars for w in ws], [w.geom["nb"] for w in ws],
                     [w.geom["edges"] for w in ws], dev)
    L = gb.layout
    sp = torch.tensor(geo["psf_sigma"], device=dev, requires_grad=True)
    sl = torch.tensor(geo["lsf_sigma"], device=dev, requires_grad=True)
    state = {"tune": ops.Tuning()}

    def rend():
        return gb.render(ssp, tg, sp, geo["psf_size"], sl, geo["wave_res"], tune=state["tune"])

    with torch.no_grad():
        tgt = rend() * 1.05
        wt = _chi2_weights(tgt)
    inputs = gb.leaves + [sp, sl]

    def eager_step(tune):
        state["tune"] = tune
        cube = rend()
        cot = 2.0 * (cube.detach() - tgt) * wt
        torch.autograd.grad(cube, inputs, grad_outputs=cot)

    fixed = ops.Tuning(persist=args.persist) if args.persist is not None else None
    if args.counters_run:
        eager_step(fixed or ops.Tuning())
        torch.cuda.synchronize()
        return
    key = autotune.problem_key(L.n_total, L.n_rows, nt, n_src, dev,
                               n_templates=grid["flux"].shape[0] * grid["flux"].shape[1])
    t0 = time.perf_counter()
    if fixed is not None:
        best, seen = fixed, {}
    elif args.no_tune:
        best, seen = ops.Tuning(), {}
    else:
        best, seen = autotune.autotune(eager_step, key, reps=1, log=log)
    log(f"autotune {time.perf_counter() - t0:.1f}s -> {best}")
    state["tune"] = best
    _lib.TIMER.start()
    eager_ms = autotune.time_ms(eager_step, best, reps=3)
    calls = _lib.TIMER.stop()
    per_call = {k: float(np.median(v[-3:])) for k, v in calls.items()}
    default_ms = autotune.time_ms(eager_step, ops.Tuning(), reps=3)

    step = GraphedStep(rend, inputs, tgt, wt)
    for _ in range(args.warmup):
        step.replay()
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gc.enable()
    ms = elapsed / args.steps * 1e3
    work = L.n_total * nt

    def loop(fn):
        """One pass over the galaxies, each its own forward + adjoint."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for g, w in enumerate(ws):
            s, r = L.stars_of(g), L.rows_of(g)
            nb = w.geom["nb"]
            leaf = {k: gb.t[k][s].detach().requires_grad_(True)
                    for k in ("mass", "age", "metallicity", "velocity")}
            out = fn(gb.t["coords"][s], leaf["velocity"], leaf["mass"], leaf["age"],
                     leaf["metallicity"], gb.edges[g], ssp, tg, nb, sp, geo["psf_size"], sl,
                     geo["wave_res"])
            cot = 2.0 * (out.detach().reshape(nb * nb, nt) - tgt[r]) * wt[r]
            torch.autograd.grad(out, list(leaf.values()) + [sp, sl],
                                grad_outputs=cot.reshape(out.shape))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    loop(render)  # warm-up of allocations
    fused_loop_ms = loop(render)
    plain_loop_ms = loop(render_plain) if args.plain_steps > 0 else None
    value = work / (elapsed / args.steps)
    counters = load_counters("counters_batched_64gal_3721") if not args.n_stars else None
    roofline = roofline_of(counters, "rbx_batch_bwd", per_call.get("rbx_batch_bwd", float("nan")),
                           L.n_total * (24 * n_src + 8 * nt))
    if roofline is not None:
        roofline["forward_launch"] = roofline_of(counters, "rbx_batch_fwd",
                                                 per_call.get("rbx_batch_fwd", float("nan")),
                                                 L.n_total * (16 * n_src + 8 * nt))
        roofline["timing"] = "launch durations from the eager step (a graph replay is one launch)"
    cpu = None
    if not args.no_cpu:
        # one galaxy of each grid size, its per-star stages scaled to the
        # stars of all galaxies of that size and its cube stages once per galaxy
        threads = len(os.sched_getaffinity(0))
        legs, total_s = [], 0.0
        for nb in sorted({w.geom["nb"] for w in ws}):
            members = [w for w in ws if w.geom["nb"] == nb]
            rep = max(members, key=lambda w: w.n_stars)
            sample = min(args.cpu_sample or 5_000, rep.n_stars)
            r = cpu_baseline(rep, sample, threads, reps=2)
            st = r["stage_s"]
            n_class = sum(w.n_stars for w in members)
            t_class = ((st["fwd_star"] + st["adj_star"]) * n_class / sample
                       + (st["fwd_cube"] + st["adj_cube"]) * len(members))
            total_s += t_class
            legs.append({"nb": nb, "galaxies": len(members), "stars": n_class,
                         "sample_stars": sample, "stage_s": st, "class_s": round(t_class, 3)})
        cpu = {"value": work / total_s, "unit": "star·λ-bins/s", "cores": threads, "kind": "port",
               "sample": (f"per grid size one galaxy's first {args.cpu_sample or 5000} stars: float32 "
                          f"torch-CPU restatement of the plain per-step path, forward + autograd "
                          f"adjoint, 1 warm-up + median of 2 passes on {threads} threads of "
                          f"{_model_name()}; per-star stages scaled to the class's stars, cube "
                          f"stages once per galaxy, summed over the batch"),
               "legs": legs, "gpu_over_cpu": round(value / (work / total_s), 1)}
    res = {
        "metric": "star·λ-bins/s fwd+adjoint, 64 ragged galaxies in one pass (HIP graph)",
        "value": value, "unit": "star·λ-bins/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "roofline": roofline, "cpu_baseline": cpu,
        "per_call_ms_eager": {k: round(v, 4) for k, v in per_call.items()},
        "config": {"workload": f"{L.n_gal} galaxies, {L.n_total} stars, {L.n_rows} cube rows x {nt}, "
                               f"SSP 6x20x{n_src}, PSF 11x11, LSF sigma 1.2 A; chi^2 forward + adjoint",
                   "n_galaxies": L.n_gal, "n_stars": L.n_total, "n_rows": L.n_rows,
                   "n_target": nt, "n_source": n_src},
        "tuning": {"chunk": best.chunk, "waves": best.waves, "persist": best.persist,
                   "eager_ms": round(eager_ms, 3), "default_tuning_eager_ms": round(default_ms, 3),
                   "candidates_ms": {k: round(v, 3) for k, v in seen.items()}},
        "loop_fused_single_galaxy_ms": round(fused_loop_ms, 3),
        "loop_plain_per_galaxy_ms": None if plain_loop_ms is None else round(plain_loop_ms, 3),
        "speedup_vs_plain_loop": None if plain_loop_ms is None else round(plain_loop_ms / ms, 3),
        "speedup_vs_fused_loop": round(fused_loop_ms / ms, 3),
    }
    print(json.dumps(res), flush=True)



def run_noise(args):
    """apply_noise forward + adjoint on a 50 x 50 x 3721 cube with positive
    flux in every spaxel (so the S2N map is non-zero everywhere and every
    element is drawn), normal noise at S/N 1, seed 0."""
    import numpy as np
    import torch

    from rubix_amd import _lib
    from rubix_amd.autograd import ApplyNoise

    dev = torch.device("cuda", 0)
    nb, nt = 50, 3721
    rng = np.random.default_rng(0)
    c_np = rng.uniform(0.5, 2.0, (nb, nb, nt)).astype(np.float32)
    g = torch.as_tensor(rng.normal(size=(nb, nb, nt)).astype(np.float32), device=dev)
    x = torch.as_tensor(c_np, device=dev).requires_grad_(True)

    def step(ev):
        if ev:
            ev[0].record()
        out = ApplyNoise.apply(x, 1.0, "normal", 0)
        if ev:
            ev[1].record()
        out.backward(g)
        if ev:
            ev[2].record()
        x.grad = None

    if args.counters_run:
        step(None)
        torch.cuda.synchronize()
        return
    _lib.TIMER.start()
    elapsed, fwd_ms, adj_ms = timed_leg(step, args.warmup, args.steps)
    calls = _lib.TIMER.stop()
    per_call = {k: float(np.median(v[-args.steps:])) for k, v in calls.items()}
    el = nb * nb * nt
    ms = elapsed / args.steps * 1e3
    # forward: cube read by the flux pass and the noise pass, out written (12 B);
    # adjoint: dout and cube read, dcube written, then dcube += dL/dF (20 B)
    fwd_b, adj_b = 12 * el, 20 * el
    dom = max(("rbx_apply_noise", fwd_b), ("rbx_apply_noise_bwd", adj_b),
              key=lambda kv: per_call.get(kv[0], 0.0))
    other = ("rbx_apply_noise", fwd_b) if dom[0] == "rbx_apply_noise_bwd" else \
        ("rbx_apply_noise_bwd", adj_b)
    # the draw (Threefry + erfinv per element) makes both calls VALU-bound
    counters = load_counters("counters_noise_50x50x3721")
    roofline = roofline_of(counters, dom[0], per_call.get(dom[0], float("nan")), dom[1])
    if roofline is not None:
        roofline["other_launch"] = roofline_of(counters, other[0],
                                               per_call.get(other[0], float("nan")), other[1])
    cpu = None
    if not args.no_cpu:
        from oracle import ref_numpy as R

        t0 = time.perf_counter()
        R.apply_noise(c_np, 1.0, "normal", 0)
        R.apply_noise_vjp(c_np, 1.0, g.cpu().numpy(), "normal", 0)
        dt = time.perf_counter() - t0
        cpu = {"value": el / dt, "unit": "cube elements/s", "cores": 1, "kind": "port",
               "sample": f"the whole {nb}x{nb}x{nt} cube: numpy restatement (float32 forward, "
                         f"float64 adjoint), one pass {dt:.1f}s on {_model_name()}"}
    res = {
        "metric": "cube elements/s, apply_noise forward + adjoint", "value": el / (ms * 1e-3),
        "unit": "cube elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"apply_noise on a {nb}x{nb}x{nt} cube (all spaxels with positive "
                               "flux), normal, S/N 1, seed 0; forward + adjoint",
                   "n_spaxel": nb * nb, "n_target": nt},
        "roofline": roofline, "cpu_baseline": cpu,
        "split": {"fwd_ms_median": round(float(np.median(fwd_ms)), 4),
                  "adj_ms_median": round(float(np.median(adj_ms)), 4)},
        "per_call_ms": {k: round(v, 4) for k, v in per_call.items()},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
