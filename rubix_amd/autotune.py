"""Autotuner of the fused kernels' launch parameters (rbx_tuning).

The tunables are the stars per reduction chunk of the cube accumulation
(more chunks = more parallel work items but more partial rows to combine),
the wavefronts per workgroup of the per-star kernels, and whether those
kernels run as one workgroup per work item or as a persistent grid of
workgroups per compute unit (8 or more: fewer leave the CUs underoccupied).  Which is best depends on the star count, the
number of spaxels (stars per spaxel sets the chunk lengths) and the
wavelength axes, so the tuner times the caller's own step on the device and
caches the winner per (device, problem shape) in a JSON file.

Search: coordinate descent from the defaults, one parameter at a time
(waves, then persist, then chunk), each candidate timed with HIP events
(median of `reps` runs after one warm-up run).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from pathlib import Path

import torch

from .ops import Tuning

CANDIDATES = {
    "waves": (1, 2, 4, 8),
    "persist": (0, 8, 16, 32),
    "chunk": (64, 128, 256, 512, 1024),
}


def cache_path() -> Path:
    p = os.environ.get("RUBIX_AMD_TUNE_CACHE")
    if p:
        return Path(p)
    return Path.home() / ".cache" / "rubix_amd" / "tuning.json"


def problem_key(n_stars: int, n_rows: int, n_target: int, n_wave: int, device=None,
                n_templates: int = 0) -> str:
    """Cache key: device name and CU count, star and row counts bucketed to
    powers of two, the two wavelength axis lengths and the number of
    templates of the grid (n_met x n_age: with n_wave it sets the template
    bytes, which decide the number of age bands of the cube's sort keys)."""
    props = torch.cuda.get_device_properties(device)
    nb = max(0, int(math.log2(max(n_stars, 1))))
    rb = max(0, int(math.log2(max(n_rows, 1))))
    return (f"{props.name}|cu{props.multi_processor_count}|n2^{nb}|rows2^{rb}"
            f"|nt{n_target}|nw{n_wave}|tpl{n_templates}")


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _store(path: Path, data: dict) -> None:
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def time_ms(run, tuning: Tuning, reps: int = 3) -> float:
    """Median device time of run(tuning) over `reps` runs after one warm-up."""
    run(tuning)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        run(tuning)
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def autotune(run, key: str, reps: int = 3, use_cache: bool = True, log=None):
    """Best Tuning for `run` (a callable taking a Tuning and running one step
    on the current stream).  Returns (tuning, {candidate: ms})."""
    path = cache_path()
    data = _load(path) if use_cache else {}
    if use_cache and key in data:
        d = data[key]
        return Tuning(d["chunk"], d["waves"], d["persist"], d.get("flags", 0)), {}
    best = Tuning()
    seen = {best: time_ms(run, best, reps)}
    for name in ("waves", "persist", "chunk"):
        cands = [dataclasses.replace(best, **{name: v}) for v in CANDIDATES[name]]
        for cand in cands:
            if cand not in seen:
                seen[cand] = time_ms(run, cand, reps)
                if log:
                    log(f"autotune {cand}: {seen[cand]:.3f} ms")
        best = min(cands, key=lambda t: seen[t])
    if use_cache:
        data = _load(path)
        data[key] = {"chunk": best.chunk, "waves": best.waves, "persist": best.persist,
                     "ms": seen[best]}
        _store(path, data)
    return best, {str(t): ms for t, ms in seen.items()}
