"""The autotuner's search and cache (host logic; the timing function is
replaced by a known cost so no GPU is needed)."""
import json

from rubix_amd import autotune as A
from rubix_amd.ops import Tuning


def test_descent_finds_separable_optimum_and_caches(tmp_path, monkeypatch):
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("RUBIX_AMD_TUNE_CACHE", str(cache))
    calls = []

    def cost(run, t, reps=3):
        calls.append(t)
        return abs(t.waves - 2) + 0.5 * abs(t.persist - 16) + abs(t.chunk - 512) / 256

    monkeypatch.setattr(A, "time_ms", cost)
    best, seen = A.autotune(None, "key-a")
    assert best == Tuning(chunk=512, waves=2, persist=16)
    # defaults + 3 other waves + 3 other persists + 4 other chunks, each timed once
    assert len(calls) == len(set(calls)) == 11
    data = json.loads(cache.read_text())
    assert data["key-a"]["chunk"] == 512 and data["key-a"]["waves"] == 2
    calls.clear()
    again, seen2 = A.autotune(None, "key-a")
    assert again == best and calls == [] and seen2 == {}


def test_unwritable_cache_is_not_fatal(tmp_path, monkeypatch):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    monkeypatch.setenv("RUBIX_AMD_TUNE_CACHE", str(blocker / "sub" / "tune.json"))
    monkeypatch.setattr(A, "time_ms", lambda run, t, reps=3: float(t.chunk))
    best, _ = A.autotune(None, "k")
    assert best.chunk == min(A.CANDIDATES["chunk"])
