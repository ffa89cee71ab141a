"""CPU checks of bench.py's measurement plumbing: the committed counter
summaries parse into the roofline object the bench lines carry, the
roofline arithmetic, the counter summariser on a small synthetic PMC output,
and the one-GPU guard."""
import csv
import json
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SUMMARIES = {
    "counters_1M_50x50x3721": ("rbx_cube_fwd", "rbx_cube_bwd"),
    "counters_dust_1M_gas1M_50x50x3721": ("rbx_cube_fwd_ext", "rbx_cube_bwd_ext"),
    "counters_10M_150x150x3721": ("rbx_cube_fwd", "rbx_cube_bwd"),
    "counters_batched_64gal_3721": ("rbx_batch_fwd", "rbx_batch_bwd"),
}


@pytest.mark.parametrize("name", sorted(SUMMARIES))
def test_committed_counter_summaries_parse(name):
    c = bench.load_counters(name)
    assert c is not None, name
    assert c["source"] == f"profiles/{name}.md"
    fwd, bwd = SUMMARIES[name]
    assert set(c["launches"]) == {fwd, bwd}
    # a 1 GiB copy reads FETCH_SIZE = 0.5 GiB and WRITE_SIZE = 1 GiB
    assert c["copy_calibration"]["fetch_scale"] == pytest.approx(2.0, rel=1e-3)
    assert c["copy_calibration"]["write_scale"] == pytest.approx(1.0, rel=1e-3)
    kern = c["launches"][bwd]["kernels"]
    assert any(k.startswith("star_bwd") for k in kern) and "star_grad_kernel" in kern
    kern = c["launches"][fwd]["kernels"]
    assert any(k.startswith("cube_fused_kernel") for k in kern)
    for L in c["launches"].values():
        for k in L["kernels"].values():
            assert k["waves"] > 0 and k["valu_per_wave"] > 0
            assert k["fetch_kib"] >= 0 and k["write_kib"] >= 0
            # a dual-issue slot holds two of the kernel's instructions
            assert 0 <= k["valu2_per_wave"] <= k["valu_per_wave"] / 2


def test_noise_counter_summary_parses():
    c = bench.load_counters("counters_noise_50x50x3721")
    assert c is not None
    assert set(c["launches"]) == {"rbx_apply_noise", "rbx_apply_noise_bwd"}
    assert set(c["launches"]["rbx_apply_noise"]["kernels"]) == set(bench.LAUNCH_KERNELS["noise_fwd"])
    assert set(c["launches"]["rbx_apply_noise_bwd"]["kernels"]) == \
        set(bench.LAUNCH_KERNELS["noise_bwd"])


def test_roofline_arithmetic(monkeypatch):
    counters = {
        "source": "profiles/x.md",
        "copy_calibration": {"fetch_scale": 2.0, "write_scale": 1.0},
        "launches": {"rbx_cube_bwd": {"kernels": {
            "a": {"waves": 1000.0, "valu_per_wave": 700.0, "valu2_per_wave": 150.0,
                  "fetch_kib": 1024.0, "write_kib": 512.0},
            "b": {"waves": 10.0, "valu_per_wave": 100.0, "fetch_kib": 0.0, "write_kib": 1024.0},
        }}},
    }
    monkeypatch.setattr(bench, "valu_peak_per_s", lambda: (1.0e9, 256))
    r = bench.roofline_of(counters, "rbx_cube_bwd", 2.0, 8.0e9)
    valu = 1000 * 700 + 10 * 100
    # a slot that issued two instructions counts once; kernel b has no dual-issue count
    slots = valu - 1000 * 150
    assert r["bound"] == "valu"
    assert r["valu_wave_instructions"] == valu
    assert r["valu_dual_issue_slots"] == 1000 * 150
    assert r["valu_issue_slots"] == slots
    # achieved in G issue slots/s (2 decimals), frac to 4 decimals
    assert r["achieved"] == pytest.approx(slots / 2e-3 / 1e9, abs=0.006)
    assert r["frac"] == pytest.approx(slots / 2e-3 / 1e9, abs=1e-4)
    traffic = (1024 * 2.0 + 512 + 1024) * 1024
    assert r["traffic"] == round(traffic)
    assert r["hbm_frac_physical"] == pytest.approx(traffic / 2e-3 / 8e12, abs=1e-4)
    assert r["effective_frac"] == pytest.approx(8e9 / 2e-3 / 8e12, abs=1e-4)
    assert bench.roofline_of(counters, "rbx_cube_fwd", 2.0, 1.0) is None
    assert bench.roofline_of(None, "rbx_cube_bwd", 2.0, 1.0) is None


def _pmc_csv(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_summarize_counters_on_synthetic_runs(tmp_path, capsys):
    k_fwd = "void rbx::cube_fused_kernel<false, 4>(rbx::SspView, ...)"
    k_bwd = "void rbx::star_bwd_reg_kernel<4, false>(rbx::SspView, ...)"
    d = {n: str(tmp_path / n) for n in ("valu", "fetch", "write", "cf", "cw")}
    # two forward launches and one adjoint launch in the run
    _pmc_csv(d["valu"], [
        {"Dispatch_Id": 1, "Kernel_Name": k_fwd, "Counter_Name": "SQ_WAVES", "Counter_Value": 100},
        {"Dispatch_Id": 1, "Kernel_Name": k_fwd, "Counter_Name": "SQ_INSTS_VALU", "Counter_Value": 5000},
        {"Dispatch_Id": 2, "Kernel_Name": k_fwd, "Counter_Name": "SQ_WAVES", "Counter_Value": 100},
        {"Dispatch_Id": 2, "Kernel_Name": k_fwd, "Counter_Name": "SQ_INSTS_VALU", "Counter_Value": 5000},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "SQ_WAVES", "Counter_Value": 40},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "SQ_INSTS_VALU", "Counter_Value": 8000},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "SQ_ACTIVE_INST_VALU2",
         "Counter_Value": 2000},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "SQ_INSTS_VALU_INT32",
         "Counter_Value": 800},
    ])
    _pmc_csv(d["fetch"], [
        {"Dispatch_Id": 1, "Kernel_Name": k_fwd, "Counter_Name": "FETCH_SIZE", "Counter_Value": 10},
        {"Dispatch_Id": 2, "Kernel_Name": k_fwd, "Counter_Name": "FETCH_SIZE", "Counter_Value": 30},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "FETCH_SIZE", "Counter_Value": 7},
    ])
    _pmc_csv(d["write"], [
        {"Dispatch_Id": 1, "Kernel_Name": k_fwd, "Counter_Name": "WRITE_SIZE", "Counter_Value": 4},
        {"Dispatch_Id": 3, "Kernel_Name": k_bwd, "Counter_Name": "WRITE_SIZE", "Counter_Value": 2},
    ])
    cb = "__amd_rocclr_copyBuffer"
    _pmc_csv(d["cf"], [{"Dispatch_Id": 1, "Kernel_Name": cb, "Counter_Name": "FETCH_SIZE",
                        "Counter_Value": 524288}])
    _pmc_csv(d["cw"], [{"Dispatch_Id": 1, "Kernel_Name": cb, "Counter_Name": "WRITE_SIZE",
                        "Counter_Value": 1048576}])
    args = bench.argparse.Namespace(
        valu_dir=d["valu"], fetch_dir=d["fetch"], write_dir=d["write"], copy_fetch_dir=d["cf"],
        copy_write_dir=d["cw"], launch=["rbx_cube_fwd=2:fwd", "rbx_cube_bwd=1:bwd"],
        workload_label="synthetic", workload="single", command=["rocprofv3 ..."])
    bench.summarize_counters(args)
    out = capsys.readouterr().out
    data = json.loads(out.split("```json\n")[1].split("\n```")[0])
    assert data["copy_calibration"]["fetch_scale"] == pytest.approx(2.0)
    assert data["copy_calibration"]["write_scale"] == pytest.approx(1.0)
    f = data["launches"]["rbx_cube_fwd"]["kernels"]["cube_fused_kernel<false, 4>"]
    assert f["waves"] == 100 and f["valu_per_wave"] == 50
    assert f["fetch_kib"] == 20 and f["write_kib"] == 2 and f["dispatches_per_launch"] == 1
    b = data["launches"]["rbx_cube_bwd"]["kernels"]["star_bwd_reg_kernel<4, false>"]
    assert b["waves"] == 40 and b["valu_per_wave"] == 200 and b["fetch_kib"] == 7
    assert b["valu2_per_wave"] == 50 and b["int32_per_wave"] == 20
    assert "valu2_per_wave" not in f
    # the table's issue-slot column: waves x (VALU - dual-issue slots) per wave
    assert "| rbx_cube_bwd | star_bwd_reg_kernel<4, false> | 40 | 200.0 | 50.0 | 6000 |" in out


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
