"""The bench's last line (the record the round's driver parses): compact JSON
of at most 4 KB with the contract's keys, the roofline and CPU-baseline
objects, built by bench.build_summary from legs of the real shapes (every
leg present, so the line is at its longest)."""
import json
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _kernels():
    return {"reduce": {"ms_total": 0.0488, "launches": 1},
            "bc": {"ms_total": 3.81, "launches": 300},
            "stage": {"ms_total": 108.1, "launches": 300},
            "chem": {"ms_total": 43.8, "launches": 100}}


def _err_row(step):
    species = ["H2", "O2", "H", "O", "OH", "HO2", "H2O2", "H2O"]
    return {"step": step, "T": {"L2": 3.4559e-07, "Linf": 5.6757e-06},
            "rho": {"L2": 7.5245e-07, "Linf": 1.1e-05}, "p": {"L2": 6.3782e-07, "Linf": 9e-06},
            "Y": {k: {"L2": 1e-6 * (i + 1) / 3.0, "Linf": 3e-6 * (i + 1)}
                  for i, k in enumerate(species)}}


def _legs():
    mixed = {"value": 2569203718.0848174, "ms_per_step": 1.5569026200000025,
             "kernels": _kernels(), "health_flags": 0}
    fp64 = {"value": 1809441270.9985764, "ms_per_step": 2.2106, "kernels": _kernels(),
            "health_flags": 0}
    roofline = {"kernel": "stage", "bound": "hbm", "achieved": 1554.353602379968,
                "peak": 8000.0, "unit": "GB/s", "frac": 0.194294200297496,
                "traffic": 847600000.0, "algorithmic_bytes": 560000000.0,
                "avg_ms": 0.3602783814072609}
    roof_valu = {"kernel": "stage", "bound": "valu-issue", "instr_per_cell": 3231.0,
                 "floor_ms": 0.32867431640625, "avg_ms": 0.3602783814072609,
                 "frac": 0.9122787637782644}
    chem = {"bound": "fp64-valu", "achieved": 25.155636306781098, "peak": 78.6,
            "unit": "TFLOP/s", "frac": 0.32004626344505216, "avg_ms": 0.4388996511697769,
            "flop_per_cell": 2760.2}
    acc = {"every": [_err_row(100), _err_row(1000)]}
    crop = {"every": [_err_row(100)]}
    adj = {"adjoint_over_fwd_step": 3.622289749615781,
           "adjoint_over_fwd_rollout": 5.603567667534413, "step_fwd_ms": 0.21953230000022472,
           "step_vjp_ms": 0.7952096000003905, "fd_directions": 32,
           "fd_max_rel_err": {"1e-06": 2.7715763121658243e-05},
           "fd_switches": {"lf_4pt": 15},
           "fd_switch_free": {
               "4pt_frozen_eps1e-6": {"directions": 17, "max_rel_err": 6.680557577221256e-08,
                                      "median_rel_err": 1.75e-08, "within_1e-6": True},
               "2pt_truncation_corrected": {"directions": 32,
                                            "max_rel_err": 7.619928319896228e-07,
                                            "median_rel_err": 1.38e-07, "within_1e-6": True}}}
    bat = {"value": 2005624356.2342854, "speedup_vs_sequential": 3.1749270816353254,
           "bit_identical_to_sequential": True}
    amr = {"workload": "blocks", "blocks_40x50": {"value": 648020188.16, "library_value":
                                                  1156240690.61, "n_blk": 2000},
           "blocks_20x20": {"value": 358180483.61, "library_value": 506728406.63,
                            "n_blk": 10000}}
    ann = {"value": 1613250220.9674447, "ms_per_step": 0.9917866299999645}
    res = [(16 * 250 * 125, 10.0 + 0.01 * w, 15.5 + 0.01 * w) for w in range(15)]
    cpu = bench._cpu_result(res, 15, 16)
    return mixed, fp64, roofline, roof_valu, chem, acc, crop, adj, bat, amr, ann, cpu


def test_summary_line_is_compact_and_complete():
    mixed, fp64, roofline, roof_valu, chem, acc, crop, adj, bat, amr, ann, cpu = _legs()
    s = bench.build_summary(100, 10, 8000, 500, mixed, fp64, roofline, roof_valu, chem, acc,
                            crop, adj, bat, amr, ann, cpu)
    line = json.dumps(s)
    assert len(line.encode()) <= 4096, len(line)
    assert "\n" not in line
    d = json.loads(line)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"):
        assert k in d, k
    assert d["metric"] == bench.METRIC
    assert d["value"] == mixed["value"] and d["n_gpus"] == 1
    assert d["steps"] == 100 and d["warmup"] == 10
    assert d["higher_is_better"] is True and d["vs_baseline"] is None
    assert d["config"]["workload"] and d["config"]["cells"] == 4000000
    r = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in r, k
    assert r["bound"] in ("hbm", "mfma")
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    c = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in c, k
    assert c["kind"] in ("reference", "port") and c["cores"] == 15
    assert d["speedup_mixed_over_fp64"] == pytest.approx(mixed["value"] / fp64["value"])
    assert d["adjoint_over_fwd_step"] == adj["adjoint_over_fwd_step"]
    assert d["mixed_err_vs_fp64_gpu"]["steps"] == 1000


def test_summary_without_optional_legs():
    mixed, fp64, roofline, roof_valu, chem = _legs()[:5]
    d = bench.build_summary(5, 1, 8000, 500, mixed, fp64, roofline, None, None, None, None,
                            None, None, None, None, None)
    for k in ("adjoint", "batch", "cpu_baseline", "mixed_err_vs_fp64_gpu",
              "fp64_annulus_4000x400"):
        assert k not in d
    assert d["value"] == mixed["value"]
