"""CPU checks of bench.py's measurement bookkeeping: the byte model the
roofline divides by, the counter-summary names its traffic figures come from,
and the one-GPU guard.  No GPU work runs here."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_of_the_headline_kernels():
    # FP64, 5 variables, no scalars: every array read or written once
    assert bench.algorithmic_bytes_per_cell("flux_x1", 5) == 5 * 8 + 5 * 8
    assert bench.algorithmic_bytes_per_cell("flux_x3", 5) == 2 * 5 * 8
    assert bench.algorithmic_bytes_per_cell("update", 5) == 4 * 5 * 8  # w, F1, F3, du
    assert bench.algorithmic_bytes_per_cell("implicit", 5, implicit_flag=9) == 5 * 5 * 8
    assert bench.algorithmic_bytes_per_cell("implicit", 5, implicit_flag=8) == 4 * 5 * 8
    assert bench.algorithmic_bytes_per_cell("prim", 5) == 2 * 5 * 8
    assert bench.algorithmic_bytes_per_cell("dt", 5) == 5 * 8
    # mixed: FP32 face fluxes, the x2/x3 faces read the FP32 copy fw
    assert bench.algorithmic_bytes_per_cell("flux_x3", 5, mixed=True) == 2 * 5 * 4
    assert bench.algorithmic_bytes_per_cell("flux_x1", 5, mixed=True) == 5 * 8 + 5 * 4 + 8
    assert bench.algorithmic_bytes_per_cell("no_such_kernel", 5) is None
    # the survey's fused-stage figure: 8 V s + three 120 B column solves
    assert bench.step_compulsory_bytes_per_cell(5, False) == 680


def test_counter_names_resolve_in_the_committed_traffic_profile():
    """Every kernel the roofline and the per-step traffic read from the
    committed FETCH_SIZE / WRITE_SIZE passes is present there under the name
    rocprofv3 gives the launched instantiation."""
    t = bench.traffic_profile()
    assert t is not None and t["fetch_scale"] > 0 and t["write_scale"] > 0
    for form in ("fp64", "mixed"):
        for kernel in bench.PMC_NAMES[form]:
            name = bench.pmc_name(kernel, form)
            assert name in t["kernels"], (form, kernel, name)
            assert t["kernels"][name]["hbm_bytes"] > 0
        per_cell = bench.step_traffic_per_cell(form, 6 * 96 * 96 * 64)
        assert per_cell is not None and 680 < per_cell < 10000
    # the tuned x1 shape selects that instantiation's entry
    assert bench.pmc_name("flux_x1", "fp64", {"x1_shape": 2}) == "k_flux_x1<2, -1>"
    assert bench.pmc_name("flux_x1", "mixed", {"x1_mix_shape": 8}) == "k_flux_x1_mix<8, 3>"
    assert "measured_at" in t


def test_relative_error_norms():
    # [panel][var][k][j][i] with 3 ghost layers on each side (the norms take
    # interior cells)
    a = np.arange(2 * 5 * 8 * 8 * 8, dtype=float).reshape(2, 5, 8, 8, 8) + 1.0
    assert np.all(bench.rel_l2(a, a) == 0) and np.all(bench.rel_linf(a, a) == 0)
    b = a.copy()
    b[:, 1] *= 1.0 + 1e-6
    e = bench.rel_l2(b, a)
    assert e.shape == (5,) and e[1] == pytest.approx(1e-6, rel=1e-6)
    assert np.all(np.delete(e, 1) == 0)


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_metric_line_keys_are_documented_in_the_contract():
    # the JSON line's fixed keys, as bench.py writes them (static check of the
    # source, since writing the line needs the GPU run)
    src = open(os.path.join(ROOT, "bench.py")).read()
    for key in ('"metric"', '"value"', '"unit"', '"n_gpus"', '"steps"', '"warmup"',
                '"ms_per_step"', '"higher_is_better"', '"scaling"', '"vs_baseline"', '"dtype"',
                '"data"', '"config"', '"roofline"', '"cpu_baseline"'):
        assert key in src, key
    base = json.load(open(os.path.join(ROOT, "BASELINE.json")))
    assert base["metric"] in src
