"""Benchmark of the cubed-sphere hydro update on one MI355X.

Workload (the headline configuration): 3D compressible ideal-gas atmosphere on
the cubed sphere, 6 panels x 96 x 96 cells x 64 levels, 5 conserved variables,
isothermal 250 K hydrostatic base (surface 1e5 Pa, top 30 km), seeded 1 %
Gaussian temperature noise and 5 m/s random-phase zonal waves; WENO5 + LMARS
horizontal fluxes, implicit 5x5 block-tridiagonal vertical solve
(implicit_flag 9), SSPRK3 with the CFL time step (CFL 0.4).

A "step" is one full SSPRK3 step (three stages) of the whole instance; the unit
of work is one interior cell advanced by one step (Athena's zone-cycle).
Inputs are resident in HBM before the timed region.  One JSON line is printed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def workload_config(ex, precision):
    return ex.make_config(nx=96, nx1=64, eos=ex.EOS_IDEAL, rsolver=ex.RS_LMARS,
                          implicit_flag=9, ntracer=0, coriolis=ex.COR_NONE, cfl_x1=1,
                          x1min=6371000.0, x1max=6371000.0 + 30000.0, gamma=1.4,
                          grav=-9.81, omega=7.292e-5, cfl=0.4, precision=precision)


IC_PARAMS = (250.0, 1.0e5, 287.0, 0.01, 5.0)


def algorithmic_bytes_per_cell(kernel, nv, s=8):
    """Compulsory HBM bytes per interior cell and launch."""
    return {
        "flux_x1": 2 * nv * s,      # read w column, write x1 face flux
        "flux_x2": 2 * nv * s,      # read w, write x2 face flux
        "flux_x3": 2 * nv * s,      # read w, write x3 face flux
        "update": 9 * nv * s,       # read u, u1, w, F1, F2, F3; write u, u1, du
        "implicit": 5 * nv * s,     # read du, w, u; write du, u
        "prim": 2 * nv * s,         # read u, write w
    }.get(kernel)


def rel_l2(a, b):
    """relative L2 error of a against b per conserved variable (interior cells)"""
    I = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    d = (a[I] - b[I]).reshape(a.shape[0], a.shape[1], -1)
    r = b[I].reshape(d.shape)
    return np.sqrt((d ** 2).sum(axis=(0, 2)) / (r ** 2).sum(axis=(0, 2)))


def run_gpu(ex, precision, steps, warmup):
    cfg = workload_config(ex, precision)
    cells = 6 * cfg.nx * cfg.nx * cfg.nx1
    u0, s0 = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    if warmup:
        g.step(warmup)
    g.sync()
    g.profile_begin()
    t0 = time.perf_counter()
    dt = g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    prof = g.profile_end()
    u, _ = g.download()
    finite = bool(np.isfinite(u).all())
    g.close()
    sec = t1 - t0
    return dict(cells=cells, sec=sec, dt=dt, prof=prof, finite=finite, nv=5, u=u)


def cpu_baseline(cfg_fn, ex):
    from oracle.oracle import Oracle, lib

    cores = int(os.environ.get("EXO_CPU_THREADS", "0")) or min(16, os.cpu_count() or 1)
    lib().or_set_threads(cores)
    cfg = cfg_fn(ex, 0)
    u0, _ = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    o = Oracle(cfg)
    o.upload(u0)
    t0 = time.perf_counter()
    o.step(1)
    sec = time.perf_counter() - t0
    o.close()
    cells = 6 * cfg.nx * cfg.nx * cfg.nx1
    return {"value": cells / sec, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"1 SSPRK3 step of the full 6x96x96x64 workload, {sec:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import exocubed_amd as ex

    r64 = run_gpu(ex, ex.PREC_FP64, args.steps, args.warmup)
    rmx = run_gpu(ex, ex.PREC_MIXED, args.steps, args.warmup)
    # the same seeded run in both forms: accuracy of the mixed form vs FP64
    l2 = rel_l2(rmx["u"], r64["u"])
    r = rmx
    ms_step = r["sec"] * 1e3 / args.steps
    value = r["cells"] * args.steps / r["sec"]
    value64 = r64["cells"] * args.steps / r64["sec"]
    # dominant kernel of the reported (mixed) form and its roofline
    prof = {k: v for k, v in r["prof"].items() if v[1] > 0}
    dom = max(prof, key=lambda k: prof[k][0])
    tot_ms, cnt = prof[dom]
    avg_s = tot_ms / cnt * 1e-3
    s_store = 4 if dom.startswith("flux") else 8
    bpc = algorithmic_bytes_per_cell(dom, r["nv"])
    if bpc and dom.startswith("flux"):
        bpc = r["nv"] * 8 + r["nv"] * s_store  # FP64 state read, FP32 flux written
    achieved = (bpc * r["cells"] / avg_s / 1e9) if bpc else None
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": None, "avg_launch_ms": tot_ms / cnt,
            "bytes_per_cell": bpc}
    kernels = {k: {"ms_per_launch": v[0] / v[1], "launches": v[1],
                   "share": v[0] / sum(x[0] for x in prof.values())} for k, v in prof.items()}
    prof64 = {k: v for k, v in r64["prof"].items() if v[1] > 0}
    kernels64 = {k: {"ms_per_launch": v[0] / v[1], "launches": v[1]} for k, v in prof64.items()}
    cpu = None if args.no_cpu else cpu_baseline(workload_config, ex)
    out = {
        "metric": "cell-updates/s mixed vs FP64 on 1 MI355X at <=1e-5 rel L2; p50/p99 run latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fluxes / f64 state, decomposition, implicit solve",
        "data": "synthetic (seeded isothermal atmosphere, 1% temperature noise)",
        "config": {"workload": "exo3 3D ideal gas, 6x96x96 cells x 64 levels, WENO5+LMARS, "
                               "implicit 5x5 vertical, SSPRK3 CFL 0.4", "precision": "mixed",
                   "cells": r["cells"], "dt_last_s": r["dt"], "state_finite": r["finite"]},
        "fp64": {"value": value64, "ms_per_step": r64["sec"] * 1e3 / args.steps,
                 "dt_last_s": r64["dt"], "kernels": kernels64},
        "mixed_over_fp64": value / value64,
        "mixed_rel_l2_vs_fp64": {n: float(x) for n, x in zip(("rho", "m1", "m2", "m3", "E"), l2)},
        "mixed_rel_l2_max": float(l2.max()),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
