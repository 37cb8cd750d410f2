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

Two further measurements ride in the same line:
  ensemble -- 256 independent 2D shallow-water instances, panel sizes drawn
              from {24..192}, mean depths 2000-8000 m, stepped as one batch
              (one launch per phase, autotuned persistent widths, each step a
              replayed HIP graph), FP64 and mixed, with the compensated mass
              drift and the batched-vs-one-at-a-time speed-up;
  serving  -- a Poisson stream (40 req/s) of 3D forward runs (N 32..128,
              32..96 levels, 5 variables + 2 tracers, base 150-300 K, 200
              steps) through the continuous-batching server: p50/p99 latency
              and aggregate cell-updates/s, FP64 and mixed.
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
# autotune launch widths and shapes before timing (off with --no-tune: a
# profiler's counter passes slow kernels unevenly and would mislead the tuner,
# so the profiled runs take the default launches)
TUNE = True


def log(msg):
    """progress on stderr (the JSON line alone goes to stdout)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def workload_config(ex, precision, implicit_flag=9):
    return ex.make_config(nx=96, nx1=64, eos=ex.EOS_IDEAL, rsolver=ex.RS_LMARS,
                          implicit_flag=implicit_flag, ntracer=0, coriolis=ex.COR_NONE, cfl_x1=1,
                          x1min=6371000.0, x1max=6371000.0 + 30000.0, gamma=1.4,
                          grav=-9.81, omega=7.292e-5, cfl=0.4, precision=precision)


IC_PARAMS = (250.0, 1.0e5, 287.0, 0.01, 5.0)


def algorithmic_bytes_per_cell(kernel, nv, mixed=False, nt=0, implicit_flag=9):
    """Compulsory HBM bytes per interior cell and launch of each kernel: every
    array it must read or write once (state FP64; face fluxes FP32 in the
    mixed ideal gas, which also stores the FP64 x1 face pressure and whose
    x2/x3 face kernels read the FP32 perturbation copy fw instead of w)."""
    f = 4 if mixed else 8
    return {
        # read w, write the x1 face flux (mixed: + the FP64 face pressure)
        "flux_x1": nv * 8 + nv * f + (8 if mixed else 0) + nt * (8 + f),
        # read w (FP64 form) or fw (mixed), write the face flux
        "flux_x2": nv * f + nv * f + nt * (f + f),
        "flux_x3": nv * f + nv * f + nt * (f + f),
        # within a stage the ideal gas forms its x2 faces and tendency in one
        # kernel (k_face2_update; the weighted average of u is formed in the
        # column step): read w, F1, F3 (mixed: also the FP32 copy fw the faces
        # read, and the FP64 x1 face pressure), write du; per scalar r (or its
        # FP32 copy), s, s1, G1, G3
        "update": (4 * nv * 8 + nt * (5 * 8) if not mixed else
                   2 * nv * 8 + nv * 4 + 2 * nv * f + 8 + nt * (4 + 2 * 8 + 2 * f)),
        # the column step of the stage (KID_IMPLICIT): k_implicit<true> (5x5,
        # implicit_flag 9) and k_implicit3<true> (3x3, implicit_flag 1) read du,
        # w, u, u1 (stages 2-3) and write u, or read du, w, u and write u, u1
        # (stage 1); k_add_du<true> (no solve) reads no w: 4 V s
        "implicit": 5 * nv * 8 if implicit_flag & 1 else 4 * nv * 8,
        "prim": 2 * nv * 8 + 2 * nt * 8 + ((nv + nt) * 4 if mixed else 0),  # read u, s;
        # write w, r (mixed: and fw)
        "dt": nv * 8,               # read w
    }.get(kernel)


# The column solve of the survey's byte table: read du and w, write du
# (15 doubles per cell); the kernel's own compulsory bytes add the stage's
# weighted average of u it forms (u, u1), algorithmic_bytes_per_cell above.
IMPLICIT_SURVEY_BYTES = 15 * 8


def step_compulsory_bytes_per_cell(nv, mixed):
    """The survey's compulsory HBM bytes per cell and step: the fused explicit
    stages (stage 1 reads u0 and writes, stages 2-3 read the stage state and
    u0 and write: 8 V s with s the stored state type, FP64 here) plus three
    column solves of 15 doubles each."""
    return 8 * nv * 8 + 3 * IMPLICIT_SURVEY_BYTES


# kernel of the live profile -> its name in the rocprofv3 PMC summary, per form
# (the trailing template arguments are the kernels' specialisations for this
# workload: equation of state 1 = ideal gas, Riemann solver 3 = LMARS, -1 = read
# per instance)
PMC_NAMES = {
    "fp64": {"implicit": "k_implicit<true>", "update": "k_face2_update<false, false, 3>",
             "flux_x1": "k_flux_x1<4, -1>", "flux_x3": "k_flux_h<2, 1, 3>",
             "edge_states": "k_edge_states", "ghost": "k_ghost",
             "prim": "k_prim<1>", "dt": "k_dt"},
    "mixed": {"implicit": "k_implicit<true>", "update": "k_face2_update<false, true, 3>",
              "flux_x1": "k_flux_x1_mix<4, 3>", "flux_x3": "k_flux_h_mix<2, false, 3>",
              "edge_states": "k_edge_states_mix", "ghost": "k_ghost", "prim": "k_prim<1>",
              "dt": "k_dt"},
}
# launches of each kernel per SSPRK3 step of the headline workload (a kernel
# a form does not launch is absent from its PMC_NAMES: the x2 faces are formed
# within the update)
STEP_LAUNCHES = {"implicit": 3, "update": 3, "flux_x1": 3, "flux_x2": 3, "flux_x3": 3,
                 "edge_states": 3, "ghost": 3, "prim": 3, "dt": 1}
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "traffic_exo3_96x96x64_lmars_implicit.json")


def traffic_profile():
    """The committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this
    workload (scaled by plain copies of known size): per-kernel HBM bytes per
    dispatch and the source revision they were measured at; None if absent."""
    try:
        with open(TRAFFIC_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def pmc_name(kernel, form, tuned=None):
    """the kernel's name in the counter summary; the x1 sweep's template
    argument is the autotuned waves per block (tuned["x1_shape"] or
    tuned["x1_mix_shape"], default 4)"""
    name = PMC_NAMES[form].get(kernel)
    if name and kernel == "flux_x1" and tuned:
        sh = tuned.get("x1_mix_shape" if form == "mixed" else "x1_shape")
        if sh:
            name = name.replace("<4,", f"<{sh},")
    return name


def measured_traffic(kernel, form, tuned=None):
    """HBM bytes per launch of one kernel from the committed counter passes"""
    t = traffic_profile()
    if not t:
        return None
    k = t["kernels"].get(pmc_name(kernel, form, tuned) or "")
    return k["hbm_bytes"] if k else None


def step_traffic_per_cell(form, cells, tuned=None):
    """Measured HBM bytes per cell and step from the counter passes: the sum
    over one step's launches (STEP_LAUNCHES) of the form's kernels (with the
    run's autotuned x1 block shape); None if the profile lacks one of them"""
    t = traffic_profile()
    if not t:
        return None
    tot = 0.0
    for k, n in STEP_LAUNCHES.items():
        if k not in PMC_NAMES[form]:
            continue
        b = t["kernels"].get(pmc_name(k, form, tuned))
        if b is None:
            return None
        tot += n * b["hbm_bytes"]
    return tot / cells


def kernel_rooflines(prof, cells, nv, mixed, nt=0, implicit_flag=9):
    """per kernel: average launch time and achieved algorithmic GB/s"""
    out = {}
    for k, (ms, n) in prof.items():
        if n == 0:
            continue
        avg = ms / n
        b = algorithmic_bytes_per_cell(k, nv, mixed, nt, implicit_flag)
        gbs = b * cells / (avg * 1e-3) / 1e9 if b else None
        out[k] = {"ms_per_launch": avg, "launches": n, "bytes_per_cell": b, "GB_s": gbs,
                  "hbm_frac": gbs / HBM_PEAK_GBS if gbs else None}
    return out


def rel_linf(a, b):
    """max |a - b| over max |b| per conserved variable (interior cells)"""
    I = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    d = np.abs(a[I] - b[I]).max(axis=(0, 2, 3, 4))
    return d / np.abs(b[I]).max(axis=(0, 2, 3, 4))


def rel_l2(a, b):
    """relative L2 error of a against b per conserved variable (interior cells)"""
    I = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    d = (a[I] - b[I]).reshape(a.shape[0], a.shape[1], -1)
    r = b[I].reshape(d.shape)
    num, den = (d ** 2).sum(axis=(0, 2)), (r ** 2).sum(axis=(0, 2))
    # a variable that is zero in b (the shallow-water x1 momentum): 0 if a
    # matches it, else inf
    return np.where(den > 0, np.sqrt(num / np.where(den > 0, den, 1.0)),
                    np.where(num > 0, np.inf, 0.0))


def run_gpu(ex, precision, steps, warmup, perturb=0.0, implicit_flag=9):
    cfg = workload_config(ex, precision, implicit_flag)
    cells = 6 * cfg.nx * cfg.nx * cfg.nx1
    u0, s0 = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    if perturb:  # relative perturbation of the initial state (sensitivity floor)
        u0 = u0 * (1.0 + perturb * np.random.default_rng(5).standard_normal(u0.shape))
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    # launch widths and block shapes from the autotuner (timed before the
    # warm-up steps; the first steps are then taken from the uploaded state)
    tuned = {k: v for k, v in g.tune(reps=3).items() if v} if TUNE else {}
    g.upload(u0, s0)
    if warmup:
        g.step(warmup)
    g.sync()
    t0 = time.perf_counter()
    dt = g.step(steps)  # each step replayed from a captured HIP graph
    g.sync()
    t1 = time.perf_counter()
    u, _ = g.download()
    # per-kernel times: a further short run with a HIP event pair per launch
    g.profile_begin()
    g.step(min(steps, 20))
    prof = g.profile_end()
    # latency of a single step taken on its own: host wall time around one
    # graph replay and the read-back of its dt, over 50 steps
    lat = []
    for _ in range(50):
        ts = time.perf_counter()
        g.step(1)
        lat.append((time.perf_counter() - ts) * 1e3)
    finite = bool(np.isfinite(u).all())
    g.close()
    sec = t1 - t0
    step_latency = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                    "steps": len(lat)}
    return dict(cells=cells, sec=sec, dt=dt, prof=prof, finite=finite, nv=5, u=u, tuned=tuned,
                step_latency_ms=step_latency)


def partial_correction(ex, steps, warmup):
    """The same workload with implicit_flag 1: the 3x3 (density, x1 momentum,
    energy) column correction of partial_correction.cpp, both forms."""
    out = {}
    runs = {}
    for name, prec in (("fp64", ex.PREC_FP64), ("mixed", ex.PREC_MIXED)):
        r = run_gpu(ex, prec, steps, warmup, implicit_flag=1)
        ms, n = r["prof"].get("implicit", (0.0, 0))
        out[name] = {"value": r["cells"] * steps / r["sec"], "ms_per_step": r["sec"] * 1e3 / steps,
                     "implicit_ms_per_launch": ms / n if n else None, "state_finite": r["finite"]}
        runs[name] = r["u"]
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed_rel_l2_vs_fp64_max"] = float(rel_l2(runs["mixed"], runs["fp64"]).max())
    out["workload"] = ("the headline workload with implicit_flag 1 (3x3 partial correction), "
                       f"{steps} steps")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(ex, cfg, steps):
    from oracle.oracle import Oracle
    u0, _ = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    o = Oracle(cfg)
    o.upload(u0)
    t0 = time.perf_counter()
    o.step(steps)
    sec = time.perf_counter() - t0
    o.close()
    return 6 * cfg.nx * cfg.nx * cfg.nx1 * steps / sec, sec


def cpu_baseline(cfg_fn, ex, steps=4, small_nx=48):
    """The C restatement (FP64, OpenMP over panel rows) timed on this host:
    `steps` SSPRK3 steps of the full workload on all usable cores, then one
    step of the same atmosphere on 6 x small_nx^2 x 64 cells on one thread."""
    from oracle.oracle import lib

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = int(os.environ.get("EXO_CPU_THREADS", "0")) or min(16, avail or 1)
    lib().or_set_threads(cores)
    cfg = cfg_fn(ex, 0)
    rate, sec = _oracle_rate(ex, cfg, steps)
    lib().or_set_threads(1)
    small = cfg_fn(ex, 0)
    small.nx = small_nx
    rate1, sec1 = _oracle_rate(ex, small, 1)
    lib().or_set_threads(cores)
    return {"value": rate, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} SSPRK3 steps of the full 6x96x96x64 workload on {cores} "
                      f"threads, {sec:.1f} s wall",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "usable_cpus": avail,
            "one_thread": {"value": rate1, "unit": "cell-updates/s", "cores": 1,
                           "sample": f"1 SSPRK3 step of 6x{small_nx}x{small_nx}x64 (same "
                                     f"atmosphere), {sec1:.1f} s wall"}}


R_EARTH = 6371000.0
ENS_SIZES = (24, 32, 48, 64, 96, 128, 192)


def ensemble(ex, precision, steps, n_inst=256, keep=16):
    """256 shallow-water instances of uneven size stepped as one batch."""
    rng = np.random.default_rng(SEED)
    Ns = rng.choice(ENS_SIZES, n_inst)
    h0 = rng.uniform(2000.0, 8000.0, n_inst)
    members = []
    for q in range(n_inst):
        cfg = ex.make_config(nx=int(Ns[q]), nx1=1, eos=ex.EOS_SHALLOW_YZ, rsolver=ex.RS_ROE_SW,
                             coriolis=ex.COR_FORCING2, cfl=0.4, x1min=R_EARTH,
                             x1max=R_EARTH + 1.0, omega=7.292e-5, precision=precision)
        u0, _ = ex.init_state(cfg, 0, SEED + q, (float(h0[q]), 50.0, 9.81))
        g = ex.Solver(cfg)
        g.upload(u0)
        members.append(g)
    cells = int(sum(6 * int(n) * int(n) for n in Ns))
    # one at a time (a launch sequence per instance) for a few steps
    t0 = time.perf_counter()
    for g in members:
        g.step(2)
    for g in members:
        g.sync()
    t_alone = time.perf_counter() - t0
    b = ex.Batch(members, graph=True)
    caps = b.tune(reps=2) if TUNE else {}
    b.step(2)  # warm-up: graph capture
    m0 = b.sums()[:, 0]
    t0 = time.perf_counter()
    b.step(steps)
    t1 = time.perf_counter()
    m1 = b.sums()[:, 0]
    b.profile_begin()
    b.step(1)
    prof = b.profile_end()
    states = [members[q].download()[0] for q in range(keep)]
    b.close()
    for g in members:
        g.close()
    return {
        "instances": n_inst, "cells": cells, "steps": steps,
        "value": cells * steps / (t1 - t0), "ms_per_step": (t1 - t0) * 1e3 / steps,
        "one_at_a_time_value": cells * 2 / t_alone,
        "mass_rel_drift_max": float(np.max(np.abs(m1 - m0) / np.abs(m0))),
        "tuned_widths": {k: v for k, v in caps.items() if v},
        "kernels_ms_per_step": {k: v[0] for k, v in prof.items() if v[1]},
    }, states


def ensemble_accuracy(ex, steps=200, keep=16):
    """Mixed against FP64 for the first `keep` ensemble instances after `steps`
    steps, with each instance's input-rounding floor: the FP64 result's change
    when the initial state is perturbed at FP32 rounding (2^-24), on the FP64
    run's dt sequence.  Shallow, finely resolved instances are dynamically
    unstable and amplify any perturbation; the bar is <= 1e-5 or <= 4x the
    floor, per instance."""
    rng = np.random.default_rng(SEED)
    Ns = rng.choice(ENS_SIZES, 256)
    h0 = rng.uniform(2000.0, 8000.0, 256)
    rows = []
    for q in range(keep):
        def cfg_of(prec):
            return ex.make_config(nx=int(Ns[q]), nx1=1, eos=ex.EOS_SHALLOW_YZ,
                                  rsolver=ex.RS_ROE_SW, coriolis=ex.COR_FORCING2, cfl=0.4,
                                  x1min=R_EARTH, x1max=R_EARTH + 1.0, omega=7.292e-5,
                                  precision=prec)
        u0, _ = ex.init_state(cfg_of(ex.PREC_FP64), 0, SEED + q, (float(h0[q]), 50.0, 9.81))

        def run(prec, u, dts=None):
            g = ex.Solver(cfg_of(prec))
            g.upload(u)
            seq = [g.step(1, dts[s] if dts else 0.0) for s in range(steps)]
            out = g.download()[0]
            g.close()
            return out, seq
        ref, dts = run(ex.PREC_FP64, u0)
        mix, _ = run(ex.PREC_MIXED, u0)
        up = u0 * (1.0 + 2.0 ** -24 * np.random.default_rng(5).standard_normal(u0.shape))
        per, _ = run(ex.PREC_FP64, up, dts)
        e = float(rel_l2(mix, ref)[[0, 2, 3]].max())
        f = float(rel_l2(per, ref)[[0, 2, 3]].max())
        rows.append({"N": int(Ns[q]), "h0": round(float(h0[q]), 1), "rel_l2": e, "floor": f})
    ok = sum(1 for r in rows if r["rel_l2"] <= 1e-5 or r["rel_l2"] <= 4 * r["floor"])
    return {"steps": steps, "instances": keep, "within_bar": ok,
            "rel_l2_max": max(r["rel_l2"] for r in rows),
            "rel_l2_median": float(np.median([r["rel_l2"] for r in rows])), "per_instance": rows}


def large(ex, steps):
    """6 x 384 x 384 x 128 rotating atmosphere, 2 tracers, HLLC, mixed form."""
    cfg = ex.make_config(nx=384, nx1=128, eos=ex.EOS_IDEAL, rsolver=ex.RS_HLLC,
                         implicit_flag=9, ntracer=2, coriolis=ex.COR_3D, cfl_x1=1,
                         x1min=R_EARTH, x1max=R_EARTH + 30000.0, gamma=1.4, grav=-9.81,
                         omega=7.292e-5, cfl=0.4, precision=ex.PREC_MIXED)
    u0, s0 = ex.init_state(cfg, 2, SEED, (300.0, 6.5e-3, 200.0, 1.0e5, 287.0, 0.5, 3.0e-3, 0.5))
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    del u0, s0
    g.step(1)
    m0 = g.sums()[0]
    g.sync()
    g.profile_begin()
    t0 = time.perf_counter()
    dt = g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    prof = g.profile_end()
    m1 = g.sums()[0]
    g.close()
    cells = 6 * 384 * 384 * 128
    return {"cells": cells, "steps": steps, "value": cells * steps / (t1 - t0),
            "unit": "cell-updates/s", "ms_per_step": (t1 - t0) * 1e3 / steps, "dt_last_s": dt,
            "mass_rel_drift": abs(m1 - m0) / abs(m0),
            "kernels": kernel_rooflines(prof, cells, 5, True, 2),
            "workload": "6x384x384 x 128 levels, 5 vars + 2 tracers, lapse-rate atmosphere "
                        "with rotation, WENO5+HLLC, implicit 5x5 FP64, mixed form"}


def output_bench(ex, reps=10):
    """The state of the headline workload written out as the reference's
    per-block NetCDF files: the device kernel that lays the float fields out
    (live HIP-event time, algorithmic bytes: each FP64 input read once, each
    float written once), the fields copied to host memory, and the six files."""
    import tempfile

    cfg = workload_config(ex, ex.PREC_FP64)
    u0, _ = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    g = ex.Solver(cfg)
    g.upload(u0)
    del u0
    g.step(1)
    cells = 6 * cfg.nx * cfg.nx * cfg.nx1
    out = {}
    for var, nin in (("prim", 5), ("uov", 2)):
        nq = len(ex.output_names(cfg, var))
        g.output_fields(var)
        g.profile_begin()
        t0 = time.perf_counter()
        for _ in range(reps):
            f = g.output_fields(var)
        t1 = time.perf_counter()
        ms, n = g.profile_end()["output"]
        bpc = nin * 8 + nq * 4
        avg = ms / n
        with tempfile.TemporaryDirectory() as d:
            t2 = time.perf_counter()
            g.write_netcdf(os.path.join(d, "out"), var, 0.0, 0)
            t3 = time.perf_counter()
            size = sum(os.path.getsize(os.path.join(d, x)) for x in os.listdir(d))
        out[var] = {"fields": nq, "kernel_ms": avg, "bytes_per_cell": bpc,
                    "kernel_GB_s": bpc * cells / (avg * 1e-3) / 1e9,
                    "kernel_hbm_frac": bpc * cells / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "to_host_ms": (t1 - t0) * 1e3 / reps, "host_MB": f.nbytes / 1e6,
                    "write_files_ms": (t3 - t2) * 1e3, "file_MB": size / 1e6}
    g.close()
    out["workload"] = "6x96x96x64 ideal-gas state after one step, interior cells, 6 files"
    return out


def recovery_bench(ex, steps=10):
    """Failure recovery on the headline workload (mixed form): the in-HBM
    snapshot a checked step takes, checked stepping against plain graph
    replay, and a checkpoint written and read back."""
    import tempfile

    cfg = workload_config(ex, ex.PREC_MIXED)
    u0, _ = ex.init_state(cfg, 1, SEED, IC_PARAMS)
    g = ex.Solver(cfg)
    g.upload(u0)
    del u0
    g.step(2)
    g.save_states()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.save_states()
    g.sync()
    snap_ms = (time.perf_counter() - t0) * 1e3 / steps
    snap_bytes = 2 * 2 * 5 * 8 * 6 * 102 * 102 * 70   # u and w with ghosts, read + write
    g.sync()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    plain = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    _, redos = g.step_redo(steps, 3)
    checked = (time.perf_counter() - t0) * 1e3 / steps
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "run.ckpt")
        t0 = time.perf_counter()
        g.save_checkpoint(path, 0.0, 0)
        t1 = time.perf_counter()
        size = os.path.getsize(path)
        h = ex.Solver(cfg)
        t2 = time.perf_counter()
        h.load_checkpoint(path)
        t3 = time.perf_counter()
        h.close()
    g.close()
    return {"snapshot_ms": snap_ms, "snapshot_GB_s": snap_bytes / (snap_ms * 1e-3) / 1e9,
            "ms_per_step_graph": plain, "ms_per_step_checked": checked, "redos": redos,
            "checkpoint_save_ms": (t1 - t0) * 1e3, "checkpoint_load_ms": (t3 - t2) * 1e3,
            "checkpoint_MB": size / 1e6,
            "workload": "6x96x96x64 mixed form; checked steps take a snapshot and a "
                        "validity read per step (max_redo 3)"}


# requests on the device at once: the --serving-sweep measurements (64
# requests per form) give the same throughput within 3 % at 4, 8, 16 and 32,
# and the lowest median latency at 4
SERVE_MAX_ACTIVE = 4


def serving(ex, precision, nreq, keep=2, max_active=SERVE_MAX_ACTIVE):
    """Poisson request stream through the continuous-batching server."""
    rng = np.random.default_rng(SEED + 3)
    arrivals = np.cumsum(rng.exponential(1.0 / 40.0, nreq))
    reqs = []
    for q in range(nreq):
        N = int(rng.integers(16, 65)) * 2
        nz = int(rng.integers(32, 97))
        T0 = float(rng.uniform(150.0, 300.0))
        cfg = ex.make_config(nx=N, nx1=nz, eos=ex.EOS_IDEAL, rsolver=ex.RS_LMARS,
                             implicit_flag=9, ntracer=2, coriolis=ex.COR_NONE, cfl_x1=1,
                             x1min=R_EARTH, x1max=R_EARTH + 30000.0, gamma=1.4, grav=-9.81,
                             omega=7.292e-5, cfl=0.4, precision=precision)
        reqs.append((cfg, (T0, 1.0e5, 287.0, 0.01, 5.0, 3.0e-3, 0.5)))
    srv = ex.Server(max_active=max_active, chunk_steps=20, graph=True, prep_threads=4)
    ids = []
    t0 = time.perf_counter()
    for q, (cfg, par) in enumerate(reqs):
        dt = t0 + arrivals[q] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(cfg, 200, kind=1, seed=SEED + 1000 + q, params=par,
                              keep_state=q < keep))
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([r[2]["latency_s"] for r in res])
    return {
        "requests": nreq, "steps_each": 200, "offered_req_per_s": 40.0,
        "max_active": max_active,
        "cells_mean": float(np.mean([r[2]["cells"] for r in res])),
        "value": st["cell_updates"] / st["span_s"], "unit": "cell-updates/s",
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "queue_p50_s": float(np.percentile([r[2]["queue_s"] for r in res], 50)),
        "peak_active": int(st["peak_active"]), "batches_formed": int(st["batches"]),
        "peak_bytes": int(st["peak_bytes"]),
        "mass_rel_drift_max": float(max(abs(r[2]["mass1"] - r[2]["mass0"]) / abs(r[2]["mass0"])
                                        for r in res)),
    }, [r[0] for r in res[:keep]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--serving-sweep", action="store_true",
                    help="only the serving stream, both forms, at max_active 4, 8, 16 and 32")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--only-large", action="store_true",
                    help="only the 6 x 384 x 384 x 128 run (--large-steps steps, for profiling)")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="only the 256-instance shallow-water batch, both forms (for profiling)")
    ap.add_argument("--only-serving", action="store_true",
                    help="only the serving stream at the default width, both forms")
    ap.add_argument("--large-steps", type=int, default=100)
    ap.add_argument("--implicit-flag", type=int, default=9, choices=(1, 9),
                    help="column correction of the headline workload: 9 full 5x5 (the "
                         "benchmark), 1 partial 3x3 (for profiling k_implicit3)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch widths and shapes, no autotuning (for counter passes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc pass)")
    args = ap.parse_args()
    global TUNE
    TUNE = not args.no_tune
    if args.copy_calibration:
        import torch
        a = torch.ones(256 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)

        def timed(fn, reps=3):
            # the last of reps runs (the first warms up), bytes read + written
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ms = []
            for _ in range(reps):
                e0.record()
                fn()
                e1.record()
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
            return min(ms[1:])

        ms16 = timed(lambda: b.copy_(a))
        # the same bytes through 8-byte loads and stores: a float64 multiply
        # on views offset by one element, which torch's elementwise kernel
        # cannot vectorise (one double per lane per access)
        n = 128 << 20
        x = torch.ones(n + 1, dtype=torch.float64, device="cuda")[1:]
        y = torch.empty(n + 1, dtype=torch.float64, device="cuda")[1:]
        ms8 = timed(lambda: torch.mul(x, 1.0, out=y))
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 4, "copies": 3,
                          "f64_scalar_copy_bytes_each_way": n * 8, "f64_copies": 3,
                          "copy_GB_s": 2 * a.numel() * 4 / ms16 / 1e6,
                          "f64_scalar_copy_GB_s": 2 * n * 8 / ms8 / 1e6}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import exocubed_amd as ex

    if args.serving_sweep:
        rows = []
        for ma in (4, 8, 16, 32):
            for name, prec in (("fp64", ex.PREC_FP64), ("mixed", ex.PREC_MIXED)):
                r, _ = serving(ex, prec, args.requests, keep=0, max_active=ma)
                rows.append({"max_active": ma, "form": name, "value": r["value"],
                             "latency_p50_s": r["latency_p50_s"],
                             "latency_p99_s": r["latency_p99_s"],
                             "peak_active": r["peak_active"], "peak_bytes": r["peak_bytes"]})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        print(json.dumps({"serving_sweep": rows, "requests": args.requests}))
        return
    if args.only_serving:
        v64, _ = serving(ex, ex.PREC_FP64, args.requests, keep=0)
        vmx, _ = serving(ex, ex.PREC_MIXED, args.requests, keep=0)
        print(json.dumps({"serving": {"fp64": v64, "mixed": vmx,
                                      "mixed_over_fp64": vmx["value"] / v64["value"]}}))
        return
    if args.only_large:
        print(json.dumps({"large": large(ex, args.large_steps)}))
        return
    if args.only_ensemble:
        e64, _ = ensemble(ex, ex.PREC_FP64, args.ensemble_steps)
        emx, _ = ensemble(ex, ex.PREC_MIXED, args.ensemble_steps)
        print(json.dumps({"ensemble": {"fp64": e64, "mixed": emx}}))
        return
    iflag = args.implicit_flag
    log(f"headline FP64, {args.steps} steps")
    r64 = run_gpu(ex, ex.PREC_FP64, args.steps, args.warmup, implicit_flag=iflag)
    log("headline mixed")
    rmx = run_gpu(ex, ex.PREC_MIXED, args.steps, args.warmup, implicit_flag=iflag)
    # the same seeded run in both forms: accuracy of the mixed form vs FP64
    l2 = rel_l2(rmx["u"], r64["u"])
    linf = rel_linf(rmx["u"], r64["u"])
    # what FP32 rounding of the input alone does to the FP64 result: the
    # noisy atmosphere amplifies any perturbation, so this is the floor any
    # FP32 computation sits on
    log("input-rounding floor")
    rfl = run_gpu(ex, ex.PREC_FP64, args.steps, args.warmup, perturb=2.0 ** -24,
                  implicit_flag=iflag)
    floor = rel_l2(rfl["u"], r64["u"])
    del rfl
    value64 = r64["cells"] * args.steps / r64["sec"]
    value_mx = rmx["cells"] * args.steps / rmx["sec"]
    # the metric's accuracy gate: the mixed form counts only at <= 1e-5
    # relative L2 against FP64 in every conserved variable after the run;
    # otherwise the headline is the FP64 form (bit-exact against the oracle)
    mixed_meets = bool(l2.max() <= 1e-5)
    r, form = (rmx, "mixed") if mixed_meets else (r64, "fp64")
    ms_step = r["sec"] * 1e3 / args.steps
    value = value_mx if mixed_meets else value64
    # dominant kernel of the reported form and its roofline
    prof = {k: v for k, v in r["prof"].items() if v[1] > 0}
    dom = max(prof, key=lambda k: prof[k][0])
    tot_ms, cnt = prof[dom]
    avg_s = tot_ms / cnt * 1e-3
    own = algorithmic_bytes_per_cell(dom, r["nv"], form == "mixed", implicit_flag=iflag)
    bpc = IMPLICIT_SURVEY_BYTES if dom == "implicit" else own
    achieved = (bpc * r["cells"] / avg_s / 1e9) if bpc else None
    own_gbs = (own * r["cells"] / avg_s / 1e9) if own else None
    traffic = measured_traffic(dom, form, r["tuned"]) if iflag == 9 else None
    tp = traffic_profile()
    roof = {"bound": "hbm", "kernel": dom, "form": form, "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "traffic_source": (f"rocprofv3 FETCH_SIZE/WRITE_SIZE passes, profiles/"
                               f"{os.path.basename(TRAFFIC_PATH)}, measured at "
                               f"{tp.get('measured_at', 'unknown')}") if traffic and tp else None,
            "avg_launch_ms": tot_ms / cnt,
            "bytes_per_cell": bpc,
            "bytes_per_cell_basis": ("the survey's column-solve figure: read du and w, write du"
                                     if dom == "implicit" else "the kernel's compulsory arrays"),
            "kernel_bytes_per_cell": own,
            "kernel_GB_s": own_gbs,
            "kernel_frac": own_gbs / HBM_PEAK_GBS if own_gbs else None}
    if dom == "implicit" and tp and tp.get("implicit_fp64_flop_per_cell"):
        fl = tp["implicit_fp64_flop_per_cell"] * r["cells"] / avg_s / 1e12
        roof["fp64_valu"] = {"achieved": fl, "peak": 78.6, "unit": "TFLOP/s",
                             "frac": fl / 78.6,
                             "flop_per_cell": tp["implicit_fp64_flop_per_cell"]}
    kernels = kernel_rooflines(prof, r["cells"], r["nv"], form == "mixed", implicit_flag=iflag)
    for k in kernels:
        kernels[k]["share"] = prof[k][0] / sum(x[0] for x in prof.values())
    kernels_mx = kernel_rooflines(rmx["prof"], rmx["cells"], rmx["nv"], True, implicit_flag=iflag)
    if not args.no_cpu:
        log("CPU baseline")
    cpu = None if args.no_cpu else cpu_baseline(workload_config, ex)
    ens = srv = None
    if not args.only_headline:
        log("ensemble FP64")
        e64, s64 = ensemble(ex, ex.PREC_FP64, args.ensemble_steps)
        log("ensemble mixed")
        emx, smx = ensemble(ex, ex.PREC_MIXED, args.ensemble_steps)
        log("ensemble accuracy")
        l2e = max(float(rel_l2(a, b)[[0, 2, 3]].max()) for a, b in zip(smx, s64))
        ens = {"fp64": e64, "mixed": emx, "mixed_over_fp64": emx["value"] / e64["value"],
               "mixed_rel_l2_vs_fp64_max": l2e,
               "accuracy_200_steps": ensemble_accuracy(ex),
               "workload": "256 shallow-water instances, N in {24..192}, 1 layer, WENO5+Roe, "
                           "Coriolis, SSPRK3 CFL 0.4"}
        log("serving FP64")
        v64, r64s = serving(ex, ex.PREC_FP64, args.requests)
        log("serving mixed")
        vmx, rmxs = serving(ex, ex.PREC_MIXED, args.requests)
        l2s = max(float(rel_l2(a, b).max()) for a, b in zip(rmxs, r64s))
        srv = {"fp64": v64, "mixed": vmx, "mixed_over_fp64": vmx["value"] / v64["value"],
               "mixed_rel_l2_vs_fp64_max": l2s,
               "workload": "Poisson 40 req/s, 3D ideal gas + 2 tracers, N 32..128, 32..96 "
                           "levels, WENO5+LMARS, implicit 5x5, 200 steps each"}
    if args.large_steps > 0 and not args.only_headline:
        log("large run")
    big = large(ex, args.large_steps) if args.large_steps > 0 and not args.only_headline else None
    if not args.only_headline:
        log("output, recovery, partial correction")
    outp = None if args.only_headline else output_bench(ex)
    recov = None if args.only_headline else recovery_bench(ex)
    part = None if args.only_headline else partial_correction(ex, min(args.steps, 50),
                                                               args.warmup)
    step_meas = step_traffic_per_cell(form, r["cells"], r["tuned"]) if iflag == 9 else None
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
        "dtype": "f64" if form == "fp64" else "f32 fluxes / f64 state, decomposition, implicit",
        "data": "synthetic (seeded isothermal atmosphere, 1% temperature noise)",
        "config": {"workload": "exo3 3D ideal gas, 6x96x96 cells x 64 levels, WENO5+LMARS, "
                               "implicit 5x5 vertical, SSPRK3 CFL 0.4", "precision": form,
                   "cells": r["cells"], "dt_last_s": r["dt"], "state_finite": r["finite"],
                   "implicit_flag": iflag},
        "headline_form": form,
        "headline_rule": ("value is the mixed form if it meets <= 1e-5 rel L2 vs FP64 in every "
                          "conserved variable after the run's steps, else the FP64 form"),
        "meets_1e-5": {"mixed": mixed_meets, "steps": args.steps + args.warmup,
                       "fp64_vs_oracle": "bit-exact (tests/test_gpu_full_size.py)"},
        "fp64": {"value": value64, "ms_per_step": r64["sec"] * 1e3 / args.steps,
                 "dt_last_s": r64["dt"], "kernels": kernel_rooflines(r64["prof"], r64["cells"],
                                                                   r64["nv"], False,
                                                                   implicit_flag=iflag),
                 "tuned_launches": r64["tuned"], "step_latency_ms": r64["step_latency_ms"]},
        "mixed": {"value": value_mx, "ms_per_step": rmx["sec"] * 1e3 / args.steps,
                  "dt_last_s": rmx["dt"], "kernels": kernels_mx,
                  "counts_for_metric": mixed_meets, "tuned_launches": rmx["tuned"],
                  "step_latency_ms": rmx["step_latency_ms"]},
        "mixed_over_fp64": value_mx / value64,
        "mixed_rel_l2_vs_fp64": {n: float(x) for n, x in zip(("rho", "m1", "m2", "m3", "E"), l2)},
        "mixed_rel_linf_vs_fp64": {n: float(x) for n, x in zip(("rho", "m1", "m2", "m3", "E"),
                                                               linf)},
        "mixed_rel_l2_max": float(l2.max()),
        "fp64_input_rounding_floor": {n: float(x) for n, x in zip(("rho", "m1", "m2", "m3", "E"),
                                                                 floor)},
        "mixed_over_floor_max": float((l2 / floor).max()),
        "step_bytes_per_cell": {
            "compulsory": step_compulsory_bytes_per_cell(r["nv"], form == "mixed"),
            "compulsory_basis": "8 V s for the fused explicit stages + 3 x 120 B column solves",
            "measured": step_meas,
            "measured_source": roof["traffic_source"]},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "serving": srv,
        "large": big,
        "output": outp,
        "recovery": recov,
        "partial_correction": part,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
