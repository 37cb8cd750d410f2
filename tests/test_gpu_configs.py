"""The batched and served benchmark workloads at their own sizes.

* Ensemble: a seeded sub-batch of the 256-instance shallow-water ensemble with
  the real panel-size mix {24, 32, 48, 64, 96, 128, 192} and mean depths
  2000-8000 m, FP64 and mixed members in one graph-replayed, width-tuned
  batch: every member bit for bit equal to stepping it alone, mass conserved
  to 1e-13 by the compensated totals.
* Serving: requests of the serving stream's sizes (N 32..128, 32..96 levels,
  5 variables + 2 tracers, base temperature 150-300 K), FP64 and mixed,
  arriving while others run: each result bit for bit equal to the same run
  alone."""
import time

import numpy as np
import pytest

from helpers import R_EARTH

pytestmark = pytest.mark.gpu
SEED = 20261202
ENS_SIZES = (24, 32, 48, 64, 96, 128, 192)


def ensemble_members(ex):
    """two members of each panel size, depths drawn as the benchmark draws
    them, alternately FP64 and mixed"""
    rng = np.random.default_rng(SEED)
    h0 = rng.uniform(2000.0, 8000.0, 2 * len(ENS_SIZES))
    out = []
    for q in range(2 * len(ENS_SIZES)):
        cfg = ex.make_config(nx=ENS_SIZES[q % len(ENS_SIZES)], nx1=1, eos=ex.EOS_SHALLOW_YZ,
                             rsolver=ex.RS_ROE_SW, coriolis=ex.COR_FORCING2, cfl=0.4,
                             x1min=R_EARTH, x1max=R_EARTH + 1.0, omega=7.292e-5,
                             precision=ex.PREC_MIXED if q % 2 else ex.PREC_FP64)
        out.append((cfg, (float(h0[q]), 50.0, 9.81), SEED + q))
    return out


def test_ensemble_sub_batch_real_sizes(ex):
    nsteps = 20
    specs = ensemble_members(ex)
    members = []
    for cfg, par, seed in specs:
        u0, _ = ex.init_state(cfg, 0, seed, par)
        g = ex.Solver(cfg)
        g.upload(u0)
        members.append(g)
    b = ex.Batch(members, graph=True)
    b.tune(reps=1)
    m0 = b.sums()[:, 0].copy()
    b.step(nsteps)
    m1 = b.sums()[:, 0]
    got = [g.download()[0] for g in members]
    b.close()
    for g in members:
        g.close()
    drift = np.abs(m1 - m0) / np.abs(m0)
    print("panel sizes", [c.nx for c, _, _ in specs], "mass drift max", drift.max())
    assert drift.max() <= 1e-13, drift
    for q, (cfg, par, seed) in enumerate(specs):
        u0, _ = ex.init_state(cfg, 0, seed, par)
        g = ex.Solver(cfg)
        g.upload(u0)
        g.step(nsteps)
        ref = g.download()[0]
        g.close()
        assert np.array_equal(got[q], ref), f"member {q} (N={cfg.nx})"


def serving_requests(ex):
    """requests drawn as the serving stream draws them, at both ends of its
    size range and in between, FP64 and mixed"""
    rng = np.random.default_rng(SEED + 3)
    out = []
    for q, (N, nz) in enumerate(((32, 96), (128, 32), (64, 64), (96, 48), (128, 96))):
        T0 = float(rng.uniform(150.0, 300.0))
        cfg = ex.make_config(nx=N, nx1=nz, eos=ex.EOS_IDEAL, rsolver=ex.RS_LMARS,
                             implicit_flag=9, ntracer=2, coriolis=ex.COR_NONE, cfl_x1=1,
                             x1min=R_EARTH, x1max=R_EARTH + 30000.0, gamma=1.4, grav=-9.81,
                             omega=7.292e-5, cfl=0.4,
                             precision=ex.PREC_MIXED if q % 2 else ex.PREC_FP64)
        out.append((cfg, (T0, 1.0e5, 287.0, 0.01, 5.0, 3.0e-3, 0.5), SEED + 1000 + q))
    return out


def test_serving_stream_sized_requests(ex):
    nsteps = 8
    reqs = serving_requests(ex)
    srv = ex.Server(max_active=3, chunk_steps=3, graph=True, prep_threads=2)
    ids = []
    for cfg, par, seed in reqs:
        ids.append(srv.submit(cfg, nsteps, kind=1, seed=seed, params=par))
        time.sleep(0.05)  # arrivals while earlier requests run
    got = [srv.wait(i, timeout=200) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    assert st["batches"] >= 2
    for q, ((cfg, par, seed), (u, s, info)) in enumerate(zip(reqs, got)):
        u0, s0 = ex.init_state(cfg, 1, seed, par)
        g = ex.Solver(cfg)
        g.upload(u0, s0)
        g.step(nsteps)
        ur, sr = g.download()
        g.close()
        assert info["steps"] == nsteps
        assert np.array_equal(u, ur), f"request {q} (N={cfg.nx}, nz={cfg.nx1})"
        assert np.array_equal(s, sr), f"request {q} scalars"
        assert abs(info["mass1"] - info["mass0"]) <= 1e-13 * abs(info["mass0"])
