"""The continuous-batching request loop (exo_server) on the GPU: requests of
mixed sizes arrive while others run, share batched launches, and each result
is bit-identical to the same run alone."""
import time

import numpy as np
import pytest

from helpers import atm_config, sw_config

pytestmark = pytest.mark.gpu


def requests(ex):
    mix = ex.PREC_MIXED
    out = [
        (sw_config(ex, nx=8, rsolver=0, nx1=1), 0, 9, ex.PREC_FP64),
        (atm_config(ex, nx=8, nx1=10, rsolver=3), 1, 7, ex.PREC_FP64),
        (atm_config(ex, nx=6, nx1=8, rsolver=2, ntracer=2, coriolis=2), 2, 5, mix),
        (sw_config(ex, nx=12, rsolver=1, nx1=1), 0, 11, mix),
        (atm_config(ex, nx=10, nx1=6, rsolver=2), 1, 8, mix),
    ]
    for c, _, _, prec in out:
        c.precision = prec
    return out


def run_alone(ex, cfg, kind, seed, nsteps):
    u0, s0 = ex.init_state(cfg, kind, seed, ())
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    g.step(nsteps)
    u, s = g.download()
    g.close()
    return u, s


def test_results_same_alone_and_in_the_mix(ex):
    reqs = requests(ex)
    srv = ex.Server(max_active=3, chunk_steps=3, graph=True)
    ids = []
    for q, (cfg, kind, nsteps, _) in enumerate(reqs):
        u0, s0 = ex.init_state(cfg, kind, 100 + q, ())
        ids.append(srv.submit(cfg, nsteps, u0=u0, s0=s0))
        time.sleep(0.02)  # arrivals while earlier requests run
    got = [srv.wait(i, timeout=120) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    assert st["batches"] >= 2  # the active set changed while running
    for q, ((cfg, kind, nsteps, _), (u, s, info)) in enumerate(zip(reqs, got)):
        ur, sr = run_alone(ex, cfg, kind, 100 + q, nsteps)
        assert info["steps"] == nsteps
        assert np.array_equal(u, ur), f"request {q}"
        if sr is not None:
            assert np.array_equal(s, sr), f"request {q} scalars"
        assert info["latency_s"] >= info["run_s"] > 0


def test_initial_state_recipe_and_admission_by_memory(ex):
    reqs = requests(ex)
    big = max(ex.footprint(c) for c, _, _, _ in reqs)
    # the largest request fits only alone; smaller ones may share the budget
    srv = ex.Server(mem_budget=int(big * 1.2), max_active=8, chunk_steps=4, graph=False)
    ids = [srv.submit(c, n, kind=k, seed=200 + q) for q, (c, k, n, _) in enumerate(reqs)]
    got = [srv.wait(i, timeout=120) for i in ids]
    st = srv.stats()
    with pytest.raises(ex.ExoError):
        srv.submit(atm_config(ex, nx=64, nx1=64), 1, kind=1, seed=1)
    srv.close()
    assert st["peak_bytes"] <= st["budget_bytes"]
    assert st["peak_active"] < len(reqs)  # memory, not slots, limited the mix
    for q, ((cfg, kind, nsteps, _), (u, _, info)) in enumerate(zip(reqs, got)):
        ur, _ = run_alone(ex, cfg, kind, 200 + q, nsteps)
        assert np.array_equal(u, ur), f"request {q}"
        # mass is conserved by every run
        assert abs(info["mass1"] - info["mass0"]) <= 1e-13 * abs(info["mass0"])


def test_dropped_state_keeps_info(ex):
    cfg = sw_config(ex, nx=8, rsolver=0, nx1=1)
    srv = ex.Server(max_active=2, chunk_steps=5)
    rid = srv.submit(cfg, 7, kind=0, seed=3, keep_state=False)
    u, s, info = srv.wait(rid, timeout=60)
    srv.close()
    assert u is None and info["steps"] == 7 and info["cells"] == 6 * 64


def test_equal_requests_queue_for_memory(ex):
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=2)
    fp = ex.footprint(cfg)
    srv = ex.Server(mem_budget=int(fp * 1.5), max_active=8, chunk_steps=2)
    ids = [srv.submit(cfg, 4, kind=1, seed=300 + q) for q in range(3)]
    got = [srv.wait(i, timeout=120) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["peak_active"] == 1 and st["completed"] == 3
    for q, (u, _, _) in enumerate(got):
        ur, _ = run_alone(ex, cfg, 1, 300 + q, 4)
        assert np.array_equal(u, ur)
