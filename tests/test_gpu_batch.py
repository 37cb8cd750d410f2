"""Batches of independent instances (exo_batch) and the compensated conserved
totals, on the GPU.

A batch runs every phase as one launch over all members (optionally persistent
and replayed from a captured HIP graph); each member's result must be
bit-identical to stepping it alone, which is itself bit-exact against the
oracle in FP64 (test_gpu_parity.py).  The members below mix equations of state
and Riemann solvers, so the batch runs the generic kernels while each member
alone runs the kernels specialised for its own: the two agree bit for bit."""
import math

import numpy as np
import pytest

from helpers import atm_config, sw_config

pytestmark = pytest.mark.gpu
SEED = 20261202


def member_specs(ex):
    """uneven sizes, both EOS, all solvers, tracers, both precisions, the full
    and the partial column correction"""
    mix = ex.PREC_MIXED
    return [
        (sw_config(ex, nx=8, rsolver=0), 0),
        (sw_config(ex, nx=12, rsolver=1), 0),
        (sw_config(ex, nx=6, rsolver=0, nx1=1), 0),
        (sw_config(ex, nx=10, rsolver=0, nx1=1), 0),
        (atm_config(ex, nx=8, nx1=8, rsolver=3), 1),
        (atm_config(ex, nx=6, nx1=12, rsolver=2, ntracer=2, coriolis=2), 2),
        (atm_config(ex, nx=10, nx1=6, rsolver=3, implicit=0), 1),
        (sw_config(ex, nx=8, rsolver=1, nx1=1), 0, mix),
        (atm_config(ex, nx=8, nx1=10, rsolver=2), 1, mix),
        (atm_config(ex, nx=6, nx1=8, rsolver=3, implicit=1), 1),
        (atm_config(ex, nx=8, nx1=6, rsolver=3, implicit=1), 1, mix),
    ]


def make(ex, spec, seed):
    cfg, kind = spec[0], spec[1]
    cfg.precision = spec[2] if len(spec) > 2 else ex.PREC_FP64
    u0, s0 = ex.init_state(cfg, kind, seed, ())
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    return g


def alone(ex, specs, nsteps):
    out = []
    for q, sp in enumerate(specs):
        g = make(ex, sp, SEED + q)
        g.step(nsteps)
        out.append(g.download())
        g.close()
    return out


def assert_same(got, ref):
    for q, ((u, s), (ur, sr)) in enumerate(zip(got, ref)):
        assert np.array_equal(u, ur), f"member {q}: max diff {np.abs(u - ur).max()}"
        if sr is not None:
            assert np.array_equal(s, sr), f"member {q} scalars"


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_alone(ex, graph):
    specs = member_specs(ex)
    ref = alone(ex, specs, 6)
    members = [make(ex, sp, SEED + q) for q, sp in enumerate(specs)]
    b = ex.Batch(members, graph=graph)
    b.step(3)
    b.step(3)  # a graph is replayed again by the second call
    got = [m.download() for m in members]
    b.close()
    assert_same(got, ref)


def test_batch_persistent_and_tuned(ex):
    specs = member_specs(ex)
    ref = alone(ex, specs, 4)
    members = [make(ex, sp, SEED + q) for q, sp in enumerate(specs)]
    b = ex.Batch(members, graph=True)
    # a very narrow persistent launch: every workgroup loops over many blocks;
    # the x1 sweeps in their non-default block shapes
    caps = {k: 7 for k in ex.LAUNCH_KINDS}
    caps.update({"x1_shape": 8, "x1_mix_shape": 2})
    b.set_caps(caps)
    b.step(2)
    caps = b.tune(reps=2)
    assert set(ex.LAUNCH_KINDS) <= set(caps)
    assert all(caps[k] in (2, 4, 8) for k in ("x1_shape", "x1_mix_shape") if k in caps)
    b.step(2)
    got = [m.download() for m in members]
    b.close()
    assert_same(got, ref)


def test_batch_of_one_equals_handle(ex):
    spec = (atm_config(ex, nx=12, nx1=12, rsolver=2), 1)
    ref = alone(ex, [spec], 5)
    m = make(ex, spec, SEED)
    b = ex.Batch([m], graph=True)
    b.step(5)
    got = [m.download()]
    b.close()
    assert_same(got, ref)


def volumes(ex, cfg):
    nc1, nc2 = cfg.nx1 + 6, cfg.nx + 6
    tab = ex.describe(cfg, 2).reshape(20, 6, nc2, nc2)
    x1 = ex.describe(cfg, 3)
    x1v, dx1f = x1[nc1 + 1:2 * nc1 + 1], x1[2 * nc1 + 1:]
    # CellVolume = dx1f * dx2f_ang * dx3f_ang * x1v^2 * sin(theta)
    return (tab[6] * tab[7] * tab[1])[:, :, :, None] * (dx1f * x1v * x1v)[None, None, None, :]


def exact_totals(ex, cfg, u):
    V = volumes(ex, cfg)
    tot, mag = [], []
    for n in range(u.shape[1]):
        a = (u[:, n] * V)[:, 3:-3, 3:-3, 3:3 + cfg.nx1].ravel()
        tot.append(math.fsum(a))  # correctly rounded
        mag.append(np.abs(a).sum())
    return np.array(tot), np.array(mag)


@pytest.mark.parametrize("which", [0, 4, 5])
def test_conserved_sums_match_exact_sum(ex, which):
    spec = member_specs(ex)[which]
    g = make(ex, spec, SEED)
    g.step(3)
    s = g.sums()
    u, _ = g.download()
    tot, mag = exact_totals(ex, spec[0], u)
    g.close()
    # compensated: within a couple of roundings of the magnitude
    assert np.all(np.abs(s - tot) <= 4 * np.finfo(float).eps * mag), (s, tot)


@pytest.mark.parametrize("which", [1, 2, 4, 5, 7, 8, 9, 10])
def test_mass_is_conserved(ex, which):
    spec = member_specs(ex)[which]
    g = make(ex, spec, SEED)
    m0 = g.sums()[0]
    g.step(20)
    m1 = g.sums()[0]
    g.close()
    assert abs(m1 - m0) <= 1e-13 * abs(m0), (m0, m1, (m1 - m0) / m0)


def test_batch_sums_equal_member_sums(ex):
    specs = member_specs(ex)
    members = [make(ex, sp, SEED + q) for q, sp in enumerate(specs)]
    b = ex.Batch(members, graph=True)
    b.step(2)
    bs = b.sums()
    for q, m in enumerate(members):
        s = m.sums()
        assert np.array_equal(bs[q, :len(s)], s)
    b.close()
