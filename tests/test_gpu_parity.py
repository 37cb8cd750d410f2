"""FP64 GPU path vs the oracle, through the C ABI.  The FP64 kernels keep the
reference's operation order and are compiled without contraction, so the
expected difference is zero; the pass bar is the 1e-12 relative bound."""
import numpy as np
import pytest

from helpers import atm_config, interior, rel_err, sw_config

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _run_both(ex, orc, cfg, kind, nsteps, seed=20261202, params=()):
    u0, s0 = ex.init_state(cfg, kind, seed, params)
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    o = orc.Oracle(cfg)
    o.upload(u0, s0)
    dg = g.step(nsteps)
    do = o.step(nsteps)
    ug, sg = g.download()
    uo, so = o.download()
    g.close()
    o.close()
    return dg, do, ug, uo, sg, so


def test_weno5_device_matches_oracle(ex, orc):
    cfg = sw_config(ex)
    g = ex.Solver(cfg)
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal((2000, 5)) * 10 ** rng.uniform(-8, 8, (2000, 1)),
                        np.array([[1., 2., 3., 4., 5.]])])
    out = np.zeros(len(x))
    assert g.L.exo_weno5_batch(g.h, np.ascontiguousarray(x), out, len(x)) == 0
    ref = np.array([orc.lib().or_weno5(*r) for r in x])
    assert out[-1] == 2.5000000000000004
    assert np.array_equal(out, ref)


def test_math_device_matches_oracle(ex, orc):
    cfg = sw_config(ex)
    g = ex.Solver(cfg)
    L = orc.lib()
    x = np.exp(np.random.default_rng(2).uniform(-30, 30, 4000))
    for op, f in [(0, lambda v: L.or_exp(v)), (1, lambda v: L.or_log(v)),
                  (2, lambda v: L.or_pow(v, 1 / 1.4))]:
        xx = np.log(x) if op == 0 else x
        out = np.zeros(len(xx))
        assert g.L.exo_math_batch(g.h, op, np.ascontiguousarray(xx), 1 / 1.4, out, len(xx)) == 0
        ref = np.array([f(v) for v in xx])
        assert np.array_equal(out, ref), op


def test_device_shared_reciprocal_division_is_ieee(ex):
    """the implicit solve's x / d from RN(1/d) plus one FMA correction equals
    the IEEE quotient on the device (random operands, hard significands)"""
    rng = np.random.default_rng(7)
    n = 400000
    x = rng.standard_normal(n) * np.exp2(rng.integers(-60, 61, n))
    d = rng.standard_normal(n) * np.exp2(rng.integers(-60, 61, n))
    hard = np.array([1.0, np.nextafter(2.0, 0.0), np.nextafter(1.0, 2.0), 1.5,
                     np.nextafter(1.0, 0.0), 3.0, 1.0 + 2.0 ** -52 * 3])
    d[::5] = rng.choice(hard, len(d[::5])) * np.exp2(rng.integers(-30, 31, len(d[::5])))
    x[1::5] = rng.choice(hard, len(x[1::5])) * np.exp2(rng.integers(-30, 31, len(x[1::5])))
    pairs = np.ascontiguousarray(np.stack([x, d], axis=1).ravel())
    g = ex.Solver(sw_config(ex))
    out = np.zeros(n)
    assert g.L.exo_math_batch(g.h, 3, pairs, 0.0, out, n) == 0
    g.close()
    assert np.array_equal(out, x / d)


def test_device_division_is_ieee(ex):
    """the FP64 kernels' quotient (reciprocal, two Newton steps, one remainder
    correction, v_div_fixup; no range scaling) equals numpy's IEEE x / d bit for
    bit: exponents over +-300, hard significands, signed zeros, infinities and
    NaN in either operand"""
    rng = np.random.default_rng(11)
    n = 600000
    x = rng.standard_normal(n) * np.exp2(rng.integers(-300, 301, n))
    d = rng.standard_normal(n) * np.exp2(rng.integers(-300, 301, n))
    hard = np.array([1.0, np.nextafter(2.0, 0.0), np.nextafter(1.0, 2.0), 1.5,
                     np.nextafter(1.0, 0.0), 3.0, 1.0 + 2.0 ** -52 * 3, 1.0 / 3.0])
    d[::5] = rng.choice(hard, len(d[::5])) * np.exp2(rng.integers(-40, 41, len(d[::5])))
    x[1::5] = rng.choice(hard, len(x[1::5])) * np.exp2(rng.integers(-40, 41, len(x[1::5])))
    x[2::5] = d[2::5] * rng.choice(hard, len(x[2::5]))  # quotients near 1
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -2.5, 1e-250, 1e250])
    sx, sd = np.meshgrid(special, special)
    x[:sx.size], d[:sx.size] = sx.ravel(), sd.ravel()
    pairs = np.ascontiguousarray(np.stack([x, d], axis=1).ravel())
    g = ex.Solver(sw_config(ex))
    out = np.zeros(n)
    assert g.L.exo_math_batch(g.h, 4, pairs, 0.0, out, n) == 0
    g.close()
    with np.errstate(divide="ignore", invalid="ignore", over="ignore", under="ignore"):
        ref = x / d  # the specials overflow and underflow on purpose
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(out), nan)
    assert np.array_equal(out[~nan].view(np.uint64), ref[~nan].view(np.uint64))


@pytest.mark.parametrize("rsolver", [0, 1])
def test_shallow_water_fluxes(ex, orc, rsolver):
    cfg = sw_config(ex, nx=8, rsolver=rsolver)
    u0, _ = ex.init_state(cfg, 0, 3)
    g = ex.Solver(cfg)
    g.upload(u0)
    o = orc.Oracle(cfg)
    o.upload(u0)
    wg, wo = g.prim(), o.prim()
    assert np.array_equal(interior(wg, 3), interior(wo, 3))
    g.calc_fluxes()
    o.calc_fluxes()
    for d in (1, 2):
        fg, fo = g.flux(d), o.flux(d)
        if d == 1:
            sl = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, 6))
        else:
            sl = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, 6))
        assert np.array_equal(fg[sl], fo[sl]), (d, np.abs(fg[sl] - fo[sl]).max())


@pytest.mark.parametrize("rsolver", [0, 1])
def test_shallow_water_steps(ex, orc, rsolver):
    cfg = sw_config(ex, nx=8, rsolver=rsolver)
    dg, do, ug, uo, _, _ = _run_both(ex, orc, cfg, 0, 4)
    assert dg == do
    e = rel_err(interior(ug, 3), interior(uo, 3))
    assert e.max() <= TOL, e


@pytest.mark.parametrize("rsolver,implicit", [(3, 9), (2, 9), (3, 0), (3, 1), (2, 1), (3, 8)])
def test_atmosphere_steps(ex, orc, rsolver, implicit):
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=rsolver, implicit=implicit)
    nsteps = 3 if implicit & 1 else 1
    dg, do, ug, uo, _, _ = _run_both(ex, orc, cfg, 1, nsteps)
    assert dg == do
    e = rel_err(interior(ug, 8), interior(uo, 8))
    assert e.max() <= TOL, e


@pytest.mark.parametrize("rsolver,nx1", [(3, 10), (2, 3), (2, 16)])
def test_partial_correction_bit_exact(ex, orc, rsolver, nx1):
    """implicit_flag 1: the 3x3 (density, x1 momentum, energy) column solve of
    partial_correction.cpp, stepped bit for bit against the oracle."""
    cfg = atm_config(ex, nx=8, nx1=nx1, rsolver=rsolver, implicit=1)
    dg, do, ug, uo, _, _ = _run_both(ex, orc, cfg, 1, 3)
    assert dg == do
    assert np.array_equal(interior(ug, nx1), interior(uo, nx1))


def test_atmosphere_fluxes_x1(ex, orc):
    cfg = atm_config(ex, nx=8, nx1=8)
    u0, _ = ex.init_state(cfg, 1, 5)
    g = ex.Solver(cfg)
    g.upload(u0)
    o = orc.Oracle(cfg)
    o.upload(u0)
    g.calc_fluxes()
    o.calc_fluxes()
    sl = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, 3 + 9))
    fg, fo = g.flux(0), o.flux(0)
    assert np.array_equal(fg[sl], fo[sl]), np.abs(fg[sl] - fo[sl]).max()
    for d in (1, 2):
        sl = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, 11))
        fg, fo = g.flux(d), o.flux(d)
        assert np.array_equal(fg[sl], fo[sl]), (d, np.abs(fg[sl] - fo[sl]).max())


def test_tracers_and_coriolis(ex, orc):
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=2, ntracer=2, coriolis=2)
    dg, do, ug, uo, sg, so = _run_both(ex, orc, cfg, 2, 2)
    assert dg == do
    assert rel_err(interior(ug, 8), interior(uo, 8)).max() <= TOL
    assert rel_err(interior(sg, 8), interior(so, 8)).max() <= TOL


def test_device_matches_golden_fixtures(ex):
    """the device WENO5 reproduces the reference's known answer bit for bit,
    and small seeded runs reproduce the committed oracle totals"""
    import json
    import math
    import os
    import sys

    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    kat = json.load(open(os.path.join(gdir, "weno_kat.json")))
    g = ex.Solver(sw_config(ex))
    for c in kat["weno5"]:
        x = np.array([c["phim2"], c["phim1"], c["phi"], c["phip1"], c["phip2"]])
        out = np.zeros(1)
        assert g.L.exo_weno5_batch(g.h, x, out, 1) == 0
        assert out[0] == c["expected"]
    g.close()
    sys.path.insert(0, gdir)
    import make_fixtures

    ref = json.load(open(os.path.join(gdir, "oracle_totals.json")))["cases"]
    for name, (make, kind, nsteps) in make_fixtures.CASES.items():
        cfg = make()
        u0, s0 = ex.init_state(cfg, kind, make_fixtures.SEED, ())
        s = ex.Solver(cfg)
        s.upload(u0, s0)
        s.step(nsteps)
        u = interior(s.download()[0], cfg.nx1)
        s.close()
        for n in range(u.shape[1]):
            assert math.fsum(u[:, n].ravel()) == ref[name]["sum"][n], (name, n)


def test_implicit_column_shim_matches_oracle(ex, orc):
    """The single-column shim of the 5x5 column solve (exo_implicit_column)
    against the oracle's FullCorrection column (or_implicit_column), bit for
    bit: stratified columns of 1, 8 and 64 levels with random density and
    pressure noise, velocities and tendencies, at three time steps."""
    import ctypes as C

    rng = np.random.default_rng(11)
    cfg = atm_config(ex, nx=6, nx1=8)   # gamma and grav of the columns
    g = ex.Solver(cfg)
    oc = orc.or_config(cfg)
    for nx1 in (1, 8, 64):
        nc1 = nx1 + 6
        dz = 30000.0 / nx1
        z = (np.arange(nc1) - 3 + 0.5) * dz
        for trial in range(3):
            w = np.zeros((5, nc1))
            w[0] = 1.2 * np.exp(-z / 7300.0) * (1 + 0.01 * rng.standard_normal(nc1))
            w[4] = 1e5 * np.exp(-z / 7300.0) * (1 + 0.01 * rng.standard_normal(nc1))
            w[1:4] = rng.normal(scale=5.0, size=(3, nc1))
            zf = np.arange(nx1 + 1) * dz
            area = 1.0e10 * (1 + zf / 6.371e6) ** 2
            vol = np.full(nx1, 1.0e10 * dz) * (1 + rng.uniform(0, 1e-3, nx1))
            du = rng.normal(scale=1e-2, size=(5, nc1)) * np.abs(w) + 1e-6
            for dt in (0.5, 20.0, 400.0):
                dg, do = du.copy(), du.copy()
                assert g.L.exo_implicit_column(g.h, nx1, np.ascontiguousarray(w), dg, area,
                                               vol, dt) == 0
                orc.lib().or_implicit_column(C.byref(oc), nx1, np.ascontiguousarray(w), do,
                                             area, vol, dt)
                assert np.array_equal(dg, do), (nx1, trial, dt, np.abs(dg - do).max())
                assert not np.array_equal(dg, du)
    g.close()
