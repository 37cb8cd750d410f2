"""Parity of the remaining C-ABI entry points against the oracle (bit for bit):
single SSPRK3 stages, the CFL time step, single-face Riemann fluxes with the
grid's face metric, fixed-dt stepping, minimum sizes and shallow water with
passive tracers."""
import numpy as np
import pytest

from helpers import atm_config, interior, sw_config

pytestmark = pytest.mark.gpu
SEED = 20261202


def pair(ex, orc, cfg, kind, seed=SEED):
    u0, s0 = ex.init_state(cfg, kind, seed, ())
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    o = orc.Oracle(cfg)
    o.upload(u0, s0)
    return g, o


@pytest.mark.parametrize("which", ["sw", "atm"])
def test_single_stages(ex, orc, which):
    if which == "sw":
        cfg, kind, nx1 = sw_config(ex, nx=8, rsolver=0), 0, 3
    else:
        cfg, kind, nx1 = atm_config(ex, nx=6, nx1=8, rsolver=2, ntracer=1), 1, 8
    g, o = pair(ex, orc, cfg, kind)
    dt = 0.5 * g.new_dt()
    assert dt == 0.5 * o.new_dt()
    for stage in (1, 2, 3):
        g.stage(stage, dt)
        o.stage(stage, dt)
        ug, sg = g.download()
        uo, so = o.download()
        assert np.array_equal(interior(ug, nx1), interior(uo, nx1)), stage
        if sg is not None:
            assert np.array_equal(interior(sg, nx1), interior(so, nx1)), stage
    g.close()
    o.close()


def test_cfl_dt_after_steps(ex, orc):
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=3)
    g, o = pair(ex, orc, cfg, 1)
    for _ in range(3):
        assert g.step(1) == o.step(1)
    assert g.new_dt() == o.new_dt()
    g.close()
    o.close()


@pytest.mark.parametrize("cfgname", ["sw_roe", "sw_lmars", "hllc", "lmars"])
def test_riemann_face_with_grid_metric(ex, orc, cfgname):
    if cfgname.startswith("sw"):
        cfg = sw_config(ex, nx=8, rsolver=0 if cfgname == "sw_roe" else 1)
    else:
        cfg = atm_config(ex, nx=8, nx1=8, rsolver=2 if cfgname == "hllc" else 3)
    g = ex.Solver(cfg)
    nc = cfg.nx + 6
    tab = ex.describe(cfg, 2).reshape(20, 6, nc, nc)
    rng = np.random.default_rng(11)
    L = orc.lib()
    oc = orc.or_config(cfg)
    nv = g.nv
    for trial in range(40):
        d = int(rng.integers(0, 3)) if nv == 5 else int(rng.integers(1, 3))
        p, k, j = int(rng.integers(0, 6)), int(rng.integers(3, nc - 3)), int(rng.integers(3, nc - 3))
        if nv == 4:
            wl = np.array([rng.uniform(5e4, 6e4), 0.0, rng.normal(0, 20), rng.normal(0, 20)])
            wr = np.array([rng.uniform(5e4, 6e4), 0.0, rng.normal(0, 20), rng.normal(0, 20)])
        else:
            wl = np.array([rng.uniform(0.5, 1.5), *rng.normal(0, 10, 3), rng.uniform(5e4, 1e5)])
            wr = np.array([rng.uniform(0.5, 1.5), *rng.normal(0, 10, 3), rng.uniform(5e4, 1e5)])
        fg = np.zeros(nv)
        assert g.L.exo_riemann_face(g.h, d, p, k, j, wl, wr, fg) == 0
        cth = tab[10 + d, p, k, j]
        cosf, sinf = (0.0, 1.0) if d == 0 else (tab[2 * d, p, k, j], tab[2 * d + 1, p, k, j])
        fo = np.zeros(nv)
        L.or_riemann_face(oc, [1, 2, 3][d], wl, wr, fo, cth, cosf, sinf, tab[0, p, k, j])
        assert np.array_equal(fg, fo), (trial, d, fg, fo)
    g.close()


def test_fixed_dt_steps(ex, orc):
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=2)
    g, o = pair(ex, orc, cfg, 1)
    assert g.step(3, 150.0) == o.step(3, 150.0) == 150.0
    assert np.array_equal(interior(g.download()[0], 8), interior(o.download()[0], 8))
    g.close()
    o.close()


@pytest.mark.parametrize("which", ["sw", "atm_explicit", "atm_implicit"])
def test_minimum_sizes(ex, orc, which):
    # the smallest grids the set-up accepts (4 x 4 panels, 3 levels); with the
    # implicit solve, 4 x 4 panels take a ~2600 s first step that the scheme
    # does not survive (the oracle goes non-finite too), so that case uses 6 x 6
    if which == "sw":
        cfg, kind, nx1 = sw_config(ex, nx=4, rsolver=0, nx1=1), 0, 1
    elif which == "atm_explicit":
        cfg, kind, nx1 = atm_config(ex, nx=4, nx1=3, rsolver=2, implicit=0), 1, 3
    else:
        cfg, kind, nx1 = atm_config(ex, nx=6, nx1=3, rsolver=2), 1, 3
    g, o = pair(ex, orc, cfg, kind)
    assert g.step(2) == o.step(2)
    ug, uo = interior(g.download()[0], nx1), interior(o.download()[0], nx1)
    assert np.all(np.isfinite(uo))
    assert np.array_equal(ug, uo)
    g.close()
    o.close()


@pytest.mark.parametrize("which,nt", [("sw", 3), ("sw", 4), ("atm", 4)])
def test_three_or_more_tracers(ex, orc, which, nt):
    # passive scalars are remapped across panel edges as scalars, whatever
    # their count (indices 2 and 3 are not a velocity pair)
    if which == "sw":
        cfg, kind, nx1 = sw_config(ex, nx=8, rsolver=1, nx1=1), 0, 1
    else:
        cfg, kind, nx1 = atm_config(ex, nx=6, nx1=6, rsolver=2), 1, 6
    cfg.ntracer = nt
    u0, _ = ex.init_state(cfg, kind, SEED, ())
    rng = np.random.default_rng(3)
    s0 = rng.uniform(0.0, 1.0, (6, nt) + u0.shape[2:]) * u0[:, :1]  # s = rho r
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    o = orc.Oracle(cfg)
    o.upload(u0, s0)
    assert g.step(3) == o.step(3)
    (ug, sg), (uo, so) = g.download(), o.download()
    assert np.array_equal(interior(ug, nx1), interior(uo, nx1))
    assert np.array_equal(interior(sg, nx1), interior(so, nx1))
    g.close()
    o.close()


@pytest.mark.parametrize("which", ["sw", "atm", "atm_tracers", "atm_partial", "atm_explicit"])
def test_stage_by_entry_points(ex, orc, which):
    """A full SSPRK3 step through the boundary's single phases
    (exo_calc_fluxes, exo_add_flux_divergence at stage 1 and
    exo_integrate_hydro at stages 2 and 3, exo_implicit_correct,
    exo_fill_ghosts), each phase against the oracle's phase, and each stage
    equal to exo_stage (whose update leaves the weighted average of u to the
    column solve or to u += du)."""
    if which == "sw":
        cfg, kind, nx1 = sw_config(ex, nx=8, rsolver=1), 0, 3
    else:
        flag = {"atm_partial": 1, "atm_explicit": 0}.get(which, 9)
        cfg = atm_config(ex, nx=6, nx1=8, rsolver=3, coriolis=2,
                         ntracer=2 if which == "atm_tracers" else 0, implicit=flag)
        kind, nx1 = 1, 8
    u0, _ = ex.init_state(cfg, kind, SEED, ())
    s0 = None
    if cfg.ntracer:
        s0 = np.random.default_rng(4).uniform(0, 1, (6, 2) + u0.shape[2:]) * u0[:, :1]
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    o = orc.Oracle(cfg)
    o.upload(u0, s0)
    f = ex.Solver(cfg)
    f.upload(u0, s0)
    dt = g.new_dt()
    beta = (1.0, 0.25, 2.0 / 3.0)
    for stage in (1, 2, 3):
        wght = beta[stage - 1] * dt
        g.calc_fluxes()
        o.calc_fluxes()
        if stage == 1:
            g.add_flux_divergence(dt)
        else:
            g.integrate_hydro(stage, dt)
        o.add_flux_divergence(wght, stage)
        if cfg.eos == ex.EOS_SHALLOW_YZ:   # the update went into u
            assert np.array_equal(interior(g.download()[0], nx1), interior(o.download()[0], nx1))
        g.implicit_correct(wght)
        o.implicit_correct(wght)
        (ug, sg), (uo, so) = g.download(), o.download()
        assert np.array_equal(interior(ug, nx1), interior(uo, nx1)), stage
        if cfg.ntracer:
            assert np.array_equal(interior(sg, nx1), interior(so, nx1)), stage
        g.fill_ghosts()
        o.fill_ghosts()
        ug, uo = g.download()[0], o.download()[0]
        edge = (slice(None), slice(None), slice(3, -3), slice(0, 3), slice(3, 3 + nx1))
        assert np.array_equal(ug[edge], uo[edge]), stage          # a ghost strip
        wg, wo = g.prim(), o.prim()
        assert np.array_equal(interior(wg, nx1), interior(wo, nx1)), stage
        f.stage(stage, dt)
        assert np.array_equal(interior(f.download()[0], nx1), interior(ug, nx1)), stage
    for h in (g, o, f):
        h.close()


@pytest.mark.parametrize("ntracer", [0, 2])
def test_mixed_stage_by_entry_points(ex, ntracer):
    """The mixed form: exo_stage (whose column solve takes the stage update in)
    equals the single phases exo_calc_fluxes, exo_add_flux_divergence,
    exo_implicit_correct, exo_fill_ghosts bit for bit."""
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=3, coriolis=2, ntracer=ntracer)
    cfg.precision = ex.PREC_MIXED
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    s0 = None
    if ntracer:
        s0 = np.random.default_rng(4).uniform(0, 1, (6, 2) + u0.shape[2:]) * u0[:, :1]
    g, f = ex.Solver(cfg), ex.Solver(cfg)
    g.upload(u0, s0)
    f.upload(u0, s0)
    dt = g.new_dt()
    g.calc_fluxes()
    g.add_flux_divergence(dt)
    g.implicit_correct(dt)
    g.fill_ghosts()
    f.stage(1, dt)
    (ug, sg), (uf, sf) = g.download(), f.download()
    assert np.array_equal(ug, uf)
    if ntracer:
        assert np.array_equal(sg, sf)
    g.close()
    f.close()
