"""Failure recovery and restart on the device.

* The redo path of the reference's main loop (main.cpp.in:126-160): an
  invalid step (negative or NaN density or pressure in the last primitive
  conversion) is undone from an in-HBM snapshot and repeated with half the
  time step; bit for bit against the oracle's restatement, same redo counts
  and time steps.
* Checkpoint / restart: a run restarted from a checkpoint continues bit for bit
  as the uninterrupted run (both precisions, with tracers); damaged or foreign
  checkpoints are refused."""
import os

import numpy as np
import pytest

from helpers import atm_config, interior, sw_config

pytestmark = pytest.mark.gpu
SEED = 20261202


def _atm(ex, rs, impl, cfl, ntracer=0):
    cfg = atm_config(ex, nx=6, nx1=6, rsolver=rs, implicit=impl, ntracer=ntracer)
    cfg.cfl = cfl
    return cfg


@pytest.mark.parametrize("rs,impl,cfl,nsteps", [(3, 9, 1.5, 3), (2, 0, 4.0, 4), (3, 9, 0.4, 3),
                                                   (3, 1, 1.5, 3)])
def test_step_redo_vs_oracle(ex, orc, rs, impl, cfl, nsteps):
    cfg = _atm(ex, rs, impl, cfl)
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    g = ex.Solver(cfg)
    g.upload(u0)
    o = orc.Oracle(cfg)
    o.upload(u0)
    dg, rg = g.step_redo(nsteps, 5)
    do, ro = o.step_redo(nsteps, 5)
    assert (dg, rg) == (do, ro)
    if cfl == 0.4:
        assert rg == 0
    else:
        assert rg > 0
    assert g.state_valid() and o.state_valid()
    assert np.array_equal(interior(g.download()[0], 6), interior(o.download()[0], 6))
    g.close()
    o.close()


def test_validity_flag_and_too_many_redos(ex, orc):
    cfg = _atm(ex, 3, 0, 8.0)
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    g = ex.Solver(cfg)
    g.upload(u0)
    assert g.state_valid()
    g.step(1)                      # no checks: the state goes invalid
    o = orc.Oracle(cfg)
    o.upload(u0)
    o.step(1)
    assert g.state_valid() == o.state_valid() is False
    g.upload(u0)
    assert g.state_valid()
    with pytest.raises(ex.ExoError, match="Too many attempts"):
        g.step_redo(4, 2)
    g.close()
    o.close()


def test_save_load_states(ex):
    cfg = atm_config(ex, nx=6, nx1=6, ntracer=2)
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    s0 = np.random.default_rng(1).uniform(0, 1, (6, 2) + u0.shape[2:]) * u0[:, :1]
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    g.step(1)
    u1, s1 = g.download()
    w1 = g.prim()
    g.save_states()
    g.step(2)
    g.load_states()
    u2, s2 = g.download()
    assert np.array_equal(u1, u2) and np.array_equal(s1, s2)
    assert np.array_equal(w1, g.prim())
    g.step(2)
    a = ex.Solver(cfg)
    a.upload(u0, s0)
    a.step(3)
    assert np.array_equal(interior(a.download()[0], 6), interior(g.download()[0], 6))
    a.close()
    g.close()


@pytest.mark.parametrize("which,precision", [("atm", 0), ("atm", 1), ("sw", 0), ("sw", 1)])
def test_checkpoint_restart_is_bit_exact(ex, tmp_path, which, precision):
    if which == "atm":
        cfg, kind, nx1 = atm_config(ex, nx=8, nx1=8, rsolver=3, ntracer=2), 1, 8
    else:
        cfg, kind, nx1 = sw_config(ex, nx=16, rsolver=0, nx1=1), 0, 1
    cfg.precision = precision
    u0, _ = ex.init_state(cfg, kind, SEED, ())
    s0 = None
    if cfg.ntracer:
        s0 = np.random.default_rng(2).uniform(0, 1, (6, 2) + u0.shape[2:]) * u0[:, :1]
    a = ex.Solver(cfg)
    a.upload(u0, s0)
    a.step(2)
    dta = a.step(3)
    b = ex.Solver(cfg)
    b.upload(u0, s0)
    b.step(2)
    path = str(tmp_path / "run.ckpt")
    b.save_checkpoint(path, time=1234.5, ncycle=2)
    assert not os.path.exists(path + ".tmp")
    b.close()
    c = ex.Solver(cfg)
    assert c.load_checkpoint(path) == (1234.5, 2)
    dtc = c.step(3)
    assert dtc == dta
    (ua, sa), (uc, sc) = a.download(), c.download()
    assert np.array_equal(interior(ua, nx1), interior(uc, nx1))
    if cfg.ntracer:
        assert np.array_equal(interior(sa, nx1), interior(sc, nx1))
    a.close()
    c.close()


def test_checkpoint_refusals_and_cross_precision(ex, tmp_path):
    cfg = atm_config(ex, nx=6, nx1=6, rsolver=3)
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    g = ex.Solver(cfg)
    g.upload(u0)
    g.step(2)
    path = str(tmp_path / "a.ckpt")
    g.save_checkpoint(path, 10.0, 2)
    # an FP64 checkpoint restarts a mixed-precision run
    cfg_m = atm_config(ex, nx=6, nx1=6, rsolver=3)
    cfg_m.precision = ex.PREC_MIXED
    m = ex.Solver(cfg_m)
    m.load_checkpoint(path)
    m.step(2)
    g.step(2)
    um, ug = interior(m.download()[0], 6), interior(g.download()[0], 6)
    assert np.isfinite(um).all()
    assert np.abs(um - ug).max() / np.abs(ug).max() < 1e-4
    # another configuration
    other = ex.Solver(atm_config(ex, nx=8, nx1=6, rsolver=3))
    with pytest.raises(ex.ExoError, match="another configuration"):
        other.load_checkpoint(path)
    # a damaged payload, a truncated file, not a checkpoint at all
    raw = open(path, "rb").read()
    bad = bytearray(raw)
    bad[len(raw) // 2] ^= 0x40
    for name, data, msg in (("flip", bytes(bad), "checksum"),
                            ("cut", raw[:len(raw) - 100], "not a complete"),
                            ("junk", b"hello world" * 10, "not a checkpoint")):
        p = str(tmp_path / name)
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(ex.ExoError, match=msg):
            g.load_checkpoint(p)
    with pytest.raises(ex.ExoError, match="cannot open"):
        g.load_checkpoint(str(tmp_path / "missing.ckpt"))
    for h in (g, m, other):
        h.close()


def test_host_restart_with_dt_state(ex):
    # a host restarting from its own restart file: upload the conserved state
    # and set the previous dt; the FP64 run continues bit for bit
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=2)
    u0, _ = ex.init_state(cfg, 1, SEED, ())
    a = ex.Solver(cfg)
    a.upload(u0)
    a.step(2)
    u_mid, _ = a.download()
    dt, dt_prev = a.dt_state()
    assert dt == dt_prev > 0
    a.step(3)
    b = ex.Solver(cfg)
    b.upload(u_mid)
    assert b.dt_state()[1] > 1e300        # no bound after a plain upload
    b.set_dt_state(dt, dt_prev)
    b.step(3)
    assert np.array_equal(interior(a.download()[0], 8), interior(b.download()[0], 8))
    with pytest.raises(ex.ExoError):
        b.set_dt_state(1.0, 0.0)
    a.close()
    b.close()
