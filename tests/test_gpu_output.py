"""Output of the state in the reference's NetCDF layout, from the device.

The "prim", "cons" and "uov" fields that the HIP output kernel writes are the
float32 roundings of the oracle's FP64 values (its primitive and conserved
state and its restatement of examples/2023-Chen-exo3/W92.cpp:115-141), bit for
bit, interior and ghost cells: x1 ghost levels are the reflecting mirrors of
the interior, and the panel-corner ghosts (never exchanged, never read) are
NaN except for lat, lon and sqrtg.  The files written from the device hold
those fields.  A mixed-precision handle's output is its own FP64 state
rounded."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from helpers import atm_config, sw_config

pytestmark = pytest.mark.gpu
SEED = 20261202


def file_layout(a, nx1, ghost):
    """(6, n, nc3, nc2, nc1) host layout -> (6, n, x1, x2, x3) float32"""
    if not ghost:
        a = a[:, :, 3:-3, 3:-3, 3:3 + nx1]
    return np.ascontiguousarray(a.transpose(0, 1, 4, 3, 2)).astype(np.float32)


def run_pair(ex, orc, cfg, kind, steps):
    u0, s0 = ex.init_state(cfg, kind, SEED, ())
    if cfg.ntracer:
        rng = np.random.default_rng(5)
        s0 = rng.uniform(0.0, 1.0, (6, cfg.ntracer) + u0.shape[2:]) * u0[:, :1]
    g = ex.Solver(cfg)
    g.upload(u0, s0)
    o = orc.Oracle(cfg)
    o.upload(u0, s0)
    assert g.step(steps) == o.step(steps)
    return g, o


def mirror_x1(a, nx1, flip):
    """x1 ghost levels as reflections of the interior, repeated while outside
    (nx1 < 3), field `flip` negated once per reflection"""
    a = a.copy()
    lo, hi = 3, 2 + nx1
    for i in list(range(3)) + list(range(hi + 1, hi + 4)):
        src, n = i, 0
        while src < lo or src > hi:
            src = 2 * lo - 1 - src if src < lo else 2 * hi + 1 - src
            n += 1
        a[..., i] = a[..., src]
        if flip is not None and n % 2:
            a[:, flip, ..., i] = -a[:, flip, ..., src]
    return a


def corners_nan(a, keep=()):
    a = a.copy()
    for q in range(a.shape[1]):
        if q in keep:
            continue
        for ks in (slice(0, 3), slice(-3, None)):
            for js in (slice(0, 3), slice(-3, None)):
                a[:, q, ks, js, :] = np.nan
    return a


def expected(o, cfg, var, ghost):
    nx1 = cfg.nx1
    idx = [0, 1, 2, 3] if cfg.eos == 0 else [0, 4, 1, 2, 3]
    if var == "uov":
        a = corners_nan(o.user_output(), keep=(0, 1, 4))
    elif var == "prim":
        parts = [o.prim()[:, idx]] + ([o.scalar_prim()] if cfg.ntracer else [])
        a = corners_nan(np.concatenate(parts, axis=1))
    else:
        u, s = o.download()
        parts = [mirror_x1(u[:, idx], nx1, 1 if cfg.eos == 0 else 2)]
        if cfg.ntracer:
            parts.append(mirror_x1(s, nx1, None))
        a = corners_nan(np.concatenate(parts, axis=1))
    return file_layout(a, nx1, ghost)


@pytest.mark.parametrize("which", ["sw", "sw1", "atm"])
def test_output_fields_vs_oracle(ex, orc, which):
    if which.startswith("sw"):
        cfg, kind = sw_config(ex, nx=8, rsolver=0, nx1=1 if which == "sw1" else 3), 0
    else:
        cfg, kind = atm_config(ex, nx=6, nx1=6, rsolver=3, ntracer=2), 1
    g, o = run_pair(ex, orc, cfg, kind, 2)
    for var in ("prim", "cons", "uov"):
        for ghost in (False, True):
            got = g.output_fields(var, ghost)
            want = expected(o, cfg, var, ghost)
            assert got.shape == want.shape
            bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
            assert not bad.any(), (var, ghost, np.argwhere(bad)[:5])
    g.close()
    o.close()


def test_output_fields_mixed_handle(ex):
    # the output kernel reads the FP64 state whatever the flux precision
    for cfg, kind in ((sw_config(ex, nx=16, rsolver=1, nx1=1), 0),
                      (atm_config(ex, nx=8, nx1=8, rsolver=3), 1)):
        cfg.precision = ex.PREC_MIXED
        g = ex.Solver(cfg)
        u0, _ = ex.init_state(cfg, kind, SEED, ())
        g.upload(u0)
        g.step(3)
        w = g.prim()
        idx = [0, 1, 2, 3] if cfg.eos == 0 else [0, 4, 1, 2, 3]
        assert np.array_equal(g.output_fields("prim"), file_layout(w[:, idx], cfg.nx1, False))
        u, _ = g.download()
        assert np.array_equal(g.output_fields("cons"), file_layout(u[:, idx], cfg.nx1, False))
        g.close()


def test_netcdf_files_from_device(ex, tmp_path):
    cfg = sw_config(ex, nx=16, rsolver=0, nx1=3)
    g = ex.Solver(cfg)
    u0, _ = ex.init_state(cfg, 0, SEED, ())
    g.upload(u0)
    dt = g.step(2)
    g.profile_begin()
    fields = g.output_fields("uov", True)
    prof = g.profile_end()
    assert prof["output"][1] == 1 and prof["output"][0] > 0.0
    names = ex.output_names(cfg, "uov")
    files = g.write_netcdf(str(tmp_path / "W92"), "uov", time=2 * dt, file_number=3,
                           blocks_per_side=2, ghost_zones=True)
    assert len(files) == 24
    nb = 8
    for gid, path in enumerate(files):
        p, z = divmod(gid, 4)
        sk, sj = z >> 1, z & 1
        with netcdf_file(path, "r", mmap=False) as f:
            assert f.variables["time"].data[0] == np.float32(2 * dt)
            for q, n in enumerate(names):
                slab = fields[p, q, :, sj * nb:sj * nb + nb + 6, sk * nb:sk * nb + nb + 6]
                assert np.array_equal(f.variables[n].data[0], slab, equal_nan=True), (gid, n)
    g.close()


def test_output_before_upload_is_refused(ex):
    g = ex.Solver(sw_config(ex, nx=8))
    with pytest.raises(ex.ExoError):
        g.output_fields("prim")
    g.close()


def test_w92_run_files_match_oracle_files(ex, orc, tmp_path):
    """Integration in the reference's style (run, write the per-block files,
    compare them with a known-good set by an L2 norm): a W92-like shallow-water
    sphere run 20 steps on the device and written by exo_write_netcdf, against
    the same run of the oracle written from its fields by the host writer."""
    cfg = sw_config(ex, nx=24, rsolver=0, nx1=3)
    g, o = run_pair(ex, orc, cfg, 0, 20)
    for var in ("prim", "uov"):
        gf = g.write_netcdf(str(tmp_path / "gpu"), var, time=3600.0, file_number=1)
        ex.write_netcdf_fields(cfg, str(tmp_path / "ref"), var, expected(o, cfg, var, False),
                               time=3600.0, file_number=1)
        of = ex.output_file_names(str(tmp_path / "ref"), var, 1)
        for a, b in zip(gf, of):
            with netcdf_file(a, "r", mmap=False) as fa, netcdf_file(b, "r", mmap=False) as fb:
                assert list(fa.variables) == list(fb.variables)
                for n in fa.variables:
                    x, y = fa.variables[n].data, fb.variables[n].data
                    assert np.sqrt(np.sum((x.astype(np.float64) - y) ** 2)) == 0.0, (a, n)
    g.close()
    o.close()
