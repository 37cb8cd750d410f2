"""The C-ABI library on the CPU: it loads, exports every function that
include/exocubed.h declares, and its host-only entry points (initial states,
set-up tables, argument checks) work without a GPU.  No kernel is launched."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import atm_config, sw_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "exocubed.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:EXO_API\s+)?(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*"
                       r"(exo_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for n in ("exo_create", "exo_destroy", "exo_calc_fluxes", "exo_add_flux_divergence",
              "exo_fill_ghosts", "exo_implicit_correct", "exo_new_dt", "exo_stage", "exo_step",
              "exo_upload_state", "exo_download_state"):
        assert n in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(ex):
    L = C.CDLL(ex.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_library_exports_nothing_else(ex):
    out = subprocess.run(["nm", "-D", "--defined-only", ex.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert exported == set(declared_functions()), exported ^ set(declared_functions())


def test_binding_covers_every_declared_symbol(ex):
    L = ex.lib()
    for n in declared_functions():
        fn = getattr(L, n)
        assert fn.argtypes is not None, f"{n} has no ctypes signature in exocubed_amd"


def test_host_only_initial_states(ex):
    cfg = sw_config(ex, nx=8)
    u, _ = ex.init_state(cfg, 0, 20261202, ())
    gh = u[:, 0, 3:-3, 3:-3, 3:6] / 9.81
    assert np.all(np.isfinite(u))
    # 5960 m mean depth plus 16 modes of 50 m peak amplitude each
    assert 5960 - 800 < gh.min() and gh.max() < 5960 + 800
    assert abs(gh.mean() - 5960) < 50
    u2, _ = ex.init_state(cfg, 0, 20261202, ())
    assert np.array_equal(u, u2)  # seeded: same seed, same state
    u3, _ = ex.init_state(cfg, 0, 7, ())
    assert not np.array_equal(u, u3)
    cfg = atm_config(ex, nx=6, nx1=8, ntracer=2)
    u, s = ex.init_state(cfg, 1, 1, ())
    rho = u[:, 0, 3:-3, 3:-3, 3:11]
    assert np.all(rho > 0) and np.all(np.diff(rho.mean(axis=(0, 1, 2))) < 0)
    assert s.shape == (6, 2, 12, 12, 14) and np.all(s[:, 1] == 0)
    assert np.all(s[:, 0, 3:-3, 3:-3, 3:11] > 0)


def test_host_only_tables(ex):
    cfg = sw_config(ex, nx=8)
    g = ex.describe(cfg, 0)
    e = ex.describe(cfg, 1)
    assert len(g) == 6 * 4 * 3 * 8 and len(e) == 6 * 4 * 8
    # every ghost has a source on another panel, and the edge map is an involution
    assert np.all(g["src_panel"] != g["tgt_panel"])
    pairs = {(r["src_panel"], r["src_dir"], r["src_k"], r["src_j"]) for r in e}
    assert len(pairs) == len(e)


def test_row_tables_give_the_cell_tables_bits(ex):
    """The one-layer shallow-water update forms tan and cos theta of a cell
    from the per-row tangent tables (x, y, C = sqrt(1 + x^2), D = sqrt(1 +
    y^2)) instead of reading the per-cell tables: with IEEE operations in the
    set-up expression's order the two must agree bit for bit."""
    for cfg in (sw_config(ex, nx=8), atm_config(ex, nx=12, nx1=4)):
        nc = cfg.nx + 6
        t1 = ex.describe(cfg, 6).reshape(8, 6, nc + 1)
        t2 = ex.describe(cfg, 2).reshape(20, 6, nc, nc)
        x, c = t1[0][:, None, :nc], t1[2][:, None, :nc]  # [p][1][j]
        y, d = t1[4][:, :nc, None], t1[6][:, :nc, None]  # [p][k][1]
        cosc = (-x * y) / (c * d)
        assert np.array_equal(cosc.view(np.uint64), t2[0].view(np.uint64))
        assert np.array_equal(np.broadcast_to(x, t2[13].shape), t2[13])  # tan x2v
        assert np.array_equal(np.broadcast_to(y, t2[14].shape), t2[14])  # tan x3v
        # C and D are the maps' sqrt(X X + 1) and sqrt(Y Y + 1) (Coriolis);
        # the centre rows have nc entries, the face rows nc + 1
        assert np.array_equal(np.sqrt(t1[0, :, :nc] * t1[0, :, :nc] + 1), t1[2, :, :nc])
        assert np.array_equal(np.sqrt(t1[4, :, :nc] * t1[4, :, :nc] + 1), t1[6, :, :nc])
        # the face rows too, which the set-up tables are formed from
        assert np.array_equal(np.sqrt(1.0 + t1[1] * t1[1]), t1[3])
        assert np.array_equal(np.sqrt(1.0 + t1[5] * t1[5]), t1[7])


def test_invalid_configs_are_rejected_without_a_gpu(ex):
    L = ex.lib()
    n = C.c_size_t()
    for bad in (dict(nx=7), dict(nx=0), dict(nx1=0), dict(eos=5), dict(rsolver=9),
                dict(ntracer=99), dict(cfl=-1.0), dict(precision=7), dict(coriolis=3),
                dict(x1max=0.0)):
        cfg = sw_config(ex, nx=8)
        for k, v in bad.items():
            setattr(cfg, k, v)
        assert L.exo_describe(C.byref(cfg), 0, None, 0, C.byref(n)) != 0, bad


def test_implicit_flags(ex):
    """x1 column correction: bit 1 solves, bit 8 picks the full system; the
    x2/x3 bits are not supported."""
    L = ex.lib()
    n = C.c_size_t()
    for flag, ok in ((0, True), (1, True), (8, True), (9, True), (2, False), (3, False),
                     (13, False), (-1, False)):
        cfg = atm_config(ex, nx=8, nx1=8, implicit=flag)
        rc = L.exo_describe(C.byref(cfg), 0, None, 0, C.byref(n))
        assert (rc == 0) == ok, flag
    full = ex.footprint(atm_config(ex, nx=8, nx1=8, implicit=9))
    part = ex.footprint(atm_config(ex, nx=8, nx1=8, implicit=1))
    none = ex.footprint(atm_config(ex, nx=8, nx1=8, implicit=0))
    assert none < part < full
