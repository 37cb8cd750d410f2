"""Per-MeshBlock NetCDF output (host side, no GPU): the files that
exo_write_netcdf_fields writes, read back with scipy's independent NetCDF-3
reader, against the layout of src/outputs/netcdf.cpp:34-469 -- file names and
block numbering, dimensions, coordinate variables and their attributes, the
fields' (time, x1, x2, x3) order and the per-block slicing of the panels."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from helpers import atm_config, sw_config


def _coords(ex, cfg):
    nc1, nc = cfg.nx1 + 6, cfg.nx + 6
    t3 = ex.describe(cfg, 3)
    x1f, x1v = t3[:nc1 + 1], t3[nc1 + 1:2 * nc1 + 1]
    x2f = ex.describe(cfg, 4).reshape(6, nc + 1)
    x3f = ex.describe(cfg, 5).reshape(6, nc + 1)
    return x1f, x1v, x2f, 0.5 * (x2f[:, 1:] + x2f[:, :-1]), x3f, 0.5 * (x3f[:, 1:] + x3f[:, :-1])


def _labelled(ex, cfg, var, ghost):
    # every value names its (panel, field, i, j, k) and is exact in float32
    shape = ex.output_shape(cfg, var, ghost)
    return np.arange(np.prod(shape), dtype=np.float64).reshape(shape).astype(np.float32)


def _zorder(z, bps):
    sk = sj = 0
    b = 0
    while (1 << b) < bps:
        sj |= ((z >> (2 * b)) & 1) << b
        sk |= ((z >> (2 * b + 1)) & 1) << b
        b += 1
    return sk, sj


@pytest.mark.parametrize("which,var,bps,ghost", [
    ("sw", "prim", 1, False), ("sw", "uov", 2, True), ("atm", "prim", 2, False),
    ("atm", "cons", 1, True), ("sw1", "prim", 1, False), ("sw1", "cons", 2, True)])
def test_block_files(ex, tmp_path, which, var, bps, ghost):
    if which == "atm":
        cfg = atm_config(ex, nx=8, nx1=4, ntracer=2)
    else:
        cfg = sw_config(ex, nx=8, nx1=1 if which == "sw1" else 3)
    names = ex.output_names(cfg, var)
    fields = _labelled(ex, cfg, var, ghost)
    base = str(tmp_path / "W92")
    ex.write_netcdf_fields(cfg, base, var, fields, time=3600.5, file_number=7,
                           blocks_per_side=bps, ghost_zones=ghost)
    files = ex.output_file_names(base, var, 7, None, bps)
    assert sorted(os.listdir(tmp_path)) == sorted(os.path.basename(f) for f in files)
    assert files[0].endswith(f"W92.block0.{var}.00007.nc")
    x1f, x1v, x2f, x2v, x3f, x3v = _coords(ex, cfg)
    G = 3 if ghost else 0
    N, nb = cfg.nx, cfg.nx // bps
    c1, c2 = cfg.nx1 + 2 * G, nb + 2 * G
    i0 = 3 - G
    for gid, path in enumerate(files):
        p, z = divmod(gid, bps * bps)
        sk, sj = _zorder(z, bps)
        lx2, lx3 = (p % 2) * bps + sj, (p // 2) * bps + sk
        j0, k0 = 3 - G + sj * nb, 3 - G + sk * nb
        with open(path, "rb") as fh:
            assert fh.read(4) == b"CDF\x02"
        with netcdf_file(path, "r", mmap=False) as f:
            assert f.NumFilesInSet == 6 * bps * bps
            dims = dict(f.dimensions)
            want = {"time": None, "x1": c1, "x2": c2, "x2f": c2 + 1, "x3": c2, "x3f": c2 + 1}
            if c1 > 1:
                want["x1f"] = c1 + 1
            assert dims == want
            assert list(f.variables)[:1] == ["time"]
            assert [v for v in f.variables if v not in want] == names
            t = f.variables["time"]
            assert t.data.tolist() == [np.float32(3600.5)]
            assert (t.axis, t.units, t.long_name) == (b"T", b"s", b"time")
            for ax, letter, cells, loc, mesh_n, cv, fc, first in (
                    ("x1", b"Z", c1, 0, cfg.nx1, x1v, x1f, i0),
                    ("x2", b"Y", c2, lx2, 2 * N, x2v[p], x2f[p], j0),
                    ("x3", b"X", c2, lx3, 3 * N, x3v[p], x3f[p], k0)):
                v = f.variables[ax]
                assert v.axis == letter and v.units == b"m"
                assert v.long_name == f"{letter.decode()}-coordinate at cell center".encode()
                assert list(v.domain_decomposition) == [1, mesh_n, cells * loc + 1,
                                                        cells * (loc + 1)]
                assert np.array_equal(v.data, cv[first:first + cells].astype(np.float32))
                if cells > 1:
                    w = f.variables[ax + "f"]
                    assert w.long_name == f"{letter.decode()}-coordinate at cell face".encode()
                    assert list(w.domain_decomposition) == [0, mesh_n, cells * loc,
                                                            cells * (loc + 1)]
                    assert np.array_equal(w.data, fc[first:first + cells + 1].astype(np.float32))
            for q, n in enumerate(names):
                v = f.variables[n]
                assert v.dimensions == ("time", "x1", "x2", "x3")
                assert v.data.dtype == np.dtype(">f4")
                want_slab = fields[p, q, :, sj * nb:sj * nb + c2, sk * nb:sk * nb + c2]
                assert np.array_equal(v.data[0], want_slab), (gid, n)
                attrs = {k: getattr(v, k) for k in ("units", "long_name") if hasattr(v, k)}
                if n == "rho":
                    assert attrs == {"units": b"kg/m^3", "long_name": b"density"}
                elif n == "press":
                    assert attrs == {"units": b"pa", "long_name": b"pressure"}
                elif n.startswith("vel"):
                    assert attrs == {"units": b"m/s", "long_name": b"velocity"}
                else:
                    assert attrs == {}


def test_output_names(ex):
    sw = sw_config(ex)
    atm = atm_config(ex, ntracer=2)
    assert ex.output_names(sw, "prim") == ["rho", "vel1", "vel2", "vel3"]
    assert ex.output_names(sw, "cons") == ["dens", "mom1", "mom2", "mom3"]
    assert ex.output_names(atm, "prim") == ["rho", "press", "vel1", "vel2", "vel3", "r0", "r1"]
    assert ex.output_names(atm, "cons") == ["dens", "Etot", "mom1", "mom2", "mom3", "s0", "s1"]
    assert ex.output_names(atm, "uov") == ["lat", "lon", "U", "V", "sqrtg"]
    with pytest.raises(ex.ExoError):
        ex.output_names(sw, "rad")


def test_write_errors(ex, tmp_path):
    cfg = sw_config(ex, nx=8)
    f = np.zeros(ex.output_shape(cfg, "prim"), dtype=np.float32)
    for bps in (0, 3, 16):   # not a power of two dividing N
        with pytest.raises(ex.ExoError):
            ex.write_netcdf_fields(cfg, str(tmp_path / "a"), "prim", f, blocks_per_side=bps)
    with pytest.raises(ex.ExoError):
        ex.write_netcdf_fields(cfg, str(tmp_path / "a"), "prim", f[:, :3])
    rc = ex.lib().exo_write_netcdf_fields(C.byref(cfg), str(tmp_path / "missing" / "a").encode(),
                                          b"prim", None, 0, 0.0, 1, 0,
                                          f.ctypes.data_as(C.c_void_p))
    assert rc == -6   # EXO_ERR_IO
    assert not os.path.exists(tmp_path / "missing")


class _Recorder:
    def __init__(self):
        self.calls = []

    def write_netcdf(self, basename, variable, time, file_number, file_id, bps, ghost):
        self.calls.append((variable, time, file_number))
        return [f"{basename}.block0.{variable}.{file_number:05d}.nc"]


def test_output_cadence(ex):
    # Athena's output cadence: a write whenever time >= next_time, then
    # file_number += 1 and next_time += dt (netcdf.cpp:459-466)
    out = ex.NetcdfOutput("W92", "uov", dt=3600.0)
    rec = _Recorder()
    for t in (0.0, 1800.0, 3599.0, 3600.0, 5000.0, 7300.0, 7400.0):
        out.maybe_write(rec, t)
    assert rec.calls == [("uov", 0.0, 0), ("uov", 3600.0, 1), ("uov", 7300.0, 2)]
    assert out.file_number == 3 and out.next_time == 10800.0
