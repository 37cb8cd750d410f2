"""MI355X-native cubed-sphere finite-volume hydro update (ExoCubed's hot path).

The compute path is the HIP library ``libexocubed.so`` built from
``exocubed_amd/csrc`` for gfx950; this module only loads it and wraps its C ABI
(``include/exocubed.h``).  There is no CPU fallback: if the library is missing
the import of :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexocubed.so")

EOS_SHALLOW_YZ, EOS_IDEAL = 0, 1
RS_ROE_SW, RS_LMARS_SW, RS_HLLC, RS_LMARS = 0, 1, 2, 3
COR_NONE, COR_FORCING2, COR_3D = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
BATCH_GRAPH = 1
KERNELS = ("prim", "ghost", "flux_x1", "edge_states", "flux_x2", "flux_x3", "update",
           "implicit", "dt", "output")
# launch kinds of the persistent-width tables (exo_tune / exo_batch_tune)
LAUNCH_KINDS = ("prim", "ghost_u", "ghost_s", "x1", "x1_mix", "edge", "edge_mix", "face",
                "face_mix", "update", "update_mix", "implicit", "add_du", "cells", "one", "dt",
                "edge_flux", "face_mix_sw", "prim_rest", "prim_ghost", "implicit3",
                "implicit_avg", "update_rs", "implicit3_avg",
                "add_du_avg", "face2_update", "face_x2", "face2_update_s",
                "face2_update_mix", "face2_update_mix_s")


def _tuning(caps):
    n = len(LAUNCH_KINDS)
    out = {k: int(c) for k, c in zip(LAUNCH_KINDS, caps[:n])}
    out.update({k + "_shape": int(c) for k, c in zip(LAUNCH_KINDS, caps[n:]) if c})
    return out


class Config(C.Structure):
    """Mirror of ``exo_config`` (include/exocubed.h)."""

    _fields_ = [
        ("nx", C.c_int), ("nx1", C.c_int), ("eos", C.c_int), ("rsolver", C.c_int),
        ("implicit_flag", C.c_int), ("ntracer", C.c_int), ("coriolis", C.c_int),
        ("cfl_x1", C.c_int), ("x1min", C.c_double), ("x1max", C.c_double),
        ("gamma", C.c_double), ("grav", C.c_double), ("omega", C.c_double),
        ("cfl", C.c_double), ("precision", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def make_config(**kw) -> Config:
    base = dict(nx=8, nx1=3, eos=EOS_SHALLOW_YZ, rsolver=RS_ROE_SW, implicit_flag=0,
                ntracer=0, coriolis=COR_NONE, cfl_x1=0, x1min=6371000.0, x1max=6371001.0,
                gamma=1.4, grav=0.0, omega=7.292e-5, cfl=0.4, precision=PREC_FP64)
    base.update(kw)
    return Config(**base)


_lib = None
_DP = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    sig = {
        "exo_create": (C.c_int, [C.POINTER(Config), C.POINTER(H)]),
        "exo_destroy": (C.c_int, [H]),
        "exo_last_error": (C.c_char_p, [H]),
        "exo_nvar": (C.c_int, [H]),
        "exo_state_doubles": (C.c_size_t, [H]),
        "exo_upload_state": (C.c_int, [H, C.c_void_p, C.c_void_p]),
        "exo_download_state": (C.c_int, [H, C.c_void_p, C.c_void_p]),
        "exo_download_prim": (C.c_int, [H, _DP]),
        "exo_calc_fluxes": (C.c_int, [H]),
        "exo_download_flux": (C.c_int, [H, C.c_int, _DP]),
        "exo_add_flux_divergence": (C.c_int, [H, C.c_double]),
        "exo_integrate_hydro": (C.c_int, [H, C.c_int, C.c_double]),
        "exo_fill_ghosts": (C.c_int, [H]),
        "exo_implicit_correct": (C.c_int, [H, C.c_double]),
        "exo_new_dt": (C.c_int, [H, C.POINTER(C.c_double)]),
        "exo_stage": (C.c_int, [H, C.c_int, C.c_double]),
        "exo_step": (C.c_int, [H, C.c_int, C.c_double, C.POINTER(C.c_double)]),
        "exo_sync": (C.c_int, [H]),
        "exo_profile_begin": (C.c_int, [H]),
        "exo_profile_end": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int]),
        "exo_init_state": (C.c_int, [H, C.c_int, C.c_uint64, C.c_void_p, C.c_int, _DP,
                                     C.c_void_p]),
        "exo_riemann_face": (C.c_int, [H, C.c_int, C.c_int, C.c_int, C.c_int, _DP, _DP, _DP]),
        "exo_implicit_column": (C.c_int, [H, C.c_int, _DP, _DP, _DP, _DP, C.c_double]),
        "exo_implicit_columns_noalloc": (C.c_int, [H, C.c_int, C.c_int, C.c_void_p,
                                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_double]),
        "exo_weno5_batch": (C.c_int, [H, _DP, _DP, C.c_int]),
        "exo_math_batch": (C.c_int, [H, C.c_int, _DP, C.c_double, _DP, C.c_int]),
        "exo_init_state_cfg": (C.c_int, [C.POINTER(Config), C.c_int, C.c_uint64, C.c_void_p,
                                         C.c_int, _DP, C.c_void_p]),
        "exo_describe": (C.c_int, [C.POINTER(Config), C.c_int, C.c_void_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
        "exo_conserved_sums": (C.c_int, [H, _DP]),
        "exo_tune": (C.c_int, [H, C.c_int, C.c_void_p, C.c_int]),
        "exo_batch_create": (C.c_int, [C.POINTER(H), C.c_int, C.c_int, C.POINTER(H)]),
        "exo_batch_destroy": (C.c_int, [H]),
        "exo_batch_last_error": (C.c_char_p, [H]),
        "exo_batch_step": (C.c_int, [H, C.c_int, C.c_double]),
        "exo_batch_sums": (C.c_int, [H, _DP]),
        "exo_batch_tune": (C.c_int, [H, C.c_int, C.c_void_p, C.c_int]),
        "exo_batch_set_caps": (C.c_int, [H, C.c_void_p, C.c_int]),
        "exo_batch_profile_begin": (C.c_int, [H]),
        "exo_batch_profile_end": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int]),
        "exo_batch_sync": (C.c_int, [H]),
        "exo_footprint": (C.c_size_t, [C.POINTER(Config)]),
        "exo_server_create": (C.c_int, [C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(H)]),
        "exo_server_submit": (C.c_int, [H, C.POINTER(Config), C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_uint64, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_uint64)]),
        "exo_server_wait": (C.c_int, [H, C.c_uint64, C.c_double, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
        "exo_server_stats": (C.c_int, [H, C.c_void_p, C.c_int]),
        "exo_server_last_error": (C.c_char_p, [H]),
        "exo_server_destroy": (C.c_int, [H]),
        "exo_state_valid": (C.c_int, [H, C.POINTER(C.c_int)]),
        "exo_save_states": (C.c_int, [H]),
        "exo_load_states": (C.c_int, [H]),
        "exo_step_redo": (C.c_int, [H, C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]),
        "exo_get_dt_state": (C.c_int, [H, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "exo_set_dt_state": (C.c_int, [H, C.c_double, C.c_double]),
        "exo_save_checkpoint": (C.c_int, [H, C.c_char_p, C.c_double, C.c_int64]),
        "exo_load_checkpoint": (C.c_int, [H, C.c_char_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int64)]),
        "exo_output_names": (C.c_int, [C.POINTER(Config), C.c_char_p, C.c_char_p, C.c_size_t,
                                       C.POINTER(C.c_int)]),
        "exo_output_fields": (C.c_int, [H, C.c_char_p, C.c_int, C.c_void_p, C.c_size_t]),
        "exo_write_netcdf": (C.c_int, [H, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int,
                                       C.c_double, C.c_int, C.c_int]),
        "exo_write_netcdf_fields": (C.c_int, [C.POINTER(Config), C.c_char_p, C.c_char_p,
                                              C.c_char_p, C.c_int, C.c_double, C.c_int,
                                              C.c_int, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def init_state(cfg: Config, kind: int, seed: int, params=()):
    """Seeded synthetic initial state (host only): conserved u and scalars s."""
    L = lib()
    nv = 4 if cfg.eos == EOS_SHALLOW_YZ else 5
    nc1, nc2 = cfg.nx1 + 6, cfg.nx + 6
    u = np.zeros((6, nv, nc2, nc2, nc1))
    s = np.zeros((6, cfg.ntracer, nc2, nc2, nc1)) if cfg.ntracer else None
    p = np.ascontiguousarray(params, dtype=np.float64)
    rc = L.exo_init_state_cfg(C.byref(cfg), kind, seed,
                              p.ctypes.data_as(C.c_void_p) if len(p) else None, len(p), u,
                              None if s is None else s.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"exo_init_state_cfg failed ({rc})")
    return u, s


def describe(cfg: Config, which: int) -> np.ndarray:
    """Set-up tables of an instance, computed on the host (see exo_describe)."""
    L = lib()
    n = C.c_size_t()
    if L.exo_describe(C.byref(cfg), which, None, 0, C.byref(n)) != 0:
        raise RuntimeError("exo_describe failed")
    if which == 0:
        dt = GHOST_DTYPE
    elif which == 1:
        dt = EDGE_DTYPE
    else:
        dt = np.float64
    out = np.zeros(n.value, dtype=dt)
    if L.exo_describe(C.byref(cfg), which, out.ctypes.data_as(C.c_void_p), out.nbytes,
                      C.byref(n)) != 0:
        raise RuntimeError("exo_describe failed")
    return out


def output_names(cfg: Config, variable: str):
    """Field names of an output variable ("prim", "cons", "uov") in file order."""
    L = lib()
    buf = C.create_string_buffer(1024)
    n = C.c_int()
    if L.exo_output_names(C.byref(cfg), variable.encode(), buf, len(buf), C.byref(n)) != 0:
        raise ExoError(f"unknown output variable {variable!r}")
    return buf.value.decode().split(",")


def output_shape(cfg: Config, variable: str, ghost_zones=False):
    g = 6 if ghost_zones else 0
    return (6, len(output_names(cfg, variable)), cfg.nx1 + g, cfg.nx + g, cfg.nx + g)


def write_netcdf_fields(cfg: Config, basename, variable, fields, time=0.0, file_number=0,
                        file_id=None, blocks_per_side=1, ghost_zones=False):
    """Per-block NetCDF files from host fields in the output layout (no GPU)."""
    f = np.ascontiguousarray(fields, dtype=np.float32)
    if f.shape != output_shape(cfg, variable, ghost_zones):
        raise ExoError(f"fields of shape {f.shape}, expected "
                       f"{output_shape(cfg, variable, ghost_zones)}")
    rc = lib().exo_write_netcdf_fields(C.byref(cfg), os.fsencode(basename), variable.encode(),
                                       None if file_id is None else file_id.encode(),
                                       file_number, time, blocks_per_side, int(ghost_zones),
                                       f.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ExoError(f"exo_write_netcdf_fields failed ({rc})")


def output_file_names(basename, variable, file_number=0, file_id=None, blocks_per_side=1):
    """Names of the files one output writes, in gid order."""
    fid = variable if file_id is None else file_id
    return [f"{basename}.block{g}.{fid}.{file_number:05d}.nc"
            for g in range(6 * blocks_per_side * blocks_per_side)]


GHOST_DTYPE = np.dtype([(f, np.int32) for f in
                        ("tgt_panel", "tgt_k", "tgt_j", "src_panel", "along_x2", "src_row",
                         "src_a0", "src_block", "tgt_block", "pad")] +
                       [(f, np.float64) for f in
                        ("A", "B", "D", "src_cy", "src_cz", "tgt_cy", "tgt_cz", "sth_s",
                         "cth_s", "sth_t", "cth_t")])
EDGE_DTYPE = np.dtype([(f, np.int32) for f in
                       ("src_panel", "src_dir", "src_k", "src_j", "tgt_panel", "tgt_dir",
                        "tgt_a", "src_block", "tgt_block", "pad")] +
                      [(f, np.float64) for f in ("x2s", "x3s", "x2d", "x3d")])


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class ExoError(RuntimeError):
    pass


class Solver:
    """One cubed-sphere instance resident on the GPU (an ``exo_handle``)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.exo_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise ExoError(f"exo_create failed ({rc})")
        self.h = h
        self.nv = 4 if cfg.eos == EOS_SHALLOW_YZ else 5
        self.nc1, self.nc2 = cfg.nx1 + 6, cfg.nx + 6
        self.shape = (6, self.nv, self.nc2, self.nc2, self.nc1)
        self.sshape = (6, cfg.ntracer, self.nc2, self.nc2, self.nc1)

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.exo_last_error(self.h).decode()
            raise ExoError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.exo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_state(self, kind, seed, params=()):
        return init_state(self.cfg, kind, seed, params)

    def upload(self, u, s=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        s = None if s is None else np.ascontiguousarray(s, dtype=np.float64)
        self._chk(self.L.exo_upload_state(self.h, _ptr(u), _ptr(s)), "exo_upload_state")

    def download(self):
        u = np.zeros(self.shape)
        s = np.zeros(self.sshape) if self.cfg.ntracer else None
        self._chk(self.L.exo_download_state(self.h, _ptr(u), _ptr(s)), "exo_download_state")
        return u, s

    def prim(self):
        w = np.zeros(self.shape)
        self._chk(self.L.exo_download_prim(self.h, w), "exo_download_prim")
        return w

    def calc_fluxes(self):
        self._chk(self.L.exo_calc_fluxes(self.h), "exo_calc_fluxes")

    def flux(self, d):
        n1 = self.nc1 + (d == 0)
        n2 = self.nc2 + (d == 1)
        n3 = self.nc2 + (d == 2)
        out = np.zeros((6, self.nv, n3, n2, n1))
        self._chk(self.L.exo_download_flux(self.h, d, out), "exo_download_flux")
        return out

    def add_flux_divergence(self, wght):
        """-div F + sources times wght (stage-1 weights) into u, or into the
        tendency du for the ideal gas"""
        self._chk(self.L.exo_add_flux_divergence(self.h, wght), "exo_add_flux_divergence")

    def integrate_hydro(self, stage, dt):
        """IntegrateHydro of SSPRK3 stage 1..3: the stage's weighted average of
        u, then -div F + sources times beta(stage) dt into u (shallow water) or
        the tendency du (ideal gas)"""
        self._chk(self.L.exo_integrate_hydro(self.h, stage, dt), "exo_integrate_hydro")

    def fill_ghosts(self):
        self._chk(self.L.exo_fill_ghosts(self.h), "exo_fill_ghosts")

    def implicit_correct(self, dt_stage):
        self._chk(self.L.exo_implicit_correct(self.h, dt_stage), "exo_implicit_correct")

    def step(self, nsteps=1, dt=0.0):
        d = C.c_double()
        self._chk(self.L.exo_step(self.h, nsteps, dt, C.byref(d)), "exo_step")
        return d.value

    def stage(self, stage, dt):
        self._chk(self.L.exo_stage(self.h, stage, dt), "exo_stage")

    def new_dt(self):
        d = C.c_double()
        self._chk(self.L.exo_new_dt(self.h, C.byref(d)), "exo_new_dt")
        return d.value

    KERNELS = KERNELS

    def profile_begin(self):
        self._chk(self.L.exo_profile_begin(self.h), "exo_profile_begin")

    def profile_end(self):
        ms = np.zeros(len(self.KERNELS))
        cnt = np.zeros(len(self.KERNELS), dtype=np.int32)
        self._chk(self.L.exo_profile_end(self.h, ms.ctypes.data_as(C.c_void_p),
                                         cnt.ctypes.data_as(C.c_void_p), len(ms)),
                  "exo_profile_end")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNELS)}

    def sync(self):
        self._chk(self.L.exo_sync(self.h), "exo_sync")

    def state_valid(self):
        v = C.c_int()
        self._chk(self.L.exo_state_valid(self.h, C.byref(v)), "exo_state_valid")
        return bool(v.value)

    def save_states(self):
        self._chk(self.L.exo_save_states(self.h), "exo_save_states")

    def load_states(self):
        self._chk(self.L.exo_load_states(self.h), "exo_load_states")

    def step_redo(self, nsteps=1, max_redo=0):
        """nsteps accepted steps, an invalid one undone and repeated with half
        the time step; returns (last dt, steps redone)"""
        d, r = C.c_double(), C.c_int()
        self._chk(self.L.exo_step_redo(self.h, nsteps, max_redo, C.byref(d), C.byref(r)),
                  "exo_step_redo")
        return d.value, r.value

    def dt_state(self):
        """(dt of the last step, previous dt bounding the next one's growth)"""
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.exo_get_dt_state(self.h, C.byref(a), C.byref(b)), "exo_get_dt_state")
        return a.value, b.value

    def set_dt_state(self, dt, dt_prev):
        self._chk(self.L.exo_set_dt_state(self.h, dt, dt_prev), "exo_set_dt_state")

    def save_checkpoint(self, path, time=0.0, ncycle=0):
        self._chk(self.L.exo_save_checkpoint(self.h, os.fsencode(path), time, ncycle),
                  "exo_save_checkpoint")

    def load_checkpoint(self, path):
        """restores a checkpoint; returns (time, ncycle)"""
        t, n = C.c_double(), C.c_int64()
        self._chk(self.L.exo_load_checkpoint(self.h, os.fsencode(path), C.byref(t), C.byref(n)),
                  "exo_load_checkpoint")
        return t.value, n.value

    def output_fields(self, variable="prim", ghost_zones=False):
        """Float fields of an output variable, [6][nvar][x1][x2][x3] (x3 fastest)."""
        out = np.zeros(output_shape(self.cfg, variable, ghost_zones), dtype=np.float32)
        self._chk(self.L.exo_output_fields(self.h, variable.encode(), int(ghost_zones),
                                           out.ctypes.data_as(C.c_void_p), out.size),
                  "exo_output_fields")
        return out

    def write_netcdf(self, basename, variable="prim", time=0.0, file_number=0, file_id=None,
                     blocks_per_side=1, ghost_zones=False):
        """The reference's per-block NetCDF files of the current state."""
        self._chk(self.L.exo_write_netcdf(self.h, os.fsencode(basename), variable.encode(),
                                          None if file_id is None else file_id.encode(),
                                          file_number, time, blocks_per_side, int(ghost_zones)),
                  "exo_write_netcdf")
        return output_file_names(basename, variable, file_number, file_id, blocks_per_side)

    def sums(self):
        """Compensated FP64 totals sum(u_n * volume) per conserved variable."""
        out = np.zeros(5)
        self._chk(self.L.exo_conserved_sums(self.h, out), "exo_conserved_sums")
        return out[:self.nv]

    def tune(self, reps=3):
        """tuned persistent widths per launch kind, and "<kind>_shape" block
        shapes (waves per x1 block, levels per update thread) where not the
        default"""
        caps = np.zeros(2 * len(LAUNCH_KINDS), dtype=np.int32)
        self._chk(self.L.exo_tune(self.h, reps, caps.ctypes.data_as(C.c_void_p), len(caps)),
                  "exo_tune")
        return _tuning(caps)


class Batch:
    """Several Solver instances advanced together (``exo_batch``): one launch
    per phase over all members, optionally replayed from a captured HIP graph.
    The members must stay alive and idle while the batch is used."""

    def __init__(self, members, graph=True):
        self.L = lib()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        rc = self.L.exo_batch_create(arr, len(self.members), BATCH_GRAPH if graph else 0,
                                     C.byref(h))
        if rc != 0:
            raise ExoError(f"exo_batch_create failed ({rc})")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.exo_batch_last_error(self.h).decode()
            raise ExoError(f"{what} failed ({rc}): {msg}")

    def step(self, nsteps=1, dt=0.0):
        self._chk(self.L.exo_batch_step(self.h, nsteps, dt), "exo_batch_step")

    def sums(self):
        out = np.zeros(len(self.members) * 5)
        self._chk(self.L.exo_batch_sums(self.h, out), "exo_batch_sums")
        return out.reshape(len(self.members), 5)

    def tune(self, reps=3):
        caps = np.zeros(2 * len(LAUNCH_KINDS), dtype=np.int32)
        self._chk(self.L.exo_batch_tune(self.h, reps, caps.ctypes.data_as(C.c_void_p),
                                        len(caps)), "exo_batch_tune")
        return _tuning(caps)

    def set_caps(self, caps):
        arr = np.array([caps.get(k, 0) for k in LAUNCH_KINDS] +
                       [caps.get(k + "_shape", 0) for k in LAUNCH_KINDS], dtype=np.int32)
        self._chk(self.L.exo_batch_set_caps(self.h, arr.ctypes.data_as(C.c_void_p), len(arr)),
                  "exo_batch_set_caps")

    def profile_begin(self):
        self._chk(self.L.exo_batch_profile_begin(self.h), "exo_batch_profile_begin")

    def profile_end(self):
        ms = np.zeros(len(KERNELS))
        cnt = np.zeros(len(KERNELS), dtype=np.int32)
        self._chk(self.L.exo_batch_profile_end(self.h, ms.ctypes.data_as(C.c_void_p),
                                               cnt.ctypes.data_as(C.c_void_p), len(ms)),
                  "exo_batch_profile_end")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNELS)}

    def sync(self):
        self._chk(self.L.exo_batch_sync(self.h), "exo_batch_sync")

    def close(self):
        if getattr(self, "h", None):
            self.L.exo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(cfg: Config) -> int:
    """Device bytes an instance of this configuration allocates (host only)."""
    return int(lib().exo_footprint(C.byref(cfg)))


SERVER_INFO = ("latency_s", "queue_s", "run_s", "dt_last", "cells", "steps", "mass0", "mass1",
               "total0", "total1")
SERVER_STATS = ("completed", "cell_updates", "span_s", "busy_s", "batches", "chunks",
                "peak_active", "peak_bytes", "budget_bytes")


class Server:
    """Continuous-batching request loop of the library (``exo_server``)."""

    def __init__(self, mem_budget=0, max_active=64, chunk_steps=20, graph=True, prep_threads=4):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.exo_server_create(int(mem_budget), max_active, chunk_steps,
                                      BATCH_GRAPH if graph else 0, prep_threads, C.byref(h))
        if rc != 0:
            raise ExoError(f"exo_server_create failed ({rc})")
        self.h = h
        self.cfgs = {}

    def _err(self, what, rc):
        msg = self.L.exo_server_last_error(self.h).decode()
        return ExoError(f"{what} failed ({rc}): {msg}")

    def submit(self, cfg, nsteps, u0=None, s0=None, kind=-1, seed=0, params=(),
               keep_state=True):
        """Queue a forward run; returns its id."""
        rid = C.c_uint64()
        p = np.ascontiguousarray(params, dtype=np.float64)
        u0 = None if u0 is None else np.ascontiguousarray(u0, dtype=np.float64)
        s0 = None if s0 is None else np.ascontiguousarray(s0, dtype=np.float64)
        rc = self.L.exo_server_submit(self.h, C.byref(cfg), _ptr(u0), _ptr(s0), kind, seed,
                                      p.ctypes.data_as(C.c_void_p) if len(p) else None, len(p),
                                      nsteps, 1 if keep_state else 0, C.byref(rid))
        if rc != 0:
            raise self._err("exo_server_submit", rc)
        self.cfgs[rid.value] = (Config(**cfg.as_dict()), keep_state)
        return rid.value

    def wait(self, rid, timeout=0.0):
        """Final state (u, s or None, None when not kept) and the info dict."""
        cfg, keep = self.cfgs.pop(rid)
        nv = 4 if cfg.eos == EOS_SHALLOW_YZ else 5
        shape = (6, nv, cfg.nx + 6, cfg.nx + 6, cfg.nx1 + 6)
        u = np.zeros(shape) if keep else None
        s = np.zeros((6, cfg.ntracer) + shape[2:]) if keep and cfg.ntracer else None
        info = np.zeros(len(SERVER_INFO))
        rc = self.L.exo_server_wait(self.h, rid, timeout, _ptr(u), _ptr(s),
                                    info.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise self._err("exo_server_wait", rc)
        return u, s, dict(zip(SERVER_INFO, info.tolist()))

    def stats(self):
        out = np.zeros(len(SERVER_STATS))
        self.L.exo_server_stats(self.h, out.ctypes.data_as(C.c_void_p), len(out))
        return dict(zip(SERVER_STATS, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            self.L.exo_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NetcdfOutput:
    """One ``<outputN>`` block of type netcdf: writes every ``dt`` of simulated
    time, numbering the files from ``file_number`` (Athena's OutputType
    cadence; src/outputs/netcdf.cpp:459-466 advances file_number and
    next_time after each write)."""

    def __init__(self, basename, variable="prim", dt=0.0, file_id=None, ghost_zones=False,
                 blocks_per_side=1, file_number=0, next_time=0.0):
        self.basename, self.variable, self.dt = basename, variable, dt
        self.file_id, self.ghost_zones, self.blocks_per_side = file_id, ghost_zones, \
            blocks_per_side
        self.file_number, self.next_time = file_number, next_time

    def due(self, time):
        return time >= self.next_time

    def write(self, solver: Solver, time):
        files = solver.write_netcdf(self.basename, self.variable, time, self.file_number,
                                    self.file_id, self.blocks_per_side, self.ghost_zones)
        self.file_number += 1
        self.next_time += self.dt
        return files

    def maybe_write(self, solver: Solver, time):
        return self.write(solver, time) if self.due(time) else []
