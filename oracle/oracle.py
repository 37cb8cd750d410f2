"""ctypes wrapper of the oracle library (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

EOS_SHALLOW_YZ, EOS_IDEAL = 0, 1


class OrConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nx1", C.c_int), ("eos", C.c_int), ("rsolver", C.c_int),
        ("implicit_flag", C.c_int), ("ntracer", C.c_int), ("coriolis", C.c_int),
        ("cfl_x1", C.c_int), ("x1min", C.c_double), ("x1max", C.c_double),
        ("gamma", C.c_double), ("grav", C.c_double), ("omega", C.c_double),
        ("cfl", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_DP = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = load(LIB)
    return _lib


def load(path):
    """The oracle library at path with its entry points typed (lib() loads the
    built one; the mutation tests load copies built with a planted mistake)."""
    L = C.CDLL(os.path.abspath(path))
    V = C.c_void_p
    sig = {
        "or_weno5": (C.c_double, [C.c_double] * 5),
        "or_weno3": (C.c_double, [C.c_double] * 3),
        "or_exp": (C.c_double, [C.c_double]),
        "or_log": (C.c_double, [C.c_double]),
        "or_pow": (C.c_double, [C.c_double, C.c_double]),
        "or_create": (V, [C.POINTER(OrConfig)]),
        "or_destroy": (None, [V]),
        "or_set_cons": (None, [V, _DP]),
        "or_get_cons": (None, [V, _DP]),
        "or_get_prim": (None, [V, _DP]),
        "or_get_scalar_prim": (None, [V, _DP]),
        "or_set_scalars": (None, [V, _DP]),
        "or_get_scalars": (None, [V, _DP]),
        "or_get_geom2d": (None, [V, C.c_int, _DP]),
        "or_get_coord": (None, [V, C.c_int, C.c_int, _DP]),
        "or_cell_volume": (C.c_double, [V, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_face_area": (C.c_double, [V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_latlon": (None, [V, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                             C.POINTER(C.c_double)]),
        "or_w92_user_out_point": (None, [C.c_double, C.c_double, C.c_int, C.c_double,
                                         C.c_double, _DP]),
        "or_user_output": (None, [V, _DP]),
        "or_fill_ghosts": (None, [V]),
        "or_cons_to_prim": (None, [V]),
        "or_calc_fluxes": (None, [V]),
        "or_get_flux": (None, [V, C.c_int, _DP]),
        "or_new_dt": (C.c_double, [V]),
        "or_stage": (None, [V, C.c_int, C.c_double]),
        "or_step": (C.c_double, [V, C.c_int, C.c_double]),
        "or_set_threads": (None, [C.c_int]),
        "or_add_flux_divergence": (None, [V, C.c_double]),
        "or_add_flux_divergence_stage": (None, [V, C.c_int, C.c_double]),
        "or_implicit_correct": (None, [V, C.c_double]),
        "or_step_redo": (C.c_int, [V, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "or_state_valid": (C.c_int, [V]),
        "or_implicit_column": (None, [C.POINTER(OrConfig), C.c_int, _DP, _DP, _DP, _DP,
                                      C.c_double]),
        "or_flux_jacobian": (None, [C.c_double, _DP, _DP]),
        "or_eigensystem": (None, [_DP, C.c_double, C.c_double, _DP, _DP, _DP]),
        "or_roe_average": (None, [_DP, _DP, C.c_double, _DP]),
        "or_block_thomas": (None, [C.c_int, _DP, _DP, _DP, _DP, _DP]),
        "or_implicit_column3": (None, [C.POINTER(OrConfig), C.c_int, _DP, _DP, _DP, _DP,
                                       C.c_double]),
        "or_block_thomas3": (None, [C.c_int, _DP, _DP, _DP, _DP, _DP]),
        "or_vec_maps": (None, [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                               C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "or_decompose_column": (None, [V, C.c_int, C.c_int, C.c_int, _DP, _DP, _DP, _DP,
                                       _DP]),
        "or_riemann_face": (None, [C.POINTER(OrConfig), C.c_int, _DP, _DP, _DP,
                                   C.c_double, C.c_double, C.c_double, C.c_double]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    return L


def or_config(cfg) -> OrConfig:
    """Build the oracle config from any object with the exo_config fields."""
    d = cfg if isinstance(cfg, dict) else {f: getattr(cfg, f) for f, _ in OrConfig._fields_}
    return OrConfig(**{f: d[f] for f, _ in OrConfig._fields_})


class Oracle:
    def __init__(self, cfg, L=None):
        self.cfg = or_config(cfg)
        self.L = L if L is not None else lib()
        self.h = self.L.or_create(C.byref(self.cfg))
        self.nv = 4 if self.cfg.eos == EOS_SHALLOW_YZ else 5
        self.nc1, self.nc2 = self.cfg.nx1 + 6, self.cfg.nx + 6
        self.shape = (6, self.nv, self.nc2, self.nc2, self.nc1)
        self.sshape = (6, self.cfg.ntracer, self.nc2, self.nc2, self.nc1)

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, u, s=None):
        if s is not None and self.cfg.ntracer:
            self.L.or_set_scalars(self.h, np.ascontiguousarray(s, dtype=np.float64))
        self.L.or_set_cons(self.h, np.ascontiguousarray(u, dtype=np.float64))

    def download(self):
        u = np.zeros(self.shape)
        self.L.or_get_cons(self.h, u)
        s = None
        if self.cfg.ntracer:
            s = np.zeros(self.sshape)
            self.L.or_get_scalars(self.h, s)
        return u, s

    def prim(self):
        w = np.zeros(self.shape)
        self.L.or_get_prim(self.h, w)
        return w

    def scalar_prim(self):
        r = np.zeros(self.sshape)
        self.L.or_get_scalar_prim(self.h, r)
        return r

    def user_output(self):
        """lat, lon, U, V, sqrtg at every cell, (6, 5, nc3, nc2, nc1)"""
        out = np.zeros((6, 5, self.nc2, self.nc2, self.nc1))
        self.L.or_user_output(self.h, out)
        return out

    def calc_fluxes(self):
        self.L.or_calc_fluxes(self.h)

    def flux(self, d):
        n1 = self.nc1 + (d == 0)
        n2 = self.nc2 + (d == 1)
        n3 = self.nc2 + (d == 2)
        out = np.zeros((6, self.nv, n3, n2, n1))
        self.L.or_get_flux(self.h, d, out)
        return out

    def step(self, nsteps=1, dt=0.0):
        return self.L.or_step(self.h, nsteps, dt)

    def add_flux_divergence(self, wght, stage=1):
        if stage == 1:
            self.L.or_add_flux_divergence(self.h, wght)
        else:
            self.L.or_add_flux_divergence_stage(self.h, stage, wght)

    def implicit_correct(self, dt_stage):
        self.L.or_implicit_correct(self.h, dt_stage)

    def fill_ghosts(self):
        self.L.or_fill_ghosts(self.h)
        self.L.or_cons_to_prim(self.h)

    def step_redo(self, nsteps=1, max_redo=0):
        d = C.c_double()
        r = self.L.or_step_redo(self.h, nsteps, max_redo, C.byref(d))
        if r < 0:
            raise RuntimeError("Too many attempts to redo the step")
        return d.value, r

    def state_valid(self):
        return bool(self.L.or_state_valid(self.h))

    def stage(self, stage, dt):
        self.L.or_stage(self.h, stage, dt)

    def new_dt(self):
        return self.L.or_new_dt(self.h)
