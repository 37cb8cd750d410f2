"""Shared configurations of the parity tests (small versions of the benchmark
configurations: shallow water exo2 and the 3D ideal-gas exo3 atmosphere)."""
import numpy as np

R_EARTH = 6371000.0


def interior(u, nx1):
    return u[:, :, 3:-3, 3:-3, 3:3 + nx1]


def rel_err(a, b):
    """max |a-b| / max |b| per conserved variable (axis 1)."""
    d = np.abs(a - b).reshape(a.shape[0], a.shape[1], -1).max(axis=(0, 2))
    s = np.abs(b).reshape(b.shape[0], b.shape[1], -1).max(axis=(0, 2))
    return d / np.where(s > 0, s, 1.0)


def sw_config(ex, nx=8, rsolver=0, nx1=3, coriolis=1):
    return ex.make_config(nx=nx, nx1=nx1, eos=ex.EOS_SHALLOW_YZ, rsolver=rsolver,
                          coriolis=coriolis, cfl=0.4, x1min=R_EARTH, x1max=R_EARTH + 1.0)


def atm_config(ex, nx=8, nx1=8, rsolver=3, implicit=9, ntracer=0, coriolis=0):
    return ex.make_config(nx=nx, nx1=nx1, eos=ex.EOS_IDEAL, rsolver=rsolver,
                          implicit_flag=implicit, ntracer=ntracer, coriolis=coriolis,
                          cfl_x1=1, cfl=0.4, x1min=R_EARTH, x1max=R_EARTH + 30e3,
                          gamma=1.4, grav=-9.81, omega=7.292e-5)
