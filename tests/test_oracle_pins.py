"""Pins of the oracle (CPU).  Each check is a golden vector from the reference,
a closed form, a library routine or a mathematical invariant chosen so that a
plausible restatement mistake (dropped term, wrong sign or index, transposed
operand) fails it."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

from helpers import R_EARTH, atm_config, interior, sw_config


# --- golden vectors: tests/test_weno.cpp:11-29 of the reference -------------
def test_weno_golden_vectors(orc):
    L = orc.lib()
    assert L.or_weno3(1.0, 2.0, 3.0) == pytest.approx(1.5, abs=1e-10)
    assert L.or_weno5(1.0, 2.0, 3.0, 4.0, 5.0) == pytest.approx(2.5000000000000004, abs=1e-10)
    # the last bit pins the summation order of the linear weights
    assert L.or_weno5(1.0, 2.0, 3.0, 4.0, 5.0) == 2.5000000000000004


def test_weno_invariants(orc):
    L = orc.lib()
    # constants are kept to the weights' rounding (0.3 + 0.6 + 0.1 = 1 - 1 ulp)
    assert abs(L.or_weno5(7.0, 7.0, 7.0, 7.0, 7.0) - 7.0) <= 2 * np.spacing(7.0)
    a = [0.3, -1.2, 2.5, 0.7, 4.1]
    r = L.or_weno5(*a)
    assert np.isfinite(r)
    # linear data: exact face value (a + b*x at x = -1/2 about cell 0)
    for s in (1e-3, 1.0, 1e5):
        v = [10.0 + s * x for x in (-2, -1, 0, 1, 2)]
        assert L.or_weno5(*v) == pytest.approx(10.0 - 0.5 * s, rel=1e-14, abs=1e-14 * s)
    # fifth-order convergence for smooth data: cell averages of sin(x)
    errs = []
    for h in (0.2, 0.1, 0.05):
        avg = [(math.cos((m - 0.5) * h) - math.cos((m + 0.5) * h)) / h for m in (-2, -1, 0, 1, 2)]
        errs.append(abs(L.or_weno5(*avg) - math.sin(-0.5 * h)))
    order = math.log(errs[1] / errs[2]) / math.log(2)
    assert order > 4.5, errs
    # a discontinuity: the stencil across the jump is down-weighted (no new extremum)
    v = L.or_weno5(0.0, 0.0, 0.0, 1.0, 1.0)
    assert -1e-6 < v < 0.05


# --- deterministic transcendental functions vs libm --------------------------
def _ulps(a, b):
    return abs(a - b) / np.spacing(abs(b))


def test_exp_log_pow_against_libm(orc):
    L = orc.lib()
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(-40, 40, 5000))
    assert max(_ulps(L.or_log(x), math.log(x)) for x in xs if abs(math.log(x)) > 1e-2) <= 2
    ys = rng.uniform(-50, 50, 5000)
    assert max(_ulps(L.or_exp(y), math.exp(y)) for y in ys) <= 2
    ps = rng.uniform(1.0, 2e5, 5000)  # hydrostatic pressures / densities
    assert max(_ulps(L.or_pow(p, 1 / 1.4), p ** (1 / 1.4)) for p in ps) <= 64
    assert L.or_exp(0.0) == 1.0 and L.or_log(1.0) == 0.0


# --- geometry: the equiangular metric ----------------------------------------
def _areas(orc, ex, n):
    o = orc.Oracle(sw_config(ex, nx=n))
    area = 0.0
    vol = 0.0
    for p in range(6):
        for k in range(3, 3 + n):
            for j in range(3, 3 + n):
                area += o.L.or_face_area(o.h, p, 0, k, j, 3)
                for i in range(3, 6):
                    vol += o.L.or_cell_volume(o.h, p, k, j, i)
    return area, vol


def test_sphere_area_converges_second_order(orc, ex):
    r1, r2 = R_EARTH, R_EARTH + 1.0
    e = []
    for n in (8, 16, 32):
        a, v = _areas(orc, ex, n)
        e.append((a / (4 * math.pi * r1 ** 2) - 1, v / (4 * math.pi / 3 * (r2 ** 3 - r1 ** 3)) - 1))
    for q in (0, 1):
        assert 3.5 < e[1][q] / e[2][q] < 4.5 and 3.5 < e[0][q] / e[1][q] < 4.5, e


def test_metric_identities(orc, ex):
    n = 8
    o = orc.Oracle(sw_config(ex, nx=n))
    nc = n + 6
    t = np.zeros((10, 6 * nc * nc))
    for w in range(10):
        o.L.or_get_geom2d(o.h, w, t[w])
    for c, s in ((0, 1), (2, 3), (4, 5)):
        assert np.allclose(t[c] ** 2 + t[s] ** 2, 1.0, atol=4e-16)
    # panel centre: orthogonal grid; cube corner (x=y=tan(pi/4)): 120 degree angle
    t = t.reshape(10, 6, nc, nc)
    assert abs(t[0, 0, 3 + n // 2, 3 + n // 2]) < 0.1
    x = np.tan(np.pi / 4 - np.pi / 4 / n)
    assert t[0, 0, 3, 3] == pytest.approx(-x * x / (1 + x * x), rel=1e-12)


def test_vector_maps_round_trip(orc):
    L = orc.lib()
    rng = np.random.default_rng(3)
    out = [C.c_double() for _ in range(4)]
    for b in range(1, 7):
        for _ in range(50):
            X, Y = rng.uniform(-1, 1, 2)
            v2, v3 = rng.normal(size=2)
            L.or_vec_maps(X, Y, b, v2, v3, *[C.byref(o) for o in out])
            assert out[2].value == pytest.approx(v2, abs=1e-13)
            assert out[3].value == pytest.approx(v3, abs=1e-13)


# --- panel-edge remap: analytic fields ----------------------------------------
def _remap_error(orc, ex, n):
    cfg = sw_config(ex, nx=n)
    o = orc.Oracle(cfg)
    nc = n + 6
    lat = C.c_double()
    lon = C.c_double()
    u = np.zeros((6, 4, nc, nc, 9))
    for p in range(6):
        for k in range(nc):
            for j in range(nc):
                o.L.or_latlon(o.h, p, k, j, C.byref(lat), C.byref(lon))
                u[p, 0, k, j, :] = 1000 + 100 * math.cos(lat.value) * math.sin(lon.value) + \
                    50 * math.sin(lat.value)
    ref = u.copy()
    for p in range(6):
        u[p, :, :3] = 0
        u[p, :, -3:] = 0
        u[p, :, :, :3] = 0
        u[p, :, :, -3:] = 0
    o.upload(u)
    got, _ = o.download()
    err = 0.0
    for p in range(6):
        for sl in [(slice(3, 3 + n), slice(0, 3)), (slice(3, 3 + n), slice(n + 3, n + 6)),
                   (slice(0, 3), slice(3, 3 + n)), (slice(n + 3, n + 6), slice(3, 3 + n))]:
            err = max(err, np.abs(got[p, 0][sl][..., 3:6] - ref[p, 0][sl][..., 3:6]).max())
    return err


def test_ghost_remap_smooth_field_second_order(orc, ex):
    e16, e32 = _remap_error(orc, ex, 16), _remap_error(orc, ex, 32)
    assert e16 < 0.2 and 3.0 < e16 / e32 < 5.0, (e16, e32)


def test_ghost_remap_keeps_constants(orc, ex):
    cfg = sw_config(ex, nx=8)
    o = orc.Oracle(cfg)
    u = np.zeros((6, 4, 14, 14, 9))
    u[:, 0] = 58467.6
    o.upload(u)
    got, _ = o.download()
    g = got[:, 0, 3:11, :3, 3:6]
    assert np.allclose(g, 58467.6, rtol=4e-16, atol=0)


# --- implicit column algebra ---------------------------------------------------
def _euler_flux(U, gm1):
    r, m1, m2, m3, E = U
    v1, v2, v3 = m1 / r, m2 / r, m3 / r
    p = gm1 * (E - 0.5 * r * (v1 * v1 + v2 * v2 + v3 * v3))
    return np.array([m1, m1 * v1 + p, m1 * v2, m1 * v3, (E + p) * v1])


def test_flux_jacobian_is_euler_flux_derivative(orc):
    L = orc.lib()
    gm1 = 0.4
    w = np.array([1.2, 30.0, -12.0, 7.0, 9.0e4])
    J = np.zeros(25)
    L.or_flux_jacobian(gm1, w, J)
    J = J.reshape(5, 5)
    r, v1, v2, v3, p = w
    U = np.array([r, r * v1, r * v2, r * v3, p / gm1 + 0.5 * r * (v1 * v1 + v2 * v2 + v3 * v3)])
    Jfd = np.zeros((5, 5))
    for q in range(5):
        h = 1e-6 * max(abs(U[q]), 1.0)
        e = np.zeros(5)
        e[q] = h
        Jfd[:, q] = (_euler_flux(U + e, gm1) - _euler_flux(U - e, gm1)) / (2 * h)
    assert np.allclose(J, Jfd, rtol=1e-6, atol=1e-6 * np.abs(Jfd).max())


def test_eigensystem(orc):
    L = orc.lib()
    gm1 = 0.4
    wl = np.array([1.1, 3.0, 2.0, -1.0, 1.0e5])
    wr = np.array([0.9, -2.0, 1.0, 0.5, 0.8e5])
    prim = np.zeros(5)
    L.or_roe_average(wl, wr, gm1, prim)
    cs = math.sqrt((1 + gm1) * prim[4] / prim[0])
    R, Ri, A = np.zeros(25), np.zeros(25), np.zeros(25)
    L.or_eigensystem(prim, cs, gm1, R, Ri, A)
    R, Ri, A = R.reshape(5, 5), Ri.reshape(5, 5), A.reshape(5, 5)
    assert np.allclose(R @ Ri, np.eye(5), atol=1e-12)
    J = np.zeros(25)
    L.or_flux_jacobian(gm1, prim, J)
    J = J.reshape(5, 5)
    lam = np.array([prim[1] - cs, prim[1], prim[1] + cs, prim[1], prim[1]])
    assert np.allclose(J @ R, R * lam, rtol=1e-10, atol=1e-8 * np.abs(J @ R).max())
    # |A| = R |Lambda| R^-1 squares to A^2
    assert np.allclose(A @ A, J @ J, rtol=1e-8, atol=1e-8 * np.abs(J @ J).max())
    # Roe average: density is the geometric mean
    assert prim[0] == pytest.approx(math.sqrt(1.1 * 0.9), rel=1e-15)


def _fma(a, b, c):
    # exactly rounded a * b + c (Fraction arithmetic, one rounding at the end)
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_abs_a_equals_dense_product(orc):
    # |A| sums only the structurally non-zero terms of R |Lambda| R^-1; each
    # skipped term is an exact zero, so the result must equal the dense
    # fused-multiply-add chain over all five terms bit for bit
    L = orc.lib()
    gm1 = 0.4
    rng = np.random.default_rng(17)
    prims = [np.array([1.0, 0.0, 0.0, 0.0, 1.0e5]), np.array([0.7, -40.0, 0.0, 12.0, 6.0e4])]
    prims += [np.array([rng.uniform(0.2, 2), *rng.normal(0, 50, 3), rng.uniform(1e4, 2e5)])
              for _ in range(30)]
    for prim in prims:
        cs = math.sqrt((1 + gm1) * prim[4] / prim[0])
        R, Ri, A = np.zeros(25), np.zeros(25), np.zeros(25)
        L.or_eigensystem(prim, cs, gm1, R, Ri, A)
        R, Ri, A = R.reshape(5, 5), Ri.reshape(5, 5), A.reshape(5, 5)
        u = prim[1]
        lam = [abs(u - cs), abs(u), abs(u + cs), abs(u), abs(u)]
        RL = [[R[a, b] * lam[b] for b in range(5)] for a in range(5)]
        for a in range(5):
            for c in range(5):
                t = RL[a][0] * Ri[0, c]
                for k in range(1, 5):
                    t = _fma(RL[a][k], Ri[k, c], t)
                assert t == A[a, c], (prim, a, c)


def test_block_thomas_vs_dense_solve(orc):
    L = orc.lib()
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 8):
        a = rng.normal(size=(n, 5, 5)) + 8 * np.eye(5)
        b = rng.normal(size=(n, 5, 5))
        c = rng.normal(size=(n, 5, 5))
        rhs = rng.normal(size=(n, 5))
        x = np.zeros((n, 5))
        L.or_block_thomas(n, a.copy(), b.copy(), c.copy(), rhs.copy(), x)
        M = np.zeros((5 * n, 5 * n))
        for i in range(n):
            M[5 * i:5 * i + 5, 5 * i:5 * i + 5] = a[i]
            if i > 0:
                M[5 * i:5 * i + 5, 5 * i - 5:5 * i] = b[i]
            if i < n - 1:
                M[5 * i:5 * i + 5, 5 * i + 5:5 * i + 10] = c[i]
        ref = np.linalg.solve(M, rhs.reshape(-1))
        assert np.allclose(x.reshape(-1), ref, rtol=1e-11, atol=1e-12)


def test_implicit_column_limits(orc, ex):
    """dt -> 0: the correction returns du; zero tendency gives zero; linear in du."""
    cfg = orc.or_config(atm_config(ex, nx1=8))
    nx1, nc1 = 8, 14
    rng = np.random.default_rng(7)
    z = (np.arange(nc1) - 3 + 0.5) * 3750.0
    w = np.zeros((5, nc1))
    w[0] = 1.2 * np.exp(-z / 7300.0)
    w[4] = 1e5 * np.exp(-z / 7300.0)
    w[1:4] = rng.normal(scale=2.0, size=(3, nc1))
    area = np.full(nx1 + 1, 1.0e10)
    vol = np.full(nx1, 3.75e13)
    du = rng.normal(size=(5, nc1))
    for dt, tol in ((1e-9, 1e-6), (1e-12, 1e-9)):
        d = du.copy()
        orc.lib().or_implicit_column(C.byref(cfg), nx1, w.copy(), d, area, vol, dt)
        assert np.allclose(d[:, 3:3 + nx1], du[:, 3:3 + nx1], rtol=tol, atol=tol)
    d0 = np.zeros_like(du)
    orc.lib().or_implicit_column(C.byref(cfg), nx1, w.copy(), d0, area, vol, 50.0)
    assert np.all(d0[:, 3:3 + nx1] == 0)
    d1, d2 = du.copy(), 3.0 * du
    orc.lib().or_implicit_column(C.byref(cfg), nx1, w.copy(), d1, area, vol, 50.0)
    orc.lib().or_implicit_column(C.byref(cfg), nx1, w.copy(), d2, area, vol, 50.0)
    assert np.allclose(d2[:, 3:3 + nx1], 3.0 * d1[:, 3:3 + nx1], rtol=1e-12, atol=1e-12)


# --- Riemann solvers: consistency F(w, w) = physical flux ---------------------
@pytest.mark.parametrize("rs", [2, 3])
def test_riemann_consistency_euler(orc, ex, rs):
    cfg = orc.or_config(atm_config(ex, rsolver=rs))
    gm1 = 0.4
    w = np.array([1.1, 12.0, -4.0, 3.0, 9.0e4])
    f = np.zeros(5)
    # x1 face, panel centre (cth = 0): the physical Euler flux in x1
    orc.lib().or_riemann_face(C.byref(cfg), 1, w.copy(), w.copy(), f, 0.0, 0.0, 1.0, 0.0)
    r, v1, v2, v3, p = w
    E = p / gm1 + 0.5 * r * (v1 * v1 + v2 * v2 + v3 * v3)
    ref = np.array([r * v1, r * v1 * v1 + p, r * v1 * v2, r * v1 * v3, (E + p) * v1])
    assert np.allclose(f, ref, rtol=1e-12)


@pytest.mark.parametrize("rs", [0, 1])
def test_riemann_consistency_shallow_water(orc, ex, rs):
    cfg = orc.or_config(sw_config(ex, rsolver=rs))
    w = np.array([5.8e4, 0.0, 20.0, -7.0])
    f = np.zeros(4)
    # x2 face at the panel centre: orthogonal metric, cth = 0
    orc.lib().or_riemann_face(C.byref(cfg), 2, w.copy(), w.copy(), f, 0.0, 0.0, 1.0, 0.0)
    h, _, v, wz = w
    ref = np.array([h * v, 0.0, h * v * v + 0.5 * h * h, h * v * wz])
    assert np.allclose(f, ref, rtol=1e-12)


# --- whole-step invariants -----------------------------------------------------
def test_shallow_water_rest_state_is_kept(orc, ex):
    cfg = sw_config(ex, nx=12, coriolis=1)
    o = orc.Oracle(cfg)
    u = np.zeros((6, 4, 18, 18, 9))
    u[:, 0] = 9.81 * 5960.0
    o.upload(u)
    o.step(10)
    got, _ = o.download()
    gi = interior(got, 3)
    assert np.abs(gi[:, 0] / (9.81 * 5960.0) - 1).max() < 1e-13
    assert np.abs(gi[:, 2:4]).max() < 1e-7 * 9.81 * 5960.0


def test_hllc_keeps_horizontally_uniform_atmosphere(orc, ex):
    """HLLC projects the metric, so a layered atmosphere at rest has no spurious
    horizontal momentum (to rounding); LMARS (lmars.cpp) does not project."""
    cfg = atm_config(ex, nx=8, nx1=8, rsolver=2)
    o = orc.Oracle(cfg)
    u, _ = ex.init_state(cfg, 1, 1, (250.0, 1e5, 287.0, 0.0, 0.0))
    o.upload(u)
    o.step(5, 100.0)
    got, _ = o.download()
    gi = interior(got, 8)
    assert np.abs(gi[:, 2:4]).max() < 1e-11


# --- W92 user outputs (examples/2023-Chen-exo3/W92.cpp:115-141) ---------------
def _uov(orc, a, b, blk, vy=0.0, vz=0.0):
    o = np.zeros(5)
    orc.lib().or_w92_user_out_point(a, b, blk, vy, vz, o)
    return o


def test_w92_user_output_known_answers(orc):
    # panel centres: the poles on blocks 1 and 5, the equator at the four
    # longitudes of blocks 2, 3, 4, 6 (after lon -> 2 pi - lon); at an
    # equatorial centre the basis is orthonormal, U = -v3 and V = -v2
    centres = {1: (90.0, 360.0), 2: (0.0, 360.0), 3: (0.0, 90.0), 4: (0.0, 270.0),
               5: (-90.0, 180.0), 6: (0.0, 180.0)}
    for blk, (lat, lon) in centres.items():
        o = _uov(orc, 0.0, 0.0, blk, 1.0, 2.0)
        assert o[0] == pytest.approx(lat, abs=1e-12) and o[1] == pytest.approx(lon, abs=1e-12)
        assert o[4] == 1.0
        if blk not in (1, 5):
            assert (o[2], o[3]) == (-2.0, -1.0)
    # off the centre of block 2: x2 runs south, x3 runs west
    o = _uov(orc, 0.1, 0.0, 2)
    assert o[0] == pytest.approx(-math.degrees(0.1), rel=1e-14) and o[1] == 360.0
    o = _uov(orc, 0.0, 0.1, 2)
    assert o[0] == pytest.approx(0.0, abs=1e-15)
    assert o[1] == pytest.approx(360.0 - math.degrees(0.1), rel=1e-14)
    # sqrtg = delta / (C D) at tangents with rational roots:
    # tan = 3/4, 0 -> 64/125; tan = 3/4, 4/3 -> 1728/12025
    for ta, tb, want in ((0.75, 0.0, Fraction(64, 125)), (0.75, 4 / 3, Fraction(1728, 12025))):
        for blk in range(1, 7):
            assert _uov(orc, math.atan(ta), math.atan(tb), blk)[4] == \
                pytest.approx(float(want), rel=1e-14)


def _unit(lat, lon):
    la, lo = math.radians(lat), math.radians(lon)
    return np.array([math.cos(la) * math.cos(lo), math.cos(la) * math.sin(lo), math.sin(la)])


def test_w92_user_output_wind_geometry(orc):
    """The lat-lon positions are a rigid map of the gnomonic panel (angles
    between points agree), the wind keeps the speed of the non-orthogonal
    (v2, v3) (|U, V|^2 = v2^2 + v3^2 + 2 v2 v3 cos theta, cos theta =
    -xy / (C D)) and a velocity along a coordinate line points along that
    line's lat-lon track."""
    rng = np.random.default_rng(7)
    h = 1e-6
    for blk in range(1, 7):
        for _ in range(40):
            a, b = rng.uniform(-0.75, 0.75, 2)
            a2, b2 = rng.uniform(-0.75, 0.75, 2)
            g1 = np.array([1.0, math.tan(a), math.tan(b)])
            g2 = np.array([1.0, math.tan(a2), math.tan(b2)])
            cos_g = g1 @ g2 / np.linalg.norm(g1) / np.linalg.norm(g2)
            o1, o2 = _uov(orc, a, b, blk), _uov(orc, a2, b2, blk)
            assert _unit(o1[0], o1[1]) @ _unit(o2[0], o2[1]) == pytest.approx(cos_g, abs=1e-13)
            vy, vz = rng.normal(size=2)
            o = _uov(orc, a, b, blk, vy, vz)
            x, y = math.tan(a), math.tan(b)
            cth = -x * y / math.sqrt((1 + x * x) * (1 + y * y))
            assert o[2] ** 2 + o[3] ** 2 == pytest.approx(vy * vy + vz * vz + 2 * vy * vz * cth,
                                                          rel=1e-12)
            if abs(o1[0]) > 80.0:   # the lat-lon track degenerates near a pole
                continue
            for e, (da, db) in enumerate(((h, 0.0), (0.0, h))):
                p0, p1 = _uov(orc, a - da, b - db, blk), _uov(orc, a + da, b + db, blk)
                dlon = (p1[1] - p0[1] + 180.0) % 360.0 - 180.0
                east = math.radians(dlon) * math.cos(math.radians(o1[0]))
                north = math.radians(p1[0] - p0[0])
                w = _uov(orc, a, b, blk, 1.0 if e == 0 else 0.0, 1.0 if e == 1 else 0.0)
                t = np.array([east, north]) / math.hypot(east, north)
                assert w[2:4] / np.linalg.norm(w[2:4]) == pytest.approx(t, abs=1e-6)


# --- redo path (main.cpp.in:126-160) -------------------------------------------
def test_redo_path(orc, ex):
    """Without an invalid step the redo loop is plain stepping; with one, the
    accepted step is a step from the restored state with the CFL dt halved once
    per redo, and too many redos fail."""
    cfg = atm_config(ex, nx=6, nx1=6, rsolver=3, implicit=9)
    u0, _ = ex.init_state(cfg, 1, 20261202, ())
    a, b = orc.Oracle(cfg), orc.Oracle(cfg)
    a.upload(u0)
    b.upload(u0)
    assert a.step_redo(3, 5) == (b.step(3), 0)
    assert np.array_equal(a.download()[0], b.download()[0])
    cfg.cfl = 1.5
    a, b = orc.Oracle(cfg), orc.Oracle(cfg)
    a.upload(u0)
    b.upload(u0)
    dt0 = b.new_dt()
    dt, nredo = a.step_redo(1, 5)
    assert nredo == 2 and a.state_valid()
    assert dt == dt0 / 4
    b.step(1, dt)
    assert np.array_equal(a.download()[0], b.download()[0])
    # the failing step itself leaves an invalid state
    c = orc.Oracle(cfg)
    c.upload(u0)
    c.step(1)
    assert not c.state_valid()
    # max_redo 0 (the reference's default) steps without checks
    a = orc.Oracle(cfg)
    a.upload(u0)
    assert a.step_redo(1, 0)[1] == 0 and not a.state_valid()
    # four explicit steps at CFL 8 need six redos; max_redo 2 allows three
    # over the run
    cfg = atm_config(ex, nx=6, nx1=6, rsolver=3, implicit=0)
    cfg.cfl = 8.0
    u0, _ = ex.init_state(cfg, 1, 20261202, ())
    a = orc.Oracle(cfg)
    a.upload(u0)
    assert a.step_redo(4, 10)[1] == 6
    a = orc.Oracle(cfg)
    a.upload(u0)
    with pytest.raises(RuntimeError, match="Too many"):
        a.step_redo(4, 2)


# --- partial (3x3) column correction: partial_correction.cpp:26-177 -----------
def test_block_thomas3_vs_dense_solve(orc):
    L = orc.lib()
    rng = np.random.default_rng(15)
    for n in (1, 2, 6):
        a = rng.normal(size=(n, 3, 3)) + 6 * np.eye(3)
        b = rng.normal(size=(n, 3, 3))
        c = rng.normal(size=(n, 3, 3))
        rhs = rng.normal(size=(n, 3))
        x = np.zeros((n, 3))
        L.or_block_thomas3(n, a.copy(), b.copy(), c.copy(), rhs.copy(), x)
        M = np.zeros((3 * n, 3 * n))
        for i in range(n):
            M[3 * i:3 * i + 3, 3 * i:3 * i + 3] = a[i]
            if i > 0:
                M[3 * i:3 * i + 3, 3 * i - 3:3 * i] = b[i]
            if i < n - 1:
                M[3 * i:3 * i + 3, 3 * i + 3:3 * i + 6] = c[i]
        ref = np.linalg.solve(M, rhs.reshape(-1))
        assert np.allclose(x.reshape(-1), ref, rtol=1e-11, atol=1e-12)


def _column(rng, nc1, transverse):
    z = (np.arange(nc1) - 3 + 0.5) * 3750.0
    w = np.zeros((5, nc1))
    w[0] = 1.2 * np.exp(-z / 7300.0)
    w[4] = 1e5 * np.exp(-z / 7300.0)
    w[1] = rng.normal(scale=2.0, size=nc1)
    if transverse:
        w[2:4] = rng.normal(scale=2.0, size=(2, nc1))
    return w


def test_partial_column_is_the_decoupled_full_column(orc, ex):
    """With no transverse velocity and no transverse tendency the 5x5 system
    decouples, and its (density, x1 momentum, energy) part is the partial
    correction's system: the two solvers agree there.  The partial solver also
    leaves the transverse tendencies untouched, returns du as dt -> 0, maps zero
    to zero and is linear."""
    cfg = orc.or_config(atm_config(ex, nx1=8))
    nx1, nc1 = 8, 14
    rng = np.random.default_rng(9)
    area = np.full(nx1 + 1, 1.0e10)
    vol = np.full(nx1, 3.75e13)
    L = orc.lib()
    w = _column(rng, nc1, transverse=False)
    du = rng.normal(size=(5, nc1))
    du[2:4] = 0.0
    full, part = du.copy(), du.copy()
    L.or_implicit_column(C.byref(cfg), nx1, w.copy(), full, area, vol, 50.0)
    L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), part, area, vol, 50.0)
    I = slice(3, 3 + nx1)
    for n in (0, 1, 4):
        assert np.allclose(part[n, I], full[n, I], rtol=1e-11, atol=1e-14), n
    assert not np.allclose(part[:, I], du[:, I])          # the correction did something
    w = _column(rng, nc1, transverse=True)
    du = rng.normal(size=(5, nc1))
    d = du.copy()
    L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), d, area, vol, 50.0)
    assert np.array_equal(d[2:4], du[2:4])
    for dt, tol in ((1e-9, 1e-6), (1e-12, 1e-9)):
        d = du.copy()
        L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), d, area, vol, dt)
        assert np.allclose(d[:, I], du[:, I], rtol=tol, atol=tol)
    d0 = np.zeros_like(du)
    L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), d0, area, vol, 50.0)
    assert np.all(d0[:, I] == 0)
    d1, d2 = du.copy(), 3.0 * du
    L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), d1, area, vol, 50.0)
    L.or_implicit_column3(C.byref(cfg), nx1, w.copy(), d2, area, vol, 50.0)
    assert np.allclose(d2[:, I], 3.0 * d1[:, I], rtol=1e-12, atol=1e-12)
