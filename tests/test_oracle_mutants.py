"""The oracle's physics pins catch plausible restatement mistakes: each
mutant below plants one mistake (a sign, a swapped index or operand, a
dropped term) in a copy of oracle/exo_oracle.c, builds it, and the named pins
of tests/oracle_checks.py must fail on it, while the unmodified copy passes
them all.  Each build runs in a process of its own (oracle_mutant_runner.py)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
RUNNER = os.path.join(HERE, "oracle_mutant_runner.py")
ALL = ("remap_vector", "lr_sync", "decomposition", "riemann", "coriolis_w2", "cfl", "ssprk3",
       "column_conserves")

# (name, text in exo_oracle.c, replacement, pins that must fail)
MUTANTS = [
    # panel-edge ghost remap (cubed_sphere_utility.cpp:92-369)
    ("remap_raise_sign", "vy = vyt / sth / sth - vzt * cth / sth / sth;",
     "vy = vyt / sth / sth + vzt * cth / sth / sth;", ("remap_vector",)),
    ("remap_lower_dropped", "vy = vyt + vzt * cth;\n            vz = vzt + vyt * cth;\n            tgt",
     "vy = vyt;\n            vz = vzt + vyt * cth;\n            tgt", ("remap_vector",)),
    ("remap_no_axis_swap", "double tmp = tgt_cy;\n              tgt_cy = tgt_cz;\n              tgt_cz = tmp;",
     "(void)0;", ("remap_vector",)),
    ("remap_weights_swapped", "tgt[ot] = ((y2 - yq) * v1 + (yq - y1) * v2) / (y2 - y1);",
     "tgt[ot] = ((yq - y1) * v1 + (y2 - yq) * v2) / (y2 - y1);", ("remap_vector",)),
    # vector maps (cubed_sphere_transform.cpp:11-61)
    ("abp_from_rll_sign", "*V2 = X * Y / delta * U - V;", "*V2 = X * Y / delta * U + V;",
     ("remap_vector", "coriolis_w2")),
    ("rll_from_abp_sign", "*V = X * Y / (C * D) * V3 - V2;", "*V = X * Y / (C * D) * V3 + V2;",
     ("remap_vector", "coriolis_w2")),
    # L/R exchange (cubed_sphere.cpp:267-687)
    ("sync_dropped", "  sync_lr(s);\n", "\n", ("lr_sync",)),
    ("sync_own_block_frame", "vec_abp_from_rll(x2d, x3d, TGBID[blockID - 1][target_dir], U, V, &vx, &vy);",
     "vec_abp_from_rll(x2d, x3d, blockID, U, V, &vx, &vy);", ("lr_sync",)),
    ("sync_no_inversion", "int k = inv ? kb2 - kk : kb1 + kk;\n              int j = inv ? jb2 - jj : jb1 + jj;",
     "int k = kb1 + kk;\n              int j = jb1 + jj;", ("lr_sync",)),
    ("sync_components_swapped", "data[off++] = n == IVY ? vx : vy;",
     "data[off++] = n == IVY ? vy : vx;", ("lr_sync",)),
    # hydrostatic decomposition (change_to_perturbation.cpp:39-190)
    ("decomp_scan_sign", "c->psf[i] = c->psf[i + 1] + grav * WC(IDN, i) * P->dx1f[i];",
     "c->psf[i] = c->psf[i + 1] - grav * WC(IDN, i) * P->dx1f[i];", ("decomposition",)),
    ("decomp_top_half_dropped", "or_exp(-grav * P->dx1f[ie] / (2 * RdTv));",
     "or_exp(-grav * P->dx1f[ie] / (RdTv));", ("decomposition",)),
    ("decomp_entropy_top_cell", "c->ent1 = or_log(WC(IPR, is)) - gamma * or_log(WC(IDN, is));",
     "c->ent1 = or_log(WC(IPR, ie)) - gamma * or_log(WC(IDN, ie));", ("decomposition",)),
    ("restore_entropy_sign", "double dsf = or_pow(c->psf[i], 1. / c->ent0) * or_exp(-c->ent1 / c->ent0);",
     "double dsf = or_pow(c->psf[i], 1. / c->ent0) * or_exp(c->ent1 / c->ent0);", ("decomposition",)),
    ("restore_face_index", "    wl[IPR * st + i] += c->psf[i];\n",
     "    wl[IPR * st + i] += c->psf[i + 1];\n", ("decomposition",)),
    # Riemann solvers (src/snap/riemann/*.cpp)
    ("roe_a1_sign", "double a1 = 0.5 * (cbar * delh - hbar * delu) / cbar;",
     "double a1 = 0.5 * (cbar * delh + hbar * delu) / cbar;", ("riemann",)),
    ("roe_speeds_swapped", "double speed[3] = {fabs(ubar - cbar), fabs(ubar), fabs(ubar + cbar)};",
     "double speed[3] = {fabs(ubar + cbar), fabs(ubar), fabs(ubar - cbar)};", ("riemann",)),
    ("lmars_sw_hbar_sign",
     "double hbar = 0.5 * (wli[IDN] + wri[IDN]) + 0.5 * cbar * (wli[ivx] - wri[ivx]);",
     "double hbar = 0.5 * (wli[IDN] + wri[IDN]) + 0.5 * cbar * (wri[ivx] - wli[ivx]);", ("riemann",)),
    ("lmars_sw_upwind_side", "if (ubar > 0.) {\n      fl[IDN] = ubar * wli[IDN];\n",
     "if (ubar < 0.) {\n      fl[IDN] = ubar * wli[IDN];\n", ("riemann",)),
    ("hllc_branch", "if (am >= 0.0) {", "if (am < 0.0) {", ("riemann",)),
    ("hllc_right_speed_sign", "double ar = wr[IVX] + cr * qr;", "double ar = wr[IVX] - cr * qr;",
     ("riemann",)),
    ("lmars_pbar_sign",
     "double pbar = 0.5 * (wli[IPR] + wri[IPR]) + 0.5 * (rhobar * cbar) * (wli[ivx] - wri[ivx]);",
     "double pbar = 0.5 * (wli[IPR] + wri[IPR]) + 0.5 * (rhobar * cbar) * (wri[ivx] - wli[ivx]);",
     ("riemann",)),
    ("lmars_ubar_sign",
     "double ubar = 0.5 * (wli[ivx] + wri[ivx]) + 0.5 / (rhobar * cbar) * (wli[IPR] - wri[IPR]);",
     "double ubar = 0.5 * (wli[ivx] + wri[ivx]) + 0.5 / (rhobar * cbar) * (wri[IPR] - wli[IPR]);",
     ("riemann",)),
    ("lmars_enthalpy_side", "fl[IEN] = ubar * wli[IDN] * hl;", "fl[IEN] = ubar * wli[IDN] * hr;",
     ("riemann",)),
    # Coriolis (W92.cpp:48-70) and the metric sources (gnomonic_equiangle.cpp:834-879)
    ("coriolis_sign", "double ll_acc_U = f * V, ll_acc_V = -f * U;",
     "double ll_acc_U = -f * V, ll_acc_V = f * U;", ("coriolis_w2",)),
    ("coriolis_not_lowered",
     "cs_contra_to_co(P, j, k, acc2, acc3, &acc2, &acc3);\n          out[C4(IVY, k, j, i)] += dt * rho * acc2;",
     "out[C4(IVY, k, j, i)] += dt * rho * acc2;", ("coriolis_w2",)),
    ("getuv_no_flip", "  *U = -*U;\n}", "}", ("coriolis_w2",)),
    ("metric_source_sign",
     "double src2 = -P->xovrD[c3] * (pr + rho * v3 * v3 * sine2) - rho * v1 * v_2 / radius;",
     "double src2 = P->xovrD[c3] * (pr + rho * v3 * v3 * sine2) - rho * v1 * v_2 / radius;",
     ("coriolis_w2",)),
    ("sw_pressure_half_dropped",
     "pr = 0.5 * P->w[C4(IDN, k, j, i)] * P->w[C4(IDN, k, j, i)];",
     "pr = P->w[C4(IDN, k, j, i)] * P->w[C4(IDN, k, j, i)];", ("coriolis_w2",)),
    # NewBlockTimeStep (patches/21.new_blockdt.patch:16-37)
    ("cfl_widths_swapped", "dt2 = P->x1v[i] * P->dx2a[c], dt3 = P->x1v[i] * P->dx3a[c];",
     "dt2 = P->x1v[i] * P->dx3a[c], dt3 = P->x1v[i] * P->dx2a[c];", ("cfl",)),
    ("cfl_speeds_swapped", "sp2 = fabs(wi[IVY]) + cs, sp3 = fabs(wi[IVZ]) + cs;",
     "sp2 = fabs(wi[IVZ]) + cs, sp3 = fabs(wi[IVY]) + cs;", ("cfl",)),
    ("cfl_implicit_x1_factor", "if ((flag & 1) && s->N > 1) dt1 /= cfl * fabs(wi[IVX]);",
     "if ((flag & 1) && s->N > 1) dt1 /= fabs(wi[IVX]);", ("cfl",)),
    ("cfl_sw_sound_speed", "? sqrt(wi[IDN])", "? sqrt(0.5 * wi[IDN])", ("cfl",)),
    # SSPRK3 and AddFluxDivergence
    ("ssprk3_g1", "static const double G1[3] = {0.0, 0.25, 2.0 / 3.0};",
     "static const double G1[3] = {0.0, 0.75, 2.0 / 3.0};", ("ssprk3",)),
    ("ssprk3_g2", "static const double G2[3] = {1.0, 0.75, 1.0 / 3.0};",
     "static const double G2[3] = {1.0, 0.75, 2.0 / 3.0};", ("ssprk3",)),
    ("ssprk3_beta3", "static const double BETA[3] = {1.0, 0.25, 2.0 / 3.0};",
     "static const double BETA[3] = {1.0, 0.25, 1.0 / 3.0};", ("ssprk3",)),
    ("ssprk3_beta1", "static const double BETA[3] = {1.0, 0.25, 2.0 / 3.0};",
     "static const double BETA[3] = {0.5, 0.25, 2.0 / 3.0};", ("ssprk3",)),
    ("divergence_x3_areas_swapped",
     "face3_area(s, P, k + 1, j, i) * flx[2][F3(n, k + 1, j, i)] -\n"
     "                               face3_area(s, P, k, j, i) * flx[2][F3(n, k, j, i)];",
     "face3_area(s, P, k, j, i) * flx[2][F3(n, k + 1, j, i)] -\n"
     "                               face3_area(s, P, k + 1, j, i) * flx[2][F3(n, k, j, i)];",
     ("ssprk3",)),
    # the implicit column corrections (full_correction.cpp, partial_correction.cpp)
    ("implicit_b_area", "b[i].a[r][q] = -(Am.a[r][q] + dfdq[i - 1].a[r][q]) * aleft / (2. * v);",
     "b[i].a[r][q] = -(Am.a[r][q] + dfdq[i - 1].a[r][q]) * aright / (2. * v);", ("column_conserves",)),
    ("implicit_c_jacobian_sign",
     "c[i].a[r][q] = -(Ap.a[r][q] - dfdq[i + 1].a[r][q]) * aright / (2. * v);",
     "c[i].a[r][q] = -(Ap.a[r][q] + dfdq[i + 1].a[r][q]) * aright / (2. * v);", ("column_conserves",)),
    ("implicit_a_area_difference", "(aright - aleft) * dfdq[i].a[r][q]",
     "(aleft - aright) * dfdq[i].a[r][q]", ("column_conserves",)),
    ("implicit_bottom_fold_dropped",
     "double sgn = q == IVX ? -1. : 1.;\n      a[is].a[r][q] += b[is].a[r][q] * sgn;",
     "double sgn = q == IVX ? -1. : 1.;\n      a[is].a[r][q] += 0. * b[is].a[r][q] * sgn;",
     ("column_conserves",)),
    ("partial_c_area", "c[i].a[r][q] = -(Ap.a[R][Q] - dfdq[i + 1].a[R][Q]) * aright / (2. * v);",
     "c[i].a[r][q] = -(Ap.a[R][Q] - dfdq[i + 1].a[R][Q]) * aleft / (2. * v);", ("column_conserves",)),
]

CFLAGS = ["-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-std=c11", "-w",
          "-shared"]


def _build(src_text, out):
    c = out + ".c"
    with open(c, "w") as f:
        f.write(src_text)
    subprocess.run(["gcc", *CFLAGS, "-I", ORACLE, "-o", out, c,
                    os.path.join(ORACLE, "exo_oracle_math.c"), "-lm"], check=True)
    return out


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("mutants"))
    with open(os.path.join(ORACLE, "exo_oracle.c")) as f:
        src = f.read()
    jobs = {"unmodified": src}
    for name, old, new, _ in MUTANTS:
        assert src.count(old) == 1, f"{name}: the planted text must occur exactly once"
        jobs[name] = src.replace(old, new)
    with ThreadPoolExecutor(max_workers=8) as pool:
        futs = {n: pool.submit(_build, t, os.path.join(d, f"lib{n}.so")) for n, t in jobs.items()}
        return {n: f.result() for n, f in futs.items()}


def _run(so, checks):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, RUNNER, so, *checks], capture_output=True, text=True,
                       env=env, timeout=300)
    return r.returncode, r.stdout + r.stderr


def test_unmodified_copy_passes_every_pin(builds):
    rc, out = _run(builds["unmodified"], ALL)
    assert rc == 0, out


@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(builds, mutant):
    name, _, _, checks = mutant
    rc, out = _run(builds[name], checks)
    assert rc == 1, f"{name}: rc {rc}\n{out}"
