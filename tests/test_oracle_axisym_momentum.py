"""Closed-form pins of the oracle's axisymmetric momentum restatement (MACSolver2DAxisym::GetRHSU,
GetRHSV, AddExternalViscosityFU/FV, GetUHat, GetVHat; reference src/mac2dAxisym.cpp:264-322,
590-910, 950-1478, which cannot be compiled here: the file needs TECIO and MKL), with planted
mistakes compiled on the CPU that the pins must catch.

Positions (base radius 0, as the reference's momentum terms compute their radii): U face i at
r = (i - g) dr, z = (j + 1/2 - g) dz; V face at r = (i + 1/2 - g) dr, z = (j - g) dz.

* viscous cross terms: v = gamma r z gives the U term mu gamma / rho; u = gamma r z and
  u* = delta r z give the V term mu (gamma + delta) / rho.  Across a planar interface the density
  is rho_L + (rho_H - rho_L) theta with theta the heavy share of the segment between the two
  level-set samples (S-N half points for U, W-E half points and S-N cells for V);
* right-hand sides: u = a r gives convection d(u^2)/dr = 2 a^2 r exactly (WENO5 is exact on
  quadratic fluxes), likewise v = b z along z; gravity -g on the faces the reference fills;
* line sweeps: every interior row of the r and z sweeps maps u = r (U faces), v = r^2 (V faces),
  z^2 (both) to the profile minus a known constant, so right-hand sides built from the closed
  forms give the profiles back; the Dirichlet sides hold the profile's ghost value (twice the
  constant where the reference adds it twice); the final U division by 1 + 2 dt mu / (rho r^2)
  follows.
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import LIB_PATH, Oracle, make_fluid, make_grid
from tests.test_oracle_pins import _mutant

G = 3
NR, NZ, DR, DZ = 9, 8, 0.11, 0.07


def _pos(nr=NR, nz=NZ, dr=DR, dz=DZ):
    i = np.arange(nr + 2 * G)[:, None]
    j = np.arange(nz + 2 * G)[None, :]
    ru, zu = (i - G) * dr + 0 * j, (j + 0.5 - G) * dz + 0 * i
    rv, zv = (i + 0.5 - G) * dr + 0 * j, (j - G) * dz + 0 * i
    rc, zc = rv, zu
    return ru, zu, rv, zv, rc, zc


U_IN = (slice(G + 1, G + NR), slice(G, G + NZ))
V_IN = (slice(G, G + NR), slice(G + 1, G + NZ))


def _theta(lo, hi, z0, heavy_above):
    """heavy share of the segment [lo, hi] for a planar interface at z0"""
    s = (hi - z0) / (hi - lo) if heavy_above else (z0 - lo) / (hi - lo)
    return np.clip(s, 0.0, 1.0)


def check_viscosity_u(O):
    ru, zu, rv, zv, rc, zc = _pos()
    gr = make_grid(NR, NZ, DR, DZ)
    gam, mu, rh, rl = 1.7, 0.3, 5.0, 2.0
    v = gam * rv * zv
    one = np.ones_like(v)
    fl = make_fluid(rh, rl, mu, mu, 0.0, 1e-3)
    dU = O.viscosity_u_axisym(gr, fl, v, one, one)
    np.testing.assert_allclose(dU[U_IN], mu * gam / rh, rtol=1e-12)
    assert not dU[:G + 1].any() and not dU[:, :G].any()
    for heavy_above in (True, False):
        z0 = 3.3 * DZ
        ls = (zc - z0) if heavy_above else (z0 - zc)
        dU = O.viscosity_u_axisym(gr, fl, v, ls, one)
        th = _theta(zu - 0.5 * DZ, zu + 0.5 * DZ, z0, heavy_above)
        np.testing.assert_allclose(dU[U_IN], (mu * gam / (rl + (rh - rl) * th))[U_IN], rtol=1e-12)


def check_viscosity_v(O):
    ru, zu, rv, zv, rc, zc = _pos()
    gr = make_grid(NR, NZ, DR, DZ)
    gam, dlt, mu, rh, rl = 1.3, -0.6, 0.25, 4.0, 1.5
    one = np.ones_like(ru)
    fl = make_fluid(rh, rl, mu, mu, 0.0, 1e-3)
    dV = O.viscosity_v_axisym(gr, fl, gam * ru * zu, dlt * ru * zu, one, one)
    np.testing.assert_allclose(dV[V_IN], mu * (gam + dlt) / rh, rtol=1e-12)
    # r-interface: only the W-E density sees it (the S-N cells share r)
    r0 = 4.4 * DR
    for heavy_above in (True, False):
        ls = (rc - r0) if heavy_above else (r0 - rc)
        dV = O.viscosity_v_axisym(gr, fl, gam * ru * zu, 0 * ru, ls, one)
        th = _theta(rv - 0.5 * DR, rv + 0.5 * DR, r0, heavy_above)
        np.testing.assert_allclose(dV[V_IN], (mu * gam / (rl + (rh - rl) * th))[V_IN], rtol=1e-12)
    # z-interface: the S-N density of the u* term from the two cells either side of the face
    z0 = 3.6 * DZ
    for heavy_above in (True, False):
        ls = (zc - z0) if heavy_above else (z0 - zc)
        dV = O.viscosity_v_axisym(gr, fl, 0 * ru, dlt * ru * zu, ls, one)
        th = _theta(zv - 0.5 * DZ, zv + 0.5 * DZ, z0, heavy_above)
        np.testing.assert_allclose(dV[V_IN], (mu * dlt / (rl + (rh - rl) * th))[V_IN], rtol=1e-12)


NEU = ("neumann",) * 4


def check_rhs(O):
    ru, zu, rv, zv, rc, zc = _pos()
    gr = make_grid(NR, NZ, DR, DZ, bc_u=NEU, bc_v=NEU)
    z = np.zeros_like(ru)
    one = np.ones_like(ru)
    dt, mu, rho = 2e-3, 0.4, 3.0
    fl = make_fluid(rho, rho, mu, mu, 0.0, dt)
    a = 0.8
    rhs = O.rhs_u_axisym(gr, fl, one, a * ru, z, one)
    np.testing.assert_allclose(rhs[U_IN], (a * ru - dt * 2 * a * a * ru)[U_IN], rtol=1e-12, atol=1e-15)
    b = -0.7
    rhs = O.rhs_v_axisym(gr, fl, one, z, z, b * zv, one)
    np.testing.assert_allclose(rhs[V_IN], (b * zv - dt * 2 * b * b * zv)[V_IN], rtol=1e-12, atol=1e-15)
    # viscous parts: the U term from v, the V term from u*
    gam, dlt = 1.1, 0.9
    rhs = O.rhs_u_axisym(gr, fl, one, z, gam * rv * zv, one)
    np.testing.assert_allclose(rhs[U_IN], dt * mu * gam / rho, rtol=1e-12)
    rhs = O.rhs_v_axisym(gr, fl, one, z, dlt * ru * zu, z, one)
    np.testing.assert_allclose(rhs[V_IN], dt * mu * dlt / rho, rtol=1e-12)
    # the ghost fill (Neumann: copies of the first interior face)
    rhs = O.rhs_u_axisym(gr, fl, one, a * ru, z, one)
    np.testing.assert_array_equal(rhs[G + 1:G + NR, G - 1], rhs[G + 1:G + NR, G])


def check_gravity(O):
    gr = make_grid(NR, NZ, DR, DZ, bc_u=NEU, bc_v=NEU)
    shp = (NR + 2 * G, NZ + 2 * G)
    one, z = np.ones(shp), np.zeros(shp)
    u0, v0, dt, g = 0.5, -0.25, 1e-3, 9.8
    for axis, fix in ((0, 0), (1, 0), (0, 1), (1, 1)):
        fl = make_fluid(2.0, 2.0, 0.0, 0.0, 0.0, dt, g=g, g_axis=axis, gravity_fix=fix)
        fl.re = 0.0  # no viscous terms (kRe <= 0)
        ru_ = O.rhs_u_axisym(gr, fl, one, u0 * one, z, one)
        np.testing.assert_allclose(ru_[U_IN], u0 + dt * (-g if axis == 0 else 0.0), rtol=1e-14)
        rv_ = O.rhs_v_axisym(gr, fl, one, z, z, v0 * one, one)
        want = np.full((NR, NZ - 1), v0)
        if fix and axis == 1:
            want += dt * -g
        elif not fix and axis == 0:
            want[1:] += dt * -g  # the U-grid gravity array: nothing on the first column
        np.testing.assert_allclose(rv_[V_IN], want, rtol=1e-14)


def check_uhat_r(O):
    """u = r: the r sweep maps r to r - 2 dt mu / (rho r) on every interior row, axis / wall
    constants hold the ghost values 0 and nr dr"""
    ru, zu, rv, zv, rc, zc = _pos()
    dt, mu, rho = 0.01, 0.7, 1.3
    fl = make_fluid(rho, rho, mu, mu, 0.0, dt)
    one = np.ones_like(ru)
    for west in ("axisym", "dirichlet", "wall"):
        gr = make_grid(NR, NZ, DR, DZ, bc_u=(west, "dirichlet", "neumann", "outlet"), c_u=(0.0, NR * DR, 0.0, 0.0))
        rhs = np.zeros_like(ru)
        rhs[U_IN] = (ru - 2 * dt * mu / (rho * ru))[U_IN]
        uh = O.uhat_axisym(gr, fl, one, rhs, one)
        want = ru / (1 + dt / rho * 2 * mu / ru ** 2)
        np.testing.assert_allclose(uh[U_IN], want[U_IN], rtol=1e-12)
        assert not uh[:G + 1].any() and not uh[G + NR:].any()


def check_uhat_z(O):
    """u = z^2 (constant in r, W / E Neumann): the z sweep maps z^2 to z^2 - 2 dt mu / rho_eff with
    rho_eff of the S-N half points; S held by the wall (axisym: mirror, z^2 is even about z = 0;
    dirichlet: the reference adds twice the constant, so c = ghost / 2), N Dirichlet"""
    ru, zu, rv, zv, rc, zc = _pos()
    dt, mu, rh, rl = 0.02, 0.5, 6.0, 1.5
    fl = make_fluid(rh, rl, mu, mu, 0.0, dt)
    one = np.ones_like(ru)
    cN = ((NZ + 0.5) * DZ) ** 2 / 2
    for south, cS in (("axisym", 0.0), ("dirichlet", (0.5 * DZ) ** 2 / 2), ("inlet", (0.5 * DZ) ** 2 / 2)):
        for heavy_above in (True, False):
            z0 = 4.25 * DZ
            ls = (zc - z0) if heavy_above else (z0 - zc)
            gr = make_grid(NR, NZ, DR, DZ, bc_u=("neumann", "outlet", south, "wall"), c_u=(0.0, 0.0, cS, cN))
            th = _theta(zu - 0.5 * DZ, zu + 0.5 * DZ, z0, heavy_above)
            rhs = np.zeros_like(ru)
            rhs[U_IN] = (zu ** 2 - 2 * dt * mu / (rl + (rh - rl) * th))[U_IN]
            uh = O.uhat_axisym(gr, fl, ls, rhs, one)
            rho_c = np.where(ls >= 0, rh, rl)
            want = zu ** 2 / (1 + dt / rho_c * 2 * mu / ru ** 2)
            np.testing.assert_allclose(uh[U_IN], want[U_IN], rtol=1e-11)


def check_vhat_r(O):
    """v = r^2 on V faces: the r sweep maps it to r^2 - 4 dt mu / rho (the axis row has no W
    coupling); E Dirichlet with c = ghost / 2 (added twice)"""
    ru, zu, rv, zv, rc, zc = _pos()
    dt, mu, rho = 0.01, 0.6, 2.5
    fl = make_fluid(rho, rho, mu, mu, 0.0, dt)
    one = np.ones_like(ru)
    cE = ((NR + 0.5) * DR) ** 2 / 2
    for west in ("axisym", "neumann", "outlet"):
        gr = make_grid(NR, NZ, DR, DZ, bc_v=(west, "wall", "neumann", "outlet"), c_v=(0.0, cE, 0.0, 0.0))
        rhs = np.zeros_like(ru)
        rhs[V_IN] = (rv ** 2 - 4 * dt * mu / rho)[V_IN]
        vh = O.vhat_axisym(gr, fl, one, rhs, one)
        np.testing.assert_allclose(vh[V_IN], (rv ** 2)[V_IN], rtol=1e-12)
        assert not vh[:, :G + 1].any() and not vh[:, G + NZ:].any()


def check_vhat_z(O):
    """v = z^2 on V faces (W / E Neumann): the z sweep maps it to z^2 - 4 dt mu / rho_eff with
    rho_eff of the two cells either side; S (z = 0) Dirichlet 0 or axisym, N Dirichlet once"""
    ru, zu, rv, zv, rc, zc = _pos()
    dt, mu, rh, rl = 0.015, 0.45, 3.0, 1.2
    fl = make_fluid(rh, rl, mu, mu, 0.0, dt)
    one = np.ones_like(ru)
    for south in ("dirichlet", "axisym", "wall"):
        for heavy_above in (True, False):
            z0 = 3.7 * DZ
            ls = (zc - z0) if heavy_above else (z0 - zc)
            gr = make_grid(NR, NZ, DR, DZ, bc_v=("neumann", "outlet", south, "inlet"),
                           c_v=(0.0, 0.0, 0.0, (NZ * DZ) ** 2))
            th = _theta(zv - 0.5 * DZ, zv + 0.5 * DZ, z0, heavy_above)
            rhs = np.zeros_like(ru)
            rhs[V_IN] = (zv ** 2 - 4 * dt * mu / (rl + (rh - rl) * th))[V_IN]
            vh = O.vhat_axisym(gr, fl, ls, rhs, one)
            np.testing.assert_allclose(vh[V_IN], (zv ** 2)[V_IN], rtol=1e-11, atol=1e-15)


def check_periodic_constant(O):
    """periodic sides take the coupling times the far end's right-hand side: a constant field stays"""
    shp = (NR + 2 * G, NZ + 2 * G)
    one = np.ones(shp)
    fl = make_fluid(2.0, 2.0, 0.3, 0.3, 0.0, 0.05)
    gr = make_grid(NR, NZ, DR, DZ, bc_u=("periodic",) * 4, bc_v=("periodic",) * 4)
    rhs = np.zeros(shp)
    rhs[U_IN] = 0.75
    ru, *_ = _pos()
    uh = O.uhat_axisym(gr, fl, one, rhs, one)
    np.testing.assert_allclose(uh[U_IN], (0.75 / (1 + 0.05 / 2.0 * 0.6 / ru ** 2))[U_IN], rtol=1e-12)
    rhs = np.zeros(shp)
    rhs[V_IN] = -0.4
    np.testing.assert_allclose(O.vhat_axisym(gr, fl, one, rhs, one)[V_IN], -0.4, rtol=1e-12)


CHECKS = {f.__name__: f for f in (check_viscosity_u, check_viscosity_v, check_rhs, check_gravity, check_uhat_r,
                                  check_uhat_z, check_vhat_r, check_vhat_z, check_periodic_constant)}


@pytest.mark.parametrize("name", list(CHECKS))
def test_axisym_momentum_pin(oracle, name):
    with np.errstate(divide="ignore"):  # closed forms over whole arrays include r = 0 on the axis face
        CHECKS[name](oracle)


MUTANTS = [
    ("heavy share from the wrong end", "else if (a < 0.0 && b >= 0.0) thH = fabs(b) / (fabs(a) + fabs(b));",
     "else if (a < 0.0 && b >= 0.0) thH = fabs(a) / (fabs(a) + fabs(b));", 0),
    ("harmonic density in the viscous U term", "A_(dU, i, j) = visZ / rhoSN;",
     "A_(dU, i, j) = visZ * (1.0 / fl->rho_h + 1.0 / fl->rho_l) * 0.5;", 0),
    ("U cross term over dz along r", "const double visZ = (muUNHalf * (vN - vW_N) / dr - muUSHalf * (vM - vW) / dr) / dz;",
     "const double visZ = (muUNHalf * (vN - vW_N) / dz - muUSHalf * (vM - vW) / dz) / dz;", 0),
    ("U cross term from the face's own column", "const double vW = A_(v, i - 1, j), vW_N = A_(v, i - 1, j + 1)",
     "const double vW = A_(v, i - 1, j), vW_N = A_(v, i - 1, j)", 0),
    ("V hoop term from u instead of u*", "* (0.5 * (A_(uhat, i, j) + A_(uhat, i + 1, j)) - 0.5 * (A_(uhat, i, j - 1)",
     "* (0.5 * (A_(u, i, j) + A_(uhat, i + 1, j)) - 0.5 * (A_(u, i, j - 1)", 0),
    ("V hoop term at the W face radius", "const double rPM = (i + 0.5 - g) * dr;\n            const double uM",
     "const double rPM = (i - g) * dr + 1e-300;\n            const double uM", 0),
    ("V densities swapped", "A_(dV, i, j) = visR / rhoWE + visZ / rhoSN;", "A_(dV, i, j) = visR / rhoSN + visZ / rhoWE;", 0),
    ("convection added", "rhs[c] = u[c] + fl->dt * (-cU[c] + gU + vU[c]);", "rhs[c] = u[c] + fl->dt * (cU[c] + gU + vU[c]);", 0),
    ("V gravity on the first column", "else gv = (gravity_active(fl, 0) && i >= g + 1) ? -fl->g : 0.0;\n            rhs[c] = v[c]",
     "else gv = gravity_active(fl, 0) ? -fl->g : 0.0;\n            rhs[c] = v[c]", 0),
    ("U r sweep W weight at the E radius", "a[i] = -dt / (rho * rUM) * (2.0 * rPW * muW) / (dr * dr);",
     "a[i] = -dt / (rho * rUM) * (2.0 * rPM * muW) / (dr * dr);", 0),
    ("U r sweep at the cell radius", "const double rUM = (i + 1.0) * dr, rPW", "const double rUM = (i + 0.5) * dr, rPW", 0),
    ("U E wall added twice", "else if (fE == 2) { d[i] -= c[i] * gr->c_u[1]; c[i] = 0.0; }",
     "else if (fE == 2) { d[i] -= c[i] * 2.0 * gr->c_u[1]; c[i] = 0.0; }", 0),
    ("U S wall added once", "else if (fS == 2) { d[j] -= a[j] * 2.0 * gr->c_u[2]; a[j] = 0.0; }",
     "else if (fS == 2) { d[j] -= a[j] * gr->c_u[2]; a[j] = 0.0; }", 0),
    ("U axis on S acts as a wall", "const int fS = sweep_fold(gr->bc_u[2], 1)", "const int fS = sweep_fold(gr->bc_u[2], 0)", 0),
    ("U z sweep density from the W / M cells", "const double rho = rho_eff_axi(fl, lsUSHalf, lsUNHalf);\n            a[j]",
     "const double rho = rho_eff_axi(fl, A_(ls, I - 1, J), A_(ls, I, J));\n            a[j]", 0),
    ("U final factor over r", "/ (1.0 + dt / rho * (A_(mu, i - 1, j) + A_(mu, i, j)) / (rUM * rUM));",
     "/ (1.0 + dt / rho * (A_(mu, i - 1, j) + A_(mu, i, j)) / rUM);", 0),
    ("V r sweep W half radius", "const double rVM = (i + 0.5) * dr, rVWHalf = i * dr,",
     "const double rVM = (i + 0.5) * dr, rVWHalf = (i + 0.5) * dr,", 0),
    ("V E wall added once", "else if (fE == 2) { d[i] -= c[i] * 2.0 * gr->c_v[1]; c[i] = 0.0; }",
     "else if (fE == 2) { d[i] -= c[i] * gr->c_v[1]; c[i] = 0.0; }", 0),
    ("V z sweep without the 2", "a[j] = -dt / rho * (2.0 * muS) / (dz * dz);", "a[j] = -dt / rho * muS / (dz * dz);", 0),
    ("V N wall added twice", "else if (fN == 2) { d[j] -= c[j] * gr->c_v[3]; c[j] = 0.0; }",
     "else if (fN == 2) { d[j] -= c[j] * 2.0 * gr->c_v[3]; c[j] = 0.0; }", 0),
    ("periodic W with the wrong sign", "else if (fW == 3) { d[i] -= a[i] * us3[j + nz * (nu - 1)]; a[i] = 0.0; }",
     "else if (fW == 3) { d[i] += a[i] * us3[j + nz * (nu - 1)]; a[i] = 0.0; }", 0),
]


def test_axisym_momentum_pins_catch_planted_mistakes(tmp_path):
    if not os.path.exists(LIB_PATH):
        pytest.skip("oracle not built")
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        libs = list(ex.map(lambda a: _mutant(str(tmp_path), *a),
                           [(k, m[1], m[2], m[3]) for k, m in enumerate(MUTANTS)]))
    missed = []
    for (desc, *_), lib in zip(MUTANTS, libs):
        O = Oracle(lib)
        caught = 0
        for check in CHECKS.values():
            try:
                with np.errstate(divide="ignore", invalid="ignore"):
                    check(O)
            except AssertionError:
                caught += 1
        if not caught:
            missed.append(desc)
    assert not missed, f"planted mistakes no pin catches: {missed}"
