"""Closed-form pins of the oracle's axisymmetric restatement (MACSolver2DAxisym, reference
src/mac2dAxisym.cpp, which cannot be compiled here: it needs TECIO and MKL), with planted
mistakes compiled on the CPU that each pin set must catch.

* single phase, no surface tension: the assembled rows are the r-weighted five-point operator,
  W / E couplings -r_{i-/+1/2} / (rho dr^2), S / N couplings -r_i / (rho dz^2), diagonal = minus
  the sum of the kept couplings; a Dirichlet wall adds twice its constant times the dropped
  coupling to the right-hand side;
* a cylindrical interface phi = R - r (linear in r): across the interface the W / E coupling uses
  the interface radius, which is R exactly, and the heavy fraction of the face; the jump term of
  a constant curvature follows;
* the divergences of u = a r, v = b z are 2a r_i + b r_i (for the Poisson right-hand side) and
  2a + b;
* the time step of uniform fields, and the tridiagonal solve against a dense solve.
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import LIB_PATH, Oracle, make_fluid, make_grid
from tests.test_oracle_pins import _mutant

G = 3


def _grid(nr, nz, dr, dz, bc_p=("axisym", "neumann", "neumann", "neumann"), c_p=(0.0,) * 4):
    return make_grid(nr, nz, dr, dz, bc_p=bc_p, c_p=c_p)


def _radii(nr, dr, base):
    i = np.arange(nr + 2 * G)
    return base + (i + 0.5 - G) * dr, base + (i - G) * dr  # centres, W faces


def check_single_phase(O):
    nr, nz, dr, dz, base, rho = 9, 7, 0.1, 0.07, 0.0, 3.0
    gr = _grid(nr, nz, dr, dz, bc_p=("axisym", "neumann", "neumann", "dirichlet"), c_p=(0, 0, 0, 0.4))
    fl = make_fluid(rho, rho, 1.0, 1.0, 0.0, 0.01)
    shp = (nr + 2 * G, nz + 2 * G)
    ls = np.ones(shp)
    z = np.zeros(shp)
    div = np.random.default_rng(1).standard_normal(shp)
    ax, ay, ae, dg, rhs = O.gfm_assemble_axisym(gr, fl, base, ls, z, z, div)
    rc, rw = _radii(nr, dr, base)
    I, J = slice(G, G + nr), slice(G, G + nz)
    W = -rw[:, None] / (rho * dr * dr) * np.ones((1, shp[1]))
    S = -rc[:, None] / (rho * dz * dz) * np.ones((1, shp[1]))
    wantW = W[I, J].copy()
    wantW[0, :] = 0.0  # axis: neighbour dropped
    wantS = S[I, J].copy()
    wantS[:, 0] = 0.0
    np.testing.assert_allclose(ax[I, J], wantW, rtol=1e-14, atol=0)
    np.testing.assert_allclose(ay[I, J], wantS, rtol=1e-14, atol=0)
    E = -(rw + dr)[:, None] / (rho * dr * dr) * np.ones((1, shp[1]))
    wantE = E[I, J].copy()
    wantE[-1, :] = 0.0
    np.testing.assert_allclose(ae[I, J], wantE, rtol=1e-14, atol=0)
    # diagonal: minus the sum of all four couplings, the Neumann / axis ones removed again;
    # the north Dirichlet coupling stays on the diagonal
    full = -(W[I, J] + E[I, J]) - 2 * S[I, J]
    full[0, :] += W[I, J][0, :]
    full[-1, :] += E[I, J][-1, :]
    full[:, 0] += S[I, J][:, 0]
    np.testing.assert_allclose(dg[I, J], full, rtol=1e-13, atol=0)
    want_rhs = -div[I, J].copy()
    want_rhs[:, -1] -= S[I, J][:, -1] * (2.0 * 0.4)
    np.testing.assert_allclose(rhs[I, J], want_rhs, rtol=1e-13, atol=1e-15)
    # a WALL side is Neumann-like in the axisymmetric assembly
    gw = _grid(nr, nz, dr, dz, bc_p=("axisym", "wall", "neumann", "dirichlet"), c_p=(0, 0.7, 0, 0.4))
    for a, b in zip(O.gfm_assemble_axisym(gw, fl, base, ls, z, z, div), (ax, ay, ae, dg, rhs)):
        np.testing.assert_array_equal(a, b)


def check_cylinder_interface(O):
    """heavy inside (phi = R - r) and heavy outside (phi = r - R): both crossing branches of the
    W and E faces"""
    nr, nz, dr, dz, base = 12, 6, 0.1, 0.1, 0.05
    R = 0.63
    gr = _grid(nr, nz, dr, dz, bc_p=("neumann",) * 4)
    rhoH, rhoL, sigma, dt, kap0 = 10.0, 1.0, 0.3, 0.01, 2.5
    fl = make_fluid(rhoH, rhoL, 1.0, 1.0, sigma, dt)
    shp = (nr + 2 * G, nz + 2 * G)
    rc, rw = _radii(nr, dr, base)
    for sgn in (1.0, -1.0):
        ls = sgn * (R - rc)[:, None] * np.ones((1, shp[1]))
        kap = np.full(shp, kap0)
        H = (ls >= 0).astype(float)
        z = np.zeros(shp)
        ax, ay, ae, dg, rhs = O.gfm_assemble_axisym(gr, fl, base, ls, kap, H, z)
        ic = int(np.argmax(rc > R))  # first padded index outside the interface
        th_in = (R - rc[ic - 1]) / (rc[ic] - rc[ic - 1])  # inner cell's share of the face
        thH = th_in if sgn > 0 else 1.0 - th_in
        rho_eff = rhoH * thH + rhoL * (1 - thH)
        rho_out = rhoL if sgn > 0 else rhoH
        j = G + 2
        np.testing.assert_allclose(ax[ic, j], -R / (rho_eff * dr * dr), rtol=1e-13)
        np.testing.assert_allclose(ae[ic - 1, j], -R / (rho_eff * dr * dr), rtol=1e-13)
        # the west neighbour's E coupling and this cell's W coupling agree to rounding (all sides
        # Neumann: the coupling into the pinned cell (g, g) is dropped from the row of (g + 1, g))
        e_west = ae[G:G + nr - 1, G:G + nz].copy()
        e_west[0, 0] = 0.0
        np.testing.assert_allclose(e_west, ax[G + 1:G + nr, G:G + nz], rtol=1e-14, atol=0)
        # one-phase faces keep the face radius
        np.testing.assert_allclose(ax[ic + 1, j], -rw[ic + 1] / (rho_out * dr * dr), rtol=1e-14)
        # jump terms of the crossing faces (H jumps by -sgn across the W face of cell ic)
        dH = -1.0 if sgn > 0 else 1.0
        FW = -dt * (-sigma * kap0) * dH * (R / rho_eff) / (dr * dr)
        np.testing.assert_allclose(rhs[ic, j], -FW, rtol=1e-12)
        FE = dt * (-sigma * kap0) * dH * (R / rho_eff) / (dr * dr)
        np.testing.assert_allclose(rhs[ic - 1, j], -FE, rtol=1e-12)


def check_axis_clamp(O):
    """an interface half a cell from the axis: the W face of the first cell across it takes the
    interface radius clamped at 0 (the ghost centre lies at r < 0)"""
    nr, nz, dr, dz = 6, 4, 0.1, 0.1
    gr = _grid(nr, nz, dr, dz, bc_p=("neumann",) * 4)
    fl = make_fluid(5.0, 1.0, 1.0, 1.0, 0.0, 0.01)
    shp = (nr + 2 * G, nz + 2 * G)
    rc, _ = _radii(nr, dr, 0.0)
    # heavy in the ghost column (r = -dr/2, phi >= 0), light from the first cell on, with the
    # interface nearer the ghost centre: r_M theta + r_W (1 - theta) < 0
    z = np.zeros(shp)
    kap = np.full(shp, 3.0)
    fl = make_fluid(5.0, 1.0, 1.0, 1.0, 0.5, 0.01)
    for sgn in (1.0, -1.0):  # both crossing branches of the W face
        ls = np.full(shp, -sgn)
        ls[G - 1, :] = 0.01 * sgn
        H = (ls >= 0).astype(float)
        ax, ay, ae, dg, rhs = O.gfm_assemble_axisym(gr, fl, 0.0, ls, kap, H, z)
        th = 0.01 / 1.01
        r_int = rc[G] * th + rc[G - 1] * (1 - th)
        assert r_int < 0
        # the W face's jump term carries the clamped radius 0; no other face of the row crosses
        np.testing.assert_array_equal(rhs[G, G + 1:G + nz - 1], 0.0)
        # the wall row's diagonal is the E and S / N part alone
        rho_s = 1.0 if sgn > 0 else 5.0  # S / N faces in the cell's phase
        want = -ae[G, G + 1] + 2 * rc[G] / (rho_s * dz * dz)
        np.testing.assert_allclose(dg[G, G + 1], want, rtol=1e-13)


def check_divergences(O):
    nr, nz, dr, dz, base = 7, 5, 0.2, 0.1, 0.3
    gr = _grid(nr, nz, dr, dz)
    a, b = 1.7, -0.6
    shp = (nr + 2 * G, nz + 2 * G)
    rc, rw = _radii(nr, dr, base)
    zc = (np.arange(shp[1]) + 0.5 - G) * dz
    zs = (np.arange(shp[1]) - G) * dz
    u = (a * rw)[:, None] * np.ones((1, shp[1]))
    v = np.ones((shp[0], 1)) * (b * zs)[None, :]
    d4p = O.divergence_axisym(gr, base, u, v, True)
    d = O.divergence_axisym(gr, base, u, v, False)
    I, J = slice(G, G + nr), slice(G, G + nz)
    np.testing.assert_allclose(d4p[I, J], ((2 * a + b) * rc[:, None] * np.ones((1, shp[1])))[I, J], rtol=1e-13)
    np.testing.assert_allclose(d[I, J], 2 * a + b, rtol=1e-13)
    assert np.all(d[:G] == 0) and np.all(d4p[:, :G] == 0)
    del zc


def check_update_dt(O):
    nr, nz, dr, dz = 8, 6, 0.05, 0.1
    gr = _grid(nr, nz, dr, dz)
    fl = make_fluid(1000.0, 10.0, 1.0, 0.1, 0.07, 0.01, g=9.8, cfl=0.4)
    shp = (nr + 2 * G, nz + 2 * G)
    u = np.full(shp, 0.3)
    v = np.full(shp, -0.8)
    kap = np.zeros(shp)
    kap[G + 2, G + 3] = -40.0
    dt = O.update_dt_axisym(gr, fl, u, v, kap)
    C = 0.5 * 0.09 / dr + 0.5 * 0.64 / dz
    Gf = max(np.sqrt(9.8 / dz), np.sqrt(9.8 / dr))
    S = np.sqrt(0.07 * 40.0 / (10.0 * min(dr, dz) ** 2))
    want = 0.4 / (0.5 * (C + np.sqrt(C * C + 4 * Gf * Gf + 4 * S * S)))
    np.testing.assert_allclose(dt, want, rtol=1e-14)


def check_tdma(O):
    rng = np.random.default_rng(4)
    for n in (2, 3, 17):
        a = rng.uniform(-1, 1, n)
        c = rng.uniform(-1, 1, n)
        b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2.0, n)
        d = rng.standard_normal(n)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        np.testing.assert_allclose(O.tdma(a, b, c, d), np.linalg.solve(M, d), rtol=1e-12, atol=1e-14)


def check_kappa_quadratic(O):
    """phi quadratic in (r, z): the central differences are exact, so interior rows follow the
    closed form (phi_rr phi_z^2 - 2 phi_rz phi_r phi_z + phi_zz phi_r^2 + phi_r |grad phi|^2 / r)
    / |grad phi|^3; on the axis row phi_r is the one-sided difference, exact at r + dr / 2, and the
    cross term sees phi_z = 0 west of the interior (the reference stores phi_z on the interior only)"""
    nr, nz, dr, dz, base = 10, 9, 0.05, 0.04, 0.0
    gr = _grid(nr, nz, dr, dz)
    p = (0.1, -0.3, 0.8, 0.2, -0.5, 0.35)  # a0, a1 r, a2 r^2, b1 z, b2 z^2, c r z
    shp = (nr + 2 * G, nz + 2 * G)
    rc, _ = _radii(nr, dr, base)
    zc = (np.arange(shp[1]) + 0.5 - G) * dz
    Rr, Zz = np.meshgrid(rc, zc, indexing="ij")
    phi = p[0] + p[1] * Rr + p[2] * Rr ** 2 + p[3] * Zz + p[4] * Zz ** 2 + p[5] * Rr * Zz
    k = O.update_kappa_axisym(gr, base, phi)
    pr = p[1] + 2 * p[2] * Rr + p[5] * Zz
    pz = p[3] + 2 * p[4] * Zz + p[5] * Rr
    S2 = pr * pr + pz * pz
    want = (2 * p[2] * pz ** 2 - 2 * p[5] * pr * pz + 2 * p[4] * pr ** 2 + pr * S2 / Rr) / S2 ** 1.5
    cap = 1.0 / min(dr, dz)
    want = np.sign(want) * np.minimum(np.abs(want), cap)
    I, J = slice(G + 1, G + nr - 1), slice(G, G + nz)
    np.testing.assert_allclose(k[I, J], want[I, J], rtol=1e-10, atol=1e-12)
    # axis row
    i = G
    prA = p[1] + p[2] * (2 * Rr[i] + dr) + p[5] * Zz[i]
    S2A = prA ** 2 + pz[i] ** 2
    cross = (pz[i + 1] - 0.0) / (2 * dr)
    wantA = (2 * p[2] * pz[i] ** 2 - 2 * cross * prA * pz[i] + 2 * p[4] * prA ** 2 + prA * S2A / Rr[i]) / S2A ** 1.5
    wantA = np.sign(wantA) * np.minimum(np.abs(wantA), cap)
    np.testing.assert_allclose(k[i, G:G + nz], wantA[G:G + nz], rtol=1e-10, atol=1e-12)


CHECKS = {f.__name__: f for f in (check_single_phase, check_cylinder_interface, check_axis_clamp,
                                  check_divergences, check_update_dt, check_tdma, check_kappa_quadratic)}


@pytest.mark.parametrize("name", list(CHECKS))
def test_axisym_pin(oracle, name):
    CHECKS[name](oracle)


MUTANTS = [
    ("W face radius across the interface", "rEff = smax(rPM * theta + rPW * (1.0 - theta), 0.0);\n            } else",
     "rEff = rPWHalf;\n            } else", 0),
    ("E face interface radius from the wrong end", "rEff = rPM * theta + rPE * (1.0 - theta);",
     "rEff = rPE * theta + rPM * (1.0 - theta);", 1),
    ("S face radius at the half index", "double pS = rPM / rhoEff;", "double pS = rPWHalf / rhoEff;", 0),
    ("no clamp at the axis", "rEff = smax(rPM * theta + rPW * (1.0 - theta), 0.0);\n            }\n            rhoEff",
     "rEff = rPM * theta + rPW * (1.0 - theta);\n            }\n            rhoEff", 0),
    ("Dirichlet wall once", "rhs[c] -= p / h2 * (-0.0 + 2.0 * gr->c_p[s]);", "rhs[c] -= p / h2 * gr->c_p[s];", 0),
    ("WALL treated as Dirichlet", "if (k == OR_BC_NEUMANN || k == OR_BC_AXISYM || k == OR_BC_WALL) {\n                    if (s == 0)",
     "if (k == OR_BC_NEUMANN || k == OR_BC_AXISYM) {\n                    if (s == 0)", 0),
    ("Poisson divergence with the centre radius", "(rPEHalf * u[IDX(gr, i + 1, j)] - rPWHalf * u[IDX(gr, i, j)]) / dr",
     "rPM * (u[IDX(gr, i + 1, j)] - u[IDX(gr, i, j)]) / dr", 0),
    ("hoop term without the 1/2", "div[IDX(gr, i, j)] = 0.5 * (u[IDX(gr, i, j)] + u[IDX(gr, i + 1, j)]) / rPM",
     "div[IDX(gr, i, j)] = (u[IDX(gr, i, j)] + u[IDX(gr, i + 1, j)]) / rPM", 0),
    ("capillary bound from the grid instead of kappa", "const double S = sqrt((fl->sigma * kA)",
     "const double S = sqrt((fl->sigma / smin(dr, dz))", 0),
    ("no hoop term in the curvature", "+ dR[c] * pow(S[c], 2.0) / rPM) / pow(S[c], 3.0);",
     ") / pow(S[c], 3.0);", 0),
    ("central r difference on the axis row", "if (i == g) dR[c] = (ls[IDX(gr, i + 1, j)] - ls[c]) / dr;", "", 0),
    ("hoop term at the W face radius", "const double rPM = base_r + (i + 0.5 - g) * dr;\n            double k =",
     "const double rPM = base_r + (i - g) * dr;\n            double k =", 0),
    ("tridiagonal forward sweep with the new c", "d[i] = (d[i] - a[i] * d[i - 1]) / (b[i] - a[i] * c[i - 1]);",
     "d[i] = (d[i] - a[i] * d[i - 1]) / (b[i] - a[i] * c[i]);", 0),
]


def test_axisym_pins_catch_planted_mistakes(tmp_path):
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
                check(O)
            except AssertionError:
                caught += 1
        if not caught:
            missed.append(desc)
    assert not missed, f"planted mistakes no pin catches: {missed}"
