"""Closed-form pins of the oracle's momentum predictor (convection, viscosity, RK stages), of the
curvature weight of the GFM jump and of the projection, and of the curvature's boundary rows
(reference src/mac2d.cpp, which cannot be compiled here: it needs TECIO), followed by a mutation
check: every pin is run against copies of the oracle with one planted mistake each, compiled on
the CPU, and each copy must fail at least one pin.

Each check_* takes an oracle object and raises AssertionError on a mismatch.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import LIB_PATH, Oracle, make_fluid, make_grid
from tests import scenarios as S

G = 3
HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "asd2_oracle.c")


def _close(got, want, atol, what):
    err = float(np.abs(np.asarray(got) - np.asarray(want)).max())
    assert err <= atol, f"{what}: max error {err:.3e} > {atol:.1e}"


def _heavy_fraction(pa, pb):
    """fraction of the segment between two samples of a planar level set (linear along the
    segment) on which phi >= 0"""
    pa = np.asarray(pa, dtype=float)
    pb = np.asarray(pb, dtype=float)
    out = np.where((pa >= 0) & (pb >= 0), 1.0, 0.0)
    cross = (pa >= 0) != (pb >= 0)
    s = np.abs(pa) + np.abs(pb)
    frac_a = np.where(cross, np.abs(pa) / np.where(s > 0, s, 1.0), 0.0)  # from a to the crossing
    return np.where(cross, np.where(pa >= 0, frac_a, 1.0 - frac_a), out)


# ----------------------------------------------------------------- convection (mac2d.cpp:308-463)
def check_convection_linear(O):
    """Linear u and v: u^2, u v-bar and v^2 are quadratic along every line, where each WENO5
    candidate is exact, so both biased derivatives equal the exact one and the LLF combine gives
    d(u^2)/dx + d(u v)/dy on U faces and d(u v)/dx + d(v^2)/dy on V faces.  The four-point
    averages of v onto U faces (and u onto V faces) are exact for linear fields.  Rows (U) and
    columns (V) within three cells of a wall are left out: the averaged velocity is zero in the
    ghost cells there (the reference fills it on interior faces only)."""
    nx, ny = 30, 24
    dx, dy = 1.0 / nx, 0.8 / ny
    gr = make_grid(nx, ny, dx, dy)
    Xu, Yu = S.faces_u(nx, ny, dx, dy)
    Xv, Yv = S.faces_v(nx, ny, dx, dy)
    a0, a1, a2 = 0.3, 0.7, -0.4
    b0, b1, b2 = -0.2, 0.5, 0.9
    u = np.ascontiguousarray(a0 + a1 * Xu + a2 * Yu)
    v = np.ascontiguousarray(b0 + b1 * Xv + b2 * Yv)
    cU = O.convection_u(gr, u, v)
    cV = O.convection_v(gr, u, v)
    v_at_u = b0 + b1 * Xu + b2 * Yu
    u_at_v = a0 + a1 * Xv + a2 * Yv
    want_u = 2 * u * a1 + (a2 * v_at_u + u * b2)
    want_v = (b1 * u_at_v + v * a1) + 2 * v * b2
    sU = (slice(G + 1, nx + G), slice(G + 3, ny + G - 3))
    sV = (slice(G + 3, nx + G - 3), slice(G + 1, ny + G))
    _close(cU[sU], want_u[sU], 1e-12, "U convection of linear fields")
    _close(cV[sV], want_v[sV], 1e-12, "V convection of linear fields")


def _kinked(t, t0, slope):
    """1 left of t0, 1 + slope (t - t0) right of it"""
    return np.where(t < t0, 1.0, 1.0 + slope * (t - t0))


def check_convection_upwind(O):
    """Transport velocity +1 (or -1) where the derivative is taken: the LLF combine with the
    signed velocity as its alpha, 1/2 (LP + LM) - alpha / 2 (LP - LM), reduces to the
    left-biased WENO derivative LM (right-biased LP for -1), i.e. the upwind one.  Profiles with a
    kink make the two biases differ; UnitHJWENO5 itself is pinned to the reference's own code by
    the golden vectors.  All four sweeps: U along x (F = u^2, alpha = u), U along y (F = u v-bar,
    alpha = v-bar), V along x (F = v u-bar, alpha = u-bar), V along y (F = v^2, alpha = v)."""
    nx, ny = 28, 26
    dx, dy = 1.0 / nx, 1.0 / ny
    gr = make_grid(nx, ny, dx, dy)
    Xu, Yu = S.faces_u(nx, ny, dx, dy)
    Xv, Yv = S.faces_v(nx, ny, dx, dy)
    lx, ly = nx + 2 * G, ny + 2 * G
    differs = []
    for s in (1.0, -1.0):
        # U along x: u = s on x < 0.5 (kink there), v = 0
        u = np.ascontiguousarray(s * _kinked(Xu, 0.5 + 0.25 * dx, 3.0))
        cU = O.convection_u(gr, u, np.zeros_like(u))
        for j in range(G, ny + G):
            FP, FM = O.unit_hjweno5(np.ascontiguousarray(u[:, j] ** 2), dx, nx)
            for i in range(G + 1, nx + G):
                if u[i, j] == s:
                    _close(cU[i, j], FM[i] if s > 0 else FP[i], 1e-12, "U convection, x sweep upwind")
                    differs.append(abs(FP[i] - FM[i]))
        # U along y: v = s (v-bar = s on interior U faces), u kinked in y, constant in x
        u = np.ascontiguousarray(0.5 + np.abs(Yu - 0.5 - 0.3 * dy))
        v = np.full((lx, ly), s)
        cU = O.convection_u(gr, u, np.ascontiguousarray(v))
        for i in range(G + 1, nx + G):
            F = np.zeros(ly)
            F[G:ny + G] = u[i, G:ny + G] * s
            FP, FM = O.unit_hjweno5(F, dy, ny)
            w = FM if s > 0 else FP
            _close(cU[i, G + 3:ny + G - 3], w[G + 3:ny + G - 3], 1e-12, "U convection, y sweep upwind")
            differs.append(np.abs(FP - FM)[G + 3:ny + G - 3].max())
        # V along x: u = s (u-bar = s on interior V faces), v kinked in x, constant in y
        v = np.ascontiguousarray(0.5 + np.abs(Xv - 0.5 - 0.3 * dx))
        cV = O.convection_v(gr, np.ascontiguousarray(np.full((lx, ly), s)), v)
        for j in range(G + 1, ny + G):
            F = np.zeros(lx)
            F[G:nx + G] = v[G:nx + G, j] * s
            FP, FM = O.unit_hjweno5(F, dx, nx)
            w = FM if s > 0 else FP
            _close(cV[G + 3:nx + G - 3, j], w[G + 3:nx + G - 3], 1e-12, "V convection, x sweep upwind")
            differs.append(np.abs(FP - FM)[G + 3:nx + G - 3].max())
        # V along y: v = s below y = 0.5 (kink there), u = 0
        v = np.ascontiguousarray(s * _kinked(Yv, 0.5 + 0.25 * dy, 3.0))
        cV = O.convection_v(gr, np.zeros_like(v), v)
        for i in range(G, nx + G):
            FP, FM = O.unit_hjweno5(np.ascontiguousarray(v[i, :] ** 2), dy, ny)
            for j in range(G + 1, ny + G):
                if v[i, j] == s:
                    _close(cV[i, j], FM[j] if s > 0 else FP[j], 1e-12, "V convection, y sweep upwind")
                    differs.append(abs(FP[j] - FM[j]))
    assert max(differs) > 1e-2, "the kinked profiles do not separate the two biases"


# ----------------------------------------------------------------- viscosity (mac2d.cpp:570-1041)
def _quad(X, Y, p):
    return p[0] + p[1] * X + p[2] * Y + p[3] * X * X + p[4] * X * Y + p[5] * Y * Y


def check_viscosity_quadratic(O):
    """Constant mu and rho, quadratic u and v: every difference quotient of the staggered stress
    divergence is exact, so the U term is mu / rho (2 u_xx + u_yy + v_xy) and the V term
    mu / rho (v_xx + 2 v_yy + u_xy) (dx != dy).  Faces next to a wall are left out (mu is zero in
    the ghost cells, as in the reference)."""
    nx, ny = 30, 24
    dx, dy = 1.0 / nx, 0.8 / ny
    gr = make_grid(nx, ny, dx, dy)
    Xu, Yu = S.faces_u(nx, ny, dx, dy)
    Xv, Yv = S.faces_v(nx, ny, dx, dy)
    A = (0.2, 0.5, -0.3, 0.7, -0.6, 0.4)
    B = (-0.1, 0.3, 0.8, -0.5, 0.25, 0.6)
    u = np.ascontiguousarray(_quad(Xu, Yu, A))
    v = np.ascontiguousarray(_quad(Xv, Yv, B))
    mu, rho = 0.3, 2.0
    fl = make_fluid(rho, rho, mu, mu, 0.0, 1e-3)
    phi = S.random_field(5, u.shape)  # any level set: both phases have the same rho and mu
    H = O.smooth_heaviside(gr, phi)
    dU = O.viscosity_u(gr, fl, u, v, phi, H)
    dV = O.viscosity_v(gr, fl, u, v, phi, H)
    uxx, uyy, uxy = 2 * A[3], 2 * A[5], A[4]
    vxx, vyy, vxy = 2 * B[3], 2 * B[5], B[4]
    _close(dU[G + 2:nx + G - 1, G + 1:ny + G - 1], mu / rho * (2 * uxx + uyy + vxy), 1e-11,
           "U viscosity of quadratic fields")
    _close(dV[G + 1:nx + G - 1, G + 2:ny + G - 1], mu / rho * (vxx + 2 * vyy + uxy), 1e-11,
           "V viscosity of quadratic fields")


def check_viscosity_face_density(O):
    """Constant mu, planar interfaces (heavy on either side), one-directional quadratics: the
    viscous term is c mu times the face factor 1 / (rho_H thetaH + rho_L (1 - thetaH)), thetaH the
    heavy fraction of the segment whose ends the reference samples: for the U term the cell
    centres either side of the face (x part, c = 4 for u = x^2) and the cell corners above and
    below it (y part, c = 2 for u = y^2); for the V term the corners either side (x part, c = 2
    for v = x^2) and the cell centres either side (y part, c = 4 for v = y^2)."""
    nx, ny = 24, 20
    dx, dy = 1.0 / nx, 1.0 / ny
    gr = make_grid(nx, ny, dx, dy)
    Xu, Yu = S.faces_u(nx, ny, dx, dy)
    Xv, Yv = S.faces_v(nx, ny, dx, dy)
    Xc, Yc = S.centres(nx, ny, dx, dy)
    mu, rH, rL = 0.2, 1000.0, 3.0
    fl = make_fluid(rH, rL, mu, mu, 0.0, 1e-3)
    x0, y0 = 0.5 + 0.37 * dx, 0.5 - 0.29 * dy
    zero = np.zeros_like(Xu)

    def factor(theta):
        return 1.0 / (rH * theta + rL * (1.0 - theta))

    for s in (1.0, -1.0):
        # vertical interface
        phi = np.ascontiguousarray(s * (x0 - Xc))
        H = O.smooth_heaviside(gr, phi)
        dU = O.viscosity_u(gr, fl, np.ascontiguousarray(Xu ** 2), zero, phi, H)
        th = _heavy_fraction(s * (x0 - (Xu - 0.5 * dx)), s * (x0 - (Xu + 0.5 * dx)))
        m = (slice(G + 2, nx + G - 1), slice(G + 1, ny + G - 1))
        _close(dU[m], (4 * mu * factor(th))[m], 1e-9, "U viscosity across a vertical interface")
        dV = O.viscosity_v(gr, fl, zero, np.ascontiguousarray(Xv ** 2), phi, H)
        th = _heavy_fraction(s * (x0 - (Xv - 0.5 * dx)), s * (x0 - (Xv + 0.5 * dx)))
        m = (slice(G + 1, nx + G - 1), slice(G + 2, ny + G - 1))
        _close(dV[m], (2 * mu * factor(th))[m], 1e-9, "V viscosity across a vertical interface")
        # horizontal interface
        phi = np.ascontiguousarray(s * (y0 - Yc))
        H = O.smooth_heaviside(gr, phi)
        dU = O.viscosity_u(gr, fl, np.ascontiguousarray(Yu ** 2), zero, phi, H)
        th = _heavy_fraction(s * (y0 - (Yu - 0.5 * dy)), s * (y0 - (Yu + 0.5 * dy)))
        m = (slice(G + 2, nx + G - 1), slice(G + 1, ny + G - 1))
        _close(dU[m], (2 * mu * factor(th))[m], 1e-9, "U viscosity across a horizontal interface")
        dV = O.viscosity_v(gr, fl, zero, np.ascontiguousarray(Yv ** 2), phi, H)
        th = _heavy_fraction(s * (y0 - (Yv - 0.5 * dy)), s * (y0 - (Yv + 0.5 * dy)))
        m = (slice(G + 1, nx + G - 1), slice(G + 2, ny + G - 1))
        _close(dV[m], (4 * mu * factor(th))[m], 1e-9, "V viscosity across a horizontal interface")
        # both pieces must have seen the interface
        assert np.unique(np.round(th[m], 12)).size > 2


def check_viscosity_symmetric(O):
    """With rho = 1 the viscous term is minus half the gradient of the discrete dissipation
    sum_cells mu (2 (du/dx)^2 + 2 (dv/dy)^2) + sum_corners mu_corner (du/dy + dv/dx)^2, so the
    map from the interior face velocities (walls and ghosts 0) to the U and V terms is a symmetric
    negative semi-definite matrix, for any mu field: this fixes which cells each corner average
    of mu takes, and that the U and V equations share them."""
    nx, ny = 6, 5
    dx, dy = 0.2, 0.17
    gr = make_grid(nx, ny, dx, dy)
    fl = make_fluid(1.0, 1.0, 2.0, 0.5, 0.0, 1e-3)
    rng = np.random.default_rng(11)
    H = np.ascontiguousarray(rng.uniform(0.0, 1.0, (nx + 2 * G, ny + 2 * G)))
    phi = np.ones_like(H)
    uk = [(i, j) for i in range(G + 1, nx + G) for j in range(G, ny + G)]
    vk = [(i, j) for i in range(G, nx + G) for j in range(G + 1, ny + G)]
    n = len(uk) + len(vk)
    L = np.zeros((n, n))
    for col in range(n):
        u = np.zeros_like(H)
        v = np.zeros_like(H)
        if col < len(uk):
            u[uk[col]] = 1.0
        else:
            v[vk[col - len(uk)]] = 1.0
        dU = O.viscosity_u(gr, fl, u, v, phi, H)
        dV = O.viscosity_v(gr, fl, u, v, phi, H)
        L[:, col] = [dU[k] for k in uk] + [dV[k] for k in vk]
    scale = np.abs(L).max()
    asym = np.abs(L - L.T).max()
    assert asym <= 1e-13 * scale, f"viscous operator not symmetric: {asym:.3e} of {scale:.3e}"
    ev = np.linalg.eigvalsh(0.5 * (L + L.T))
    assert ev.max() <= 1e-12 * scale, f"viscous operator not negative semi-definite: {ev.max():.3e}"


# ----------------------------------------------------------------- RK stages (mac2d.cpp:1081-1239)
def check_momentum_rk_linear(O):
    """With convection negligible (velocities of size 1e-7) and no gravity the right-hand side is
    the linear viscous operator L (walls and ghosts from the no-slip BCs, also linear), and each
    time scheme reproduces its Taylor polynomial exactly: Euler u + dt L u, Heun's RK2 adds
    dt^2/2 L^2 u, SSP-RK3 adds dt^3/6 L^3 u as well."""
    nx, ny = 12, 10
    dx, dy = 1.0 / nx, 1.0 / ny
    gr = make_grid(nx, ny, dx, dy)
    fl = make_fluid(1.0, 1.0, 1.0, 1.0, 0.0, 3e-4)
    phi = np.ones((nx + 2 * G, ny + 2 * G))
    Hs = O.smooth_heaviside(gr, phi)
    rng = np.random.default_rng(3)
    u = np.ascontiguousarray(rng.standard_normal(phi.shape) * 1e-7)
    v = np.ascontiguousarray(rng.standard_normal(phi.shape) * 1e-7)
    O.bc_u(gr, u)
    O.bc_v(gr, v)

    def Lop(a, b):
        a = O.bc_u(gr, np.ascontiguousarray(a.copy()))
        b = O.bc_v(gr, np.ascontiguousarray(b.copy()))
        return O.viscosity_u(gr, fl, a, b, phi, Hs), O.viscosity_v(gr, fl, a, b, phi, Hs)

    L1 = Lop(u, v)
    L2 = Lop(*L1)
    L3 = Lop(*L2)
    dt = fl.dt
    sU = (slice(G + 1, nx + G), slice(G, ny + G))
    sV = (slice(G, nx + G), slice(G + 1, ny + G))
    scale = max(np.abs(dt * L1[0]).max(), np.abs(dt * L1[1]).max())
    assert 0.1 < scale / 1e-7 < 10  # dt L is of order one: every Taylor term counts
    terms = {0: (1.0, 0.0, 0.0), 1: (1.0, 0.5, 0.0), 2: (1.0, 0.5, 1.0 / 6.0)}
    for order, (c1, c2, c3) in terms.items():
        uh, vh = O.intermediate_vel(gr, fl, order, phi, u, v, Hs)
        want_u = u + c1 * dt * L1[0] + c2 * dt ** 2 * L2[0] + c3 * dt ** 3 * L3[0]
        want_v = v + c1 * dt * L1[1] + c2 * dt ** 2 * L2[1] + c3 * dt ** 3 * L3[1]
        _close(uh[sU], want_u[sU], 1e-6 * scale, f"U after time scheme {order}")
        _close(vh[sV], want_v[sV], 1e-6 * scale, f"V after time scheme {order}")


# ----------------------------------------------------------------- curvature weight (mac2d.cpp:1345-1476, 1740-1849)
def _planar_case(n, s, vertical):
    dx = 1.0 / n
    gr = make_grid(n, n, dx, dx)
    Xc, Yc = S.centres(n, n, dx, dx)
    x0, y0 = 0.5 + 0.3 * dx, 0.5 - 0.2 * dx
    phi = np.ascontiguousarray(s * (x0 - Xc) if vertical else s * (y0 - Yc))
    kap_lin = np.ascontiguousarray(0.7 + 1.3 * Xc - 0.9 * Yc)
    return gr, dx, Xc, Yc, x0, y0, phi, kap_lin


def check_gfm_kappa_weight(O):
    """Linear curvature, planar interface between cell centres (heavy on either side, vertical and
    horizontal): the curvature of each crossed face's jump term is the curvature at the interface
    point, kappa(x_Gamma) (linear interpolation to the point where phi vanishes).  Measured as the
    ratio of the right-hand side with that curvature to the one with curvature 1, which leaves out
    every other factor of the jump term (those are pinned by the constant-curvature solve)."""
    n = 16
    for vertical in (True, False):
        for s in (1.0, -1.0):
            gr, dx, Xc, Yc, x0, y0, phi, kap = _planar_case(n, s, vertical)
            H = O.heaviside(gr, phi)
            fl = make_fluid(1000.0, 2.0, 1.0, 1.0, 0.5, 1e-3)
            zero = np.zeros_like(phi)
            _, _, _, r_lin = O.gfm_assemble(gr, fl, phi, kap, H, zero)
            _, _, _, r_one = O.gfm_assemble(gr, fl, phi, np.ones_like(phi), H, zero)
            a = G + int(np.floor((x0 if vertical else y0) / dx - 0.5))  # last cell before the interface
            for k in (a, a + 1):
                for m in range(G + 1, n + G - 1):
                    c = (k, m) if vertical else (m, k)
                    want = 0.7 + 1.3 * x0 - 0.9 * Yc[c] if vertical else 0.7 + 1.3 * Xc[c] - 0.9 * y0
                    assert r_one[c] != 0.0
                    _close(r_lin[c] / r_one[c], want, 1e-12, "GFM jump curvature at the interface")


def check_update_vel_kappa_weight(O):
    """The projection's jump term on the crossed faces carries the same interface curvature:
    with u* = v* = 0 and p = 0 the new face velocity is the jump term alone, and its ratio between
    a linear curvature and curvature 1 is kappa(x_Gamma)."""
    n = 16
    for vertical in (True, False):
        for s in (1.0, -1.0):
            gr, dx, Xc, Yc, x0, y0, phi, kap = _planar_case(n, s, vertical)
            H = O.heaviside(gr, phi)
            fl = make_fluid(1000.0, 2.0, 1.0, 1.0, 0.5, 1e-3)
            zero = np.zeros_like(phi)
            out = []
            for kk in (kap, np.ones_like(phi)):
                u, v = np.zeros_like(phi), np.zeros_like(phi)
                O.update_vel(gr, fl, u, v, zero, zero, zero, phi, np.ascontiguousarray(kk), H)
                out.append(u if vertical else v)
            f = G + int(np.floor((x0 if vertical else y0) / dx - 0.5)) + 1  # the crossed face
            for m in range(G + 1, n + G - 1):
                c = (f, m) if vertical else (m, f)
                want = 0.7 + 1.3 * x0 - 0.9 * Yc[c] if vertical else 0.7 + 1.3 * Xc[c] - 0.9 * y0
                assert out[1][c] != 0.0
                _close(out[0][c] / out[1][c], want, 1e-12, "projection jump curvature at the interface")


# ----------------------------------------------------------------- curvature (mac2d.cpp:221-252)
def check_kappa_quadratic(O):
    """phi quadratic over the whole padded array: all central differences are exact, so away from
    the first and last rows kappa = (phi_x^2 phi_yy - 2 phi_x phi_y phi_xy + phi_y^2 phi_xx) /
    |grad phi|^3 exactly.  On the rows j = g and j = ny + g - 1 the reference takes phi_xy from
    stored phi_x values one row out, which are zero outside the interior: phi_xy there is
    phi_x(y + dy) / (2 dy) and -phi_x(y - dy) / (2 dy).  Values beyond 1 / min(dx, dy) are clipped
    to it, keeping their sign."""
    nx, ny = 20, 16
    dx, dy = 1.0 / nx, 0.9 / ny
    gr = make_grid(nx, ny, dx, dy)
    X, Y = S.centres(nx, ny, dx, dy)
    p = (0.3, 0.8, -0.5, 0.4, -0.3, 0.25)
    phi = np.ascontiguousarray(_quad(X, Y, p))
    kap = O.update_kappa(gr, phi)

    def px(x, y):
        return p[1] + 2 * p[3] * x + p[4] * y

    def py(x, y):
        return p[2] + p[4] * x + 2 * p[5] * y

    fxx, fyy, fxy = 2 * p[3], 2 * p[5], p[4]
    fx, fy = px(X, Y), py(X, Y)

    def curv(fxy_):
        return (fx * fx * fyy - 2 * fx * fy * fxy_ + fy * fy * fxx) / (fx * fx + fy * fy) ** 1.5

    want = curv(fxy)
    assert np.abs(want[G:nx + G, G:ny + G]).max() < 1.0 / max(dx, dy)
    _close(kap[G:nx + G, G + 1:ny + G - 1], want[G:nx + G, G + 1:ny + G - 1], 1e-10, "interior curvature")
    low = curv(px(X, Y + dy) / (2 * dy))
    high = curv(-px(X, Y - dy) / (2 * dy))
    _close(kap[G:nx + G, G], low[G:nx + G, G], 1e-10, "curvature on the first row")
    _close(kap[G:nx + G, ny + G - 1], high[G:nx + G, ny + G - 1], 1e-10, "curvature on the last row")
    assert np.all(kap[:G] == 0) and np.all(kap[:, :G] == 0)  # only the interior is written
    # clipping: phi = |x - c|^2 has level-set curvature 1 / |x - c| exactly in the discrete
    # formula; within 0.8 of c it exceeds 1 / min(dx, dy) = 1 / 0.8 and is clipped to it
    nx2, ny2 = 12, 10
    gr2 = make_grid(nx2, ny2, 1.0, 0.8)
    X2, Y2 = S.centres(nx2, ny2, 1.0, 0.8)
    rr = np.sqrt((X2 - 6.0) ** 2 + (Y2 - 4.0) ** 2)
    k2 = O.update_kappa(gr2, np.ascontiguousarray(rr ** 2))
    m = (slice(G, nx2 + G), slice(G + 1, ny2 + G - 1))
    want2 = np.minimum(1.0 / rr, 1.0 / 0.8)
    assert np.sum(1.0 / rr[m] > 1.0 / 0.8) >= 4
    _close(k2[m], want2[m], 1e-12, "clipped curvature")


CHECKS = {
    "convection_linear": check_convection_linear,
    "convection_upwind": check_convection_upwind,
    "viscosity_quadratic": check_viscosity_quadratic,
    "viscosity_face_density": check_viscosity_face_density,
    "viscosity_symmetric": check_viscosity_symmetric,
    "momentum_rk_linear": check_momentum_rk_linear,
    "gfm_kappa_weight": check_gfm_kappa_weight,
    "update_vel_kappa_weight": check_update_vel_kappa_weight,
    "kappa_quadratic": check_kappa_quadratic,
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_oracle_pin(oracle, name):
    CHECKS[name](oracle)


# ----------------------------------------------------------------- planted mistakes
# (description, text in oracle/asd2_oracle.c, replacement, which occurrence (0-based))
MUTANTS = [
    ("LLF sign of the x bias, U convection", "- aX * (0.5 * (LXP[c] - LXM[c]))",
     "+ aX * (0.5 * (LXP[c] - LXM[c]))", 0),
    ("LLF sign of the y bias, V convection", "- aY * (0.5 * (LYP[c] - LYM[c]))",
     "+ aY * (0.5 * (LYP[c] - LYM[c]))", 1),
    ("v averaged from the wrong column onto U faces", "v[IDX(gr, i - 1, j + 1)] + v[IDX(gr, i, j + 1)]) * 0.25",
     "v[IDX(gr, i + 1, j + 1)] + v[IDX(gr, i, j + 1)]) * 0.25", 0),
    ("u averaged from the wrong row onto V faces", "u[IDX(gr, i + 1, j - 1)] + u[IDX(gr, i + 1, j)]) * 0.25",
     "u[IDX(gr, i + 1, j + 1)] + u[IDX(gr, i + 1, j)]) * 0.25", 0),
    ("normal stress without its factor 2, U viscosity", "const double visX = 2.0 * (muU_X_E - muU_X_W) / dx;",
     "const double visX = (muU_X_E - muU_X_W) / dx;", 0),
    ("north corner mu from the wrong cell, U viscosity",
     "0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j + 1) + M_(i, j + 1)) * (uN - uM) / dy",
     "0.25 * (M_(i, j) + M_(i - 1, j) + M_(i + 1, j + 1) + M_(i, j + 1)) * (uN - uM) / dy", 0),
    ("east corner mu from the wrong cell, V viscosity",
     "0.25 * (M_(i, j) + M_(i + 1, j) + M_(i + 1, j - 1) + M_(i, j - 1)) * (vE - vM) / dx",
     "0.25 * (M_(i, j) + M_(i + 1, j) + M_(i + 1, j + 1) + M_(i, j - 1)) * (vE - vM) / dx", 0),
    ("x density used for the y part, V viscosity", "dV[IDX(gr, i, j)] = visX * iWE + visY * iSN;",
     "dV[IDX(gr, i, j)] = visX * iWE + visY * iWE;", 0),
    ("heavy fraction measured from the light end (U viscosity face density)",
     "const double th = fabs(a) / (fabs(a) + fabs(b));\n        return 1.0 / (rL * rH)",
     "const double th = fabs(b) / (fabs(a) + fabs(b));\n        return 1.0 / (rL * rH)", 0),
    ("heavy fraction measured from the light end (V viscosity face density)",
     "const double th = fabs(a) / (fabs(a) + fabs(b));\n        return rH * th",
     "const double th = fabs(b) / (fabs(a) + fabs(b));\n        return rH * th", 0),
    ("harmonic branch weights swapped", "return 1.0 / (rL * rH) / (1.0 / rL * th + 1.0 / rH * (1.0 - th));",
     "return 1.0 / (rL * rH) / (1.0 / rH * th + 1.0 / rL * (1.0 - th));", 0),
    ("RK3 second-stage weights", "0.75 * u[C_] + 0.25 * FU2[C_]", "0.5 * u[C_] + 0.5 * FU2[C_]", 0),
    ("RK2 final average", "UINT uhat[C_] = 0.5 * u[C_] + 0.5 * FU2[C_];",
     "UINT uhat[C_] = 0.5 * u[C_] + 0.5 * FU1[C_];", 0),
    ("curvature weight taken from the far side", "*theta = nb_first ? fabs(lsA) / s : fabs(lsB) / s;",
     "*theta = nb_first ? fabs(lsB) / s : fabs(lsA) / s;", 0),
    ("curvature weight on the east face from the wrong neighbour", "kE = K_(i, j) * th + K_(i + 1, j) * (1.0 - th);",
     "kE = K_(i, j) * (1.0 - th) + K_(i + 1, j) * th;", 0),
    ("projection curvature weight of the south face",
     "const double kEff = kappa[IDX(gr, i, j)] * th + kappa[IDX(gr, i, j - 1)] * (1.0 - th);",
     "const double kEff = kappa[IDX(gr, i, j - 1)] * th + kappa[IDX(gr, i, j)] * (1.0 - th);", 0),
    ("sign of the curvature cross term", "- 2.0 * dX[c] * dY[c] * (dX[IDX(gr, i, j + 1)]",
     "+ 2.0 * dX[c] * dY[c] * (dX[IDX(gr, i, j + 1)]", 0),
    ("phi_x also stored outside the interior (no boundary artefact)",
     "for (int j = g; j < ny + g; j++) {\n            const int64_t c = IDX(gr, i, j);\n            dX[c]",
     "for (int j = g - 1; j < ny + g + 1; j++) {\n            const int64_t c = IDX(gr, i, j);\n            dX[c]", 0),
]


def _mutant(tmp, k, old, new, occ):
    with open(ORACLE_SRC) as f:
        text = f.read()
    pos = -1
    for _ in range(occ + 1):
        pos = text.find(old, pos + 1)
        assert pos >= 0, f"mutation site not found: {old!r}"
    text = text[:pos] + new + text[pos + len(old):]
    d = os.path.join(tmp, f"m{k}")
    os.makedirs(d)
    src = os.path.join(d, "asd2_oracle.c")
    with open(src, "w") as f:
        f.write(text)
    shutil.copy(os.path.join(os.path.dirname(ORACLE_SRC), "asd2_oracle.h"), d)
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-D_DEFAULT_SOURCE",
                    "-shared", "-o", lib, src, "-lm"], check=True)
    return lib


def test_pins_catch_planted_mistakes(tmp_path):
    """Every planted mistake makes at least one pin fail; the unmodified source passes them all
    (test_oracle_pin)."""
    if not os.path.exists(LIB_PATH):
        pytest.skip("oracle not built")
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        libs = list(ex.map(lambda a: _mutant(str(tmp_path), *a), [(k, m[1], m[2], m[3]) for k, m in enumerate(MUTANTS)]))
    missed = []
    for (desc, *_), lib in zip(MUTANTS, libs):
        O = Oracle(lib)
        caught = []
        for name, check in CHECKS.items():
            try:
                check(O)
            except AssertionError:
                caught.append(name)
        if not caught:
            missed.append(desc)
    assert not missed, f"planted mistakes no pin catches: {missed}"
