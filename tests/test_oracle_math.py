"""Pins of the FP64 oracle by mathematics where the reference cannot be run here
(src/mac2d.cpp needs TECIO, src/poisson2d.cpp needs MKL): closed forms, invariants and
direct solves, each chosen so that a dropped term, a wrong sign or a transposed index
fails it.  Level-set functions get closed-form pins in addition to the golden vectors."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import make_fluid, make_grid
from tests import scenarios as S
from tests.gfm_common import gfm_case, interior, random_spd_csr, rhs_vector, scatter, stencil_to_csr


# ----------------------------------------------------------------- WENO5 / advection
def test_weno5_exact_on_cubics(oracle):
    """Every HJ-WENO5 candidate stencil is exact for cubic F, so FP = FM = F'(x_i)."""
    n, d = 50, 0.02
    x = (np.arange(n + 6) - 3.0) * d
    F = np.ascontiguousarray(0.3 * x ** 3 - 1.2 * x ** 2 + 0.7 * x + 2.0)
    FP, FM = oracle.unit_hjweno5(F, d, n)
    xs = x[3:n + 3]
    exact = 0.9 * xs ** 2 - 2.4 * xs + 0.7
    np.testing.assert_allclose(FP[3:n + 3], exact, rtol=0, atol=1e-11)
    np.testing.assert_allclose(FM[3:n + 3], exact, rtol=0, atol=1e-11)
    assert np.all(FP[:3] == 0) and np.all(FM[n + 3:] == 0)  # only [g, n + g) is written


def test_weno5_fifth_order(oracle):
    errs = []
    for n in (40, 80, 160):
        d = 1.0 / n
        x = (np.arange(n + 6) - 3.0) * d
        F = np.ascontiguousarray(np.sin(2 * np.pi * x) + 0.5 * np.cos(np.pi * x))
        FP, FM = oracle.unit_hjweno5(F, d, n)
        xs = x[3:n + 3]
        exact = 2 * np.pi * np.cos(2 * np.pi * xs) - 0.5 * np.pi * np.sin(np.pi * xs)
        errs.append(max(np.abs(FP[3:n + 3] - exact).max(), np.abs(FM[3:n + 3] - exact).max()))
    rate = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert min(rate) > 4.0, (errs, rate)


def test_rk3_step_on_linear_level_set(oracle):
    """phi = a x + b y + c with uniform (u, v): L = u a + v b exactly, so one SSP-RK3 step
    gives phi - dt (u a + v b) away from the walls (pins the stage weights)."""
    nx = ny = 40
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    X, Y = S.centres(nx, ny, dx, dx)
    a, b, c0 = 0.6, -0.8, 0.1
    phi = np.ascontiguousarray(a * X + b * Y + c0)
    u = np.full(phi.shape, 1.0)  # the LLF combine uses the signed velocity: |u| = 1 keeps it upwind
    v = np.full(phi.shape, -1.0)
    dt = 0.3 * dx
    out = oracle.solve_levelset(gr, phi.copy(), u, v, dt)
    m = slice(12, 28)
    np.testing.assert_allclose(out[m, m], (phi - dt * (1.0 * a - 1.0 * b))[m, m], atol=1e-13)


def test_translation_matches_exact_shift(oracle):
    nx = ny = 64
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    X, Y = S.centres(nx, ny, dx, dx)
    phi = np.ascontiguousarray(0.05 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y))
    u = np.full(phi.shape, 1.0)
    v = np.zeros(phi.shape)
    dt = 0.25 * dx
    out = phi.copy()
    for _ in range(8):
        oracle.solve_levelset(gr, out, u, v, dt)
    exact = 0.05 * np.sin(2 * np.pi * (X - 8 * dt)) * np.cos(2 * np.pi * Y)
    m = slice(20, 44)
    assert np.abs(out[m, m] - exact[m, m]).max() < 1e-5


# ----------------------------------------------------------------- reinitialisation
def test_subcell_reinit_keeps_planar_sdf(oracle):
    """|grad phi| = 1 for a tilted plane: the subcell reinit leaves it unchanged away
    from the walls (the Neumann mirror bends it there)."""
    nx = ny = 48
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    X, Y = S.centres(nx, ny, dx, dx)
    n = np.array([0.6, 0.8])
    phi = np.ascontiguousarray(n[0] * (X - 0.43) + n[1] * (Y - 0.51))
    out = oracle.reinit_minrk2(gr, phi.copy(), 3)
    m = slice(3 + 16, 3 + 32)
    assert np.abs(out[m, m] - phi[m, m]).max() < 1e-13


def test_reinit_restores_distance_and_keeps_interface(oracle):
    nx = ny = 96
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    sdf = S.circle_sdf(nx, ny, dx, dx, 0.5, 0.5, 0.3)
    phi = np.ascontiguousarray(sdf * (1.0 + 0.5 * np.exp(-((sdf / 0.1) ** 2))))  # same zero set
    out = phi.copy()
    for _ in range(6):
        oracle.reinit_minrk2(gr, out, 5)
    band = np.abs(sdf) < 3 * dx
    g = oracle.eno_absgrad(gr, out, out)
    assert np.abs(g[band] - 1.0).max() < 0.05
    # interface location: phi and the exact SDF agree near the zero set to O(dx^2)
    assert np.abs(out[band] - sdf[band]).max() < 0.1 * dx


def test_sussman_keeps_area(oracle):
    nx = ny = 64
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    sdf = S.circle_sdf(nx, ny, dx, dx, 0.5, 0.5, 0.27)
    phi = np.ascontiguousarray(sdf * (1.0 + 0.4 * np.sin(8 * np.pi * sdf)))
    area0 = (interior(gr, phi) > 0).sum()
    out = oracle.reinit_sussman(gr, phi.copy(), 5)
    assert abs(int((interior(gr, out) > 0).sum()) - int(area0)) <= 2


# ----------------------------------------------------------------- curvature / Heaviside
def test_kappa_of_circle(oracle):
    """div(grad phi/|grad phi|) of phi = r - d is -1/d: second-order convergence of the
    curvature at the cell centres nearest the interface, clip at 1/min(dx, dy)."""
    r = 0.3
    errs = []
    for nx in (64, 128, 256):
        dx = 1.0 / nx
        gr = make_grid(nx, nx, dx, dx)
        phi = S.circle_sdf(nx, nx, dx, dx, 0.5, 0.5, r)  # heavy inside
        oracle.bc_p(gr, phi)
        kap = oracle.update_kappa(gr, phi)
        X, Y = S.centres(nx, nx, dx, dx)
        d = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
        band = np.abs(phi) < dx
        errs.append(np.abs(kap[band] + 1.0 / d[band]).max() * r)
        assert np.abs(kap).max() <= 1.0 / dx + 1e-12  # mac2d.cpp:246
    assert errs[-1] < 0.01
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs


def test_heavisides(oracle):
    gr = make_grid(20, 20, 0.05, 0.05)
    phi = np.ascontiguousarray(np.linspace(-0.2, 0.2, 26 * 26).reshape(26, 26))
    H = oracle.heaviside(gr, phi)
    assert set(np.unique(H)) <= {0.0, 1.0} and np.all((H == 1.0) == (phi >= 0))
    Hs = oracle.smooth_heaviside(gr, phi)
    eps = 1.5 * 0.05
    assert np.all(Hs[phi > eps] == 1.0) and np.all(Hs[phi < -eps] == 0.0)
    mid = np.abs(phi) <= eps
    # the smoothed Heaviside is odd about 1/2: Hs(phi) + Hs(-phi) = 1
    sym = np.ascontiguousarray(np.concatenate([phi[mid], -phi[mid]]).reshape(1, -1))
    grs = make_grid(1, sym.shape[1] - 6, 0.05, 0.05)
    Hsym = oracle.smooth_heaviside(grs, np.ascontiguousarray(np.repeat(sym, 7, axis=0)))[0]
    half = mid.sum()
    np.testing.assert_allclose(Hsym[:half] + Hsym[half:], 1.0, rtol=0, atol=2e-16)
    assert np.any(np.abs(Hsym[:half] - 0.5) > 0.1)
    z = oracle.smooth_heaviside(gr, np.zeros((26, 26)))
    assert np.all(z == 0.5)


def test_divergence_of_linear_field(oracle):
    nx, ny, dx = 20, 16, 0.1
    gr = make_grid(nx, ny, dx, dx)
    Xu, Yu = S.faces_u(nx, ny, dx, dx)
    Xv, Yv = S.faces_v(nx, ny, dx, dx)
    div = oracle.divergence(gr, np.ascontiguousarray(2 * Xu), np.ascontiguousarray(-0.5 * Yv))
    np.testing.assert_allclose(interior(gr, div), 1.5, atol=1e-12)


# ----------------------------------------------------------------- GFM assembly
def test_gfm_reduces_to_laplacian(oracle):
    """rho_H = rho_L = rho and sigma = 0: the operator is the 5-point -Laplacian / rho with
    Neumann walls and the rhs is -div."""
    rho = 3.0
    gr, fl, phi, kap, H, div = gfm_case(oracle, 24, 20, seed=3, rho=(rho, rho), sigma=0.0)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    h2 = gr.dx * gr.dx
    AX, AY, D = interior(gr, ax), interior(gr, ay), interior(gr, dg)
    np.testing.assert_allclose(AX[1, 1:], -1.0 / (rho * h2), rtol=1e-15)  # row 1 but the pinned cell's face
    np.testing.assert_allclose(AX[2:, :], -1.0 / (rho * h2), rtol=1e-15)
    np.testing.assert_allclose(AX[0, :], 0.0)  # wall faces dropped
    np.testing.assert_allclose(AY[:, 0], 0.0)
    np.testing.assert_allclose(D[1:-1, 1:-1], 4.0 / (rho * h2), rtol=1e-14)
    np.testing.assert_allclose(D[1:-1, 0], 3.0 / (rho * h2), rtol=1e-14)
    np.testing.assert_allclose(D[0, 0], 2.0 / (rho * h2), rtol=1e-14)
    assert AX[1, 0] == 0.0 and AY[0, 1] == 0.0  # coupling to the pinned cell removed
    R = interior(gr, rhs)
    np.testing.assert_allclose(R.ravel()[1:], -interior(gr, div).ravel()[1:], rtol=1e-15)
    assert R[0, 0] == 0.0


def test_gfm_density_weighting_across_interface(oracle):
    """A vertical planar interface between cells: the face coefficient is
    1/(rho_H thetaH + rho_L (1 - thetaH)) with thetaH the heavy fraction."""
    nx = ny = 16
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    X, _ = S.centres(nx, ny, dx, dx)
    x0 = 0.5 + 0.3 * dx
    phi = np.ascontiguousarray(x0 - X)  # heavy on the left
    kap = np.zeros_like(phi)
    H = oracle.heaviside(gr, phi)
    fl = make_fluid(1000.0, 1.0, 1.0, 1.0, 0.0, 1e-3)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, np.zeros_like(phi))
    i = 3 + int(np.floor(x0 / dx - 0.5)) + 1  # first light cell
    lsW, lsM = phi[i - 1, 8], phi[i, 8]
    thH = abs(lsW) / (abs(lsW) + abs(lsM))
    beta = 1.0 / (1000.0 * thH + 1.0 * (1 - thH))
    assert abs(ax[i, 8] + beta / (dx * dx)) <= 1e-12 * beta / (dx * dx)
    assert abs(ax[i - 3, 8] + 1.0 / (1000 * dx * dx)) < 1e-15 / dx / dx
    assert abs(ax[i + 3, 8] + 1.0 / (dx * dx)) < 1e-12 / dx / dx


def test_gfm_constant_curvature_closed_form(oracle):
    k0, sigma, dt = 3.0, 0.5, 1e-3
    gr, fl, phi, kap, H, div = gfm_case(oracle, 32, 32, seed=9, sigma=sigma, dt=dt, kappa_const=k0)
    div[:] = 0.0
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    x, it, rr, st = oracle.pcg(gr, ax, ay, dg, rhs, 1e-13, 10000, 1)
    assert st == 0
    np.testing.assert_allclose(interior(gr, x), -dt * sigma * k0 * interior(gr, H), atol=1e-9 * dt * sigma * k0)


@pytest.mark.parametrize("bc,c_p", [(("neumann",) * 4, (0.0,) * 4),
                                    (("dirichlet", "neumann", "neumann", "outlet"), (0.2, 0.0, 0.0, -0.1))])
def test_gfm_operator_spd(oracle, bc, c_p):
    gr, fl, phi, kap, H, div = gfm_case(oracle, 20, 18, seed=4, bc_p=bc, c_p=c_p)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    pin = all(b == "neumann" for b in bc)
    A, ids = stencil_to_csr(gr, ax, ay, dg, pin=pin)
    assert abs(A - A.T).max() == 0.0
    assert spla.eigsh(A, k=1, which="SA", return_eigenvectors=False)[0] > 0


# ----------------------------------------------------------------- Krylov
@pytest.mark.parametrize("mode", [0, 1])
def test_pcg_against_direct_solve(oracle, mode):
    gr, fl, phi, kap, H, div = gfm_case(oracle, 32, 28, seed=12)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    A, ids = stencil_to_csr(gr, ax, ay, dg)
    xd = scatter(gr, spla.spsolve(A.tocsc(), rhs_vector(gr, rhs, ids)), ids)
    tol = 1e-6 if mode == 0 else 1e-12
    x, it, rr, st = oracle.pcg(gr, ax, ay, dg, rhs, tol, 100000, mode)
    assert st == 0
    err = np.linalg.norm(interior(gr, x - xd)) / np.linalg.norm(interior(gr, xd))
    assert err < (1e-3 if mode == 0 else 1e-9)


def test_pcg_finite_termination_tiny_grid(oracle):
    """CG on n unknowns terminates in at most n steps in exact arithmetic."""
    gr, fl, phi, kap, H, div = gfm_case(oracle, 3, 3, seed=2, rho=(5.0, 1.0))
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    rhs[4, 4] += 1.0
    x, it, rr, st = oracle.pcg(gr, ax, ay, dg, rhs, 1e-12, 100, 1)
    assert st == 0 and it <= 8 + 1


def test_reference_cg_recurrence_stalls_on_tight_tolerance(oracle):
    """The +1e-12 in alpha's denominator (poisson2d.cpp:373) biases small steps: the
    reference recurrence reaches its own 1e-6 but not 1e-12 on this system."""
    gr, fl, phi, kap, H, div = gfm_case(oracle, 32, 32, seed=13)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    _, _, _, st6 = oracle.pcg(gr, ax, ay, dg, rhs, 1e-6, 20000, 0)
    _, _, rr, st12 = oracle.pcg(gr, ax, ay, dg, rhs, 1e-12, 20000, 0)
    assert st6 == 0 and st12 == 1 and rr > 1e-12


def test_bicgstab_against_direct_solve(oracle):
    gr, fl, phi, kap, H, div = gfm_case(oracle, 30, 26, seed=14)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    A, ids = stencil_to_csr(gr, ax, ay, dg)
    xd = scatter(gr, spla.spsolve(A.tocsc(), rhs_vector(gr, rhs, ids)), ids)
    x, it, rr, st = oracle.bicgstab(gr, ax, ay, dg, rhs, 1e-12, 10000)
    assert st == 0
    assert np.linalg.norm(interior(gr, x - xd)) / np.linalg.norm(interior(gr, xd)) < 1e-9


def test_csr_pcg_equals_stencil_pcg(oracle):
    gr, fl, phi, kap, H, div = gfm_case(oracle, 24, 24, seed=15)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    A, ids = stencil_to_csr(gr, ax, ay, dg)
    b = rhs_vector(gr, rhs, ids)
    xc, itc, _, st = oracle.pcg_csr(A.indptr, A.indices, A.data, A.diagonal(), b, 1e-12, 10000, 1)
    xs, its, _, st2 = oracle.pcg(gr, ax, ay, dg, rhs, 1e-12, 10000, 1)
    assert st == 0 and st2 == 0 and abs(itc - its) <= 2
    np.testing.assert_allclose(scatter(gr, xc, ids), xs, atol=1e-9 * np.abs(xs).max())


# ----------------------------------------------------------------- projection / CFL / momentum
def test_update_vel_balanced_jump(oracle):
    """With kappa = k0 and p* = -dt sigma k0 H the GFM jump term cancels the pressure
    gradient on every face: u = u* (mac2d.cpp:1797-1802 against :1369-1370)."""
    k0, sigma, dt = 3.0, 0.5, 1e-3
    gr, fl, phi, kap, H, div = gfm_case(oracle, 32, 32, seed=9, sigma=sigma, dt=dt, kappa_const=k0)
    ps = np.ascontiguousarray(-dt * sigma * k0 * H)
    us, vs = S.smooth_velocity(3, 32, 32, gr.dx, gr.dy)
    u = np.zeros_like(us); v = np.zeros_like(vs)
    oracle.update_vel(gr, fl, u, v, us, vs, ps, phi, kap, H)
    np.testing.assert_allclose(u[4:35, 3:35], us[4:35, 3:35], atol=1e-14)
    np.testing.assert_allclose(v[3:35, 4:35], vs[3:35, 4:35], atol=1e-14)


def test_update_dt_product_quirk(oracle):
    """sigma = mu = g = 0 and u = U: dt = cfl dx / (U^2 / 2) (the product u_{i+1} u_i / 2
    of mac2d.cpp:1860)."""
    nx, dx, U = 16, 0.1, 3.0
    gr = make_grid(nx, nx, dx, dx)
    fl = make_fluid(1.0, 1.0, 0.0, 0.0, 0.0, 1.0, cfl=0.4)
    u = np.full((nx + 6, nx + 6), U)
    v = np.zeros_like(u)
    assert abs(oracle.update_dt(gr, fl, u, v) - 0.4 * dx / (U * U / 2)) < 1e-15


def test_momentum_uniform_flow_is_steady(oracle):
    nx = ny = 24
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx, bc_u=("neumann",) * 4, bc_v=("neumann",) * 4)
    U = 0.7
    u = np.full((nx + 6, ny + 6), U)
    v = np.zeros_like(u)
    phi = np.full_like(u, 1.0)
    Hs = oracle.smooth_heaviside(gr, phi)
    fl = make_fluid(1.0, 1.0, 0.01, 0.01, 0.0, 1e-3)
    uh, vh = oracle.intermediate_vel(gr, fl, 2, phi, u, v, Hs)
    np.testing.assert_allclose(uh[6:nx, 6:ny], U, atol=1e-13)
    np.testing.assert_allclose(vh[6:nx, 6:ny], 0.0, atol=1e-13)


def test_momentum_gravity(oracle):
    nx = ny = 40
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx)
    phi = np.full((nx + 6, ny + 6), 1.0)
    Hs = oracle.smooth_heaviside(gr, phi)
    u = np.zeros_like(phi); v = np.zeros_like(phi)
    dt = 1e-3
    fl = make_fluid(1.0, 1.0, 0.01, 0.01, 0.0, dt, g=9.8, g_axis=1, gravity_fix=1)
    for order in (0, 1, 2):  # every RK variant integrates a constant source exactly
        uh, vh = oracle.intermediate_vel(gr, fl, order, phi, u, v, Hs)
        # away from the walls, where the no-slip ghosts and viscosity act
        # beyond the reach (3 RK stages x 3 cells) of the no-slip walls and their viscosity
        np.testing.assert_allclose(vh[14:nx - 8, 14:ny - 8], -9.8 * dt, rtol=1e-12)
        assert np.abs(uh[14:nx - 8, 14:ny - 8]).max() == 0.0
    fl_q = make_fluid(1.0, 1.0, 0.01, 0.01, 0.0, dt, g=9.8, g_axis=1, gravity_fix=0)
    uh, vh = oracle.intermediate_vel(gr, fl_q, 2, phi, u, v, Hs)
    assert np.abs(vh).max() == 0.0  # mac2d.cpp:294: the V equation never sees gravity along y


def test_ls_diagnostics_closed_forms(oracle):
    """Run diagnostics on a circle SDF (radius 0.25 at (0.5, 0.5) on 128^2, heavy inside): the
    mass count and centroid equal an independent numpy recount over the interior; the area, the
    interface length (2 pi r) and the circularity (1) match the circle to the grid's accuracy; the
    centroid is exactly the centre (the cell set is mirror-symmetric and the coordinates are exact
    binary fractions); against phi - d the error norms are the closed forms d, d N and
    d / sqrt(N) (the drivers' L2 normalisation sqrt(sum e^2) / (nx ny)) and the mass loss is the
    relative change of the positive-cell count."""
    n, r = 128, 0.25
    dx = 1.0 / n
    gr = make_grid(n, n, dx, dx)
    phi = S.circle_sdf(n, n, dx, dx, 0.5, 0.5, r)
    d = oracle.ls_diagnostics(gr, phi, base=(0.0, 0.0))
    inner = phi[3:-3, 3:-3]
    pos = inner > 0.0
    assert d["cells"] == pos.sum()
    assert d["mass"] == pos.sum() * dx * dx
    xc = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    assert d["centroid_x"] == pytest.approx(X[pos].mean(), abs=1e-14)
    assert d["centroid_y"] == pytest.approx(Y[pos].mean(), abs=1e-14)
    assert d["centroid_x"] == pytest.approx(0.5, abs=1e-13) and d["centroid_y"] == pytest.approx(0.5, abs=1e-13)
    assert d["mass"] == pytest.approx(np.pi * r * r, rel=1e-2)
    assert d["length"] == pytest.approx(2.0 * np.pi * r, rel=1e-2)
    assert d["circularity"] == pytest.approx(1.0, rel=1e-2)
    # offset base coordinates move the centroid by the offset
    d_off = oracle.ls_diagnostics(gr, phi, base=(1.25, -0.5))
    assert d_off["centroid_x"] == pytest.approx(1.75, abs=1e-13) and d_off["centroid_y"] == pytest.approx(0.0, abs=1e-13)
    # error norms against a uniformly shifted field
    delta = 0.01
    ref = phi - delta
    e = oracle.ls_diagnostics(gr, phi, ref=ref)
    N = n * n
    assert e["err_linf"] == pytest.approx(delta, rel=1e-12)
    assert e["err_l1"] == pytest.approx(delta * N, rel=1e-12)
    assert e["err_l2"] == pytest.approx(delta / np.sqrt(N), rel=1e-12)
    ref_pos = (ref[3:-3, 3:-3] > 0.0).sum()
    assert e["ref_cells"] == ref_pos
    assert e["mass_loss_pct"] == pytest.approx(abs(pos.sum() - ref_pos) / ref_pos * 100.0, rel=1e-14)
    # an ellipse (semi-axes 0.3, 0.15) is less circular
    x = (np.arange(n + 6) - 3 + 0.5) * dx
    Xp, Yp = np.meshgrid(x, x, indexing="ij")
    ell = 1.0 - np.sqrt(((Xp - 0.5) / 0.3) ** 2 + ((Yp - 0.5) / 0.15) ** 2)
    ell *= 0.15  # roughly distance-scaled near the minor axis
    assert oracle.ls_diagnostics(gr, np.ascontiguousarray(ell))["circularity"] < 0.95


def test_gs_one_sweep_by_hand(oracle):
    """GS_2FUniform_2D (src/poisson2d.cpp:7-88) on [[4, 1], [1, 3]] x = [1, 2] from 0: the sweep
    gives 1/4 and (2 - 1/4)/3 = 7/12, the relaxation 1.2 times that: (0.3, 0.7)."""
    x, it, err = oracle.gs_csr([0, 2, 4], [0, 1, 0, 1], np.array([4.0, 1.0, 1.0, 3.0]), np.array([1.0, 2.0]),
                               np.zeros(2), 1)
    assert it == 1 and err > 1e6
    np.testing.assert_allclose(x, [0.3, 0.7], rtol=1e-15)


@pytest.mark.parametrize("seed", [1, 2])
def test_gs_sweeps_equal_triangular_solves(oracle, seed):
    """Each sweep is the forward substitution (D + L) x' = b - U x (a dense triangular solve here),
    then x = x + 1.2 (x' - x); and a converged run is close to the direct solution."""
    import scipy.linalg as sl
    A = random_spd_csr(40, seed)
    rng = np.random.default_rng(seed + 10)
    b = rng.standard_normal(40)
    x0 = rng.standard_normal(40)
    Ad = A.toarray()
    L, U = np.tril(Ad), np.triu(Ad, 1)
    x = x0.copy()
    for _ in range(5):
        xs = sl.solve_triangular(L, b - U @ x, lower=True)
        x = x + 1.2 * (xs - x)
    got, it, _ = oracle.gs_csr(A.indptr, A.indices, A.data, b, x0, 5)
    assert it == 5
    np.testing.assert_allclose(got, x, rtol=1e-12, atol=1e-12 * np.abs(x).max())
    xc, itc, errc = oracle.gs_csr(A.indptr, A.indices, A.data, b, x0, 100000)
    assert errc <= 1e-6 and itc < 100000
    xd = np.linalg.solve(Ad, b)
    assert np.linalg.norm(xc - xd) / np.linalg.norm(xd) <= 1e-4
