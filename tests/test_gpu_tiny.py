"""The smallest grids the library takes (g = 3 cells per side: every ghost mirror of the P fill,
ApplyBC_P_2D src/bc2d.cpp:346-353, and of the fused ghost writes reaches an interior cell) and
barely wider ones, against the FP64 oracle through the C ABI; thinner grids are refused with an
error.  Covered: P and U/V ghost fills (exact), one advection step and MinRK2
reinitialisation (max |dphi| <= 1e-12), GFM assembly (1e-14) and the FP64 Jacobi-PCG solve (rel
L2 <= 1e-10 against the oracle's PCG)."""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_fluid, make_grid
from tests import scenarios as S
from tests.gfm_common import interior

pytestmark = pytest.mark.gpu

SIZES = [(3, 3), (4, 3), (3, 5), (5, 4)]
P_BCS = [("neumann",) * 4, ("dirichlet", "neumann", "outlet", "wall"), ("lsfree", "neumann", "neumann", "lsfree")]
C_P = (0.1, -0.2, 0.3, 0.4)


def _ids(v):
    return "x".join(map(str, v)) if isinstance(v[0], int) else "-".join(v)


@pytest.mark.parametrize("bc", P_BCS, ids=_ids)
@pytest.mark.parametrize("nx,ny", SIZES)
def test_bc_p_tiny(oracle, nx, ny, bc):
    dx = 0.1
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=C_P)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_P, bc, C_P)
    a = S.random_field(nx * 10 + ny, (nx + 6, ny + 6))
    d = ctx.array(a)
    ctx.bc_apply(A.GRID_P, d)
    np.testing.assert_array_equal(d.download(), oracle.bc_p(gr, a.copy()))


@pytest.mark.parametrize("kind", ["neumann", "dirichlet", "wall", "outlet"])
@pytest.mark.parametrize("nx,ny", SIZES)
def test_bc_uv_tiny(oracle, nx, ny, kind):
    dx = 0.1
    cu, cv = (0.1, 0.2, 0.3, 0.4), (0.5, -0.6, 0.7, 0.8)
    gr = make_grid(nx, ny, dx, dx, bc_u=(kind,) * 4, c_u=cu, bc_v=(kind,) * 4, c_v=cv)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_U, (kind,) * 4, cu)
    ctx.set_bc(A.GRID_V, (kind,) * 4, cv)
    for gk, seed, fn in ((A.GRID_U, 2, oracle.bc_u), (A.GRID_V, 3, oracle.bc_v)):
        a = S.random_field(seed, (nx + 6, ny + 6))
        d = ctx.array(a)
        ctx.bc_apply(gk, d)
        np.testing.assert_array_equal(d.download(), fn(gr, a.copy()))


@pytest.mark.parametrize("bc", P_BCS[:2], ids=_ids)
@pytest.mark.parametrize("nx,ny", SIZES)
def test_levelset_tiny(oracle, nx, ny, bc):
    dx = 0.1
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=C_P)
    phi = S.circle_sdf(nx, ny, dx, dx, 0.5 * nx * dx, 0.5 * ny * dx, 0.3 * max(nx, ny) * dx)
    phi = np.ascontiguousarray(phi + 0.01 * S.random_field(4, phi.shape))
    u = 0.5 + 0.2 * S.random_field(5, phi.shape)
    v = -0.3 + 0.2 * S.random_field(6, phi.shape)
    dt = 0.2 * dx
    ref = oracle.solve_levelset(gr, phi.copy(), u, v, dt)
    ref = oracle.reinit_minrk2(gr, ref, 2)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_P, bc, C_P)
    d_phi = ctx.array(phi)
    ctx.ls_advect(d_phi, ctx.array(u), ctx.array(v), dt)
    ctx.ls_reinit_minrk2(d_phi, 2)
    out = d_phi.download()
    g = 3
    err = float(np.abs(out[g:g + nx, g:g + ny] - ref[g:g + nx, g:g + ny]).max())
    print(f"{nx}x{ny} {bc}: max|dphi| = {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("nx,ny", SIZES)
def test_pressure_tiny(oracle, nx, ny):
    dx = 0.1
    bc = ("dirichlet", "neumann", "neumann", "neumann")  # one Dirichlet side: no pinned cell
    c_p = (0.2, 0.0, 0.0, 0.0)
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=c_p)
    phi = S.circle_sdf(nx, ny, dx, dx, 0.4 * nx * dx, 0.6 * ny * dx, 0.35 * max(nx, ny) * dx)
    oracle.bc_p(gr, phi)
    kap = oracle.update_kappa(gr, phi)
    oracle.bc_p(gr, kap)
    H = oracle.heaviside(gr, phi)
    uu, vv = S.smooth_velocity(1, nx, ny, dx, dx)
    div = oracle.divergence(gr, uu, vv)
    fl = make_fluid(1000.0, 1.0, 1.0, 0.01, 1.0, 1e-3)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    x_ref, _, _, _ = oracle.pcg(gr, ax, ay, dg, rhs, 1e-13, 1000, 1)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_P, bc, c_p)
    flg = A.make_fluid(fl.rho_h, fl.rho_l, fl.mu_h, fl.mu_l, fl.sigma, fl.dt)
    ctx.gfm_assemble(flg, *[ctx.array(a) for a in (phi, kap, H, div)])
    for name, got, exp in zip(("ax", "ay", "diag", "rhs"), ctx.gfm_export(), (ax, ay, dg, rhs)):
        assert np.abs(got - exp).max() <= 1e-14 * max(np.abs(exp).max(), 1e-300), name
    ps = ctx.array()
    st = ctx.poisson_solve(ps, ctx.opts(tol=1e-13))
    x = interior(gr, ps.download())
    xr = interior(gr, x_ref)
    err = float(np.linalg.norm(x - xr) / max(np.linalg.norm(xr), 1e-300))
    print(f"{nx}x{ny}: {st.iterations} iterations, rel L2 {err:.2e}")
    assert st.converged == 1 and err <= 1e-10


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 3), (3, 2), (1, 8)])
def test_grid_below_ghost_layer_refused(nx, ny):
    with pytest.raises(A.Asd2Error):
        A.Context(nx, ny, 0.1, 0.1)
