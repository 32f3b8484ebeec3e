"""GPU parity of the MAC helpers around the pressure solve (reference src/mac2d.cpp:87-122,
221-252, 254-1239, 1729-1880) against the FP64 oracle, through the C ABI.

Bars: Heaviside fields exact; curvature, divergence, velocity correction and the time step
to FP64 rounding (relative 1e-12 of the field's scale); momentum terms (WENO5 convection
with the same one-division weight evaluation as the level-set kernels, viscous fluxes,
gravity) to relative 1e-11 of the term's scale."""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_fluid, make_grid
from tests import scenarios as S

pytestmark = pytest.mark.gpu

U_BCS = [(("dirichlet",) * 4, (0.0, 0.0, 0.0, 0.0)),
         (("neumann", "outlet", "wall", "inlet"), (0.1, 0.0, 0.0, 0.2)),
         (("periodic",) * 4, (0.0,) * 4)]


def _fluid_pair(rho=(1000.0, 1.0), mu=(1e-2, 1e-4), sigma=1.0, dt=1e-4, g=0.0, g_axis=1, gravity_fix=1):
    fo = make_fluid(rho[0], rho[1], mu[0], mu[1], sigma, dt, g=g, g_axis=g_axis, gravity_fix=gravity_fix)
    fa = A.make_fluid(rho[0], rho[1], mu[0], mu[1], sigma, dt, g=g, g_axis=g_axis, gravity_fix=gravity_fix)
    return fo, fa


def _case(oracle, nx, ny, seed, bc_u=U_BCS[0], bc_p=("neumann",) * 4):
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx, bc_p=bc_p, bc_u=bc_u[0], c_u=bc_u[1], bc_v=bc_u[0], c_v=bc_u[1][::-1])
    phi, _ = S.random_droplets(seed, nx, ny, dx, dx, n_range=(2, 5), r_range=(0.1, 0.22), lx=nx * dx, ly=ny * dx)
    oracle.bc_p(gr, phi)
    u, v = S.smooth_velocity(seed, nx, ny, dx, dx)
    oracle.bc_u(gr, u)
    oracle.bc_v(gr, v)
    ctx = A.Context(nx, ny, dx, dx)
    ctx.set_bc(A.GRID_P, bc_p)
    ctx.set_bc(A.GRID_U, bc_u[0], bc_u[1])
    ctx.set_bc(A.GRID_V, bc_u[0], bc_u[1][::-1])
    return gr, ctx, np.ascontiguousarray(phi), np.ascontiguousarray(u), np.ascontiguousarray(v)


def _close(out, ref, rel, what):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(out - ref).max() / scale
    print(f"{what}: max rel err {err:.3e} (scale {scale:.3e})")
    assert err <= rel, (what, err)


@pytest.mark.parametrize("nx,ny", [(64, 64), (57, 43)])
def test_kappa(oracle, nx, ny):
    gr, ctx, phi, u, v = _case(oracle, nx, ny, seed=11)
    ref = oracle.update_kappa(gr, phi)
    d_phi, d_k = ctx.array(phi), ctx.array()
    ctx.kappa(d_phi, d_k)
    _close(d_k.download(), ref, 1e-12, "kappa")


@pytest.mark.parametrize("smooth", [False, True])
def test_heaviside(oracle, smooth):
    gr, ctx, phi, u, v = _case(oracle, 61, 50, seed=12)
    ref = oracle.smooth_heaviside(gr, phi) if smooth else oracle.heaviside(gr, phi)
    d_phi, d_H = ctx.array(phi), ctx.array()
    ctx.heaviside(d_phi, d_H, smooth=smooth)
    out = d_H.download()
    if smooth:
        _close(out, ref, 1e-14, "smooth heaviside")
    else:
        np.testing.assert_array_equal(out, ref)


def test_divergence(oracle):
    gr, ctx, phi, u, v = _case(oracle, 64, 48, seed=13)
    ref = oracle.divergence(gr, u, v)
    d_u, d_v, d_d = ctx.array(u), ctx.array(v), ctx.array()
    ctx.divergence(d_u, d_v, d_d)
    _close(d_d.download(), ref, 1e-12, "divergence")


@pytest.mark.parametrize("bc_u", U_BCS[:2], ids=["dirichlet", "mixed"])
def test_update_vel(oracle, bc_u):
    gr, ctx, phi, u, v = _case(oracle, 56, 48, seed=14, bc_u=bc_u)
    fo, fa = _fluid_pair(dt=2e-4)
    kap = oracle.update_kappa(gr, phi)
    oracle.bc_p(gr, kap)
    H = oracle.heaviside(gr, phi)
    ps = np.ascontiguousarray(S.random_field(3, phi.shape) * 1e-3)
    oracle.bc_p(gr, ps)
    us, vs = u.copy(), v.copy()
    ru, rv = oracle.update_vel(gr, fo, np.zeros_like(u), np.zeros_like(v), us, vs, ps, phi, kap, H)
    d = {k: ctx.array(a) for k, a in dict(u=np.zeros_like(u), v=np.zeros_like(v), us=us, vs=vs, ps=ps,
                                              ls=phi, kap=kap, H=H).items()}
    ctx.update_vel(fa, d["u"], d["v"], d["us"], d["vs"], d["ps"], d["ls"], d["kap"], d["H"])
    _close(d["u"].download(), ru, 1e-12, "u")
    _close(d["v"].download(), rv, 1e-12, "v")


def test_update_dt(oracle):
    gr, ctx, phi, u, v = _case(oracle, 64, 64, seed=15)
    fo, fa = _fluid_pair(g=9.8)
    ref = oracle.update_dt(gr, fo, u, v)
    out = ctx.update_dt(fa, ctx.array(u), ctx.array(v))
    assert abs(out - ref) <= 1e-14 * abs(ref), (out, ref)


@pytest.mark.parametrize("grid_kind", [A.GRID_U, A.GRID_V], ids=["U", "V"])
@pytest.mark.parametrize("term", [A.MOM_CONVECTION, A.MOM_VISCOSITY, A.MOM_RHS],
                         ids=["convection", "viscosity", "rhs"])
def test_momentum_terms(oracle, grid_kind, term):
    gr, ctx, phi, u, v = _case(oracle, 48, 40, seed=16)
    fo, fa = _fluid_pair(g=9.8)
    Hs = oracle.smooth_heaviside(gr, phi)
    if grid_kind == A.GRID_U:
        conv, visc = oracle.convection_u(gr, u, v), oracle.viscosity_u(gr, fo, u, v, phi, Hs)
        grav = 0.0
    else:
        conv, visc = oracle.convection_v(gr, u, v), oracle.viscosity_v(gr, fo, u, v, phi, Hs)
        grav = -fo.g
    if term == A.MOM_CONVECTION:
        ref = conv
    elif term == A.MOM_VISCOSITY:
        ref = visc
    else:
        ref = -conv + visc
        mask = np.zeros_like(ref, dtype=bool)
        g = gr.g
        if grid_kind == A.GRID_U:
            mask[g + 1:g + gr.nx, g:g + gr.ny] = True
        else:
            mask[g:g + gr.nx, g + 1:g + gr.ny] = True
        ref = np.where(mask, ref + grav, 0.0)
    d = [ctx.array(a) for a in (phi, u, v, Hs)]
    out = ctx.array()
    ctx.momentum_term(fa, grid_kind, term, d[0], d[1], d[2], d[3], out)
    _close(out.download(), ref, 1e-11, f"momentum term {term} on grid {grid_kind}")


@pytest.mark.parametrize("order", [0, 1, 2], ids=["euler", "rk2", "rk3"])
@pytest.mark.parametrize("bc_u", [U_BCS[0], U_BCS[2]], ids=["dirichlet", "periodic"])
def test_intermediate_vel(oracle, order, bc_u):
    gr, ctx, phi, u, v = _case(oracle, 48, 48, seed=17, bc_u=bc_u)
    fo, fa = _fluid_pair(g=9.8, dt=5e-4)
    Hs = oracle.smooth_heaviside(gr, phi)
    ru, rv = oracle.intermediate_vel(gr, fo, order, phi, u, v, Hs)
    d = [ctx.array(a) for a in (phi, u, v)]
    uh, vh = ctx.array(), ctx.array()
    ctx.intermediate_vel(fa, order, d[0], d[1], d[2], uh, vh, ctx.array(Hs))
    _close(uh.download(), ru, 1e-11, "uhat")
    _close(vh.download(), rv, 1e-11, "vhat")
