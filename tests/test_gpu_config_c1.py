"""The stationary-droplet configuration exactly as stated for the GFM benchmark case (the physics
of the reference driver tests/LargeTests/test_mac2d_large.cpp:377-570 at benchmark parameters):
128 x 128 unit box, droplet r = 0.25 at (0.5, 0.5), heavy inside, density 1000 : 1, viscosity
100 : 1 (1e-2 : 1e-4), sigma = 1, no gravity, dt = 1e-4 fixed (no CFL update), 10 steps of
HJ-WENO5 / SSP-RK3 advection, 3 MinRK2 reinitialisation iterations, curvature and Heaviside, the
RK3 momentum predictor, divergence, GFM assembly, pressure by Jacobi-PCG to a relative residual
of 1e-10, projection; no-slip walls, Neumann pressure.

The GPU step (C ABI) and the oracle step run side by side from the same state and are compared
after every step: level set max |diff| <= 1e-12, pressure relative L2 <= 1e-8, velocities max
|diff| <= 1e-8 of the projection's velocity scale max |ps| / (rho_L dx)."""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_fluid, make_grid
from tests import scenarios as S
from tests.gfm_common import interior

pytestmark = pytest.mark.gpu

N, STEPS, DT = 128, 10, 1e-4
RHO_H, RHO_L, MU_H, MU_L, SIGMA = 1000.0, 1.0, 1e-2, 1e-4, 1.0
REINIT = 3
ORDER_RK3 = 2


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def test_c1_stationary_droplet_10_steps(oracle):
    dx = 1.0 / N
    gr = make_grid(N, N, dx, dx, bc_p=("neumann",) * 4, bc_u=("dirichlet",) * 4, bc_v=("dirichlet",) * 4)
    fl = make_fluid(RHO_H, RHO_L, MU_H, MU_L, SIGMA, DT, g=0.0)
    ctx = A.Context(N, N, dx, dx)
    ctx.set_bc(A.GRID_P, ("neumann",) * 4)
    ctx.set_bc(A.GRID_U, ("dirichlet",) * 4)
    ctx.set_bc(A.GRID_V, ("dirichlet",) * 4)
    flg = A.make_fluid(RHO_H, RHO_L, MU_H, MU_L, SIGMA, DT)
    opts = ctx.opts(tol=1e-10, precision=A.PREC_FP64, mode=A.CG_STANDARD)

    ls = S.circle_sdf(N, N, dx, dx, 0.5, 0.5, 0.25)
    oracle.bc_p(gr, ls)
    oracle.reinit_minrk2(gr, ls, REINIT)
    oracle.bc_p(gr, ls)
    shp = ls.shape
    u, v, ps, kap = np.zeros(shp), np.zeros(shp), np.zeros(shp), np.zeros(shp)

    d_ls = ctx.array(S.circle_sdf(N, N, dx, dx, 0.5, 0.5, 0.25))
    ctx.bc_apply(A.GRID_P, d_ls)
    ctx.ls_reinit_minrk2(d_ls, REINIT)
    ctx.bc_apply(A.GRID_P, d_ls)
    d_u, d_v, d_uh, d_vh = ctx.array(), ctx.array(), ctx.array(), ctx.array()
    d_k, d_H, d_Hs, d_div, d_ps = (ctx.array() for _ in range(5))
    assert np.abs(d_ls.download() - ls).max() <= 1e-12

    worst = {"ls": 0.0, "ps": 0.0, "u": 0.0, "v": 0.0}
    for step in range(STEPS):
        # oracle
        ls = oracle.solve_levelset(gr, ls, u, v, DT)
        oracle.bc_p(gr, ls)
        oracle.reinit_minrk2(gr, ls, REINIT)
        oracle.bc_p(gr, ls)
        oracle.update_kappa(gr, ls, kap)
        H = oracle.heaviside(gr, ls)
        Hs = oracle.smooth_heaviside(gr, ls)
        uh, vh = oracle.intermediate_vel(gr, fl, ORDER_RK3, ls, u, v, Hs)
        oracle.bc_u(gr, uh)
        oracle.bc_v(gr, vh)
        div = oracle.divergence(gr, uh, vh)
        oracle.bc_p(gr, kap)
        ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, ls, kap, H, div)
        x, it, _, st = oracle.pcg(gr, ax, ay, dg, rhs, 1e-10, 100000, 1)
        assert st == 0
        ps = np.zeros(shp)
        interior(gr, ps)[:] = interior(gr, x)
        oracle.bc_p(gr, ps)
        oracle.update_vel(gr, fl, u, v, uh, vh, ps, ls, kap, H)

        # GPU
        ctx.ls_advect(d_ls, d_u, d_v, DT)
        ctx.bc_apply(A.GRID_P, d_ls)
        ctx.ls_reinit_minrk2(d_ls, REINIT)
        ctx.bc_apply(A.GRID_P, d_ls)
        ctx.kappa(d_ls, d_k)
        ctx.heaviside(d_ls, d_H)
        ctx.heaviside(d_ls, d_Hs, smooth=True)
        ctx.intermediate_vel(flg, ORDER_RK3, d_ls, d_u, d_v, d_uh, d_vh, d_Hs)
        ctx.bc_apply(A.GRID_U, d_uh)
        ctx.bc_apply(A.GRID_V, d_vh)
        ctx.divergence(d_uh, d_vh, d_div)
        ctx.bc_apply(A.GRID_P, d_k)
        ctx.gfm_assemble(flg, d_ls, d_k, d_H, d_div)
        d_ps.upload(np.zeros(shp))
        s = ctx.poisson_solve(d_ps, opts)
        ctx.bc_apply(A.GRID_P, d_ps)
        ctx.update_vel(flg, d_u, d_v, d_uh, d_vh, d_ps, d_ls, d_k, d_H)

        # the droplet is at rest: u is the small remainder of the pressure gradient and the surface
        # tension jump cancelling each other, so velocity differences are measured against the
        # projection's velocity scale max |ps| / (rho_L dx), not against |u| itself
        scale = np.abs(ps).max() / (RHO_L * dx)
        e = {"ls": float(np.abs(d_ls.download() - ls).max()),
             "ps": _rel(interior(gr, d_ps.download()), interior(gr, ps)),
             "u": float(np.abs(d_u.download() - u).max() / scale),
             "v": float(np.abs(d_v.download() - v).max() / scale)}
        for k in worst:
            worst[k] = max(worst[k], e[k])
        print(f"step {step + 1}: GPU PCG {s.iterations} iters (oracle {it}); ls {e['ls']:.1e}, ps {e['ps']:.1e}, "
              f"u {e['u']:.1e}, v {e['v']:.1e}", flush=True)
        assert s.converged == 1 and s.true_rel_residual <= 1e-10
        assert e["ls"] <= 1e-12, e
        assert e["ps"] <= 1e-8 and e["u"] <= 1e-8 and e["v"] <= 1e-8, e
    print("worst over the run", worst)
