"""The axisymmetric (r, z) path (MACSolver2DAxisym, reference src/mac2dAxisym.cpp) against the
oracle's restatement of it (oracle/asd2_oracle.c, pinned by tests/test_oracle_axisym.py and
tests/test_oracle_axisym_momentum.py): the curvature with its hoop term, the two divergences, the
r-weighted ghost-fluid assembly, the pressure solve on that system with the library's solvers,
the time step, the batched tridiagonal solve, and the momentum step (right-hand sides, viscous
terms, semi-implicit line sweeps).  A drop sits on the axis (a sphere in 3D), heavy inside; the W side is the
axis.  Bars: curvature, divergences, assembly, time step 1e-13 relative (one pow may round
differently on the device); pressure rel L2 1e-8 at a true relative residual of 1e-10."""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_fluid, make_grid

pytestmark = pytest.mark.gpu

G = 3


def _case(nr, nz, base, bc_p, c_p, seed):
    rng = np.random.default_rng(seed)
    lr, lz = 1.0, nz / nr
    dr, dz = lr / nr, lz / nz
    gr = make_grid(nr, nz, dr, dz, bc_p=bc_p, c_p=c_p)
    shp = (nr + 2 * G, nz + 2 * G)
    rc = base + (np.arange(shp[0]) + 0.5 - G) * dr
    zc = (np.arange(shp[1]) + 0.5 - G) * dz
    Rr, Zz = np.meshgrid(rc, zc, indexing="ij")
    R0, z0 = rng.uniform(0.2, 0.35), rng.uniform(0.4, 0.6) * lz
    ls = R0 - np.sqrt(Rr ** 2 + (Zz - z0) ** 2)
    u = rng.standard_normal(shp) * 0.1
    v = rng.standard_normal(shp) * 0.1
    ctx = A.Context(nr, nz, dr, dz, base_x=base)
    ctx.set_bc(A.GRID_P, bc_p, c_p)
    return gr, ctx, ls, u, v


BCS = [(("axisym", "neumann", "neumann", "neumann"), (0.0,) * 4),
       (("axisym", "wall", "neumann", "dirichlet"), (0.0, 0.0, 0.0, 0.3))]


def _rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _close(out, ref, rel, what):
    scale = max(np.abs(ref).max(), 1e-300)
    err = float(np.abs(out - ref).max() / scale)
    print(f"{what}: max rel err {err:.3e} (scale {scale:.3e})")
    assert err <= rel, (what, err)


@pytest.mark.parametrize("bc", range(len(BCS)))
@pytest.mark.parametrize("nr,nz,base", [(48, 64, 0.0), (37, 29, 0.0), (64, 64, 0.25)])
def test_axisym_pressure_path(oracle, nr, nz, base, bc):
    bc_p, c_p = BCS[bc]
    gr, ctx, ls, u, v = _case(nr, nz, base, bc_p, c_p, seed=nr + nz + bc)
    fl = make_fluid(1000.0, 1.0, 1.0, 0.01, 0.5, 1e-3)
    flg = A.make_fluid(1000.0, 1.0, 1.0, 0.01, 0.5, 1e-3)
    d_ls, d_u, d_v = ctx.array(ls), ctx.array(u), ctx.array(v)
    # curvature
    kap_want = oracle.update_kappa_axisym(gr, base, ls)
    d_k = ctx.array(np.zeros_like(ls))
    ctx.kappa_axisym(d_ls, d_k)
    kap = d_k.download()
    inner = (slice(G, G + nr), slice(G, G + nz))
    assert _rel_max(kap[inner], kap_want[inner]) <= 1e-13
    # divergences
    for fp in (True, False):
        d = ctx.array()
        ctx.divergence_axisym(d_u, d_v, d, fp)
        np.testing.assert_allclose(d.download(), oracle.divergence_axisym(gr, base, u, v, fp), rtol=1e-14, atol=1e-14)
    div4p = oracle.divergence_axisym(gr, base, u, v, True)
    # assembly with the oracle's curvature (ghosts by the P rule, as SolvePoisson applies them)
    oracle.bc_p(gr, kap_want)
    H = (ls >= 0).astype(np.float64)
    ax, ay, ae, dg, rhs = oracle.gfm_assemble_axisym(gr, fl, base, ls, kap_want, H, div4p)
    ctx.gfm_assemble_axisym(flg, d_ls, ctx.array(kap_want), ctx.array(H), ctx.array(div4p))
    gax, gay, gdg, grhs = ctx.gfm_export()
    for name, got, want in (("ax", gax, ax), ("ay", gay, ay), ("diag", gdg, dg), ("rhs", grhs, rhs)):
        err = _rel_max(got[inner], want[inner])
        assert err <= 1e-13, (name, err)
    # the stored operator is the symmetric one: the reference's own E value of each row agrees
    # with the east neighbour's W value to rounding
    e_west = ae[G:G + nr - 1, G:G + nz].copy()
    if bc == 0:
        e_west[0, 0] = 0.0  # the coupling into the pinned cell
    np.testing.assert_allclose(e_west, ax[G + 1:G + nr, G:G + nz], rtol=1e-13, atol=0)
    # pressure: the oracle's FP64 PCG on its own arrays against the library's solvers
    x_ref, it_ref, _, st = oracle.pcg(gr, ax, ay, dg, rhs, 1e-13, 200000, 1)
    assert st == 0
    # the multilevel hierarchy needs sides it can coarsen; odd sizes take the mixed Jacobi-PCG
    coarse = A.PRECOND_MULTILEVEL if nr % 8 == 0 and nz % 8 == 0 else A.PRECOND_JACOBI
    for precision, precond in ((A.PREC_FP64, A.PRECOND_JACOBI), (A.PREC_MIXED, coarse)):
        ps = ctx.array()
        s = ctx.poisson_solve(ps, ctx.opts(tol=1e-10, precision=precision, precond=precond))
        err = float(np.linalg.norm(ps.download()[inner] - x_ref[inner]) / np.linalg.norm(x_ref[inner]))
        print(f"{nr}x{nz} base {base} bc {bc}: precision {precision} precond {precond}: {s.iterations} iters "
              f"(oracle FP64 PCG {it_ref}), rel L2 {err:.2e}")
        assert s.converged == 1 and s.true_rel_residual <= 1e-10 and err <= 1e-8
    # time step
    fl_dt = make_fluid(1000.0, 1.0, 1.0, 0.01, 0.5, 1e-3, g=9.8, cfl=0.3)
    flg_dt = A.make_fluid(1000.0, 1.0, 1.0, 0.01, 0.5, 1e-3, g=9.8, cfl=0.3)
    dt = ctx.update_dt_axisym(flg_dt, d_u, d_v, ctx.array(kap_want))
    assert abs(dt - oracle.update_dt_axisym(gr, fl_dt, u, v, kap_want)) <= 1e-15 * dt


def test_axisym_single_phase_projection(oracle):
    """one density, no surface tension: after the solve and the planar UpdateVel kernel (the
    reference's axisymmetric UpdateVel computes the same face updates), the r-weighted divergence
    of the new velocity vanishes to the solver tolerance"""
    nr, nz = 64, 96
    # closed box (no-slip walls, the axis on W): the r-weighted divergence sums to zero, so the
    # pinned all-Neumann system is compatible and every cell's equation holds
    gr, ctx, ls, u, v = _case(nr, nz, 0.0, ("axisym", "neumann", "neumann", "neumann"), (0.0,) * 4, seed=5)
    ctx.set_bc(A.GRID_U, ("dirichlet",) * 4)
    ctx.set_bc(A.GRID_V, ("neumann", "dirichlet", "dirichlet", "dirichlet"))
    flg = A.make_fluid(1.0, 1.0, 1.0, 1.0, 0.0, 1e-2)
    one = np.ones_like(ls)
    d_u, d_v = ctx.array(u), ctx.array(v)
    ctx.bc_apply(A.GRID_U, d_u)
    ctx.bc_apply(A.GRID_V, d_v)
    d4 = ctx.array()
    ctx.divergence_axisym(d_u, d_v, d4, True)
    z = ctx.array()
    ctx.gfm_assemble_axisym(flg, ctx.array(one), z, ctx.array(one), d4)
    ps = ctx.array()
    ctx.poisson_solve(ps, ctx.opts(tol=1e-12))
    ctx.bc_apply(A.GRID_P, ps)
    un, vn = ctx.array(), ctx.array()
    ctx.update_vel(flg, un, vn, d_u, d_v, ps, ctx.array(one), z, ctx.array(one))
    after = ctx.array()
    ctx.divergence_axisym(un, vn, after, True)
    before = d4.download()[G:G + nr, G:G + nz]
    res = after.download()[G:G + nr, G:G + nz]
    print(f"r-weighted divergence: before {np.abs(before).max():.2e}, after {np.abs(res).max():.2e}")
    assert np.abs(res).max() <= 1e-9 * np.abs(before).max()


@pytest.mark.parametrize("n,count", [(2, 5), (17, 300), (1000, 2049)])
def test_tdma_batched(oracle, n, count):
    rng = np.random.default_rng(n + count)
    a = rng.uniform(-1, 1, (n, count))
    c = rng.uniform(-1, 1, (n, count))
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2.0, (n, count))
    d = rng.standard_normal((n, count))
    ctx = A.Context(8, 8, 1.0, 1.0)
    da, db, dc, dd = (A.DeviceArray.from_host(np.ascontiguousarray(x)) for x in (a, b, c, d))
    ctx.tdma(da, db, dc, dd)
    got = dd.download()
    for s in range(0, count, max(1, count // 7)):
        np.testing.assert_allclose(got[:, s], oracle.tdma(a[:, s], b[:, s], c[:, s], d[:, s]), rtol=1e-13, atol=1e-15)


def test_tdma_rejects_single_unknown():
    ctx = A.Context(8, 8, 1.0, 1.0)
    x = A.DeviceArray.from_host(np.ones((1, 4)))
    with pytest.raises(A.Asd2Error):
        ctx.tdma(x, x, x, x)


# axisymmetric momentum: U / V ghost rules (kinds, constants) of each case
MOM_BCS = [
    (("axisym", "wall", "neumann", "outlet"), (0.0, 0.0, 0.0, 0.0),
     ("axisym", "wall", "wall", "outlet"), (0.0, 0.0, 0.0, 0.0)),
    (("dirichlet", "neumann", "dirichlet", "inlet"), (0.0, 0.0, 0.1, -0.2),
     ("neumann", "dirichlet", "axisym", "dirichlet"), (0.0, 0.05, 0.0, 0.3)),
    (("periodic",) * 4, (0.0,) * 4, ("periodic",) * 4, (0.0,) * 4),
]


@pytest.mark.parametrize("bc", range(len(MOM_BCS)))
@pytest.mark.parametrize("nr,nz,gaxis,fix", [(64, 64, 1, 1), (37, 52, 0, 0), (48, 29, 0, 1)])
def test_axisym_momentum(oracle, nr, nz, gaxis, fix, bc):
    """GetRHSU -> GetUHat -> U ghosts -> GetRHSV -> GetVHat -> V ghosts (the reference driver's
    order, tests/LargeTests/test_mac2d_large.cpp:1282-1290) against the oracle, each piece on the
    oracle's own inputs and then the chain on the device.  Bars: viscous terms and both line sweeps
    bit for bit (the same operations in the same order; dr = 1/64 takes the exact-product path);
    the right-hand sides, whose convection is the momentum kernels' WENO5, relative 1e-11 of their
    scale as for the planar momentum terms; the chained velocities relative 1e-10."""
    import tests.scenarios as S
    bu, cu, bv, cv = MOM_BCS[bc]
    dr, dz = 1.0 / nr, 1.0 / nr
    gr = make_grid(nr, nz, dr, dz, bc_p=("axisym", "neumann", "neumann", "neumann"),
                   bc_u=bu, c_u=cu, bc_v=bv, c_v=cv)
    rng = np.random.default_rng(nr + nz + bc)
    rc = (np.arange(nr + 2 * G) + 0.5 - G) * dr
    zc = (np.arange(nz + 2 * G) + 0.5 - G) * dz
    Rr, Zz = np.meshgrid(rc, zc, indexing="ij")
    R0, z0 = rng.uniform(0.2, 0.3), rng.uniform(0.4, 0.6) * nz * dz
    ls = np.ascontiguousarray(R0 - np.sqrt(Rr ** 2 + (Zz - z0) ** 2))
    H = oracle.smooth_heaviside(gr, ls)
    u, v = S.smooth_velocity(nr + nz, nr, nz, dr, dz)
    u, v = u * 0.5, v * 0.5
    oracle.bc_u(gr, u)
    oracle.bc_v(gr, v)
    fl = make_fluid(1000.0, 1.2, 1e-2, 1.8e-3, 0.0, 2e-3, g=9.8, g_axis=gaxis, gravity_fix=fix)
    flg = A.make_fluid(1000.0, 1.2, 1e-2, 1.8e-3, 0.0, 2e-3, g=9.8, g_axis=gaxis, gravity_fix=fix)
    ctx = A.Context(nr, nz, dr, dz)
    ctx.set_bc(A.GRID_U, bu, cu)
    ctx.set_bc(A.GRID_V, bv, cv)
    d_ls, d_H, d_u, d_v = ctx.array(ls), ctx.array(H), ctx.array(u), ctx.array(v)
    tag = f"{nr}x{nz} bc {bc} g_axis {gaxis} fix {fix}"
    print(tag, flush=True)
    # U: viscous term, right-hand side, line sweeps
    out = ctx.array()
    ctx.momentum_term_axisym(flg, A.GRID_U, A.MOM_VISCOSITY, d_ls, d_u, None, d_v, d_H, out)
    np.testing.assert_array_equal(out.download(), oracle.viscosity_u_axisym(gr, fl, v, ls, H))
    rhs_u = oracle.rhs_u_axisym(gr, fl, ls, u, v, H)
    ctx.rhs_u_axisym(flg, d_ls, d_u, d_v, d_H, out)
    _close(out.download(), rhs_u, 1e-11, f"{tag}: U right-hand side")
    uh_ref = oracle.uhat_axisym(gr, fl, ls, rhs_u, H)
    d_uh = ctx.array()
    ctx.uhat_axisym(flg, d_ls, ctx.array(rhs_u), d_H, d_uh)
    np.testing.assert_array_equal(d_uh.download(), uh_ref)
    oracle.bc_u(gr, uh_ref)
    # V: viscous term with u*, right-hand side, line sweeps
    d_uhr = ctx.array(uh_ref)
    ctx.momentum_term_axisym(flg, A.GRID_V, A.MOM_VISCOSITY, d_ls, d_u, d_uhr, d_v, d_H, out)
    np.testing.assert_array_equal(out.download(), oracle.viscosity_v_axisym(gr, fl, u, uh_ref, ls, H))
    rhs_v = oracle.rhs_v_axisym(gr, fl, ls, u, uh_ref, v, H)
    ctx.rhs_v_axisym(flg, d_ls, d_u, d_uhr, d_v, d_H, out)
    _close(out.download(), rhs_v, 1e-11, f"{tag}: V right-hand side")
    vh_ref = oracle.vhat_axisym(gr, fl, ls, rhs_v, H)
    d_vh = ctx.array()
    ctx.vhat_axisym(flg, d_ls, ctx.array(rhs_v), d_H, d_vh)
    np.testing.assert_array_equal(d_vh.download(), vh_ref)
    oracle.bc_v(gr, vh_ref)
    # the chain on the device
    d_ru, d_rv, d_us, d_vs = ctx.array(), ctx.array(), ctx.array(), ctx.array()
    ctx.rhs_u_axisym(flg, d_ls, d_u, d_v, d_H, d_ru)
    ctx.uhat_axisym(flg, d_ls, d_ru, d_H, d_us)
    ctx.bc_apply(A.GRID_U, d_us)
    ctx.rhs_v_axisym(flg, d_ls, d_u, d_us, d_v, d_H, d_rv)
    ctx.vhat_axisym(flg, d_ls, d_rv, d_H, d_vs)
    ctx.bc_apply(A.GRID_V, d_vs)
    _close(d_us.download(), uh_ref, 1e-10, f"{tag}: chained u*")
    _close(d_vs.download(), vh_ref, 1e-10, f"{tag}: chained v*")


def test_axisym_momentum_refusals():
    fl = A.make_fluid(1.0, 1.0, 1.0, 1.0, 0.0, 1e-3)
    ctx = A.Context(8, 8, 0.125, 0.125)
    x, y = ctx.array(), ctx.array()
    with pytest.raises(A.Asd2Error):
        ctx.vhat_axisym(fl, x, y, x, y)  # output aliases the right-hand side
    with pytest.raises(A.Asd2Error):
        ctx.momentum_term_axisym(fl, A.GRID_V, A.MOM_VISCOSITY, x, x, None, x, x, y)  # V needs u*
