"""GPU parity of the level-set path (reference src/levelset2d.cpp, src/bc2d.cpp) against the
FP64 oracle, called through the C ABI.  BC fills and the HJ-WENO5 / SSP-RK3 advection are
bit-exact on every grid (the stage kernels evaluate the reference's operation order); the
reinitialisations are bit-exact where the spacing is a power of two and within max |delta phi|
<= 1e-12 (the parity bar of the level set) elsewhere."""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_grid
from tests import scenarios as S

pytestmark = pytest.mark.gpu

TOL_PHI = 1e-12


def _bar(dx):
    """The parity bar: bit-exact where the spacing is a power of two (the stencils then evaluate the
    reference's operation order exactly, divisions by dx included), 1e-12 otherwise."""
    m, _ = np.frexp(dx)
    return 0.0 if m == 0.5 else TOL_PHI

GRIDS = [(64, 64), (67, 45), (128, 128)]
P_BCS = [("neumann",) * 4, ("dirichlet",) * 4, ("neumann", "dirichlet", "wall", "outlet"),
         ("neumann", "neumann", "custom", "lsfree")]


def _ctx(nx, ny, dx, dy, bc_p=("neumann",) * 4, c_p=(0.0,) * 4):
    c = A.Context(nx, ny, dx, dy)
    c.set_bc(A.GRID_P, bc_p, c_p)
    return c


def _level_set(nx, ny, seed=0):
    dx = 1.0 / max(nx, ny)
    phi = S.circle_sdf(nx, ny, dx, dx, 0.5 * nx * dx, 0.45 * ny * dx, 0.22 * min(nx, ny) * dx)
    phi += 0.02 * dx * S.random_field(seed, phi.shape)
    return np.ascontiguousarray(phi), dx


@pytest.mark.parametrize("bc", P_BCS, ids=lambda b: "-".join(b))
def test_bc_p_exact(oracle, bc):
    nx, ny, dx = 37, 29, 0.03
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=(0.1, -0.2, 0.3, 0.4))
    a = S.random_field(1, (nx + 6, ny + 6))
    ctx = _ctx(nx, ny, dx, dx, bc, (0.1, -0.2, 0.3, 0.4))
    d = ctx.array(a)
    ctx.bc_apply(A.GRID_P, d)
    np.testing.assert_array_equal(d.download(), oracle.bc_p(gr, a.copy()))


@pytest.mark.parametrize("kind", ["neumann", "dirichlet", "periodic", "axisym", "wall", "inlet", "outlet"])
def test_bc_uv_exact(oracle, kind):
    nx, ny, dx = 31, 26, 0.03
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


@pytest.mark.parametrize("nx,ny", GRIDS)
@pytest.mark.parametrize("bc", P_BCS[:3], ids=lambda b: "-".join(b))
def test_advect_one_step(oracle, nx, ny, bc):
    phi, dx = _level_set(nx, ny)
    u, v = S.single_vortex(nx, ny, dx, dx)
    dt = 0.5 * dx
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=(0.05, 0.0, -0.05, 0.1))
    ref = oracle.solve_levelset(gr, phi.copy(), u, v, dt)
    ctx = _ctx(nx, ny, dx, dx, bc, (0.05, 0.0, -0.05, 0.1))
    d_phi, d_u, d_v = ctx.array(phi), ctx.array(u), ctx.array(v)
    ctx.ls_advect(d_phi, d_u, d_v, dt)
    out = d_phi.download()
    err = np.abs(out - ref).max()
    print(f"advect {nx}x{ny} {bc}: max|dphi| = {err:.3e}")
    assert err == 0.0
    assert np.abs(out - phi).max() > 1e-4  # the step moved the level set


def test_advect_custom_and_lsfree(oracle):
    nx, ny = 48, 40
    phi, dx = _level_set(nx, ny, 3)
    u, v = S.single_vortex(nx, ny, dx, dx)
    bc = P_BCS[3]
    gr = make_grid(nx, ny, dx, dx, bc_p=bc)
    ref = oracle.solve_levelset(gr, phi.copy(), u, v, 0.3 * dx)
    ctx = _ctx(nx, ny, dx, dx, bc)
    d_phi = ctx.array(phi)
    ctx.ls_advect(d_phi, ctx.array(u), ctx.array(v), 0.3 * dx)
    assert np.abs(d_phi.download() - ref).max() == 0.0


def test_weno_rhs(oracle):
    nx, ny = 53, 47
    phi, dx = _level_set(nx, ny, 4)
    uc = S.random_field(5, phi.shape)
    vc = S.random_field(6, phi.shape)
    gr = make_grid(nx, ny, dx, dx)
    ref = oracle.hjweno5_ls(gr, phi.copy(), uc, vc)
    ctx = _ctx(nx, ny, dx, dx)
    d_phi, d_L = ctx.array(phi), ctx.array()
    ctx.ls_weno_rhs(d_phi, ctx.array(uc), ctx.array(vc), d_L)
    np.testing.assert_array_equal(d_L.download(), ref)


@pytest.mark.parametrize("subcell", [True, False])
def test_absgrad(oracle, subcell):
    nx, ny = 61, 58
    phi, dx = _level_set(nx, ny, 7)
    gr = make_grid(nx, ny, dx, dx)
    oracle.bc_p(gr, phi)
    phi0 = phi.copy()
    phi1 = phi + 0.3 * dx * S.random_field(8, phi.shape)
    fn = oracle.subcell_absgrad if subcell else oracle.eno_absgrad
    ref = fn(gr, phi1, phi0)
    ctx = _ctx(nx, ny, dx, dx)
    d_out = ctx.array()
    if subcell:
        ctx.ls_subcell_absgrad(ctx.array(phi1), ctx.array(phi0), d_out)
    else:
        ctx.ls_eno_absgrad(ctx.array(phi1), ctx.array(phi0), d_out)
    assert np.abs(d_out.download() - ref).max() <= 1e-12


@pytest.mark.parametrize("iters", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("nx,ny", GRIDS)
def test_reinit_minrk2(oracle, nx, ny, iters):
    # one launch per iteration with phi ping-ponged between two arrays: even and odd counts
    phi, dx = _level_set(nx, ny, 9)
    phi = np.ascontiguousarray(phi * (1.0 + 0.3 * np.sin(7 * phi / dx)))  # not a distance function
    gr = make_grid(nx, ny, dx, dx)
    ref = oracle.reinit_minrk2(gr, phi.copy(), iters)
    ctx = _ctx(nx, ny, dx, dx)
    d = ctx.array(phi)
    ctx.ls_reinit_minrk2(d, iters)
    err = np.abs(d.download() - ref).max()
    print(f"minrk2 {nx}x{ny} K={iters}: max|dphi| = {err:.3e}")
    assert err <= _bar(dx)


@pytest.mark.parametrize("iters", [2, 3])
@pytest.mark.parametrize("bc", P_BCS, ids=lambda b: "-".join(b))
def test_reinit_minrk2_sides(oracle, bc, iters):
    # every ghost rule of the P grid on the substage-A ring (mirror, -mirror + 2c); CUSTOM and
    # LS-free sides run the two-launch path; the whole padded array is compared, ghosts included
    nx, ny = 67, 45
    c_p = (0.05, -0.02, 0.0, 0.1)
    phi, dx = _level_set(nx, ny, 21)
    phi = np.ascontiguousarray(phi * (1.0 + 0.3 * np.sin(5 * phi / dx)))
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=c_p)
    ref = oracle.reinit_minrk2(gr, phi.copy(), iters)
    ctx = _ctx(nx, ny, dx, dx, bc, c_p)
    d = ctx.array(phi)
    ctx.ls_reinit_minrk2(d, iters)
    err = np.abs(d.download() - ref).max()
    print(f"minrk2 {bc} K={iters}: max|dphi| = {err:.3e}")
    assert err <= TOL_PHI


@pytest.mark.parametrize("nx,ny", GRIDS[:2])
def test_reinit_sussman(oracle, nx, ny):
    phi, dx = _level_set(nx, ny, 10)
    phi = np.ascontiguousarray(phi * (1.0 + 0.2 * np.cos(5 * phi / dx)))
    gr = make_grid(nx, ny, dx, dx)
    ref = oracle.reinit_sussman(gr, phi.copy(), 5)
    ctx = _ctx(nx, ny, dx, dx)
    d = ctx.array(phi)
    ctx.ls_reinit_sussman(d, 5)
    err = np.abs(d.download() - ref).max()
    print(f"sussman {nx}x{ny}: max|dphi| = {err:.3e}")
    assert err <= TOL_PHI


@pytest.mark.parametrize("bc", P_BCS[:3], ids=lambda b: "-".join(b))
def test_reinit_original(oracle, bc):
    nx, ny = 67, 45
    phi, dx = _level_set(nx, ny, 12)
    phi = np.ascontiguousarray(phi * (1.0 + 0.25 * np.sin(3 * phi / dx)))
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=(0.02, -0.01, 0.0, 0.03))
    ref = oracle.reinit_original(gr, phi.copy(), 5)
    ctx = _ctx(nx, ny, dx, dx, bc, (0.02, -0.01, 0.0, 0.03))
    d = ctx.array(phi)
    ctx.ls_reinit_original(d, 5)
    err = np.abs(d.download() - ref).max()
    print(f"original {bc}: max|dphi| = {err:.3e}")
    assert err <= TOL_PHI


def test_reinit_first_time(oracle):
    nx, ny = 40, 36
    phi, dx = _level_set(nx, ny, 16)
    phi = np.ascontiguousarray(phi * 1.8)
    gr = make_grid(nx, ny, dx, dx)
    iters = oracle.reinit_iterations(dx, max(dx * nx, dx * ny))
    ref = oracle.reinit_first_time(gr, phi.copy(), iters)
    ctx = _ctx(nx, ny, dx, dx)
    d = ctx.array(phi)
    ctx.ls_reinit_first_time(d, iters)
    err = np.abs(d.download() - ref).max()
    print(f"first-time sussman ({iters} iterations): max|dphi| = {err:.3e}")
    # ~2 nx pseudo-time iterations: rounding differences of the reorganised ENO2 accumulate
    assert err <= 1e-10


def test_weno_absgrad(oracle):
    nx, ny = 59, 62
    phi, dx = _level_set(nx, ny, 17)
    gr = make_grid(nx, ny, dx, dx)
    oracle.bc_p(gr, phi)
    phi1 = np.ascontiguousarray(phi + 0.3 * dx * S.random_field(18, phi.shape))
    ref = oracle.weno_absgrad(gr, phi1, phi)
    ctx = _ctx(nx, ny, dx, dx)
    out = ctx.array()
    ctx.ls_weno_absgrad(ctx.array(phi1), ctx.array(phi), out)
    assert np.abs(out.download() - ref).max() <= 1e-12


def test_pointwise_sign_and_delta(oracle):
    nx, ny = 50, 41
    phi, dx = _level_set(nx, ny, 13)
    gr = make_grid(nx, ny, dx, dx)
    ctx = _ctx(nx, ny, dx, dx)
    d_phi, d_out = ctx.array(phi), ctx.array()
    ctx.ls_smoothed_sign(d_phi, d_out)
    ref = oracle.smoothed_sign(gr, phi)
    assert np.abs(d_out.download() - ref).max() <= 1e-15
    ctx.ls_heaviside_deriv(d_phi, d_out)
    ref = oracle.heaviside_deriv(gr, phi)
    out = d_out.download()
    np.testing.assert_array_equal(out == 0.0, ref == 0.0)  # the mask the constraint uses
    assert np.abs(out - ref).max() <= 1e-14 * np.abs(ref).max()


def test_sussman_constraint(oracle):
    nx, ny = 47, 52
    phi, dx = _level_set(nx, ny, 14)
    gr = make_grid(nx, ny, dx, dx)
    oracle.bc_p(gr, phi)
    phi0 = phi.copy()
    phi1 = np.ascontiguousarray(phi + 0.2 * dx * S.random_field(15, phi.shape))
    oracle.bc_p(gr, phi1)
    hd = oracle.heaviside_deriv(gr, phi1)
    ref = oracle.sussman_constraint(gr, phi1, phi0, hd)
    ctx = _ctx(nx, ny, dx, dx)
    out = ctx.array()
    ctx.ls_sussman_constraint(ctx.array(phi1), ctx.array(phi0), ctx.array(hd), out)
    err = np.abs(out.download() - ref).max()
    print(f"sussman constraint: max err {err:.3e} (scale {np.abs(ref).max():.3e})")
    assert err <= 1e-13


@pytest.mark.parametrize("n", [7, 64, 301])
def test_unit_weno5(oracle, n):
    g, d = 3, 0.37
    F = np.ascontiguousarray(np.sin(np.linspace(0, 7, n + 2 * g)) + 0.1 * S.random_field(n, (n + 2 * g,)))
    rp, rm = oracle.unit_hjweno5(F, d, n, g)
    ctx = _ctx(8, 8, 0.1, 0.1)
    fp, fm = ctx.ls_unit_weno5(F, d, n, g)
    np.testing.assert_array_equal(fp[g:n + g], rp[g:n + g])  # the reference's operation order: bit-exact
    np.testing.assert_array_equal(fm[g:n + g], rm[g:n + g])


def test_reinit_iterations_match_time_loop(oracle):
    lib = A.load()
    for eps in (1.0 / 128, 0.04 / 128, 1.0 / 2048):
        for mt in (1.25 * eps, 2.25 * eps, 1.5 * 1.5 * eps, 256 * eps):
            assert lib.asd2_ls_reinit_iterations(eps, mt) == oracle.reinit_iterations(eps, mt)


PERIODIC_BCS = [("periodic",) * 4, ("periodic", "periodic", "neumann", "dirichlet"),
                ("neumann", "dirichlet", "periodic", "periodic")]


@pytest.mark.parametrize("bc", PERIODIC_BCS, ids=lambda b: "-".join(b))
def test_periodic_p_level_set(oracle, bc):
    """Periodic P sides (src/bc2d.cpp:786-808, reproduced with the east fill that writes interior
    rows) through the ghost fill, one advection and every reinitialisation, against the oracle,
    which the CPU tests compare with the reference's own code on the same sides."""
    nx, ny = 45, 38
    c_p = (0.02, -0.01, 0.03, 0.0)
    phi, dx = _level_set(nx, ny, 31)
    gr = make_grid(nx, ny, dx, dx, bc_p=bc, c_p=c_p)
    ctx = _ctx(nx, ny, dx, dx, bc, c_p)
    a = S.random_field(5, (nx + 6, ny + 6))
    d = ctx.array(a)
    ctx.bc_apply(A.GRID_P, d)
    np.testing.assert_array_equal(d.download(), oracle.bc_p(gr, a.copy()))
    u, v = S.single_vortex(nx, ny, dx, dx)
    d = ctx.array(phi)
    ctx.ls_advect(d, ctx.array(u), ctx.array(v), 0.4 * dx)
    ref = oracle.solve_levelset(gr, phi.copy(), u, v, 0.4 * dx)
    assert np.abs(d.download() - ref).max() == 0.0
    wavy = np.ascontiguousarray(phi * (1.0 + 0.3 * np.sin(5 * phi / dx)))
    for name, k in (("minrk2", 3), ("original", 4), ("sussman", 3), ("first_time", 6)):
        d = ctx.array(wavy)
        getattr(ctx, "ls_reinit_" + name)(d, k)
        ref = getattr(oracle, "reinit_" + name)(gr, wavy.copy(), k)
        err = np.abs(d.download() - ref).max()
        print(f"{name} {bc}: max|dphi| = {err:.3e}")
        assert err <= TOL_PHI


def test_periodic_p_refused_by_gfm():
    """The ghost-fluid assembly keeps refusing periodic P sides (the reference's periodic rows,
    src/mac2d.cpp:1529-1531, store a density reciprocal as the coupling without the 1/h^2 factor and
    keep the wrapped column index)."""
    n = 16
    dx = 1.0 / n
    ctx = _ctx(n, n, dx, dx, ("periodic",) * 4)
    z = ctx.array()
    fl = A.make_fluid(1000.0, 1.0, 1e-2, 1e-4, 1.0, 1e-4)
    with pytest.raises(A.Asd2Error) as e:
        ctx.gfm_assemble(fl, z, z, z, z)
    assert e.value.code == A.ERR_UNSUPPORTED


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(96, 80), (256, 256)])
def test_ls_diagnostics_vs_oracle(oracle, nx, ny):
    """asd2_ls_diagnostics against the oracle on a perturbed circle with a reference field and
    offset base coordinates: counts and the maximum exact, sums to rel 1e-12 (block order)."""
    phi, dx = _level_set(nx, ny)
    ref = np.ascontiguousarray(phi + 0.3 * dx * S.random_field(7, phi.shape))
    gr = make_grid(nx, ny, dx, dx)
    ctx = A.Context(nx, ny, dx, dx, base_x=0.75, base_y=-1.5)
    got = ctx.ls_diagnostics(ctx.array(phi), ctx.array(ref))
    exp = oracle.ls_diagnostics(gr, phi, ref=ref, base=(0.75, -1.5))
    print(got)
    for k in ("cells", "ref_cells", "err_linf"):
        assert got[k] == exp[k], k
    for k in ("mass", "centroid_x", "centroid_y", "length", "circularity", "err_l1", "err_l2", "mass_loss_pct"):
        assert got[k] == pytest.approx(exp[k], rel=1e-12, abs=1e-300), k
    alone = ctx.ls_diagnostics(ctx.array(phi))
    assert alone["err_l1"] == 0.0 and alone["cells"] == exp["cells"]
