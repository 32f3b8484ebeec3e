"""Parity at the full problem sizes, through properties that do not need the oracle to run the
whole problem.

* Level set, 4096 x 4096 single vortex (circle r = 0.15 at (0.5, 0.75), CFL 0.5, one HJ-WENO5 +
  TVD-RK3 step and 5 MinRK2 subcell-fix reinitialisation iterations; reference
  LevelSetSolver2D::Solve_LevelSet_2D, src/levelset2d.cpp:43-86, and Reinit_MinRK2_2D, :433-486).
  Every stencil of the step is local: the advection reaches 3 cells per RK stage (9 per step) and
  each reinitialisation iteration 4 cells (20 for 5), so a cell's result depends only on the
  input within 29 cells.  The oracle runs on windows of the full field cut out with a 40-cell
  margin; inside the margin its result must equal the GPU's full-grid result (bar 1e-12, the
  level-set bar).
* GFM pressure, 2048 x 2048 seeded droplet field (1-20 circles, radii U[0.02, 0.1], density ratio
  log-uniform in [10, 1000], rhs = divergence of a seeded smooth velocity, Neumann walls).  The
  assembly is compared coefficient by coefficient with the oracle's at full size; the multilevel
  mixed-precision solution must satisfy the oracle's own system to the requested true residual
  (recomputed on the host with the oracle's stencil apply) and match the FP64 Jacobi-PCG solution
  to relative L2 <= 1e-8 at true residual 1e-10.  The FP64 Jacobi-PCG run (the reference's
  algorithm, src/poisson2d.cpp:285-401) stops on its recursive residual, asked for 1e-11; after
  its ~15k iterations the true residual has drifted above that, so it is held to 1e-10.
"""
import numpy as np
import pytest

import asd2_amd as A
from oracle import make_fluid, make_grid
from tests import scenarios as S
from tests.gfm_common import interior

pytestmark = pytest.mark.gpu

G = 3
MARGIN = 40  # > 29, the dependency radius of one advection step and 5 reinitialisation iterations


def _window(a, i0, j0, w):
    """Padded sub-array of the interior cells [i0, i0 + w) x [j0, j0 + w) with a G-cell ghost rim
    taken from the full field (interior indices from 0)."""
    return np.ascontiguousarray(a[i0:i0 + w + 2 * G, j0:j0 + w + 2 * G])


def test_levelset_4096_windows_match_oracle(oracle):
    n = 4096
    dx = 1.0 / n
    phi = S.circle_sdf(n, n, dx, dx, 0.5, 0.75, 0.15, heavy_inside=True)
    u, v = S.single_vortex(n, n, dx, dx)
    dt = 0.5 * dx  # CFL 0.5 with max |u| = 1
    ctx = A.Context(n, n, dx, dx)
    d_phi, d_u, d_v = ctx.array(phi), ctx.array(u), ctx.array(v)
    ctx.ls_advect(d_phi, d_u, d_v, dt)
    adv = d_phi.download()
    ctx.ls_reinit_minrk2(d_phi, 5)
    out = d_phi.download()
    assert np.isfinite(out).all()
    assert np.abs(out - phi).max() > 1e-5  # the step moved the level set
    core = 160
    w = core + 2 * MARGIN
    a = G + MARGIN
    worst = 0.0
    # windows on the interface (top, left, lower right) and one far from it
    for cx, cy in ((0.5, 0.9), (0.35, 0.75), (0.5 + 0.15 * np.cos(-0.8), 0.75 + 0.15 * np.sin(-0.8)), (0.2, 0.2)):
        i0 = int(cx * n) - w // 2
        j0 = int(cy * n) - w // 2
        gr = make_grid(w, w, dx, dx)
        ref_adv = oracle.solve_levelset(gr, _window(phi, i0, j0, w), _window(u, i0, j0, w), _window(v, i0, j0, w), dt)
        e_adv = float(np.abs(adv[i0 + a:i0 + a + core, j0 + a:j0 + a + core] - ref_adv[a:a + core, a:a + core]).max())
        ref = oracle.reinit_minrk2(gr, ref_adv.copy(), 5)
        exp = ref[a:a + core, a:a + core]
        err = float(np.abs(out[i0 + a:i0 + a + core, j0 + a:j0 + a + core] - exp).max())
        # the reinitialisation of the GPU's own advected window: separates the two kernels' shares
        own = oracle.reinit_minrk2(gr, _window(adv, i0, j0, w), 5)
        e_re = float(np.abs(out[i0 + a:i0 + a + core, j0 + a:j0 + a + core] - own[a:a + core, a:a + core]).max())
        crosses = bool((exp.min() < 0.0) and (exp.max() > 0.0))
        print(f"window at ({cx:.3f}, {cy:.3f}) interface={crosses}: advection max|dphi| = {e_adv:.3e}, "
              f"reinit of the same input {e_re:.3e}, step {err:.3e}")
        worst = max(worst, err, e_adv, e_re)
    # h = 1/4096 is a power of two: advection and reinitialisation evaluate the reference's operation
    # order exactly, so the step is bit-exact (the subcell fix amplifies last-bit differences of the
    # advected field by up to 1e6 near the interface, so nothing short of exact would be safe here)
    assert worst == 0.0


def _relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_pressure_2048_against_oracle_system(oracle):
    n, seed = 2048, 3
    dx = 1.0 / n
    gr = make_grid(n, n, dx, dx)
    phi, circles = S.random_droplets(seed, n, n, dx, dx)
    oracle.bc_p(gr, phi)
    kap = oracle.update_kappa(gr, phi)
    oracle.bc_p(gr, kap)
    H = oracle.heaviside(gr, phi)
    uu, vv = S.smooth_velocity(seed, n, n, dx, dx)
    div = oracle.divergence(gr, uu, vv)
    ratio = S.density_ratio(seed)
    fl = make_fluid(ratio, 1.0, 1.0, 0.01, 1.0, 1e-4)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)

    ctx = A.Context(n, n, dx, dx)
    ctx.set_bc(A.GRID_P, ("neumann",) * 4)
    flg = A.make_fluid(fl.rho_h, fl.rho_l, fl.mu_h, fl.mu_l, fl.sigma, fl.dt)
    ctx.gfm_assemble(flg, *[ctx.array(a) for a in (phi, kap, H, div)])
    for name, got, exp in zip(("ax", "ay", "diag", "rhs"), ctx.gfm_export(), (ax, ay, dg, rhs)):
        err = float(np.abs(got - exp).max() / max(np.abs(exp).max(), 1e-300))
        assert err <= 1e-14, (name, err)

    bn = np.linalg.norm(interior(gr, rhs))

    def true_res(x):
        return float(np.linalg.norm(interior(gr, rhs - oracle.stencil_apply(gr, ax, ay, dg, x))) / bn)

    d_ref = ctx.array()
    st_ref = ctx.poisson_solve(d_ref, ctx.opts(tol=1e-11))
    x_ref = d_ref.download()
    d_ml = ctx.array()
    st_ml = ctx.poisson_solve(d_ml, ctx.opts(tol=1e-10, precond=A.PRECOND_MULTILEVEL, precision=A.PREC_MIXED))
    x_ml = d_ml.download()
    r_ref, r_ml = true_res(x_ref), true_res(x_ml)
    err = _relL2(interior(gr, x_ml), interior(gr, x_ref))
    print(f"2048^2, {len(circles)} droplets, density ratio {ratio:.1f}: FP64 Jacobi-PCG {st_ref.iterations} iters "
          f"(oracle-system residual {r_ref:.2e}); multilevel {st_ml.iterations} iters, {st_ml.reliable_updates} "
          f"reliable updates (oracle-system residual {r_ml:.2e}); rel L2 {err:.2e}")
    assert st_ref.converged == 1 and st_ml.converged == 1 and st_ml.fallbacks == 0
    assert r_ref <= 1e-10 and r_ml <= 1.05e-10
    assert abs(st_ml.true_rel_residual - r_ml) <= 1e-3 * r_ml  # the solver's own FP64 residual
    assert err <= 1e-8
    assert st_ml.iterations * 20 <= st_ref.iterations
