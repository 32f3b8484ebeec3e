"""The rising-bubble run (bubble r = 0.25 at (0.5, 0.5) in the [0, 1] x [0, 2] box, density
1000/100, viscosity 10/1, sigma = 24.5, g = 0.98, no-slip walls, Neumann pressure) at 256 x 512
for 20 steps and at the full 1024 x 2048 for 5 steps, once with the mixed-precision multilevel
pressure path and once with the FP64 Jacobi-PCG path (the reference's method, src/poisson2d.cpp:285-401), both to a relative FP64
residual of 1e-8.  The step is the driver loop of the reference (tests/LargeTests/
test_mac2d_large.cpp:376-540: advection, 3 MinRK2 reinitialisation iterations, curvature,
Heaviside, momentum predictor, GFM assembly, pressure, projection, CFL step), gravity on v
applied.  The two runs must agree on the bubble's mass, centroid and circularity
(asd2_ls_diagnostics) to 1e-6 relative and on the level set and velocity fields to 1e-6 of their
maxima: the mixed-precision path does not change the physics.  The bubble rises (mean v > 0 over
its cells)."""
import numpy as np
import pytest

import asd2_amd as A
from tests import scenarios as S

pytestmark = pytest.mark.gpu


class Bubble:
    RHO_H, RHO_L, MU_H, MU_L, SIGMA, G, CFL = 1000.0, 100.0, 10.0, 1.0, 24.5, 0.98, 0.3

    def __init__(self, NX, NY):
        self.dx = 1.0 / NX
        c = A.Context(NX, NY, self.dx, self.dx)
        c.set_bc(A.GRID_P, ("neumann",) * 4)
        c.set_bc(A.GRID_U, ("dirichlet",) * 4)
        c.set_bc(A.GRID_V, ("dirichlet",) * 4)
        self.c = c
        phi = S.circle_sdf(NX, NY, self.dx, self.dx, 0.5, 0.5, 0.25, heavy_inside=False)  # liquid outside
        self.ls = c.array(phi)
        self.u, self.v, self.uh, self.vh = c.array(), c.array(), c.array(), c.array()
        self.kap, self.H, self.Hs, self.div, self.ps = c.array(), c.array(), c.array(), c.array(), c.array()
        c.bc_apply(A.GRID_P, self.ls)
        c.ls_reinit_minrk2(self.ls, 3)
        self.dt = self.CFL * self.dx

    def fluid(self):
        return A.make_fluid(self.RHO_H, self.RHO_L, self.MU_H, self.MU_L, self.SIGMA, self.dt, g=self.G,
                            cfl=self.CFL)

    def step(self, opts):
        c, fl = self.c, self.fluid()
        c.ls_advect(self.ls, self.u, self.v, self.dt)
        c.ls_reinit_minrk2(self.ls, 3)
        c.kappa(self.ls, self.kap)
        c.heaviside(self.ls, self.H)
        c.heaviside(self.ls, self.Hs, smooth=True)
        c.intermediate_vel(fl, 2, self.ls, self.u, self.v, self.uh, self.vh, self.Hs)
        c.bc_apply(A.GRID_U, self.uh)
        c.bc_apply(A.GRID_V, self.vh)
        c.divergence(self.uh, self.vh, self.div)
        c.bc_apply(A.GRID_P, self.kap)
        c.gfm_assemble(fl, self.ls, self.kap, self.H, self.div)
        st = c.poisson_solve(self.ps, opts)
        assert st.converged == 1 and st.true_rel_residual <= opts.tol * 1.0000001, st
        c.bc_apply(A.GRID_P, self.ps)
        c.update_vel(fl, self.u, self.v, self.uh, self.vh, self.ps, self.ls, self.kap, self.H)
        self.dt = c.update_dt(fl, self.u, self.v)
        return st

    def light_phase(self):
        """diagnostics of the bubble (the phi < 0 phase): run on -phi"""
        neg = self.c.array(-self.ls.download())
        return self.c.ls_diagnostics(neg)


def _run(opts, NX, NY, STEPS):
    import time
    b = Bubble(NX, NY)
    d0 = b.light_phase()
    iters = []
    t0 = time.perf_counter()
    for _ in range(STEPS):
        iters.append(b.step(opts).iterations)
    b.c.sync()
    ms = (time.perf_counter() - t0) * 1e3 / STEPS
    return d0, b.light_phase(), iters, ms, b.ls.download(), b.v.download()


@pytest.mark.parametrize("NX,NY,STEPS", [(256, 512, 20), (1024, 2048, 5)])
def test_rising_bubble_mixed_matches_fp64(NX, NY, STEPS):
    tol = 1e-8
    d0, dm, it_m, ms_m, phi_m, v_m = _run(A.Context.opts(tol=tol, precision=A.PREC_MIXED,
                                                          precond=A.PRECOND_MULTILEVEL), NX, NY, STEPS)
    _, df, it_f, ms_f, phi_f, v_f = _run(A.Context.opts(tol=tol, precision=A.PREC_FP64, mode=A.CG_STANDARD),
                                         NX, NY, STEPS)
    g = 3
    inner = (slice(g, g + NX), slice(g, g + NY))
    e_phi = float(np.abs(phi_m[inner] - phi_f[inner]).max() / np.abs(phi_f[inner]).max())
    e_v = float(np.abs(v_m[inner] - v_f[inner]).max() / np.abs(v_f[inner]).max())
    rise_m = float(v_m[inner][phi_m[inner] < 0].mean())
    rise_f = float(v_f[inner][phi_f[inner] < 0].mean())
    print(f"{NX}x{NY}, {STEPS} steps: multilevel {ms_m:.2f} ms/step ({np.mean(it_m):.0f} PCG iterations), "
          f"FP64 Jacobi-PCG {ms_f:.2f} ms/step ({np.mean(it_f):.0f})")
    for k in ("mass", "centroid_x", "centroid_y", "circularity", "length"):
        print(f"  {k}: start {d0[k]:.9f}  multilevel {dm[k]:.12f}  FP64 {df[k]:.12f}")
    print(f"  max|dphi| / max|phi| = {e_phi:.2e}, max|dv| / max|v| = {e_v:.2e}, mean v in the bubble "
          f"{rise_m:.6e} (FP64 {rise_f:.6e})")
    assert rise_m > 0 and rise_f > 0  # the bubble rises
    assert e_phi <= 1e-6 and e_v <= 1e-6
    for k in ("centroid_x", "centroid_y", "circularity", "length"):
        assert abs(dm[k] - df[k]) <= 1e-6 * abs(df[k]), k
    assert abs(dm["mass"] - df["mass"]) <= 2 * (1.0 / NX) ** 2  # at most two cells change sign
    assert np.mean(it_m) * 10 <= np.mean(it_f)
