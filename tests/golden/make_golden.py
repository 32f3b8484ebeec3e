"""Writes the level-set / BC golden vectors of tests/golden/*.npz.

The expected outputs come from the reference's own sources (src/levelset2d.cpp and
src/bc2d.cpp, compiled unmodified into oracle/_ref/libasd2ref.so by oracle/Makefile);
the inputs are seeded synthetic fields.  Run from the repository root:

    make -C oracle all ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import BC, Reference, make_grid  # noqa: E402
from tests import scenarios as S  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def grid_dict(gr):
    return {"nx": gr.nx, "ny": gr.ny, "g": gr.g, "dx": gr.dx, "dy": gr.dy,
            "bc_p": np.array(list(gr.bc_p)), "c_p": np.array(list(gr.c_p)),
            "bc_u": np.array(list(gr.bc_u)), "c_u": np.array(list(gr.c_u)),
            "bc_v": np.array(list(gr.bc_v)), "c_v": np.array(list(gr.c_v))}


def level_set(nx, ny, dx, seed):
    phi = S.circle_sdf(nx, ny, dx, dx, 0.52 * nx * dx, 0.47 * ny * dx, 0.3 * min(nx, ny) * dx)
    phi = phi * (1.0 + 0.25 * np.sin(3.0 * phi / dx)) + 0.05 * dx * S.random_field(seed, phi.shape)
    return np.ascontiguousarray(phi)


def main():
    R = Reference()
    cases = {}
    nx, ny, dx = 24, 20, 1.0 / 24

    # boundary conditions
    for name, kinds in (("neumann", ("neumann",) * 4), ("dirichlet", ("dirichlet",) * 4),
                        ("mixed", ("neumann", "dirichlet", "lsfree", "wall"))):
        gr = make_grid(nx, ny, dx, dx, bc_p=kinds, c_p=(0.1, -0.2, 0.3, 0.4))
        a = S.random_field(100, (nx + 6, ny + 6))
        cases[f"bc_p_{name}"] = dict(grid=gr, a=a, out=R.bc_p(gr, a.copy()))
    for kind in ("neumann", "dirichlet", "outlet", "wall"):
        gr = make_grid(nx, ny, dx, dx, bc_u=(kind,) * 4, c_u=(0.1, 0.2, 0.3, 0.4),
                       bc_v=(kind,) * 4, c_v=(0.5, 0.6, -0.7, 0.8))
        a = S.random_field(101, (nx + 6, ny + 6))
        b = S.random_field(102, (nx + 6, ny + 6))
        cases[f"bc_uv_{kind}"] = dict(grid=gr, a=a, b=b, out_u=R.bc_u(gr, a.copy()), out_v=R.bc_v(gr, b.copy()))

    # WENO5 line derivative
    n = 40
    F = np.ascontiguousarray(np.sin(np.linspace(0, 4, n + 6)) + 0.1 * S.random_field(103, (n + 6,)))
    FP, FM = R.unit_hjweno5(F, 0.1, n, 3)
    cases["unit_hjweno5"] = dict(F=F, d=np.array(0.1), n=np.array(n), FP=FP, FM=FM)

    gr = make_grid(nx, ny, dx, dx)
    phi = level_set(nx, ny, dx, 104)
    uc, vc = S.random_field(105, phi.shape), S.random_field(106, phi.shape)
    p2 = phi.copy()
    L = R.hjweno5_ls(gr, p2, uc, vc)
    cases["hjweno5_ls"] = dict(grid=gr, phi=phi, uc=uc, vc=vc, L=L, phi_bc=p2)

    # one SSP-RK3 step of Solve_LevelSet_2D, two BC sets
    for name, kinds in (("neumann", ("neumann",) * 4), ("dirichlet", ("dirichlet",) * 4)):
        gr = make_grid(nx, ny, dx, dx, bc_p=kinds, c_p=(0.02, -0.02, 0.01, 0.0))
        u, v = S.single_vortex(nx, ny, dx, dx)
        phi = level_set(nx, ny, dx, 107)
        cases[f"advect_{name}"] = dict(grid=gr, phi=phi, u=u, v=v, dt=np.array(0.5 * dx),
                                      out=R.solve_levelset(gr, phi.copy(), u, v, 0.5 * dx))

    # |grad phi| with and without the subcell fix, smoothed sign, Heaviside derivative
    gr = make_grid(nx, ny, dx, dx)
    phi0 = level_set(nx, ny, dx, 108)
    R.bc_p(gr, phi0)
    phi1 = np.ascontiguousarray(phi0 + 0.2 * dx * S.random_field(109, phi0.shape))
    cases["absgrad"] = dict(grid=gr, phi=phi1, phi0=phi0, eno=R.eno_absgrad(gr, phi1, phi0),
                            subcell=R.subcell_absgrad(gr, phi1, phi0), sign=R.smoothed_sign(gr, phi0),
                            hderiv=R.heaviside_deriv(gr, phi1))
    cases["weno_absgrad"] = dict(grid=gr, phi=phi1, phi0=phi0, out=R.weno_absgrad(gr, phi1, phi0))
    hd = R.heaviside_deriv(gr, phi1)
    cases["sussman_constraint"] = dict(grid=gr, phi=phi1, phi0=phi0, hd=hd,
                                       out=R.sussman_constraint(gr, phi1, phi0, hd))

    # reinitialisation: MinRK2 with 3 and 5 iterations, Sussman with 5
    for K in (3, 5):
        phi = level_set(nx, ny, dx, 110 + K)
        cases[f"minrk2_{K}"] = dict(grid=gr, phi=phi, iters=np.array(K),
                                    out=R.reinit_minrk2(gr, phi.copy(), (K - 0.5) * 0.5 * dx))
    phi = level_set(nx, ny, dx, 120)
    cases["sussman_5"] = dict(grid=gr, phi=phi, iters=np.array(5), out=R.reinit_sussman(gr, phi.copy(), 2.25 * dx))
    cases["original_5"] = dict(grid=gr, phi=phi, iters=np.array(5),
                               out=R.reinit_original(gr, phi.copy(), 2.25 * dx))
    # first-time Sussman: the loop runs until the pseudo time reaches max(dx nx, dy ny)
    cases["first_time"] = dict(grid=gr, phi=phi, out=R.reinit_first_time(gr, phi.copy()))

    # a 64^2 droplet (config-1 geometry) after one advection step + 3 MinRK2 iterations
    n64 = 64
    d64 = 1.0 / n64
    gr = make_grid(n64, n64, d64, d64)
    phi = S.circle_sdf(n64, n64, d64, d64, 0.5, 0.5, 0.25)
    u, v = S.single_vortex(n64, n64, d64, d64)
    out = R.solve_levelset(gr, phi.copy(), u, v, 0.5 * d64)
    out = R.reinit_minrk2(gr, out, 1.25 * d64)
    cases["droplet64_step"] = dict(grid=gr, phi=phi, u=u, v=v, dt=np.array(0.5 * d64), iters=np.array(3), out=out)

    for name, c in cases.items():
        flat = {}
        for k, v in c.items():
            if k == "grid":
                for gk, gv in grid_dict(v).items():
                    flat["grid_" + gk] = np.asarray(gv)
            else:
                flat[k] = np.asarray(v)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **flat)
    print(f"wrote {len(cases)} fixtures to {OUT}")


if __name__ == "__main__":
    main()
