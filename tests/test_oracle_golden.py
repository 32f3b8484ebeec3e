"""The FP64 oracle against the reference: the committed golden vectors (produced by the
reference's own src/levelset2d.cpp + src/bc2d.cpp, see tests/golden/make_golden.py) and,
where oracle/_ref is built, live comparisons on further seeded inputs.  Bar: bit-exact."""
import glob
import os

import numpy as np
import pytest

from oracle import Grid, make_grid
from tests import scenarios as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    d = {k: z[k] for k in z.files}
    gr = None
    if "grid_nx" in d:
        gr = Grid()
        gr.nx, gr.ny, gr.g = int(d["grid_nx"]), int(d["grid_ny"]), int(d["grid_g"])
        gr.dx, gr.dy = float(d["grid_dx"]), float(d["grid_dy"])
        for k in range(4):
            gr.bc_p[k], gr.c_p[k] = int(d["grid_bc_p"][k]), float(d["grid_c_p"][k])
            gr.bc_u[k], gr.c_u[k] = int(d["grid_bc_u"][k]), float(d["grid_c_u"][k])
            gr.bc_v[k], gr.c_v[k] = int(d["grid_bc_v"][k]), float(d["grid_c_v"][k])
    return gr, d


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def test_all_fixtures_present():
    names = {os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz"))}
    assert {"advect_neumann", "minrk2_3", "sussman_5", "original_5", "first_time", "weno_absgrad", "droplet64_step", "bc_p_mixed"} <= names


@pytest.mark.parametrize("name", ["bc_p_neumann", "bc_p_dirichlet", "bc_p_mixed"])
def test_bc_p(oracle, name):
    gr, d = load(name)
    np.testing.assert_array_equal(oracle.bc_p(gr, c(d["a"])), d["out"])


@pytest.mark.parametrize("name", ["bc_uv_neumann", "bc_uv_dirichlet", "bc_uv_outlet", "bc_uv_wall"])
def test_bc_uv(oracle, name):
    gr, d = load(name)
    np.testing.assert_array_equal(oracle.bc_u(gr, c(d["a"])), d["out_u"])
    np.testing.assert_array_equal(oracle.bc_v(gr, c(d["b"])), d["out_v"])


def test_unit_hjweno5(oracle):
    _, d = load("unit_hjweno5")
    FP, FM = oracle.unit_hjweno5(c(d["F"]), float(d["d"]), int(d["n"]))
    np.testing.assert_array_equal(FP, d["FP"])
    np.testing.assert_array_equal(FM, d["FM"])


def test_hjweno5_ls(oracle):
    gr, d = load("hjweno5_ls")
    phi = c(d["phi"])
    L = oracle.hjweno5_ls(gr, phi, c(d["uc"]), c(d["vc"]))
    np.testing.assert_array_equal(L, d["L"])
    np.testing.assert_array_equal(phi, d["phi_bc"])  # the P BC is applied to the input


@pytest.mark.parametrize("name", ["advect_neumann", "advect_dirichlet"])
def test_advect(oracle, name):
    gr, d = load(name)
    out = oracle.solve_levelset(gr, c(d["phi"]), c(d["u"]), c(d["v"]), float(d["dt"]))
    np.testing.assert_array_equal(out, d["out"])


def test_absgrad_sign_hderiv(oracle):
    gr, d = load("absgrad")
    np.testing.assert_array_equal(oracle.eno_absgrad(gr, c(d["phi"]), c(d["phi0"])), d["eno"])
    np.testing.assert_array_equal(oracle.subcell_absgrad(gr, c(d["phi"]), c(d["phi0"])), d["subcell"])
    np.testing.assert_array_equal(oracle.smoothed_sign(gr, c(d["phi0"])), d["sign"])
    np.testing.assert_array_equal(oracle.heaviside_deriv(gr, c(d["phi"])), d["hderiv"])


def test_sussman_constraint(oracle):
    gr, d = load("sussman_constraint")
    out = oracle.sussman_constraint(gr, c(d["phi"]), c(d["phi0"]), c(d["hd"]))
    np.testing.assert_array_equal(out, d["out"])


@pytest.mark.parametrize("name", ["minrk2_3", "minrk2_5"])
def test_minrk2(oracle, name):
    gr, d = load(name)
    np.testing.assert_array_equal(oracle.reinit_minrk2(gr, c(d["phi"]), int(d["iters"])), d["out"])


def test_sussman(oracle):
    gr, d = load("sussman_5")
    np.testing.assert_array_equal(oracle.reinit_sussman(gr, c(d["phi"]), int(d["iters"])), d["out"])


def test_original(oracle):
    gr, d = load("original_5")
    np.testing.assert_array_equal(oracle.reinit_original(gr, c(d["phi"]), int(d["iters"])), d["out"])


def test_first_time(oracle):
    gr, d = load("first_time")
    iters = oracle.reinit_iterations(min(gr.dx, gr.dy), max(gr.dx * gr.nx, gr.dy * gr.ny))
    np.testing.assert_array_equal(oracle.reinit_first_time(gr, c(d["phi"]), iters), d["out"])


def test_weno_absgrad(oracle):
    gr, d = load("weno_absgrad")
    np.testing.assert_array_equal(oracle.weno_absgrad(gr, c(d["phi"]), c(d["phi0"])), d["out"])


def test_droplet_step(oracle):
    gr, d = load("droplet64_step")
    out = oracle.solve_levelset(gr, c(d["phi"]), c(d["u"]), c(d["v"]), float(d["dt"]))
    out = oracle.reinit_minrk2(gr, out, int(d["iters"]))
    np.testing.assert_array_equal(out, d["out"])


# ----------------------------------------------------------------- live, against oracle/_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("kinds", [("neumann",) * 4, ("dirichlet", "neumann", "outlet", "wall"), ("periodic",) * 4,
                                   ("periodic", "periodic", "neumann", "dirichlet")])
def test_live_levelset_path(oracle, reference, seed, kinds):
    nx, ny = 33 + seed, 29 + 2 * seed
    dx = 1.0 / nx
    gr = make_grid(nx, ny, dx, dx, bc_p=kinds, c_p=(0.01 * seed, 0.0, -0.01, 0.02))
    phi = S.circle_sdf(nx, ny, dx, dx, 0.5, 0.45 * ny * dx, 0.3)
    phi = np.ascontiguousarray(phi * (1 + 0.2 * np.sin(seed * phi / dx)))
    u, v = S.smooth_velocity(seed, nx, ny, dx, dx)
    a, b = phi.copy(), phi.copy()
    np.testing.assert_array_equal(oracle.solve_levelset(gr, a, u, v, 0.4 * dx),
                                  reference.solve_levelset(gr, b, u, v, 0.4 * dx))
    np.testing.assert_array_equal(oracle.reinit_minrk2(gr, a, 5), reference.reinit_minrk2(gr, b, 2.25 * dx))
    np.testing.assert_array_equal(oracle.reinit_sussman(gr, a, 3), reference.reinit_sussman(gr, b, 1.25 * dx))
    np.testing.assert_array_equal(oracle.reinit_original(gr, a, 5), reference.reinit_original(gr, b, 2.25 * dx))
    np.testing.assert_array_equal(oracle.weno_absgrad(gr, a, phi), reference.weno_absgrad(gr, b, phi))
    iters = oracle.reinit_iterations(dx, max(dx * nx, dx * ny))
    np.testing.assert_array_equal(oracle.reinit_first_time(gr, a, iters), reference.reinit_first_time(gr, b))


def test_live_iteration_count(oracle):
    # kMaxATime defaults to 1.5 * kThickness = 2.25 eps (levelset2d.cpp:11) -> 5 iterations
    for eps in (1 / 64, 1 / 128, 0.04 / 128):
        assert oracle.reinit_iterations(eps, 2.25 * eps) == 5
        assert oracle.reinit_iterations(eps, 1.25 * eps) == 3
