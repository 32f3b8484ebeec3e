"""Pins of the CPU restatement of PoissonSolver2D::MKL_2FUniform_2D (oracle/fastpoisson.py,
reference src/poisson2d.cpp:90-228).  MKL itself is absent, so the restatement is pinned by what
the mathematics fixes about the discrete problem it states:

* the five-point difference is exact on cubics (and the mirror closure on quadratics even about
  the Neumann node), and cos(pi k j / M) (Neumann-Neumann) or the sine / quarter-wave variants
  are exact eigenvectors of the 1D difference operator with the mirror closure, so
  u = p(x) e_l(y) is reproduced to rounding for each pairing of side kinds, with its Dirichlet
  values supplied through ps exactly where the reference reads them;
* the Helmholtz term shifts the eigenvalue;
* the degenerate all-Neumann problem returns the compatible projection's solution with zero
  weighted mean;
* every planted mistake below (a wrong closure factor, spacing, boundary read, sign or corner)
  fails one of these checks."""
import types

import numpy as np
import pytest

from oracle import fastpoisson as FP

G = 3


def _nodes(n, d, length):
    """node coordinates of the reference's MKL mesh: ax + i h, ax = -d / 2"""
    h = FP.mesh_step(length, d, n)
    return -0.5 * d + np.arange(n + 2) * h, h


def _eig(kind_lo, kind_hi, M, l, j):
    """the l-th eigenvector of the 1D operator at nodes j, and its theta (units of pi)"""
    if kind_lo == "N" and kind_hi == "N":
        th = l / M
        return np.cos(np.pi * th * j), th
    if kind_lo == "D" and kind_hi == "D":
        th = l / M
        return np.sin(np.pi * th * j), th
    th = (l - 0.5) / M
    return (np.cos if kind_lo == "N" else np.sin)(np.pi * th * j), th


def _manufactured(mod, nx, ny, bc_x, bc_y, l, q, seed, x_poly=True):
    """u = p(x) e_l(y) (x_poly) or e_l(x) p(y); Dirichlet sides only along the polynomial axis.
    Returns (computed, exact) over all nodes (nx + 2) x (ny + 2)."""
    rng = np.random.default_rng(seed)
    lx, ly = 1.0, 1.3
    dx, dy = lx / nx, ly / ny
    xs, hx = _nodes(nx, dx, lx)
    ys, hy = _nodes(ny, dy, ly)
    a = rng.uniform(-1, 1, 4)

    def poly(t, sides):
        """a cubic (Dirichlet at both ends), or a quadratic even about the Neumann end node (the
        mirror closure holds for it exactly); value and second derivative"""
        if sides == "DD":
            return a[0] + a[1] * t + a[2] * t ** 2 + a[3] * t ** 3, 2 * a[2] + 6 * a[3] * t
        t0 = t[-1] if sides == "DN" else t[0]
        return a[0] + a[2] * (t - t0) ** 2, np.full_like(t, 2 * a[2])

    if x_poly:
        p, pdd = poly(xs, bc_x)
        e, th = _eig(bc_y[0], bc_y[1], ny + 1, l, np.arange(ny + 2))
        lam = 4 * np.sin(0.5 * np.pi * th) ** 2 / hy ** 2
        U = np.outer(p, e)
        F = np.outer(-pdd + lam * p + q * p, e)
    else:
        p, pdd = poly(ys, bc_y)
        e, th = _eig(bc_x[0], bc_x[1], nx + 1, l, np.arange(nx + 2))
        lam = 4 * np.sin(0.5 * np.pi * th) ** 2 / hx ** 2
        U = np.outer(e, p)
        F = np.outer(e, -pdd + lam * p + q * p)
    bc = bc_x + bc_y
    c = rng.uniform(-1, 1, 4)
    rhs = rng.standard_normal((nx + 2 * G, ny + 2 * G))
    rhs[G - 1:nx + G + 1, G - 1:ny + G + 1] = F
    ps = rng.standard_normal((nx + 2 * G, ny + 2 * G))
    # Dirichlet values where the reference reads them: bd = 2c - ps(...) (src/poisson2d.cpp:150-180)
    j = np.arange(ny + 2)
    i = np.arange(nx + 2)
    if bc[0] == "D":
        ps[G, j + G] = 2 * c[0] - U[0, :]
    if bc[1] == "D":
        ps[nx + G, j + G] = 2 * c[1] - U[-1, :]
    if bc[2] == "D":
        ps[i + G, G] = 2 * c[2] - U[:, 0]
    if bc[3] == "D":
        ps[i + G, ny + G] = 2 * c[3] - U[:, -1]
    out = mod.mkl_2funiform(ps, rhs, lx, ly, dx, dy, bc, c, g=G, q=q)
    return out[G - 1:nx + G + 1, G - 1:ny + G + 1], U


CASES = [
    # (bc along x, bc along y, polynomial along x, mode, q)
    ("DD", "NN", True, 3, 0.0),
    ("DN", "NN", True, 2, 0.0),
    ("ND", "NN", True, 4, 0.0),
    ("DD", "NN", True, 0, 0.0),
    ("NN", "DD", False, 2, 0.0),
    ("NN", "DN", False, 3, 0.0),
    ("NN", "ND", False, 1, 0.0),
    ("DD", "DD", True, 2, 0.0),
    ("DD", "DN", True, 5, 0.0),
    ("DD", "NN", True, 3, 7.5),
    ("NN", "DD", False, 2, 2.0),
]


def _check_manufactured(mod):
    for k, (bx, by, xpoly, l, q) in enumerate(CASES):
        if bx == "DD" and by in ("DD", "DN"):
            continue  # Dirichlet on both axes: the corner reads overlap; see the eigenmode check
        got, want = _manufactured(mod, 13, 10, bx, by, l, q, seed=40 + k, x_poly=xpoly)
        assert np.allclose(got, want, rtol=0, atol=1e-9 * np.abs(want).max()), (bx, by, l, q)


def _check_dirichlet_eigenmode(mod):
    """Dirichlet on all four sides with zero boundary values (ps = 2c where read): the product of
    two sine modes solves with eigenvalue lam_x + lam_y (+ q)"""
    nx, ny, lx, ly = 12, 9, 1.0, 0.75
    dx, dy = lx / nx, ly / ny
    _, hx = _nodes(nx, dx, lx)
    _, hy = _nodes(ny, dy, ly)
    for kx, ky, q in ((2, 3, 0.0), (5, 1, 3.0)):
        ex = np.sin(np.pi * kx * np.arange(nx + 2) / (nx + 1))
        ey = np.sin(np.pi * ky * np.arange(ny + 2) / (ny + 1))
        lam = 4 * np.sin(0.5 * np.pi * kx / (nx + 1)) ** 2 / hx ** 2 + 4 * np.sin(0.5 * np.pi * ky / (ny + 1)) ** 2 / hy ** 2
        U = np.outer(ex, ey)
        # ps = 1 and c = 1/2 everywhere: every boundary value 2c - ps is 0 wherever it is read
        c = np.full(4, 0.5)
        ps = np.ones((nx + 2 * G, ny + 2 * G))
        rhs = np.zeros_like(ps)
        rhs[G - 1:nx + G + 1, G - 1:ny + G + 1] = (lam + q) * U
        out = mod.mkl_2funiform(ps, rhs, lx, ly, dx, dy, "DDDD", c, g=G, q=q)
        got = out[G - 1:nx + G + 1, G - 1:ny + G + 1]
        assert np.allclose(got, U, rtol=0, atol=1e-10), (kx, ky, q)


def _check_neumann_eigenmode(mod):
    nx, ny, lx, ly = 11, 14, 1.0, 1.2
    dx, dy = lx / nx, ly / ny
    _, hx = _nodes(nx, dx, lx)
    _, hy = _nodes(ny, dy, ly)
    for kx, ky, q in ((3, 2, 0.0), (0, 4, 0.0), (1, 0, 0.0), (2, 5, 1.5), (0, 0, 4.0)):
        U = np.outer(np.cos(np.pi * kx * np.arange(nx + 2) / (nx + 1)), np.cos(np.pi * ky * np.arange(ny + 2) / (ny + 1)))
        lam = 4 * np.sin(0.5 * np.pi * kx / (nx + 1)) ** 2 / hx ** 2 + 4 * np.sin(0.5 * np.pi * ky / (ny + 1)) ** 2 / hy ** 2
        rhs = np.zeros((nx + 2 * G, ny + 2 * G))
        rhs[G - 1:nx + G + 1, G - 1:ny + G + 1] = (lam + q) * U
        ps = np.random.default_rng(kx + 10 * ky).standard_normal(rhs.shape)
        out = mod.mkl_2funiform(ps, rhs, lx, ly, dx, dy, "NNNN", [0, 0, 0, 0], g=G, q=q)
        assert np.allclose(out[G - 1:nx + G + 1, G - 1:ny + G + 1], U, rtol=0, atol=1e-10), (kx, ky, q)
        # cells outside the node mesh keep the caller's values
        assert np.array_equal(out[:G - 1], ps[:G - 1]) and np.array_equal(out[:, ny + G + 1:], ps[:, ny + G + 1:])


def _check_degenerate(mod):
    """q = 0, Neumann everywhere, a right-hand side with nonzero mean: the result has zero weighted
    mean and satisfies the equations with the projected right-hand side"""
    nx, ny = 10, 7
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((nx + 2 * G, ny + 2 * G)) + 0.7
    ps = np.zeros_like(rhs)
    out = mod.mkl_2funiform(ps, rhs, 1.0, 0.7, 1.0 / nx, 0.7 / ny, "NNNN", [0] * 4, g=G)
    w = np.ones(nx + 2)
    w[[0, -1]] = 0.5
    v = np.ones(ny + 2)
    v[[0, -1]] = 0.5
    W = np.outer(w, v)
    U = out[G - 1:nx + G + 1, G - 1:ny + G + 1]
    F = rhs[G - 1:nx + G + 1, G - 1:ny + G + 1]
    assert abs((W * U).sum()) <= 1e-12 * np.abs(U).max()
    Fp = F - (W * F).sum() / W.sum()
    L = FP.apply_operator(out, 1.0, 0.7, 1.0 / nx, 0.7 / ny, "NNNN", g=G)
    assert np.allclose(L, Fp, rtol=0, atol=1e-9 * np.abs(Fp).max())


CHECKS = [_check_manufactured, _check_dirichlet_eigenmode, _check_neumann_eigenmode, _check_degenerate]


@pytest.mark.parametrize("check", CHECKS, ids=lambda f: f.__name__)
def test_fastpoisson_pins(check):
    check(FP)


def test_equations_hold_for_random_input():
    """the returned nodes satisfy the stated five-point equations with the right-hand side at the
    unknown nodes, and the Dirichlet nodes hold 2c - ps read where the reference reads it"""
    rng = np.random.default_rng(9)
    nx, ny = 9, 12
    for bc in ("DDNN", "NNDD", "DNND", "DDDD", "NDDN"):
        rhs = rng.standard_normal((nx + 2 * G, ny + 2 * G))
        ps = rng.standard_normal(rhs.shape)
        c = rng.uniform(-1, 1, 4)
        out = FP.mkl_2funiform(ps, rhs, 1.0, 1.1, 1.0 / nx, 1.1 / ny, bc, c, g=G, q=0.5)
        L = FP.apply_operator(out, 1.0, 1.1, 1.0 / nx, 1.1 / ny, bc, g=G, q=0.5)
        ix = FP.unknown_range(bc[0], bc[1], nx + 1)
        jy = FP.unknown_range(bc[2], bc[3], ny + 1)
        assert np.allclose(L, rhs[np.ix_(ix + G - 1, jy + G - 1)], rtol=0, atol=1e-9)
        if bc[0] == "D":
            assert np.array_equal(out[G - 1, G - 1:ny + G + 1], 2 * c[0] - ps[G, G:ny + G + 2])


# planted mistakes: each is a plausible slip in the restatement
MUTANTS = [
    ("Neumann closure factor 1", "upper[0] = -2.0", "upper[0] = -1.0"),
    ("Neumann closure on the wrong end", "lower[-1] = -2.0", "lower[0] = -2.0"),
    ("spacing over n intervals", "return (bx - ax) / (n + 1)", "return (bx - ax) / n"),
    ("west value read one row off", "2.0 * c[0] - ps[g, j + g]", "2.0 * c[0] - ps[g, j + g - 1]"),
    ("east value read at the last interior cell", "2.0 * c[1] - ps[nx + g, j + g]", "2.0 * c[1] - ps[nx + g - 1, j + g]"),
    ("Dirichlet value without the factor 2", "2.0 * c[2] - ps[i + g, g]", "c[2] - ps[i + g, g]"),
    ("boundary term with the wrong sign", "f[0, :] += bax[jy] / (hx * hx)", "f[0, :] -= bax[jy] / (hx * hx)"),
    ("Helmholtz term with the wrong sign", "A = A + q * sp.identity(A.shape[0])", "A = A - q * sp.identity(A.shape[0])"),
    ("unweighted mean", "w[0] = 0.5", "w[0] = 1.0"),
    ("north values on the wrong row", "out[np.arange(nx + 2) + g - 1, ny + g] = bby", "out[np.arange(nx + 2) + g - 1, ny + g - 1] = bby"),
]


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistakes(name, old, new):
    src = open(FP.__file__).read()
    assert src.count(old) == 1, name
    mod = types.ModuleType("fastpoisson_mutant")
    exec(compile(src.replace(old, new), "fastpoisson_mutant", "exec"), mod.__dict__)
    failed = 0
    for check in CHECKS + [lambda m: _check_read_positions(m)]:
        try:
            check(mod)
        except AssertionError:
            failed += 1
    assert failed > 0, f"no pin catches: {name}"


def _check_read_positions(mod):
    """the Dirichlet nodes hold 2c - ps at the reference's read positions"""
    rng = np.random.default_rng(11)
    nx, ny = 6, 5
    rhs = rng.standard_normal((nx + 2 * G, ny + 2 * G))
    ps = rng.standard_normal(rhs.shape)
    c = rng.uniform(-1, 1, 4)
    out = mod.mkl_2funiform(ps, rhs, 1.0, 1.0, 1.0 / nx, 1.0 / ny, "DDDD", c, g=G)
    j = np.arange(1, ny + 1)
    i = np.arange(1, nx + 1)
    assert np.array_equal(out[G - 1, j + G - 1], 2 * c[0] - ps[G, j + G])
    assert np.array_equal(out[nx + G, j + G - 1], 2 * c[1] - ps[nx + G, j + G])
    assert np.array_equal(out[i + G - 1, G - 1], 2 * c[2] - ps[i + G, G])
    assert np.array_equal(out[i + G - 1, ny + G], 2 * c[3] - ps[i + G, ny + G])
