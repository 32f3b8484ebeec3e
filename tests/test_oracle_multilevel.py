"""Pins of the multilevel restatement (oracle/multilevel.py), the CPU statement of the GPU's
multilevel coarse correction: the V-cycle and the whole fine-level map are symmetric positive
definite linear maps, one level is the textbook symmetric two-grid formula (damped Jacobi, exact
coarse solve, damped Jacobi), the split dots the GPU sums equal the direct ones, the smoother is the
Chebyshev polynomial with its minimax error on [2 / 32, 2], and the flexible CG with it converges
to the direct solution.  Iteration counts on seeded systems are recorded for the GPU comparison
(tests/test_gpu_multilevel.py)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import multilevel as ML
from tests.gfm_common import gfm_case, seeded_droplet_system

G = 3


def _system(oracle, nx, ny, seed):
    gr, fl, phi, kap, H, div = gfm_case(oracle, nx, ny, seed=seed, rho=(1000.0, 1.0), sigma=0.3)
    ax, ay, dg, rhs = oracle.gfm_assemble(gr, fl, phi, kap, H, div)
    A = ML.stencil_matrix(ax, ay, dg, nx, ny)
    b = rhs[G:G + nx, G:G + ny].ravel().copy()
    return A, b


def test_vcycle_is_symmetric_positive_definite(oracle):
    A, _ = _system(oracle, 256, 192, 3)
    M = ML.Multilevel(A, 256, 192)
    assert M.c0 == 8 and len(M.levels) == 1 and M.coarsest.shape[0] == 8 * 6
    n0 = M.levels[0][0].shape[0]
    V = np.stack([M.vcycle(e) for e in np.eye(n0)], axis=1)
    assert np.abs(V - V.T).max() <= 1e-12 * np.abs(V).max()
    assert np.linalg.eigvalsh(0.5 * (V + V.T)).min() > 0


def test_one_level_is_the_two_level_formula(oracle):
    A, b = _system(oracle, 128, 96, 4)
    M = ML.Multilevel(A, 128, 96, levels=0)
    A0 = (M.P0.T @ A @ M.P0).tocsc()
    r = np.random.default_rng(0).standard_normal(b.size)
    w = 0.8
    z = w * r / A.diagonal()
    z = z + M.P0 @ spla.spsolve(A0, M.P0.T @ (r - A @ z))
    want = z + w * (r - A @ z) / A.diagonal()
    np.testing.assert_allclose(M(r), want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_fine_level_map_is_spd(oracle):
    """r -> z of the whole preconditioner, as a matrix on a small system: symmetric (the two damped
    Jacobi steps mirror each other around a symmetric coarse correction) and positive definite"""
    nx, ny = 48, 40
    A, _ = _system(oracle, nx, ny, 7)
    M = ML.Multilevel(A, nx, ny)
    Z = np.stack([M(e) for e in np.eye(nx * ny)], axis=1)
    assert np.abs(Z - Z.T).max() <= 1e-10 * np.abs(Z).max()
    assert np.linalg.eigvalsh(0.5 * (Z + Z.T)).min() > 0


def test_split_dots_equal_direct(oracle):
    """the GPU forms z.r and z.q_1 from separately summed fine and coarse parts before z exists"""
    A, b = _system(oracle, 128, 96, 5)
    M = ML.Multilevel(A, 128, 96)
    rng = np.random.default_rng(1)
    r, q = rng.standard_normal(b.size), rng.standard_normal(b.size)
    z = M(r)
    zr, zq = M.split_dots(r, q)
    assert zr == pytest.approx(z @ r, rel=1e-10)
    assert zq == pytest.approx(z @ q, rel=1e-9, abs=1e-12 * np.linalg.norm(z) * np.linalg.norm(q))


def test_chebyshev_smoother_is_minimax():
    """On D^-1 A = diag(lam) the four steps from y = 0 leave the error (1 - lam q(lam)) e, and the
    Chebyshev polynomial's largest value on [lo, hi] is 1 / T_4((hi + lo) / (hi - lo))."""
    lo, hi = ML.CHEB_LMAX / ML.CHEB_ALPHA, ML.CHEB_LMAX
    lam = np.linspace(lo, hi, 4001)
    c0, a, b = ML.chebyshev()
    y = np.zeros_like(lam)
    d = None
    for k in range(ML.CHEB_STEPS):
        z = 1.0 - lam * y  # D^-1 (g - A y) for g = 1, D = 1
        d = c0 * z if k == 0 else a[k] * d + b[k] * z
        y = y + d
    err = np.abs(1.0 - lam * y)
    s = (hi + lo) / (hi - lo)
    T4 = np.cosh(4 * np.arccosh(s))
    assert err.max() == pytest.approx(1.0 / T4, rel=1e-9)
    assert np.all(np.abs(1.0 - np.linspace(0, lo, 50) * np.interp(np.linspace(0, lo, 50), lam, y)) <= 1.0 + 1e-12)


def test_fcg_converges_to_direct_solution(oracle):
    A, b = _system(oracle, 128, 128, 6)
    M = ML.Multilevel(A, 128, 128)
    x, it, ok = ML.fcg(A, b, M, tol=1e-12)
    xd = spla.spsolve(A.tocsc(), b)
    assert ok and it < 300
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-8


def test_seeded_droplet_count(oracle):
    """The restatement's iteration count on a seeded-droplet system at 256^2 (1e-8), against
    Jacobi-PCG on the same system."""
    gr, (ax, ay, dg, rhs) = seeded_droplet_system(oracle, 256, 901)
    A = ML.stencil_matrix(ax, ay, dg, 256, 256)
    b = rhs[G:G + 256, G:G + 256].ravel().copy()
    _, it, ok = ML.fcg(A, b, ML.Multilevel(A, 256, 256), tol=1e-8)
    _, itj, _, stj = oracle.pcg(gr, ax, ay, dg, rhs, 1e-8, 100000, 1)
    print("multilevel", it, "jacobi", itj)
    assert ok and stj == 0 and it * 10 < itj
