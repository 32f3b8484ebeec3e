"""The constant-coefficient fast Poisson solve (asd2_fast_poisson, the GPU replacement of the MKL
Helmholtz call in PoissonSolver2D::MKL_2FUniform_2D, reference src/poisson2d.cpp:90-228) against
the oracle's sparse direct solve of the same discrete problem (oracle/fastpoisson.py), for every
pairing of Neumann and Dirichlet sides, with and without the Helmholtz term, on random right-hand
sides (ghost cells included, as the reference reads them) and random caller ps (from which the
reference reads its Dirichlet values).  Bar: max |GPU - oracle| <= 1e-10 max |oracle| on the
node mesh, and every other cell of ps unchanged.  At 2048^2 the residual of the stated equations
is recomputed on the host."""
import time

import numpy as np
import pytest

import asd2_amd as A
from oracle import fastpoisson as FP

pytestmark = pytest.mark.gpu

G = 3
KINDS = {"N": "neumann", "D": "dirichlet"}


def _run(nx, ny, bc, q, seed, lx=1.0, ly=None):
    ly = ly if ly is not None else lx * ny / nx
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    c = rng.uniform(-1, 1, 4)
    rhs = rng.standard_normal((nx + 2 * G, ny + 2 * G))
    ps = rng.standard_normal(rhs.shape)
    ctx = A.Context(nx, ny, dx, dy)
    ctx.set_bc(A.GRID_P, tuple(KINDS[k] for k in bc), tuple(c))
    d_rhs, d_ps = ctx.array(rhs), ctx.array(ps)
    ctx.fast_poisson(d_rhs, d_ps, lx, ly, dx, dy, q)
    ctx.sync()
    return ctx, rhs, ps, c, d_ps.download(), (lx, ly, dx, dy)


@pytest.mark.parametrize("nx,ny", [(16, 16), (48, 40), (129, 77), (256, 256)])
@pytest.mark.parametrize("bc", ["NNNN", "DDDD", "DNND", "NDDN", "DDNN", "NNDD"])
@pytest.mark.parametrize("q", [0.0, 3.0])
def test_fast_poisson_matches_oracle(nx, ny, bc, q):
    ctx, rhs, ps, c, got, (lx, ly, dx, dy) = _run(nx, ny, bc, q, seed=nx * 7 + ny + len(bc) + int(q))
    want = FP.mkl_2funiform(ps, rhs, lx, ly, dx, dy, bc, c, g=G, q=q)
    mesh = (slice(G - 1, nx + G + 1), slice(G - 1, ny + G + 1))
    err = np.abs(got[mesh] - want[mesh]).max() / np.abs(want[mesh]).max()
    print(f"{nx}x{ny} {bc} q={q}: max rel diff {err:.2e}")
    assert err <= 1e-10
    outside = np.ones(ps.shape, bool)
    outside[mesh] = False
    assert np.array_equal(got[outside], ps[outside])


def test_fast_poisson_rejects_other_sides():
    ctx = A.Context(16, 16, 1 / 16, 1 / 16)
    ctx.set_bc(A.GRID_P, ("neumann", "neumann", "wall", "neumann"))
    with pytest.raises(A.Asd2Error):
        ctx.fast_poisson(ctx.array(), ctx.array(), 1.0, 1.0, 1 / 16, 1 / 16)


class _GpuSolve:
    """the oracle's call signature, solved on the GPU"""

    @staticmethod
    def mkl_2funiform(ps, rhs, lx, ly, dx, dy, bc, c, g=G, q=0.0):
        nx, ny = ps.shape[0] - 2 * g, ps.shape[1] - 2 * g
        ctx = A.Context(nx, ny, dx, dy)
        ctx.set_bc(A.GRID_P, tuple(KINDS[k] for k in bc), tuple(float(v) for v in c))
        d_ps = ctx.array(ps)
        ctx.fast_poisson(ctx.array(rhs), d_ps, lx, ly, dx, dy, q)
        return d_ps.download()


@pytest.mark.parametrize("bc_x,bc_y,xpoly,mode", [("DD", "NN", True, 3), ("NN", "DN", False, 2), ("ND", "NN", True, 5)])
def test_fast_poisson_2048_manufactured(bc_x, bc_y, xpoly, mode):
    """u = p(x) e_l(y) (p a cubic, or a quadratic even about a Neumann node; e_l a discrete
    eigenvector), exact for the five-point problem, reproduced at 2048^2 (the construction of the
    oracle's own pins, tests/test_oracle_fastpoisson.py)"""
    from tests.test_oracle_fastpoisson import _manufactured
    t0 = time.perf_counter()
    got, want = _manufactured(_GpuSolve, 2048, 2048, bc_x, bc_y, mode, 0.0, seed=77, x_poly=xpoly)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"2048^2 {bc_x}{bc_y} mode {mode}: max rel error {err:.2e} ({(time.perf_counter() - t0) * 1e3:.0f} ms "
          f"incl. set-up)")
    assert err <= 1e-10


def test_fast_poisson_2048_neumann_residual():
    """four Neumann sides at 2048^2: zero weighted mean, and the stated equations hold with the
    projected right-hand side to the rounding the operator's condition number (~4e6) allows"""
    n = 2048
    ctx, rhs, ps, c, got, (lx, ly, dx, dy) = _run(n, n, "NNNN", 0.0, seed=3)
    L = FP.apply_operator(got, lx, ly, dx, dy, "NNNN", g=G)
    F = rhs[G - 1:n + G + 1, G - 1:n + G + 1]
    w = np.ones(n + 2)
    w[[0, -1]] = 0.5
    W = np.outer(w, w)
    F = F - (W * F).sum() / W.sum()
    U = got[G - 1:n + G + 1, G - 1:n + G + 1]
    res = np.abs(L - F).max() / np.abs(F).max()
    print(f"2048^2 NNNN: residual {res:.2e}, weighted mean {(W * U).sum() / W.sum():.2e}")
    assert abs((W * U).sum()) / W.sum() <= 1e-12 * np.abs(U).max()
    assert res <= 1e-6
