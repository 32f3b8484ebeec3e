"""Host logic of the coarse-level training data (asd2_amd/learned.py): the level operator built from
the fine padded arrays is the Galerkin product P^T A P of the multilevel hierarchy, and its matrix
form applies as the network's training kernels apply it."""
import numpy as np
import pytest

from asd2_amd import learned as LN
from oracle import multilevel as ML

G = 3


def _padded_system(rng, nx, ny):
    ax = np.zeros((nx + 2 * G, ny + 2 * G))
    ay = np.zeros_like(ax)
    ax[G + 1:G + nx, G:G + ny] = -rng.uniform(0.1, 10.0, (nx - 1, ny))
    ay[G:G + nx, G + 1:G + ny] = -rng.uniform(0.1, 10.0, (nx, ny - 1))
    dg = np.zeros_like(ax)
    s = -(ax[G:G + nx, G:G + ny] + ax[G + 1:G + nx + 1, G:G + ny] + ay[G:G + nx, G:G + ny]
          + ay[G:G + nx, G + 1:G + ny + 1])
    dg[G:G + nx, G:G + ny] = s + rng.uniform(0.0, 0.5, (nx, ny)) * (rng.random((nx, ny)) < 0.1)
    dg[G, G] += 1.0
    return ax, ay, dg


def _apply_like_kernel(acx, acy, dc, y):
    """the level operator as the training kernels apply it (W / S face coefficient of each cell)"""
    H, W = y.shape
    v = dc[:H, :W] * y
    v[1:, :] += acx[1:H, :W] * y[:-1, :]
    v[:-1, :] += acx[1:H, :W] * y[1:, :]
    v[:, 1:] += acy[:H, 1:W] * y[:, :-1]
    v[:, :-1] += acy[:H, 1:W] * y[:, 1:]
    return v


@pytest.mark.parametrize("c", [2, 4, 8])
def test_coarse_operator_is_galerkin(c):
    rng = np.random.default_rng(c)
    nx, ny = 4 * c, 3 * c
    ax, ay, dg = _padded_system(rng, nx, ny)
    A = ML.stencil_matrix(ax, ay, dg, nx, ny)
    P = ML.aggregation(nx, ny, c)
    ref = (P.T @ A @ P).toarray()
    op = LN.coarse_operator(ax, ay, dg, c)
    M = LN.coarse_matrix(*op).toarray()
    assert np.allclose(M, ref, rtol=1e-13, atol=1e-12 * np.abs(ref).max())
    y = rng.standard_normal((nx // c, ny // c))
    assert np.allclose(_apply_like_kernel(*op, y).ravel(), ref @ y.ravel(), rtol=1e-12, atol=1e-12)
    # the energy of a sample: g.A^-1 g > 0 for the SPD level operator, = y.A y for g = A y
    g = ref @ y.ravel()
    assert np.isclose(g @ np.linalg.solve(ref, g), y.ravel() @ ref @ y.ravel(), rtol=1e-10)


def test_coarse_operator_padding_row_is_zero():
    """the level arrays are (H + 1) x (W + 1) like the solver's: the extra row and column hold zeros"""
    rng = np.random.default_rng(3)
    ax, ay, dg = _padded_system(rng, 8, 8)
    acx, acy, dc = LN.coarse_operator(ax, ay, dg, 4)
    assert acx.shape == acy.shape == dc.shape == (3, 3)
    assert np.all(acx[2, :] == 0) and np.all(acy[:, 2] == 0) and np.all(dc[2, :] == 0)
