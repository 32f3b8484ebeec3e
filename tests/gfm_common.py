"""Shared GFM test scenario and CSR helpers (host side, numpy/scipy)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import make_fluid, make_grid
from tests import scenarios as S


def gfm_case(oracle, nx, ny, seed=0, rho=(1000.0, 1.0), sigma=1.0, dt=1e-4, bc_p=("neumann",) * 4,
             c_p=(0.0,) * 4, kappa_const=None):
    """Droplets + curvature + Heaviside + divergence of a smooth velocity (the inputs of
    MACSolver2D::SolvePoisson, reference src/mac2d.cpp:1249)."""
    lx = 1.0
    dx = lx / nx
    dy = dx
    gr = make_grid(nx, ny, dx, dy, bc_p=bc_p, c_p=c_p)
    phi, _ = S.random_droplets(seed, nx, ny, dx, dy, n_range=(2, 6), r_range=(0.08, 0.2),
                               lx=nx * dx, ly=ny * dy)
    oracle.bc_p(gr, phi)
    if kappa_const is None:
        kap = oracle.update_kappa(gr, phi)
        oracle.bc_p(gr, kap)
    else:
        kap = np.full(phi.shape, float(kappa_const))
    H = oracle.heaviside(gr, phi)
    u, v = S.smooth_velocity(seed, nx, ny, dx, dy)
    div = oracle.divergence(gr, u, v)
    fl = make_fluid(rho[0], rho[1], 1.0, 0.01, sigma, dt)
    return gr, fl, phi, kap, H, div


def interior(gr, a):
    g = gr.g
    return a[g:g + gr.nx, g:g + gr.ny]


def stencil_to_csr(gr, ax, ay, diag, pin=True):
    """CSR of the matrix-free operator in the reference's unknown order (x fastest,
    L = (i - g) + (j - g) nx, the pinned cell removed; mac2d.cpp:1479-1514)."""
    nx, ny, g = gr.nx, gr.ny, gr.g
    ids = -np.ones((nx, ny), dtype=np.int64)
    k = 0
    for j in range(ny):
        for i in range(nx):
            if pin and i == 0 and j == 0:
                continue
            ids[i, j] = k
            k += 1
    rows, cols, vals = [], [], []
    AX, AY, D = interior(gr, ax), interior(gr, ay), interior(gr, diag)
    AXe = ax[g + 1:g + 1 + nx, g:g + ny]
    AYn = ay[g:g + nx, g + 1:g + 1 + ny]
    for j in range(ny):
        for i in range(nx):
            r = ids[i, j]
            if r < 0:
                continue
            for (ii, jj, val) in ((i, j - 1, AY[i, j]), (i - 1, j, AX[i, j]), (i, j, D[i, j]),
                                  (i + 1, j, AXe[i, j]), (i, j + 1, AYn[i, j])):
                if 0 <= ii < nx and 0 <= jj < ny and ids[ii, jj] >= 0 and (val != 0.0 or (ii, jj) == (i, j)):
                    rows.append(r)
                    cols.append(ids[ii, jj])
                    vals.append(val)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(k, k))
    A.sort_indices()
    return A, ids


def rhs_vector(gr, rhs, ids):
    b = np.zeros(ids.max() + 1)
    R = interior(gr, rhs)
    m = ids >= 0
    b[ids[m]] = R[m]
    return b


def scatter(gr, x, ids):
    out = np.zeros((gr.nx + 2 * gr.g, gr.ny + 2 * gr.g))
    I = interior(gr, out)
    m = ids >= 0
    I[m] = x[ids[m]]
    return out


def random_spd_csr(n, seed):
    """Irregular sparse SPD matrix: a random graph Laplacian plus a positive diagonal shift."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(n):
        for c in rng.choice(n, size=3, replace=False):
            if c != r:
                rows += [r, c]
                cols += [c, r]
    W = sp.csr_matrix((rng.uniform(0.5, 2.0, len(rows)), (rows, cols)), shape=(n, n))
    W = (W + W.T) * 0.5
    A = sp.diags(np.asarray(W.sum(axis=1)).ravel() + rng.uniform(0.1, 1.0, n)) - W
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A



def seeded_droplet_system(oracle, n, seed):
    """The system asd2_amd.learned.build_system assembles on the GPU for this seed, rebuilt with the
    oracle from the same random draws: droplets (1-20 circles, radius in [0.02, 0.1]), smooth
    velocity with no-slip walls, density ratio log-uniform in [10, 1000] (rho_h = 1000), no surface
    tension, dt = 0.3 dx."""
    import asd2_amd.learned as LN
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    gr = make_grid(n, n, dx, dx)
    phi = LN.droplets(rng, n, dx)
    oracle.bc_p(gr, phi)
    kap = oracle.update_kappa(gr, phi)
    oracle.bc_p(gr, kap)
    H = oracle.heaviside(gr, phi)
    u, v = LN.smooth_velocity(rng, n, dx)
    oracle.bc_u(gr, u)
    oracle.bc_v(gr, v)
    div = oracle.divergence(gr, u, v)
    rl = 1000.0 / float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
    fl = make_fluid(1000.0, rl, 1.0, 0.01, float(rng.uniform(0.0, 0.0)), 0.3 * dx)
    return gr, oracle.gfm_assemble(gr, fl, phi, kap, H, div)
