"""FP64 CPU restatement of the multilevel flexible PCG of the GPU library (test infrastructure only).

No reference counterpart: the reference preconditions its CG with the Jacobi diagonal only
(src/poisson2d.cpp:345-396).  This module states the algorithm the GPU solver
(``ASD2_PRECOND_MULTILEVEL``) implements, in numpy/scipy, so that its iteration counts and its
algebra can be checked:

* first coarsening: c0 x c0 blocks, c0 the largest power of two <= 8 dividing both sides;
  A_0 = P^T A P with P the piecewise-constant prolongation (the Galerkin operator);
* further levels: blocks of 4 x 4 (2 x 2 when a side is not divisible by 4 or shorter than 8) until
  both sides are at most 16; the coarsest system is solved with its inverse;
* smoothing on every level above the coarsest, before and after the coarse correction: the
  degree-4 Chebyshev polynomial in D^-1 A for the interval [2 / 32, 2];
* fine level: a damped Jacobi step (omega = 0.8) on either side of the coarse correction,
  z1 = omega D^-1 r + P y with y = V-cycle(P^T (r - omega A D^-1 r)), z = z1 + omega D^-1 (r - A z1)
  (a symmetric multiplicative two-grid step on the fine level);
* Krylov: flexible CG with one direction kept, beta = -(z.q_old) / (p_old.q_old),
  alpha = (z.r) / (p.q), stopping at ||r|| <= tol ||b||.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

G = 3
CHEB_LMAX, CHEB_ALPHA, CHEB_STEPS = 2.0, 32.0, 4
OMEGA = 0.8  # damping of the fine level's Jacobi steps


def stencil_matrix(ax, ay, dg, nx, ny, g=G):
    """The matrix-free 5-point operator (padded arrays, ax / ay = coefficient of the W / S face of
    each cell) as CSR, unknown k = i * ny + j over the interior cells; a cell with no stored
    coupling (the pinned cell of an all-Neumann system) keeps its diagonal row."""
    AX, AY, D = ax[g:g + nx, g:g + ny], ay[g:g + nx, g:g + ny], dg[g:g + nx, g:g + ny]
    idx = np.arange(nx * ny).reshape(nx, ny)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [D.ravel()]
    for w, a, b in ((AX[1:, :], idx[1:, :], idx[:-1, :]), (AY[:, 1:], idx[:, 1:], idx[:, :-1])):
        m = w != 0
        rows += [a[m], b[m]]
        cols += [b[m], a[m]]
        vals += [w[m], w[m]]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nx * ny,) * 2)
    A.sum_duplicates()
    return A


def aggregation(H, W, c):
    """piecewise-constant prolongation from (H / c) x (W / c) blocks to H x W cells (j fastest)"""
    Hc, Wc = H // c, W // c
    blk = ((np.arange(H) // c)[:, None] * Wc + (np.arange(W) // c)[None, :]).ravel()
    return sp.csr_matrix((np.ones(H * W), (np.arange(H * W), blk)), shape=(H * W, Hc * Wc))


def chebyshev(lmax=CHEB_LMAX, alpha=CHEB_ALPHA, steps=CHEB_STEPS):
    """coefficients of the Chebyshev smoother on [lmax / alpha, lmax]: step 0 d = z / theta, step k
    d = a_k d + b_k z (z = D^-1 (g - A y)), each followed by y += d"""
    lo, hi = lmax / alpha, lmax
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    rho = 1.0 / sigma
    a, b = [0.0] * steps, [0.0] * steps
    for k in range(1, steps):
        rn = 1.0 / (2.0 * sigma - rho)
        a[k] = rn * rho
        b[k] = 2.0 * rn / delta
        rho = rn
    return 1.0 / theta, a, b


def first_coarsening(nx, ny):
    c = 1
    while c < 8 and nx % (2 * c) == 0 and ny % (2 * c) == 0:
        c *= 2
    return c


def level_plan(H, W):
    """block factors of the levels below the first coarsening, down to at most 16 x 16"""
    f = []
    while (H > 16 or W > 16) and H % 2 == 0 and W % 2 == 0 and len(f) + 1 < 8:
        k = 4 if (H % 4 == 0 and W % 4 == 0 and H >= 8 and W >= 8) else 2
        f.append(k)
        H //= k
        W //= k
    return f, H, W


class Multilevel:
    """z = M(r) of the GPU's multilevel coarse correction (see the module docstring)"""

    def __init__(self, A, nx, ny, levels=None):
        self.A = A.tocsr()
        self.dinv = 1.0 / self.A.diagonal()
        self.c0 = first_coarsening(nx, ny)
        H, W = nx // self.c0, ny // self.c0
        self.P0 = aggregation(nx, ny, self.c0)
        Al = (self.P0.T @ self.A @ self.P0).tocsr()
        factors, _, _ = level_plan(H, W)
        if levels is not None:
            factors = factors[:levels]
        self.levels = []  # (A_l, diag_l, P_l to the next level)
        for f in factors:
            P = aggregation(H, W, f)
            self.levels.append((Al, Al.diagonal(), P))
            Al = (P.T @ Al @ P).tocsr()
            H //= f
            W //= f
        self.coarsest = Al
        self.Minv = np.linalg.inv(Al.toarray())
        self.c0_, self.ca, self.cb = chebyshev()

    def smooth(self, A, dc, g, y):
        d = None
        for k in range(CHEB_STEPS):
            z = (g - A @ y) / dc
            d = self.c0_ * z if k == 0 else self.ca[k] * d + self.cb[k] * z
            y = y + d
        return y

    def vcycle(self, g, l=0):
        if l == len(self.levels):
            return self.Minv @ g
        A, dc, P = self.levels[l]
        y = self.smooth(A, dc, g, np.zeros_like(g))
        y = y + P @ self.vcycle(P.T @ (g - A @ y), l + 1)
        return self.smooth(A, dc, g, y)

    def __call__(self, r):
        z = OMEGA * self.dinv * r
        z = z + self.P0 @ self.vcycle(self.P0.T @ (r - self.A @ z))
        return z + OMEGA * self.dinv * (r - self.A @ z)

    def split_dots(self, r, q):
        """(z.r, z.q) of z = M(r) from the parts the GPU sums separately: with s = D^-1 r,
        g = P^T (r - omega A s) and y = V-cycle(g),
        z.r = 2 omega s.r - omega^2 (A s).s + y.g,
        z.q = 2 omega s.q - omega^2 (A s).D^-1 q + y.P^T q - omega y.P^T A D^-1 q"""
        w = OMEGA
        s = self.dinv * r
        As = self.A @ s
        g = self.P0.T @ (r - w * As)
        y = self.vcycle(g)
        dq = self.dinv * q
        zr = 2 * w * (s @ r) - w * w * (As @ s) + y @ g
        zq = 2 * w * (s @ q) - w * w * (As @ dq) + y @ (self.P0.T @ q) - w * (y @ (self.P0.T @ (self.A @ dq)))
        return zr, zq


def fcg(A, b, M, tol=1e-8, maxit=10000):
    """flexible CG with one kept direction; returns (x, iterations, converged)"""
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b)
    if nb == 0:
        return x, 0, True
    z = M(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        q = A @ p
        pq = p @ q
        al = rz / pq
        x += al * p
        r -= al * q
        if np.linalg.norm(r) <= tol * nb:
            return x, it, True
        z = M(r)
        p = z - (z @ q) / pq * p
        rz = r @ z
    return x, maxit, False
