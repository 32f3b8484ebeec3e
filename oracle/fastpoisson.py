"""FP64 CPU restatement of PoissonSolver2D::MKL_2FUniform_2D (test infrastructure only).

Reference: src/poisson2d.cpp:90-228, called from MACSolver2D::SolvePoisson with rhs = -div on the
interior and ps zeroed (src/mac2d.cpp:1324-1328, 1658-1689).

Third-party boundary: the reference hands the solve to the Intel MKL Poisson Library
(d_init_Helmholtz_2D / d_commit_Helmholtz_2D / d_Helmholtz_2D / free_Helmholtz_2D,
src/poisson2d.cpp:188-212).  MKL is not vendored and no version is pinned (the project file names
the MKL of Intel C++ Compiler XE 14.0).  Its published problem statement is restated here:

* the Helmholtz problem -u_xx - u_yy + q u = f on the rectangle [ax, bx] x [ay, by] with nx x ny
  mesh intervals, nodes x_i = ax + i hx (i = 0..nx, hx = (bx - ax) / nx), y likewise;
* the standard five-point difference  (-u_{i-1,j} + 2u_{ij} - u_{i+1,j}) / hx^2
  + (-u_{i,j-1} + 2u_{ij} - u_{i,j+1}) / hy^2 + q u_ij = f_ij  at every node that is not on a
  Dirichlet side;
* a Dirichlet side: the side's nodes hold the given values (bd_ax, bd_bx, bd_ay, bd_by);
* a Neumann side: the side's nodes are unknowns and the normal derivative is the central
  difference across the side, u_{-1} = u_1 for the zero derivative the reference always passes
  (src/poisson2d.cpp:150-180 set the Neumann values to 0);
* q = 0 with Neumann on every side is degenerate (MKL returns stat = 1, which the reference
  accepts at :206).  MKL's additive constant is not documented: this restatement returns the
  solution of the compatible projection of the system with zero weighted mean, weights
  w_x(i) w_y(j) with w = 1/2 on a Neumann side's nodes and 1 elsewhere (the weights that make the
  operator symmetric).  Parity with MKL's constant is unpinned; the pressure differences that
  UpdateVel uses are not affected.

The reference's data flow around the library call, which this module follows:

* the mesh spans the cell centres from the first ghost to the last: ax = -dx/2,
  bx = lenX + dx/2, nx + 1 intervals (src/poisson2d.cpp:99, 141-144), so node (i, j) is padded
  cell (i + g - 1, j + g - 1), ghosts included (:182-186, 214-218);
* f at every node is rhs at that cell, ghost cells included (:182-186);
* Dirichlet values, read from the ps the caller passes: bd_ax[j] = 2 c_W - ps(g, j + g),
  bd_bx[j] = 2 c_E - ps(nx + g, j + g), bd_ay[i] = 2 c_S - ps(i + g, g),
  bd_by[i] = 2 c_N - ps(i + g, ny + g) (:150-180; note the node-to-cell offset of one and the
  east/north reads of the first ghost, kept as written);
* the solution at every node is written back to ps (:214-218).  At a corner node of two
  Dirichlet sides this restatement writes the x side's value (MKL's choice there is unpinned).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

NEUMANN, DIRICHLET = "N", "D"


def mesh_step(length, d, n):
    """MKL's node spacing for the reference's rectangle: (bx - ax) / intervals"""
    ax = -0.5 * d
    bx = length + 0.5 * d
    return (bx - ax) / (n + 1)


def unknown_range(lo, hi, M):
    """node indices that are unknowns along one axis (nodes 0..M)"""
    return np.arange(0 if lo == NEUMANN else 1, M + 1 if hi == NEUMANN else M)


def boundary_values(ps, g, nx, ny, bc, c):
    """bd_ax, bd_bx (length ny + 2) and bd_ay, bd_by (length nx + 2) as the reference fills them
    (0 on Neumann sides)"""
    j = np.arange(ny + 2)
    i = np.arange(nx + 2)
    bax = 2.0 * c[0] - ps[g, j + g] if bc[0] == DIRICHLET else np.zeros(ny + 2)
    bbx = 2.0 * c[1] - ps[nx + g, j + g] if bc[1] == DIRICHLET else np.zeros(ny + 2)
    bay = 2.0 * c[2] - ps[i + g, g] if bc[2] == DIRICHLET else np.zeros(nx + 2)
    bby = 2.0 * c[3] - ps[i + g, ny + g] if bc[3] == DIRICHLET else np.zeros(nx + 2)
    return bax, bbx, bay, bby


def _axis_operator(lo, hi, M, h):
    idx = unknown_range(lo, hi, M)
    m = len(idx)
    main = np.full(m, 2.0)
    upper = np.full(m - 1, -1.0)
    lower = np.full(m - 1, -1.0)
    if lo == NEUMANN and m > 1:
        upper[0] = -2.0
    if hi == NEUMANN and m > 1:
        lower[-1] = -2.0
    T = sp.diags([lower, main, upper], [-1, 0, 1], shape=(m, m), format="csr") / (h * h)
    w = np.ones(m)
    if lo == NEUMANN:
        w[0] = 0.5
    if hi == NEUMANN:
        w[-1] = 0.5
    return idx, T, w


def mkl_2funiform(ps, rhs, len_x, len_y, dx, dy, bc, c, g=3, q=0.0):
    """ps after PoissonSolver2D::MKL_2FUniform_2D(ps, rhs, lenX, lenY, dx, dy, PBC); bc = 4
    characters 'N' / 'D' for the W, E, S, N sides, c their Dirichlet constants; ps, rhs padded
    (nx + 2g) x (ny + 2g) arrays.  Returns a new array."""
    nx, ny = ps.shape[0] - 2 * g, ps.shape[1] - 2 * g
    Mx, My = nx + 1, ny + 1
    hx, hy = mesh_step(len_x, dx, nx), mesh_step(len_y, dy, ny)
    bax, bbx, bay, bby = boundary_values(ps, g, nx, ny, bc, c)
    ix, Tx, wx = _axis_operator(bc[0], bc[1], Mx, hx)
    jy, Ty, wy = _axis_operator(bc[2], bc[3], My, hy)
    A = (sp.kron(Tx, sp.identity(len(jy))) + sp.kron(sp.identity(len(ix)), Ty)).tocsr()
    if q != 0.0:
        A = A + q * sp.identity(A.shape[0])
    f = rhs[np.ix_(ix + g - 1, jy + g - 1)].copy()
    # known boundary nodes next to the unknowns
    if bc[0] == DIRICHLET:
        f[0, :] += bax[jy] / (hx * hx)
    if bc[1] == DIRICHLET:
        f[-1, :] += bbx[jy] / (hx * hx)
    if bc[2] == DIRICHLET:
        f[:, 0] += bay[ix] / (hy * hy)
    if bc[3] == DIRICHLET:
        f[:, -1] += bby[ix] / (hy * hy)
    b = f.ravel()
    degenerate = q == 0.0 and all(s == NEUMANN for s in bc)
    if degenerate:
        W = np.outer(wx, wy).ravel()
        b = b - (W @ b) / W.sum()
        keep = np.arange(1, A.shape[0])
        u = np.zeros(A.shape[0])
        u[1:] = spla.spsolve(A[keep][:, keep].tocsc(), b[1:])
        u -= (W @ u) / W.sum()
    else:
        u = spla.spsolve(A.tocsc(), b)
    out = ps.copy()
    out[np.ix_(ix + g - 1, jy + g - 1)] = u.reshape(len(ix), len(jy))
    # Dirichlet nodes (x sides last, so they hold the corners)
    if bc[2] == DIRICHLET:
        out[np.arange(nx + 2) + g - 1, g - 1] = bay
    if bc[3] == DIRICHLET:
        out[np.arange(nx + 2) + g - 1, ny + g] = bby
    if bc[0] == DIRICHLET:
        out[g - 1, np.arange(ny + 2) + g - 1] = bax
    if bc[1] == DIRICHLET:
        out[nx + g, np.arange(ny + 2) + g - 1] = bbx
    return out


def apply_operator(u, len_x, len_y, dx, dy, bc, g=3, q=0.0):
    """-Lap_h u + q u at the unknown nodes, u the padded output of mkl_2funiform (its boundary
    nodes hold the Dirichlet values; Neumann sides mirror across the boundary node)"""
    nx, ny = u.shape[0] - 2 * g, u.shape[1] - 2 * g
    hx, hy = mesh_step(len_x, dx, nx), mesh_step(len_y, dy, ny)
    U = u[g - 1:nx + g + 1, g - 1:ny + g + 1]  # nodes 0..M
    P = np.pad(U, 1)
    P[0, :] = P[2, :]
    P[-1, :] = P[-3, :]
    P[:, 0] = P[:, 2]
    P[:, -1] = P[:, -3]
    L = ((2 * P[1:-1, 1:-1] - P[:-2, 1:-1] - P[2:, 1:-1]) / hx ** 2
         + (2 * P[1:-1, 1:-1] - P[1:-1, :-2] - P[1:-1, 2:]) / hy ** 2 + q * P[1:-1, 1:-1])
    ix = unknown_range(bc[0], bc[1], nx + 1)
    jy = unknown_range(bc[2], bc[3], ny + 1)
    return L[np.ix_(ix, jy)]
