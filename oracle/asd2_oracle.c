/*
 * asd2_oracle.c -- FP64 CPU restatement of the ASD2 2D two-phase hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function follows the cited reference loop nest and keeps its floating-point
 * operation order (the build compiles this file with -ffp-contract=off), so that the
 * level-set functions reproduce the reference bit for bit.  How each function is
 * pinned:
 *   - BC, UnitHJWENO5, HJWENO5_LS_2D, Solve_LevelSet_2D, ENO/Subcell |grad phi|,
 *     smoothed sign, Heaviside derivative, Sussman constraint, Reinit_MinRK2_2D,
 *     Reinit_Sussman_2D: compared bit-for-bit with the reference sources
 *     (src/levelset2d.cpp, src/bc2d.cpp) compiled into oracle/_ref, plus closed forms
 *     (WENO5 exact on cubics, planar SDF fixed point of the subcell reinit).
 *   - UpdateKappa: closed form on a circle SDF (kappa -> 1/r), sign and clip checks.
 *   - GFM assembly: closed forms (rho_H = rho_L, sigma = 0 gives the 5-point
 *     Laplacian / rho; a constant curvature gives p* = -dt sigma kappa H exactly),
 *     symmetry and positive definiteness of the intended operator.
 *   - CG / BiCGStab: scipy sparse direct solve of the same operator, finite
 *     termination on tiny grids.
 *   - Gauss-Seidel (GS_2FUniform_2D): one sweep of a 2 x 2 system against hand
 *     arithmetic; fixed point = the direct solve.
 *   - UpdateVel: constant-curvature droplet leaves u* unchanged (balanced jump).
 *   - Momentum predictor (convection, viscosity, gravity, RK): closed forms
 *     (WENO5 exact on cubic fluxes, quadratic velocities with constant mu, planar
 *     interface face densities, RK stage weights on a linear operator) with planted
 *     mistakes (tests/test_oracle_pins.py); src/mac2d.cpp needs TECIO headers and
 *     cannot be compiled here.
 *   - Axisymmetric pieces (src/mac2dAxisym.cpp, not compilable here either): closed
 *     forms with planted mistakes (tests/test_oracle_axisym.py,
 *     tests/test_oracle_axisym_momentum.py).
 */
#include "asd2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(gr, i, j) ((int64_t)(j) + (int64_t)((gr)->ny + 2 * (gr)->g) * (int64_t)(i))

static double sq(double x) { return x * x; }
/* std::max / std::min semantics (include <algorithm>) */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }
static int isign(double v) { return (v > 0) - (v < 0); }

int64_t or_size(const or_grid* gr) {
    return (int64_t)(gr->nx + 2 * gr->g) * (int64_t)(gr->ny + 2 * gr->g);
}

static double* zalloc(const or_grid* gr) {
    return (double*)calloc((size_t)or_size(gr), sizeof(double));
}

/* ------------------------------------------------------------------ BC (P grid) */
/* src/bc2d.cpp:491-573 dispatch, :786-832 periodic/Neumann, :868-894 Dirichlet,
 * :896-934 LS free extrapolation.  Corner ghosts are never written. */
static int p_family(int k) {
    if (k == OR_BC_PERIODIC) return 0;
    if (k == OR_BC_NEUMANN || k == OR_BC_AXISYM || k == OR_BC_WALL) return 1;
    if (k == OR_BC_DIRICHLET || k == OR_BC_INLET || k == OR_BC_OUTLET || k == OR_BC_PRESSURE) return 2;
    if (k == OR_BC_LSFREE) return 3;
    return -1;
}

void or_bc_p(const or_grid* gr, double* a) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    int f;
    /* W */
    f = p_family(gr->bc_p[0]);
    if (f == 0) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, nx + i, j)];
    } else if (f == 1) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, 2 * g - i - 1, j)];
    } else if (f == 2) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, 2 * g - i - 1, j)] + 2.0 * gr->c_p[0];
    } else if (f == 3) {
        for (int i = g - 1; i >= 0; i--) for (int j = g; j < ny + g; j++)
            a[IDX(gr, i, j)] = a[IDX(gr, i + 1, j)] + dx * (a[IDX(gr, i + 1, j)] - a[IDX(gr, i + 2, j)]) / dx;
    }
    /* E */
    f = p_family(gr->bc_p[1]);
    if (f == 0) {
        /* bc2d.cpp:792-796 writes idx(kNx + i, j) (inside the domain); reproduced as written */
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++) a[IDX(gr, nx + i, j)] = a[IDX(gr, i + g, j)];
    } else if (f == 1) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, nx + 2 * g - i - 1, j)] = a[IDX(gr, nx + i, j)];
    } else if (f == 2) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, nx + 2 * g - i - 1, j)] = -a[IDX(gr, nx + i, j)] + 2.0 * gr->c_p[1];
    } else if (f == 3) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, nx + g + i, j)] = a[IDX(gr, nx + g + i - 1, j)]
                + dx * (a[IDX(gr, nx + g + i - 1, j)] - a[IDX(gr, nx + g + i - 2, j)]) / dx;
    }
    /* S */
    f = p_family(gr->bc_p[2]);
    if (f == 0) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, ny + j)];
    } else if (f == 1) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, 2 * g - j - 1)];
    } else if (f == 2) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, i, 2 * g - j - 1)] + 2.0 * gr->c_p[2];
    } else if (f == 3) {
        for (int i = g; i < nx + g; i++) for (int j = g - 1; j >= 0; j--)
            a[IDX(gr, i, j)] = a[IDX(gr, i, j + 1)] + dy * (a[IDX(gr, i, j + 1)] - a[IDX(gr, i, j + 2)]) / dy;
    }
    /* N */
    f = p_family(gr->bc_p[3]);
    if (f == 0) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, ny + j + g)] = a[IDX(gr, i, j + g)];
    } else if (f == 1) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, ny + 2 * g - j - 1)] = a[IDX(gr, i, ny + j)];
    } else if (f == 2) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, ny + 2 * g - j - 1)] = -a[IDX(gr, i, ny + j)] + 2.0 * gr->c_p[3];
    } else if (f == 3) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, ny + g + j)] = a[IDX(gr, i, ny + g + j - 1)]
                + dy * (a[IDX(gr, i, ny + g + j - 1)] - a[IDX(gr, i, ny + g + j - 2)]) / dy;
    }
}

/* ------------------------------------------------------------------ BC (U grid) */
/* src/bc2d.cpp:123-130 (order W,E,S,N), dispatch :355-421, kernels :575-679 */
void or_bc_u(const or_grid* gr, double* a) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int* k = gr->bc_u;
    const double* c = gr->c_u;
    /* W: periodic / neumann,outlet / dirichlet,axisym,wall,inlet */
    if (k[0] == OR_BC_PERIODIC) {
        for (int i = 0; i <= g; i++) for (int j = g; j < ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, nx - 1 + i, j)];
    } else if (k[0] == OR_BC_NEUMANN || k[0] == OR_BC_OUTLET) {
        for (int j = 0; j < ny + 2 * g; j++) a[IDX(gr, g, j)] = a[IDX(gr, g + 1, j)];
        for (int i = 0; i < g; i++) for (int j = g; j <= ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, 2 * g - i, j)];
    } else if (k[0] == OR_BC_AXISYM || k[0] == OR_BC_WALL || k[0] == OR_BC_DIRICHLET || k[0] == OR_BC_INLET) {
        for (int j = 0; j < ny + 2 * g; j++) a[IDX(gr, g, j)] = c[0];
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, 2 * g - i, j)] + 2.0 * c[0];
    }
    /* E */
    if (k[1] == OR_BC_PERIODIC) {
        for (int i = 0; i < g; i++) for (int j = g; j < ny + g; j++) a[IDX(gr, nx + i + g, j)] = a[IDX(gr, i + g, j)];
    } else if (k[1] == OR_BC_NEUMANN || k[1] == OR_BC_OUTLET) {
        for (int j = 0; j < ny + 2 * g; j++) a[IDX(gr, nx + g, j)] = a[IDX(gr, nx + g - 1, j)];
        for (int i = 1; i < g; i++) for (int j = g; j <= ny + g; j++) a[IDX(gr, nx + 2 * g - i, j)] = a[IDX(gr, nx + i, j)];
    } else if (k[1] == OR_BC_AXISYM || k[1] == OR_BC_WALL || k[1] == OR_BC_DIRICHLET || k[1] == OR_BC_INLET) {
        for (int j = 0; j < ny + 2 * g; j++) a[IDX(gr, nx + g, j)] = c[1];
        for (int i = 1; i < g; i++) for (int j = g; j < ny + g; j++)
            a[IDX(gr, nx + 2 * g - i, j)] = -a[IDX(gr, nx + i, j)] + 2.0 * c[1];
    }
    /* S: periodic / neumann,axisym,outlet / wall,dirichlet,inlet */
    if (k[2] == OR_BC_PERIODIC) {
        for (int i = g + 1; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, ny + j)];
    } else if (k[2] == OR_BC_NEUMANN || k[2] == OR_BC_AXISYM || k[2] == OR_BC_OUTLET) {
        for (int i = g + 1; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, 2 * g - j - 1)];
    } else if (k[2] == OR_BC_WALL || k[2] == OR_BC_DIRICHLET || k[2] == OR_BC_INLET) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, i, 2 * g - j - 1)] + 2.0 * c[2];
    }
    /* N */
    if (k[3] == OR_BC_PERIODIC) {
        for (int i = g + 1; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, ny + j + g)] = a[IDX(gr, i, j + g)];
    } else if (k[3] == OR_BC_NEUMANN || k[3] == OR_BC_AXISYM || k[3] == OR_BC_OUTLET) {
        for (int i = g + 1; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, ny + 2 * g - j - 1)] = a[IDX(gr, i, ny + j)];
    } else if (k[3] == OR_BC_WALL || k[3] == OR_BC_DIRICHLET || k[3] == OR_BC_INLET) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, ny + 2 * g - j - 1)] = -a[IDX(gr, i, ny + j)] + 2.0 * c[3];
    }
}

/* ------------------------------------------------------------------ BC (V grid) */
/* src/bc2d.cpp:239-246 (order W,E,S,N), dispatch :423-489, kernels :681-784 */
void or_bc_v(const or_grid* gr, double* a) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int* k = gr->bc_v;
    const double* c = gr->c_v;
    /* W: periodic / neumann,axisym,outlet / wall,dirichlet,inlet */
    if (k[0] == OR_BC_PERIODIC) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, nx + i, j)];
    } else if (k[0] == OR_BC_NEUMANN || k[0] == OR_BC_AXISYM || k[0] == OR_BC_OUTLET) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++) a[IDX(gr, i, j)] = a[IDX(gr, 2 * g - i - 1, j)];
    } else if (k[0] == OR_BC_WALL || k[0] == OR_BC_DIRICHLET || k[0] == OR_BC_INLET) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, 2 * g - i - 1, j)] + 2.0 * c[0];
    }
    /* E */
    if (k[1] == OR_BC_PERIODIC) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++) a[IDX(gr, nx + i + g, j)] = a[IDX(gr, i + g, j)];
    } else if (k[1] == OR_BC_NEUMANN || k[1] == OR_BC_AXISYM || k[1] == OR_BC_OUTLET) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++)
            a[IDX(gr, nx + 2 * g - i - 1, j)] = a[IDX(gr, nx + i, j)];
    } else if (k[1] == OR_BC_WALL || k[1] == OR_BC_DIRICHLET || k[1] == OR_BC_INLET) {
        for (int i = 0; i < g; i++) for (int j = g + 1; j < ny + g; j++)
            a[IDX(gr, nx + 2 * g - i - 1, j)] = -a[IDX(gr, nx + i, j)] + 2.0 * c[1];
    }
    /* S: periodic / neumann,outlet / dirichlet,axisym,wall,inlet */
    if (k[2] == OR_BC_PERIODIC) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j <= g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, ny - 1 + j)];
    } else if (k[2] == OR_BC_NEUMANN || k[2] == OR_BC_OUTLET) {
        for (int i = 0; i < nx + 2 * g; i++) a[IDX(gr, i, g)] = a[IDX(gr, i, g + 1)];
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, j)] = a[IDX(gr, i, 2 * g - j)];
    } else if (k[2] == OR_BC_AXISYM || k[2] == OR_BC_WALL || k[2] == OR_BC_DIRICHLET || k[2] == OR_BC_INLET) {
        for (int i = 0; i < nx + 2 * g; i++) a[IDX(gr, i, g)] = c[2];
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++)
            a[IDX(gr, i, j)] = -a[IDX(gr, i, 2 * g - j)] + 2.0 * c[2];
    }
    /* N */
    if (k[3] == OR_BC_PERIODIC) {
        for (int i = g; i < nx + g; i++) for (int j = 0; j < g; j++) a[IDX(gr, i, ny + j + g)] = a[IDX(gr, i, j + g)];
    } else if (k[3] == OR_BC_NEUMANN || k[3] == OR_BC_OUTLET) {
        for (int i = 0; i < nx + 2 * g; i++) a[IDX(gr, i, ny + g)] = a[IDX(gr, i, ny + g - 1)];
        for (int i = g; i < nx + g; i++) for (int j = 1; j < g; j++) a[IDX(gr, i, ny + 2 * g - j)] = a[IDX(gr, i, ny + j)];
    } else if (k[3] == OR_BC_AXISYM || k[3] == OR_BC_WALL || k[3] == OR_BC_DIRICHLET || k[3] == OR_BC_INLET) {
        for (int i = 0; i < nx + 2 * g; i++) a[IDX(gr, i, ny + g)] = c[3];
        for (int i = g; i < nx + g; i++) for (int j = 1; j < g; j++)
            a[IDX(gr, i, ny + 2 * g - j)] = -a[IDX(gr, i, ny + j)] + 2.0 * c[3];
    }
}

/* ------------------------------------------------------------------ WENO5 */
/* src/levelset2d.cpp:159-262.  FP, FM are written on [g, n+g) only. */
void or_unit_hjweno5(const double* F, double* FP, double* FM, double d, int n, int g) {
    const double kWeightEps = 1e-6;
    const int len = n + 2 * g;
    double* D = (double*)calloc((size_t)len, sizeof(double));
    for (int i = 0; i < len - 1; i++) D[i] = F[i + 1] - F[i];
    for (int i = g; i < n + g; i++) {
        double IS0 = 13.0 / 12.0 * sq((D[i - 3] - 2.0 * D[i - 2] + D[i - 1]) / d)
                   + 1.0 / 4.0 * sq((D[i - 3] - 4.0 * D[i - 2] + 3.0 * D[i - 1]) / d);
        double IS1 = 13.0 / 12.0 * sq((D[i - 2] - 2.0 * D[i - 1] + D[i]) / d)
                   + 1.0 / 4.0 * sq((D[i - 2] - D[i]) / d);
        double IS2 = 13.0 / 12.0 * sq((D[i - 1] - 2.0 * D[i] + D[i + 1]) / d)
                   + 1.0 / 4.0 * sq((3.0 * D[i - 1] - 4.0 * D[i] + D[i + 1]) / d);
        double a0 = 0.1 / sq(kWeightEps + IS0);
        double a1 = 0.6 / sq(kWeightEps + IS1);
        double a2 = 0.3 / sq(kWeightEps + IS2);
        double w0 = a0 / (a0 + a1 + a2);
        double w1 = a1 / (a0 + a1 + a2);
        double w2 = a2 / (a0 + a1 + a2);
        FM[i] = w0 * (1.0 / 3.0 * D[i - 3] - 7.0 / 6.0 * D[i - 2] + 11.0 / 6.0 * D[i - 1]) / d
              + w1 * (-1.0 / 6.0 * D[i - 2] + 5.0 / 6.0 * D[i - 1] + 1.0 / 3.0 * D[i]) / d
              + w2 * (1.0 / 3.0 * D[i - 1] + 5.0 / 6.0 * D[i] - 1.0 / 6.0 * D[i + 1]) / d;
    }
    for (int i = g; i < n + g; i++) {
        double IS0 = 13.0 / 12.0 * sq((D[i + 2] - 2.0 * D[i + 1] + D[i]) / d)
                   + 1.0 / 4.0 * sq((D[i + 2] - 4.0 * D[i + 1] + 3.0 * D[i]) / d);
        double IS1 = 13.0 / 12.0 * sq((D[i + 1] - 2.0 * D[i] + D[i - 1]) / d)
                   + 1.0 / 4.0 * sq((D[i + 1] - D[i - 1]) / d);
        double IS2 = 13.0 / 12.0 * sq((D[i] - 2.0 * D[i - 1] + D[i - 2]) / d)
                   + 1.0 / 4.0 * sq((3.0 * D[i] - 4.0 * D[i - 1] + D[i - 2]) / d);
        double a0 = 0.1 / sq(kWeightEps + IS0);
        double a1 = 0.6 / sq(kWeightEps + IS1);
        double a2 = 0.3 / sq(kWeightEps + IS2);
        double w0 = a0 / (a0 + a1 + a2);
        double w1 = a1 / (a0 + a1 + a2);
        double w2 = a2 / (a0 + a1 + a2);
        FP[i] = w0 * (1.0 / 3.0 * D[i + 2] - 7.0 / 6.0 * D[i + 1] + 11.0 / 6.0 * D[i]) / d
              + w1 * (-1.0 / 6.0 * D[i + 1] + 5.0 / 6.0 * D[i] + 1.0 / 3.0 * D[i - 1]) / d
              + w2 * (1.0 / 3.0 * D[i] + 5.0 / 6.0 * D[i - 1] - 1.0 / 6.0 * D[i - 2]) / d;
    }
    free(D);
}

/* src/levelset2d.cpp:88-157.  Applies the P BC to ls (mutates its ghosts). */
void or_hjweno5_ls(const or_grid* gr, double* ls, const double* uc, const double* vc, double* L) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int64_t n = or_size(gr);
    double* LXP = zalloc(gr); double* LXM = zalloc(gr);
    double* LYP = zalloc(gr); double* LYM = zalloc(gr);
    const int lx = nx + 2 * g, ly = ny + 2 * g;
    double* F = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FP = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FM = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    or_bc_p(gr, ls);
    for (int j = g; j < ny + g; j++) {
        for (int i = 0; i < lx; i++) F[i] = ls[IDX(gr, i, j)] * uc[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)lx); memset(FM, 0, sizeof(double) * (size_t)lx);
        or_unit_hjweno5(F, FP, FM, gr->dx, nx, g);
        for (int i = 0; i < lx; i++) { LXP[IDX(gr, i, j)] = FP[i]; LXM[IDX(gr, i, j)] = FM[i]; }
    }
    for (int i = g; i < nx + g; i++) {
        for (int j = 0; j < ly; j++) F[j] = ls[IDX(gr, i, j)] * vc[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)ly); memset(FM, 0, sizeof(double) * (size_t)ly);
        or_unit_hjweno5(F, FP, FM, gr->dy, ny, g);
        for (int j = 0; j < ly; j++) { LYP[IDX(gr, i, j)] = FP[j]; LYM[IDX(gr, i, j)] = FM[j]; }
    }
    for (int64_t k = 0; k < n; k++) L[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            const double alphaX = uc[c], alphaY = vc[c];
            L[c] = 0.5 * (LXP[c] + LXM[c]) + 0.5 * (LYP[c] + LYM[c])
                 - alphaX * (0.5 * (LXP[c] - LXM[c])) - alphaY * (0.5 * (LYP[c] - LYM[c]));
        }
    free(LXP); free(LXM); free(LYP); free(LYM); free(F); free(FP); free(FM);
}

/* src/levelset2d.cpp:43-86 (SSP-RK3, cell-centred velocity frozen over the step) */
void or_solve_levelset(const or_grid* gr, double* ls, const double* u, const double* v, double dt) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    double* tU = zalloc(gr); double* tV = zalloc(gr);
    double* LS1 = zalloc(gr); double* LS2 = zalloc(gr); double* DL = zalloc(gr);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            tU[IDX(gr, i, j)] = (u[IDX(gr, i, j)] + u[IDX(gr, i + 1, j)]) * 0.5;
            tV[IDX(gr, i, j)] = (v[IDX(gr, i, j)] + v[IDX(gr, i, j + 1)]) * 0.5;
        }
    or_bc_p(gr, tU); or_bc_p(gr, tV);
    or_hjweno5_ls(gr, ls, tU, tV, DL);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) LS1[IDX(gr, i, j)] = ls[IDX(gr, i, j)] - dt * DL[IDX(gr, i, j)];
    or_bc_p(gr, LS1);
    or_hjweno5_ls(gr, LS1, tU, tV, DL);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++)
            LS2[IDX(gr, i, j)] = 0.75 * ls[IDX(gr, i, j)] + 0.25 * LS1[IDX(gr, i, j)] - 0.25 * dt * DL[IDX(gr, i, j)];
    or_bc_p(gr, LS2);
    or_hjweno5_ls(gr, LS2, tU, tV, DL);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++)
            ls[IDX(gr, i, j)] = 1.0 / 3.0 * ls[IDX(gr, i, j)] + 2.0 / 3.0 * LS2[IDX(gr, i, j)]
                              - 2.0 / 3.0 * dt * DL[IDX(gr, i, j)];
    or_bc_p(gr, ls);
    free(tU); free(tV); free(LS1); free(LS2); free(DL);
}

/* ------------------------------------------------------------------ |grad phi| */
/* src/levelset2d.cpp:1036-1046 */
static double minmod(double v1, double v2) {
    if ((fabs(v1) <= fabs(v2)) && (v1 * v2 > 0)) return v1;
    else if ((fabs(v1) > fabs(v2)) && (v1 * v2 > 0)) return v2;
    else if (v1 * v2 <= 0) return 0.0;
    return NAN;
}

static double godunov(double ls0, double dPX, double dMX, double dPY, double dMY) {
    if (ls0 >= 0)
        return sqrt(smax(sq(smin(dPX, 0.0)), sq(smax(dMX, 0.0)))
                  + smax(sq(smin(dPY, 0.0)), sq(smax(dMY, 0.0))));
    return sqrt(smax(sq(smax(dPX, 0.0)), sq(smin(dMX, 0.0)))
              + smax(sq(smax(dPY, 0.0)), sq(smin(dMY, 0.0))));
}

/* src/levelset2d.cpp:559-613 */
void or_eno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
#define L_(a, b) ls[IDX(gr, a, b)]
            double dPX = (L_(i + 1, j) - L_(i, j)) / dx - dx * 0.5 * minmod(
                (L_(i - 1, j) - 2.0 * L_(i, j) + L_(i + 1, j)) / (dx * dx),
                (L_(i, j) - 2.0 * L_(i + 1, j) + L_(i + 2, j)) / (dx * dx));
            double dMX = (L_(i, j) - L_(i - 1, j)) / dx + dx * 0.5 * minmod(
                (L_(i - 1, j) - 2.0 * L_(i, j) + L_(i + 1, j)) / (dx * dx),
                (L_(i - 2, j) - 2.0 * L_(i - 1, j) + L_(i, j)) / (dx * dx));
            double dPY = (L_(i, j + 1) - L_(i, j)) / dy - dy * 0.5 * minmod(
                (L_(i, j - 1) - 2.0 * L_(i, j) + L_(i, j + 1)) / (dy * dy),
                (L_(i, j) - 2.0 * L_(i, j + 1) + L_(i, j + 2)) / (dy * dy));
            double dMY = (L_(i, j) - L_(i, j - 1)) / dy + dy * 0.5 * minmod(
                (L_(i, j - 1) - 2.0 * L_(i, j) + L_(i, j + 1)) / (dy * dy),
                (L_(i, j - 2) - 2.0 * L_(i, j - 1) + L_(i, j)) / (dy * dy));
#undef L_
            out[IDX(gr, i, j)] = godunov(ls0[IDX(gr, i, j)], dPX, dMX, dPY, dMY);
        }
}

/* src/levelset2d.cpp:488-557: UnitHJWENO5 along every interior row and column of phi,
 * Godunov |grad phi| with the branch of phi0 >= 0 */
void or_weno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int lx = nx + 2 * g, ly = ny + 2 * g;
    double* dPX = zalloc(gr); double* dMX = zalloc(gr);
    double* dPY = zalloc(gr); double* dMY = zalloc(gr);
    const int lmax = lx > ly ? lx : ly;
    double* F = (double*)calloc((size_t)lmax, sizeof(double));
    double* FP = (double*)calloc((size_t)lmax, sizeof(double));
    double* FM = (double*)calloc((size_t)lmax, sizeof(double));
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = 0.0;
    for (int j = g; j < ny + g; j++) {
        for (int i = 0; i < lx; i++) { F[i] = ls[IDX(gr, i, j)]; FP[i] = 0.0; FM[i] = 0.0; }
        or_unit_hjweno5(F, FP, FM, gr->dx, nx, g);
        for (int i = 0; i < lx; i++) { dPX[IDX(gr, i, j)] = FP[i]; dMX[IDX(gr, i, j)] = FM[i]; }
    }
    for (int i = g; i < nx + g; i++) {
        for (int j = 0; j < ly; j++) { F[j] = ls[IDX(gr, i, j)]; FP[j] = 0.0; FM[j] = 0.0; }
        or_unit_hjweno5(F, FP, FM, gr->dy, ny, g);
        for (int j = 0; j < ly; j++) { dPY[IDX(gr, i, j)] = FP[j]; dMY[IDX(gr, i, j)] = FM[j]; }
    }
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            out[c] = godunov(ls0[c], dPX[c], dMX[c], dPY[c], dMY[c]);
        }
    free(dPX); free(dMX); free(dPY); free(dMY); free(F); free(FP); free(FM);
}

/* src/levelset2d.cpp:615-785 (Min-Gibou subcell fix on top of ENO2) */
void or_subcell_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    const double kSmallValue = 1.0e-10;
    const int64_t n = or_size(gr);
    double* dPX = zalloc(gr); double* dMX = zalloc(gr);
    double* dPY = zalloc(gr); double* dMY = zalloc(gr);
#define L_(a, b) ls[IDX(gr, a, b)]
#define I_(a, b) ls0[IDX(gr, a, b)]
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            dPX[c] = (L_(i + 1, j) - L_(i, j)) / dx - dx * 0.5 * minmod(
                (L_(i - 1, j) - 2.0 * L_(i, j) + L_(i + 1, j)) / (dx * dx),
                (L_(i, j) - 2.0 * L_(i + 1, j) + L_(i + 2, j)) / (dx * dx));
            dMX[c] = (L_(i, j) - L_(i - 1, j)) / dx + dx * 0.5 * minmod(
                (L_(i - 1, j) - 2.0 * L_(i, j) + L_(i + 1, j)) / (dx * dx),
                (L_(i - 2, j) - 2.0 * L_(i - 1, j) + L_(i, j)) / (dx * dx));
            if (I_(i, j) * I_(i + 1, j) < 0) {
                double DDX = minmod(
                    (I_(i - 1, j) - 2.0 * I_(i, j) + I_(i + 1, j)) / (dx * dx),
                    (I_(i, j) - 2.0 * I_(i + 1, j) + I_(i + 2, j)) / (dx * dx));
                double c2 = 0.5 * DDX;
                double c1 = (I_(i + 1, j) - I_(i, j)) / dx;
                double c0 = (I_(i + 1, j) + I_(i, j)) * 0.5 - c2 * dx * dx * 0.25;
                double nd;
                if (fabs(c2) > kSmallValue)
                    nd = 0.5 * dx + (-c1 - isign(I_(i, j)) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
                else
                    nd = 0.5 * dx - c0 / c1;
                dPX[c] = -L_(i, j) / nd - 0.5 * nd * DDX;
                if (fabs(nd) < dx * 0.01) dPX[c] = -0.5 * nd * DDX;
            }
            if (I_(i, j) * I_(i - 1, j) < 0) {
                double DDX = minmod(
                    (I_(i - 1, j) - 2.0 * I_(i, j) + I_(i + 1, j)) / (dx * dx),
                    (I_(i - 2, j) - 2.0 * I_(i - 1, j) + I_(i, j)) / (dx * dx));
                double c2 = 0.5 * DDX;
                double c1 = (I_(i, j) - I_(i - 1, j)) / dx;
                double c0 = (I_(i, j) + I_(i - 1, j)) * 0.5 - c2 * dx * dx * 0.25;
                double nd;
                if (fabs(c2) > kSmallValue)
                    nd = 0.5 * dx - (-c1 + isign(I_(i, j)) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
                else
                    nd = 0.5 * dx + c0 / c1;
                dMX[c] = L_(i, j) / nd + 0.5 * nd * DDX;
                if (fabs(nd) < dx * 0.01) dMX[c] = 0.5 * nd * DDX;
            }
        }
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            dPY[c] = (L_(i, j + 1) - L_(i, j)) / dy - dy * 0.5 * minmod(
                (L_(i, j - 1) - 2.0 * L_(i, j) + L_(i, j + 1)) / (dy * dy),
                (L_(i, j) - 2.0 * L_(i, j + 1) + L_(i, j + 2)) / (dy * dy));
            dMY[c] = (L_(i, j) - L_(i, j - 1)) / dy + dy * 0.5 * minmod(
                (L_(i, j - 1) - 2.0 * L_(i, j) + L_(i, j + 1)) / (dy * dy),
                (L_(i, j - 2) - 2.0 * L_(i, j - 1) + L_(i, j)) / (dy * dy));
            if (I_(i, j) * I_(i, j + 1) < 0) {
                double DDY = minmod(
                    (I_(i, j - 1) - 2.0 * I_(i, j) + I_(i, j + 1)) / (dy * dy),
                    (I_(i, j) - 2.0 * I_(i, j + 1) + I_(i, j + 2)) / (dy * dy));
                double c2 = 0.5 * DDY;
                double c1 = (I_(i, j + 1) - I_(i, j)) / dy;
                double c0 = (I_(i, j + 1) + I_(i, j)) * 0.5 - c2 * dy * dy * 0.25;
                double nd;
                if (fabs(c2) > kSmallValue)
                    nd = 0.5 * dy + (-c1 - isign(I_(i, j)) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
                else
                    nd = 0.5 * dy - c0 / c1;
                dPY[c] = -L_(i, j) / nd - 0.5 * nd * DDY;
                if (fabs(nd) < dy * 0.01) dPY[c] = -0.5 * nd * DDY;
            }
            if (I_(i, j) * I_(i, j - 1) < 0) {
                double DDY = minmod(
                    (I_(i, j - 1) - 2.0 * I_(i, j) + I_(i, j + 1)) / (dy * dy),
                    (I_(i, j - 2) - 2.0 * I_(i, j - 1) + I_(i, j)) / (dy * dy));
                double c2 = 0.5 * DDY;
                double c1 = (I_(i, j) - I_(i, j - 1)) / dy;
                double c0 = (I_(i, j) + I_(i, j - 1)) * 0.5 - c2 * dy * dy * 0.25;
                double nd;
                if (fabs(c2) > kSmallValue)
                    nd = 0.5 * dy - (-c1 + isign(I_(i, j)) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
                else
                    nd = 0.5 * dy + c0 / c1;
                dMY[c] = L_(i, j) / nd + 0.5 * nd * DDY;
                if (fabs(nd) < dy * 0.01) dMY[c] = 0.5 * nd * DDY;
            }
        }
#undef L_
#undef I_
    for (int64_t k = 0; k < n; k++) out[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            out[c] = godunov(ls0[c], dPX[c], dMX[c], dPY[c], dMY[c]);
        }
    free(dPX); free(dMX); free(dPY); free(dMY);
}

/* src/levelset2d.cpp:914-923 */
void or_smoothed_sign(const or_grid* gr, const double* ls, double* out) {
    const double eps = smin(gr->dx, gr->dy);
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = ls[k] / sqrt(sq(ls[k]) + eps * eps);
}

/* src/levelset2d.cpp:27-41 */
void or_heaviside_deriv(const or_grid* gr, const double* ls, double* out) {
    const double eps = smin(gr->dx, gr->dy);
    for (int64_t k = 0; k < or_size(gr); k++) {
        if (fabs(ls[k]) > eps) out[k] = 0.0;
        else out[k] = 0.5 * (1.0 / eps + cos(M_PI * ls[k] / eps));
    }
}

/* src/levelset2d.cpp:787-893 (9-point Simpson volume constraint, quirks kept) */
void or_sussman_constraint(const or_grid* gr, const double* ls, const double* ls0,
                           const double* hd, double* out) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    const double eps = smin(dx, dy), thick = 1.5 * eps, adt = eps * 0.5;
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = 0.0;
#define L_(a, b) ls[IDX(gr, a, b)]
#define I_(a, b) ls0[IDX(gr, a, b)]
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            if (hd[IDX(gr, i, j)] == 0.0) continue;
            double F[3][3], IF[3][3], Hp[3][3], Lm[3][3], dA[3][3], f[3][3];
            F[0][0] = 0.25 * (L_(i, j) + L_(i - 1, j) + L_(i - 1, j - 1) + L_(i, j - 1));
            F[0][1] = 0.5 * (L_(i, j - 1) + L_(i, j));
            F[0][2] = 0.25 * (L_(i, j) + L_(i - 1, j) + L_(i - 1, j + 1) + L_(i, j + 1));
            F[1][0] = 0.5 * (L_(i, j - 1) + L_(i, j));
            F[1][1] = L_(i, j);
            F[1][2] = 0.5 * (L_(i, j) + L_(i, j + 1));
            F[2][0] = 0.25 * (L_(i, j) + L_(i + 1, j) + L_(i + 1, j - 1) + L_(i, j - 1));
            F[2][1] = 0.5 * (L_(i, j) + L_(i + 1, j));
            F[2][2] = 0.25 * (L_(i, j) + L_(i + 1, j) + L_(i + 1, j + 1) + L_(i, j + 1));
            IF[0][0] = 0.25 * (I_(i, j) + I_(i - 1, j) + I_(i - 1, j - 1) + I_(i, j - 1));
            IF[0][1] = 0.5 * (I_(i, j - 1) + I_(i, j));
            IF[0][2] = 0.25 * (I_(i, j) + I_(i - 1, j) + I_(i - 1, j + 1) + I_(i, j + 1));
            IF[1][0] = 0.5 * (I_(i, j - 1) + I_(i, j));
            IF[1][1] = I_(i, j);
            IF[1][2] = 0.5 * (I_(i, j) + I_(i, j + 1));
            IF[2][0] = 0.25 * (I_(i, j) + I_(i + 1, j) + I_(i + 1, j - 1) + I_(i, j - 1));
            IF[2][1] = 0.5 * (I_(i, j) + I_(i + 1, j));
            IF[2][2] = 0.25 * (I_(i, j) + I_(i + 1, j) + I_(i + 1, j + 1) + I_(i, j + 1));
            dA[0][0] = sqrt(sq(((I_(i, j) + I_(i, j - 1)) * 0.5 - 0.5 * (I_(i - 1, j) + I_(i - 1, j - 1))) / dx)
                          + sq((IF[0][1] - 0.5 * (I_(i, j - 1) + I_(i - 1, j - 1))) / dy));
            dA[0][1] = sqrt(sq((IF[1][1] - I_(i - 1, j)) / dx) + sq((IF[2][0] - IF[0][0]) / dy));
            dA[0][2] = sqrt(sq((IF[1][2] - -0.5 * (I_(i - 1, j) + I_(i - 1, j + 1))) / dx)
                          + sq((0.5 * (I_(i - 1, j + 1) + I_(i, j + 1)) - IF[0][1]) / dy));
            dA[1][0] = sqrt(sq((IF[2][0] - IF[0][0]) / dx) + sq((IF[1][1] - 0.5 * (I_(i, j - 1) + I_(i, j))) / dy));
            dA[1][1] = sqrt(sq((IF[2][1] - IF[0][1]) / dx) + sq((IF[1][2] - IF[1][0]) / dy));
            dA[1][2] = sqrt(sq((IF[2][2] - IF[0][2]) / dx) + sq((I_(i, j + 1) - IF[1][1]) / dy));
            dA[2][0] = sqrt(sq((0.5 * (I_(i + 1, j) + I_(i + 1, j - 1)) - IF[1][0]) / dx)
                          + sq((IF[2][1] - 0.5 * (I_(i, j - 1) + I_(i + 1, j - 1))) / dy));
            dA[2][1] = sqrt(sq((I_(i + 1, j) - IF[1][1]) / (2.0 * dx)) + sq((IF[2][2] - IF[2][0]) / dy));
            dA[2][2] = sqrt(sq((0.5 * (I_(i + 1, j) + I_(i + 1, j + 1)) - IF[1][1]) / dx)
                          + sq((0.5 * (I_(i, j + 1) + I_(i + 1, j + 1)) - IF[2][1]) / dy));
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    if (fabs(F[a][b]) > thick) Hp[a][b] = 0.0;
                    else Hp[a][b] = 0.5 / eps + 0.5 / eps * cos(M_PI * F[a][b] / eps);
                    Lm[a][b] = (F[a][b] - IF[a][b]) / adt;
                    f[a][b] = Hp[a][b] * dA[a][b];
                }
            double s1 = 0.0, s2 = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) { s1 += Hp[a][b] * Lm[a][b]; s2 += Hp[a][b] * f[a][b]; }
            s1 += 15.0 * Hp[1][1] * Lm[1][1];
            s2 += 15.0 * Hp[1][1] * f[1][1];
            out[IDX(gr, i, j)] -= adt * f[1][1] * s1 / s2;
        }
#undef L_
#undef I_
}

/* iteration count of the pseudo-time loop, src/levelset2d.cpp:447-448 */
int or_reinit_iterations(double eps, double max_time) {
    const double adt = eps * 0.5;
    double t = 0.0;
    int k = 0;
    while (t < max_time) { t += adt; k++; }
    return k;
}

/* src/levelset2d.cpp:433-486 */
void or_reinit_minrk2(const or_grid* gr, double* ls, int iters) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double eps = smin(gr->dx, gr->dy), adt = eps * 0.5;
    const int64_t n = or_size(gr);
    or_bc_p(gr, ls);
    double* ls0 = zalloc(gr); double* S = zalloc(gr); double* a1 = zalloc(gr); double* a2 = zalloc(gr);
    double* LS1 = zalloc(gr); double* LS2 = zalloc(gr); double* tmp = zalloc(gr);
    memcpy(ls0, ls, sizeof(double) * (size_t)n);
    or_smoothed_sign(gr, ls0, S);
    for (int it = 0; it < iters; it++) {
        or_bc_p(gr, ls);
        or_subcell_absgrad(gr, ls, ls0, a1);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS1[c] = ls[c] - adt * S[c] * (a1[c] - 1.0);
            }
        or_bc_p(gr, LS1);
        or_subcell_absgrad(gr, LS1, ls0, a2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS2[c] = LS1[c] - adt * S[c] * (a2[c] - 1.0);
            }
        or_bc_p(gr, LS2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                tmp[c] = 0.5 * (ls[c] + LS2[c]);
            }
        or_bc_p(gr, tmp);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) ls[IDX(gr, i, j)] = tmp[IDX(gr, i, j)];
    }
    or_bc_p(gr, ls);
    free(ls0); free(S); free(a1); free(a2); free(LS1); free(LS2); free(tmp);
}

/* src/levelset2d.cpp:264-316: Heun with plain ENO2 and the integer sign of the input
 * (captured before any BC), BC on every stage; the Heaviside-derivative mask computed there
 * (:307) does not change the result (its test is commented out). */
void or_reinit_original(const or_grid* gr, double* ls, int iters) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double eps = smin(gr->dx, gr->dy), adt = eps * 0.5;
    const int64_t n = or_size(gr);
    double* ls0 = zalloc(gr); double* a1 = zalloc(gr); double* a2 = zalloc(gr);
    double* LS1 = zalloc(gr); double* LS2 = zalloc(gr); double* tmp = zalloc(gr);
    memcpy(ls0, ls, sizeof(double) * (size_t)n);
    for (int it = 0; it < iters; it++) {
        or_bc_p(gr, ls);
        or_eno_absgrad(gr, ls, ls0, a1);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS1[c] = ls[c] - adt * isign(ls0[c]) * (a1[c] - 1.0);
            }
        or_bc_p(gr, LS1);
        or_eno_absgrad(gr, LS1, ls0, a2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS2[c] = LS1[c] - adt * isign(ls0[c]) * (a2[c] - 1.0);
            }
        /* the BCs of LS2 and tmpLS (levelset2d.cpp:299, 305): they change interior rows where the
         * east P side is periodic (bc2d.cpp:792-796 writes inside the domain) */
        or_bc_p(gr, LS2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                tmp[c] = 0.5 * (ls[c] + LS2[c]);
            }
        or_bc_p(gr, tmp);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) ls[IDX(gr, i, j)] = tmp[IDX(gr, i, j)];
    }
    or_bc_p(gr, ls);
    free(ls0); free(a1); free(a2); free(LS1); free(LS2); free(tmp);
}

/* src/levelset2d.cpp:377-431: phi0 normalised by its ENO2 |grad phi0| where that is
 * positive (then BC), Heun whose second stage restarts from phi, until the pseudo time
 * reaches max(dx nx, dy ny) (`iters` = that loop count). */
void or_reinit_first_time(const or_grid* gr, double* ls, int iters) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double eps = smin(gr->dx, gr->dy), adt = eps * 0.5;
    const int64_t n = or_size(gr);
    double* ls0 = zalloc(gr); double* a1 = zalloc(gr); double* a2 = zalloc(gr);
    double* LS1 = zalloc(gr); double* LS2 = zalloc(gr);
    memcpy(ls0, ls, sizeof(double) * (size_t)n);
    or_eno_absgrad(gr, ls0, ls0, a1);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            if (a1[c] > 0.0) ls0[c] = ls0[c] / a1[c];
        }
    or_bc_p(gr, ls0);
    for (int it = 0; it < iters; it++) {
        or_bc_p(gr, ls);
        or_eno_absgrad(gr, ls, ls0, a1);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS1[c] = ls[c] - adt * isign(ls0[c]) * (a1[c] - 1.0);
            }
        or_bc_p(gr, LS1);
        or_eno_absgrad(gr, LS1, ls0, a2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS2[c] = ls[c] - adt * isign(ls0[c]) * (a2[c] - 1.0);
            }
        or_bc_p(gr, LS2);  /* levelset2d.cpp:419: changes interior rows where the east P side is periodic */
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                ls[c] = 0.5 * (ls[c] + LS2[c]);
            }
    }
    or_bc_p(gr, ls);
    free(ls0); free(a1); free(a2); free(LS1); free(LS2);
}

/* src/levelset2d.cpp:318-375 (LS1 gets no BC: its ghosts stay 0) */
void or_reinit_sussman(const or_grid* gr, double* ls, int iters) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double eps = smin(gr->dx, gr->dy), adt = eps * 0.5;
    const int64_t n = or_size(gr);
    double* ls0 = zalloc(gr); double* a1 = zalloc(gr); double* a2 = zalloc(gr);
    double* LS1 = zalloc(gr); double* LS2 = zalloc(gr); double* tmp = zalloc(gr);
    double* hd = zalloc(gr); double* cons = zalloc(gr);
    memcpy(ls0, ls, sizeof(double) * (size_t)n);
    for (int it = 0; it < iters; it++) {
        or_bc_p(gr, ls);
        or_eno_absgrad(gr, ls, ls0, a1);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS1[c] = ls[c] - adt * isign(ls0[c]) * (a1[c] - 1.0);
            }
        or_eno_absgrad(gr, LS1, ls0, a2);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                LS2[c] = LS1[c] - adt * isign(ls0[c]) * (a2[c] - 1.0);
            }
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) {
                const int64_t c = IDX(gr, i, j);
                tmp[c] = 0.5 * (ls[c] + LS2[c]);
            }
        or_bc_p(gr, tmp);
        or_heaviside_deriv(gr, tmp, hd);
        or_sussman_constraint(gr, tmp, ls0, hd, cons);
        for (int i = g; i < nx + g; i++)
            for (int j = g; j < ny + g; j++) ls[IDX(gr, i, j)] += cons[IDX(gr, i, j)];
    }
    or_bc_p(gr, ls);
    free(ls0); free(a1); free(a2); free(LS1); free(LS2); free(tmp); free(hd); free(cons);
}

/* ------------------------------------------------------------------ MAC helpers */
/* src/mac2d.cpp:221-252.  Writes the interior only. */
void or_update_kappa(const or_grid* gr, const double* ls, double* kappa) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy, eps = 1.0e-100;
    double* dX = zalloc(gr); double* dY = zalloc(gr); double* S = zalloc(gr);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            dX[c] = (ls[IDX(gr, i + 1, j)] - ls[IDX(gr, i - 1, j)]) / (2.0 * dx);
            dY[c] = (ls[IDX(gr, i, j + 1)] - ls[IDX(gr, i, j - 1)]) / (2.0 * dy);
            S[c] = sqrt(dX[c] * dX[c] + dY[c] * dY[c] + eps);
        }
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            double k = (dX[c] * dX[c] * (ls[IDX(gr, i, j - 1)] - 2.0 * ls[c] + ls[IDX(gr, i, j + 1)]) / (dy * dy)
                      - 2.0 * dX[c] * dY[c] * (dX[IDX(gr, i, j + 1)] - dX[IDX(gr, i, j - 1)]) / (2.0 * dy)
                      + dY[c] * dY[c] * (ls[IDX(gr, i - 1, j)] - 2.0 * ls[c] + ls[IDX(gr, i + 1, j)]) / (dx * dx))
                     / pow(S[c], 3.0);
            kappa[c] = isign(k) * fabs(smin(fabs(k), 1.0 / smin(dx, dy)));
        }
    free(dX); free(dY); free(S);
}

/* src/mac2d.cpp:87-101 */
void or_heaviside(const or_grid* gr, const double* ls, double* H) {
    for (int64_t k = 0; k < or_size(gr); k++) H[k] = (ls[k] >= 0.0) ? 1.0 : 0.0;
}

/* src/mac2d.cpp:103-122 */
void or_smooth_heaviside(const or_grid* gr, const double* ls, double* H) {
    const double eps = smin(gr->dx, gr->dy) * 1.5;
    for (int64_t k = 0; k < or_size(gr); k++) {
        if (ls[k] > eps) H[k] = 1.0;
        else if (ls[k] < -eps) H[k] = 0.0;
        else H[k] = 0.5 * (1.0 + ls[k] / eps + 1.0 / M_PI * sin(M_PI * ls[k] / eps));
    }
}

/* src/mac2d.cpp:1729-1738 */
void or_divergence(const or_grid* gr, const double* u, const double* v, double* div) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    for (int64_t k = 0; k < or_size(gr); k++) div[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++)
            div[IDX(gr, i, j)] = (u[IDX(gr, i + 1, j)] - u[IDX(gr, i, j)]) / gr->dx
                               + (v[IDX(gr, i, j + 1)] - v[IDX(gr, i, j)]) / gr->dy;
}

/* face fraction of the heavy phase and the curvature weight, for the face between
 * a cell "m" and its neighbour "nb"; first = 1 when nb is the west/south neighbour.
 * src/mac2d.cpp:1345-1358 (W), :1375-1388 (E) */
static void face_theta(double lsA, double lsB, int nb_first, double* thetaH, double* theta) {
    /* lsA = west/south value, lsB = east/north value of the face */
    const double s = fabs(lsA) + fabs(lsB);
    if (lsA >= 0.0 && lsB >= 0.0) { *thetaH = 1.0; *theta = 0.0; }
    else if (lsA < 0.0 && lsB < 0.0) { *thetaH = 0.0; *theta = 0.0; }
    else if (lsA >= 0.0 && lsB < 0.0) {
        *thetaH = fabs(lsA) / s;
        *theta = nb_first ? fabs(lsA) / s : fabs(lsB) / s;
    } else {
        *thetaH = fabs(lsB) / s;
        *theta = nb_first ? fabs(lsA) / s : fabs(lsB) / s;
    }
}

/* src/mac2d.cpp:1740-1849 */
void or_update_vel(const or_grid* gr, const or_fluid* fl, double* u, double* v,
                   const double* us, const double* vs, const double* ps,
                   const double* ls, const double* kappa, const double* H) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    for (int i = g + 1; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            double thH, th;
            face_theta(ls[IDX(gr, i - 1, j)], ls[IDX(gr, i, j)], 1, &thH, &th);
            const double rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iRho = 1.0 / rhoEff;
            const double kEff = kappa[IDX(gr, i, j)] * th + kappa[IDX(gr, i - 1, j)] * (1.0 - th);
            double val = fl->dt * (-fl->sigma * kEff) * (H[IDX(gr, i, j)] - H[IDX(gr, i - 1, j)]);
            val *= iRho / dx;
            val += us[IDX(gr, i, j)] - iRho * (ps[IDX(gr, i, j)] - ps[IDX(gr, i - 1, j)]) / dx;
            u[IDX(gr, i, j)] = val;
        }
    for (int i = g; i < nx + g; i++)
        for (int j = g + 1; j < ny + g; j++) {
            double thH, th;
            face_theta(ls[IDX(gr, i, j - 1)], ls[IDX(gr, i, j)], 1, &thH, &th);
            const double rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iRho = 1.0 / rhoEff;
            const double kEff = kappa[IDX(gr, i, j)] * th + kappa[IDX(gr, i, j - 1)] * (1.0 - th);
            double val = fl->dt * (-fl->sigma * kEff) * (H[IDX(gr, i, j)] - H[IDX(gr, i, j - 1)]);
            val *= iRho / dy;
            val += vs[IDX(gr, i, j)] - iRho * (ps[IDX(gr, i, j)] - ps[IDX(gr, i, j - 1)]) / dy;
            v[IDX(gr, i, j)] = val;
        }
    or_bc_u(gr, u);
    or_bc_v(gr, v);
}

/* src/mac2d.cpp:1851-1880 */
double or_update_dt(const or_grid* gr, const or_fluid* fl, const double* u, const double* v) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    double uA = 0.0, vA = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            uA = smax(uA, fabs((u[IDX(gr, i + 1, j)] * u[IDX(gr, i, j)]) * 0.5));
            vA = smax(vA, fabs((v[IDX(gr, i, j + 1)] * v[IDX(gr, i, j)]) * 0.5));
        }
    const double C = uA / dx + vA / dy;
    const double V = smax(fl->mu_h, fl->mu_l) * (2.0 / (dx * dx + dy * dy));
    const double G = smax(sqrt(fabs(fl->g) / dy), sqrt(fabs(fl->g) / dx));
    const double S = sqrt((fl->sigma / smin(dx, dy)) / (smin(fl->rho_h, fl->rho_l) * pow(smin(dx, dy), 2.0)));
    double dt = 1.7976931348623157e308;
    dt = smin(dt, fl->cfl / (0.5 * (C + V + sqrt(pow(C + V, 2.0) + 4.0 * G * G + 4.0 * S * S))));
    return dt;
}

/* ------------------------------------------------------------------ momentum */
/* src/mac2d.cpp:308-387 */
void or_convection_u(const or_grid* gr, const double* u, const double* v, double* cU) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int lx = nx + 2 * g, ly = ny + 2 * g;
    double* tV = zalloc(gr);
    double* LXP = zalloc(gr); double* LXM = zalloc(gr); double* LYP = zalloc(gr); double* LYM = zalloc(gr);
    double* F = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FP = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FM = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    for (int j = g; j < ny + g; j++)
        for (int i = g + 1; i < nx + g; i++)
            tV[IDX(gr, i, j)] = (v[IDX(gr, i, j)] + v[IDX(gr, i - 1, j)] + v[IDX(gr, i - 1, j + 1)] + v[IDX(gr, i, j + 1)]) * 0.25;
    for (int j = g; j < ny + g; j++) {
        for (int i = 0; i < lx; i++) F[i] = u[IDX(gr, i, j)] * u[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)lx); memset(FM, 0, sizeof(double) * (size_t)lx);
        or_unit_hjweno5(F, FP, FM, gr->dx, nx, g);
        for (int i = 0; i < lx; i++) { LXP[IDX(gr, i, j)] = FP[i]; LXM[IDX(gr, i, j)] = FM[i]; }
    }
    for (int i = g; i < nx + g; i++) {
        for (int j = 0; j < ly; j++) F[j] = u[IDX(gr, i, j)] * tV[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)ly); memset(FM, 0, sizeof(double) * (size_t)ly);
        or_unit_hjweno5(F, FP, FM, gr->dy, ny, g);
        for (int j = 0; j < ly; j++) { LYP[IDX(gr, i, j)] = FP[j]; LYM[IDX(gr, i, j)] = FM[j]; }
    }
    for (int64_t k = 0; k < or_size(gr); k++) cU[k] = 0.0;
    for (int i = g + 1; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            const double aX = u[c], aY = tV[c];
            cU[c] = (0.5 * (LXP[c] + LXM[c]) + 0.5 * (LYP[c] + LYM[c])
                   - aX * (0.5 * (LXP[c] - LXM[c])) - aY * (0.5 * (LYP[c] - LYM[c])));
        }
    free(tV); free(LXP); free(LXM); free(LYP); free(LYM); free(F); free(FP); free(FM);
}

/* src/mac2d.cpp:389-463 */
void or_convection_v(const or_grid* gr, const double* u, const double* v, double* cV) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int lx = nx + 2 * g, ly = ny + 2 * g;
    double* tU = zalloc(gr);
    double* LXP = zalloc(gr); double* LXM = zalloc(gr); double* LYP = zalloc(gr); double* LYM = zalloc(gr);
    double* F = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FP = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    double* FM = (double*)calloc((size_t)(lx > ly ? lx : ly), sizeof(double));
    for (int j = g + 1; j < ny + g; j++)
        for (int i = g; i < nx + g; i++)
            tU[IDX(gr, i, j)] = (u[IDX(gr, i, j)] + u[IDX(gr, i, j - 1)] + u[IDX(gr, i + 1, j - 1)] + u[IDX(gr, i + 1, j)]) * 0.25;
    for (int j = g; j < ny + g; j++) {
        for (int i = 0; i < lx; i++) F[i] = v[IDX(gr, i, j)] * tU[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)lx); memset(FM, 0, sizeof(double) * (size_t)lx);
        or_unit_hjweno5(F, FP, FM, gr->dx, nx, g);
        for (int i = 0; i < lx; i++) { LXP[IDX(gr, i, j)] = FP[i]; LXM[IDX(gr, i, j)] = FM[i]; }
    }
    for (int i = g; i < nx + g; i++) {
        for (int j = 0; j < ly; j++) F[j] = v[IDX(gr, i, j)] * v[IDX(gr, i, j)];
        memset(FP, 0, sizeof(double) * (size_t)ly); memset(FM, 0, sizeof(double) * (size_t)ly);
        or_unit_hjweno5(F, FP, FM, gr->dy, ny, g);
        for (int j = 0; j < ly; j++) { LYP[IDX(gr, i, j)] = FP[j]; LYM[IDX(gr, i, j)] = FM[j]; }
    }
    for (int64_t k = 0; k < or_size(gr); k++) cV[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g + 1; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            const double aX = tU[c], aY = v[c];
            cV[c] = (0.5 * (LXP[c] + LXM[c]) + 0.5 * (LYP[c] + LYM[c])
                   - aX * (0.5 * (LXP[c] - LXM[c])) - aY * (0.5 * (LYP[c] - LYM[c])));
        }
    free(tU); free(LXP); free(LXM); free(LYP); free(LYM); free(F); free(FP); free(FM);
}

/* harmonic-average inverse density at a U/V face, src/mac2d.cpp:729-781 */
static double irho_harmonic(const or_fluid* fl, double a, double b) {
    const double rH = fl->rho_h, rL = fl->rho_l;
    if (a >= 0 && b >= 0) return 1.0 / rH;
    else if (a <= 0 && b <= 0) return 1.0 / rL;
    else if (a > 0 && b <= 0) {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return 1.0 / (rL * rH) / (1.0 / rL * th + 1.0 / rH * (1.0 - th));
    } else {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return 1.0 / (rH * rL) / (1.0 / rH * th + 1.0 / rL * (1.0 - th));
    }
}

/* arithmetic-average density at a V face, src/mac2d.cpp:947-999 */
static double rho_arith(const or_fluid* fl, double a, double b) {
    const double rH = fl->rho_h, rL = fl->rho_l;
    if (a >= 0 && b >= 0) return rH;
    else if (a <= 0 && b <= 0) return rL;
    else if (a > 0 && b <= 0) {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return rH * th + rL * (1.0 - th);
    } else {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return rL * th + rH * (1.0 - th);
    }
}

/* src/mac2d.cpp:570-809 */
void or_viscosity_u(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                    const double* ls, const double* H, double* dU) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) dU[k] = 0.0;
    if (fl->re <= 0.0) return;
    double* mu = zalloc(gr);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) mu[IDX(gr, i, j)] = fl->mu_l + (fl->mu_h - fl->mu_l) * H[IDX(gr, i, j)];
#define M_(a, b) mu[IDX(gr, a, b)]
#define LS_(a, b) ls[IDX(gr, a, b)]
    for (int i = g + 1; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const double uM = u[IDX(gr, i, j)], uW = u[IDX(gr, i - 1, j)], uE = u[IDX(gr, i + 1, j)];
            const double uS = u[IDX(gr, i, j - 1)], uN = u[IDX(gr, i, j + 1)];
            const double vW = v[IDX(gr, i - 1, j)], vW_N = v[IDX(gr, i - 1, j + 1)];
            const double vM = v[IDX(gr, i, j)], vN = v[IDX(gr, i, j + 1)];
            const double lsW = LS_(i - 1, j), lsM = LS_(i, j);
            const double lsUNHalf = 0.25 * (LS_(i, j) + LS_(i - 1, j) + LS_(i - 1, j + 1) + LS_(i, j + 1));
            const double lsUSHalf = 0.25 * (LS_(i, j) + LS_(i - 1, j) + LS_(i - 1, j - 1) + LS_(i, j - 1));
            const double iWE = irho_harmonic(fl, lsW, lsM);
            const double iSN = irho_harmonic(fl, lsUSHalf, lsUNHalf);
            const double muU_X_W = M_(i - 1, j) * (uM - uW) / dx;
            const double muU_X_E = M_(i, j) * (uE - uM) / dx;
            const double muU_Y_S = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j - 1) + M_(i, j - 1)) * (uM - uS) / dy;
            const double muU_Y_N = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j + 1) + M_(i, j + 1)) * (uN - uM) / dy;
            const double muV_X_S = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j - 1) + M_(i, j - 1)) * (vM - vW) / dx;
            const double muV_X_N = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j + 1) + M_(i, j + 1)) * (vN - vW_N) / dx;
            const double visX = 2.0 * (muU_X_E - muU_X_W) / dx;
            const double visY = (muU_Y_N - muU_Y_S) / dy + (muV_X_N - muV_X_S) / dy;
            dU[IDX(gr, i, j)] = visX * iWE + visY * iSN;
        }
#undef M_
#undef LS_
    free(mu);
}

/* src/mac2d.cpp:811-1041 */
void or_viscosity_v(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                    const double* ls, const double* H, double* dV) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) dV[k] = 0.0;
    if (fl->re <= 0.0) return;
    double* mu = zalloc(gr);
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) mu[IDX(gr, i, j)] = fl->mu_l + (fl->mu_h - fl->mu_l) * H[IDX(gr, i, j)];
#define M_(a, b) mu[IDX(gr, a, b)]
#define LS_(a, b) ls[IDX(gr, a, b)]
    for (int i = g; i < nx + g; i++)
        for (int j = g + 1; j < ny + g; j++) {
            const double vW = v[IDX(gr, i - 1, j)], vE = v[IDX(gr, i + 1, j)], vS = v[IDX(gr, i, j - 1)];
            const double vN = v[IDX(gr, i, j + 1)], vM = v[IDX(gr, i, j)];
            const double uM = u[IDX(gr, i, j)], uE = u[IDX(gr, i + 1, j)];
            const double uS = u[IDX(gr, i, j - 1)], uE_S = u[IDX(gr, i + 1, j - 1)];
            const double lsVWHalf = 0.25 * (LS_(i, j) + LS_(i - 1, j) + LS_(i - 1, j - 1) + LS_(i, j - 1));
            const double lsVEHalf = 0.25 * (LS_(i, j) + LS_(i + 1, j) + LS_(i + 1, j - 1) + LS_(i, j - 1));
            const double lsM = LS_(i, j), lsS = LS_(i, j - 1);
            const double iWE = 1.0 / rho_arith(fl, lsVWHalf, lsVEHalf);
            const double iSN = 1.0 / rho_arith(fl, lsS, lsM);
            const double muU_Y_W = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j - 1) + M_(i, j - 1)) * (uM - uS) / dy;
            const double muU_Y_E = 0.25 * (M_(i, j) + M_(i + 1, j) + M_(i + 1, j - 1) + M_(i, j - 1)) * (uE - uE_S) / dy;
            const double muV_X_W = 0.25 * (M_(i, j) + M_(i - 1, j) + M_(i - 1, j - 1) + M_(i, j - 1)) * (vM - vW) / dx;
            const double muV_X_E = 0.25 * (M_(i, j) + M_(i + 1, j) + M_(i + 1, j - 1) + M_(i, j - 1)) * (vE - vM) / dx;
            const double muV_Y_S = M_(i, j - 1) * (vM - vS) / dy;
            const double muV_Y_N = M_(i, j) * (vN - vM) / dy;
            const double visX = (muU_Y_E - muU_Y_W) / dx + (muV_X_E - muV_X_W) / dx;
            const double visY = 2.0 * (muV_Y_N - muV_Y_S) / dy;
            dV[IDX(gr, i, j)] = visX * iWE + visY * iSN;
        }
#undef M_
#undef LS_
    free(mu);
}

static int gravity_active(const or_fluid* fl, int axis) {
    return (fl->fr != 0 || fl->g != 0.0) && !isnan(fl->fr) && !isinf(fl->fr) && fl->g_axis == axis;
}

/* src/mac2d.cpp:254-306 (UpdateFU/UpdateFV) with gravity :1043-1079 */
static void rhs_u(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u,
                  const double* v, const double* H, double* out, double* t1, double* t2) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    or_convection_u(gr, u, v, t1);
    or_viscosity_u(gr, fl, u, v, ls, H, t2);
    const double gU = gravity_active(fl, 0) ? -fl->g : 0.0;
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = 0.0;
    for (int i = g + 1; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            out[c] = -t1[c] + t2[c] + gU;
        }
}

static void rhs_v(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u,
                  const double* v, const double* H, double* out, double* t1, double* t2) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    or_convection_v(gr, u, v, t1);
    or_viscosity_v(gr, fl, u, v, ls, H, t2);
    for (int64_t k = 0; k < or_size(gr); k++) out[k] = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g + 1; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            double gv;
            if (fl->gravity_fix) gv = gravity_active(fl, 1) ? -fl->g : 0.0;
            else /* mac2d.cpp:294 uses the U-grid gravity array here */
                gv = (gravity_active(fl, 0) && i >= g + 1) ? -fl->g : 0.0;
            out[c] = -t1[c] + t2[c] + gv;
        }
}

/* src/mac2d.cpp:1081-1239; time_order 0 = Euler, 1 = RK2, 2 = RK3 */
void or_intermediate_vel(const or_grid* gr, const or_fluid* fl, int time_order,
                         const double* ls, const double* u, const double* v,
                         double* uhat, double* vhat, const double* Hs) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const int64_t n = or_size(gr);
    const double dt = fl->dt;
    double* FU1 = zalloc(gr); double* FV1 = zalloc(gr); double* FU2 = zalloc(gr); double* FV2 = zalloc(gr);
    double* t1 = zalloc(gr); double* t2 = zalloc(gr);
#define UINT for (int i = g + 1; i < nx + g; i++) for (int j = g; j < ny + g; j++)
#define VINT for (int i = g; i < nx + g; i++) for (int j = g + 1; j < ny + g; j++)
#define C_ IDX(gr, i, j)
    if (time_order == 0) {
        rhs_u(gr, fl, ls, u, v, Hs, FU1, t1, t2);
        rhs_v(gr, fl, ls, u, v, Hs, FV1, t1, t2);
        UINT uhat[C_] = u[C_] + dt * FU1[C_];
        VINT vhat[C_] = v[C_] + dt * FV1[C_];
    } else if (time_order == 1) {
        rhs_u(gr, fl, ls, u, v, Hs, FU1, t1, t2);
        rhs_v(gr, fl, ls, u, v, Hs, FV1, t1, t2);
        UINT FU2[C_] = u[C_] + dt * FU1[C_];
        VINT FV2[C_] = v[C_] + dt * FV1[C_];
        or_bc_u(gr, FU2); or_bc_v(gr, FV2);
        rhs_u(gr, fl, ls, FU2, FV2, Hs, FU1, t1, t2);
        rhs_v(gr, fl, ls, FU2, FV2, Hs, FV1, t1, t2);
        UINT FU2[C_] = FU2[C_] + dt * FU1[C_];
        VINT FV2[C_] = FV2[C_] + dt * FV1[C_];
        UINT uhat[C_] = 0.5 * u[C_] + 0.5 * FU2[C_];
        VINT vhat[C_] = 0.5 * v[C_] + 0.5 * FV2[C_];
    } else {
        rhs_u(gr, fl, ls, u, v, Hs, FU1, t1, t2);
        rhs_v(gr, fl, ls, u, v, Hs, FV1, t1, t2);
        UINT FU2[C_] = u[C_] + dt * FU1[C_];
        VINT FV2[C_] = v[C_] + dt * FV1[C_];
        or_bc_u(gr, FU2); or_bc_v(gr, FV2);
        rhs_u(gr, fl, ls, FU2, FV2, Hs, FU1, t1, t2);
        rhs_v(gr, fl, ls, FU2, FV2, Hs, FV1, t1, t2);
        UINT FU2[C_] = FU2[C_] + dt * FU1[C_];
        VINT FV2[C_] = FV2[C_] + dt * FV1[C_];
        memset(FU1, 0, sizeof(double) * (size_t)n); memset(FV1, 0, sizeof(double) * (size_t)n);
        UINT FU1[C_] = 0.75 * u[C_] + 0.25 * FU2[C_];
        VINT FV1[C_] = 0.75 * v[C_] + 0.25 * FV2[C_];
        or_bc_u(gr, FU1); or_bc_v(gr, FV1);
        rhs_u(gr, fl, ls, FU1, FV1, Hs, FU2, t1, t2);
        rhs_v(gr, fl, ls, FU1, FV1, Hs, FV2, t1, t2);
        UINT FU1[C_] = FU1[C_] + dt * FU2[C_];
        VINT FV1[C_] = FV1[C_] + dt * FV2[C_];
        UINT uhat[C_] = 1.0 / 3.0 * u[C_] + 2.0 / 3.0 * FU1[C_];
        VINT vhat[C_] = 1.0 / 3.0 * v[C_] + 2.0 / 3.0 * FV1[C_];
    }
#undef UINT
#undef VINT
#undef C_
    free(FU1); free(FV1); free(FU2); free(FV2); free(t1); free(t2);
}

/* ------------------------------------------------------------------ GFM assembly */
static int all_neumann_family(const or_grid* gr) {
    for (int s = 0; s < 4; s++) {
        const int k = gr->bc_p[s];
        if (!(k == OR_BC_NEUMANN || k == OR_BC_AXISYM || k == OR_BC_WALL)) return 0;
    }
    return 1;
}

/* src/mac2d.cpp:1330-1476 (face coefficients and jump rhs) and :1478-1656 (matrix,
 * BC handling, pin).  The matrix is stored matrix-free: ax/ay hold the CSR
 * off-diagonal value of the west/south face of each cell (0 where the reference
 * drops the neighbour), diag the CSR centre value. */
void or_gfm_assemble_stencil(const or_grid* gr, const or_fluid* fl, const double* ls,
                             const double* kappa, const double* H, const double* div,
                             double* ax, double* ay, double* diag, double* rhs) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy, dt = fl->dt, sigma = fl->sigma;
    const int pin = all_neumann_family(gr);
    for (int64_t k = 0; k < or_size(gr); k++) { ax[k] = 0.0; ay[k] = 0.0; diag[k] = 0.0; rhs[k] = 0.0; }
#define LS_(a, b) ls[IDX(gr, a, b)]
#define K_(a, b) kappa[IDX(gr, a, b)]
#define H_(a, b) H[IDX(gr, a, b)]
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            const double lsW = LS_(i - 1, j), lsE = LS_(i + 1, j), lsM = LS_(i, j);
            const double lsS = LS_(i, j - 1), lsN = LS_(i, j + 1);
            double thH, th, rhoEff, kE;
            face_theta(lsW, lsM, 1, &thH, &th);
            rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iW = 1.0 / rhoEff;
            kE = K_(i, j) * th + K_(i - 1, j) * (1.0 - th);
            double FW = -dt * (-sigma * kE) * (H_(i, j) - H_(i - 1, j));
            FW *= iW / (dx * dx);
            face_theta(lsM, lsE, 0, &thH, &th);
            rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iE = 1.0 / rhoEff;
            kE = K_(i, j) * th + K_(i + 1, j) * (1.0 - th);
            double FE = dt * (-sigma * kE) * (H_(i + 1, j) - H_(i, j));
            FE *= iE / (dx * dx);
            face_theta(lsS, lsM, 1, &thH, &th);
            rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iS = 1.0 / rhoEff;
            kE = K_(i, j) * th + K_(i, j - 1) * (1.0 - th);
            double FS = -dt * (-sigma * kE) * (H_(i, j) - H_(i, j - 1));
            FS *= iS / (dy * dy);
            face_theta(lsM, lsN, 0, &thH, &th);
            rhoEff = fl->rho_h * thH + fl->rho_l * (1.0 - thH);
            const double iN = 1.0 / rhoEff;
            kE = K_(i, j) * th + K_(i, j + 1) * (1.0 - th);
            double FN = dt * (-sigma * kE) * (H_(i, j + 1) - H_(i, j));
            FN *= iN / (dy * dy);
            double r = 0.0;
            r -= FW + FE + FS + FN;
            r -= div[c];
            /* matrix row, mac2d.cpp:1498-1583 */
            const double nS = -iS, nW = -iW, nE = -iE, nN = -iN;
            double C = -(nW + nE) / (dx * dx) - (nS + nN) / (dy * dy);
            double aW = nW / (dx * dx), aS = nS / (dy * dy);
            const int kw = gr->bc_p[0], ke = gr->bc_p[1], ks = gr->bc_p[2], kn = gr->bc_p[3];
            if (i == g) {
                if (kw == OR_BC_NEUMANN) { aW = 0.0; C += nW / (dx * dx); }
                else if (kw == OR_BC_DIRICHLET || kw == OR_BC_INLET || kw == OR_BC_OUTLET ||
                         kw == OR_BC_PRESSURE || kw == OR_BC_WALL) { aW = 0.0; r -= nW / (dx * dx) * gr->c_p[0]; }
            }
            if (i == g + nx - 1) {
                if (ke == OR_BC_NEUMANN) { C += nE / (dx * dx); }
                else if (ke == OR_BC_DIRICHLET || ke == OR_BC_INLET || ke == OR_BC_OUTLET ||
                         ke == OR_BC_PRESSURE || ke == OR_BC_WALL) { r -= nE / (dx * dx) * gr->c_p[1]; }
            }
            if (j == g) {
                if (ks == OR_BC_NEUMANN) { aS = 0.0; C += nS / (dy * dy); }
                else if (ks == OR_BC_DIRICHLET || ks == OR_BC_INLET || ks == OR_BC_OUTLET ||
                         ks == OR_BC_PRESSURE || ks == OR_BC_WALL) { aS = 0.0; r -= nS / (dy * dy) * gr->c_p[2]; }
            }
            if (j == g + ny - 1) {
                if (kn == OR_BC_NEUMANN) { C += nN / (dy * dy); }
                else if (kn == OR_BC_DIRICHLET || kn == OR_BC_INLET || kn == OR_BC_OUTLET ||
                         kn == OR_BC_PRESSURE || kn == OR_BC_WALL) { r -= nN / (dy * dy) * gr->c_p[3]; }
            }
            if (pin) {
                /* the pinned cell (g,g) is removed; its neighbours see p = 0 there */
                if (i == g + 1 && j == g) aW = 0.0;
                else if (i == g && j == g + 1) aS = 0.0;
            }
            ax[c] = aW;
            ay[c] = aS;
            diag[c] = C;
            rhs[c] = r;
            if (pin && i == g && j == g) rhs[c] = 0.0;
        }
#undef LS_
#undef K_
#undef H_
}

/* (A p) for the matrix-free operator; interior cells only */
void or_stencil_apply(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                      const double* p, double* q) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const int64_t c = IDX(gr, i, j);
            q[c] = diag[c] * p[c] + ax[c] * p[IDX(gr, i - 1, j)] + ax[IDX(gr, i + 1, j)] * p[IDX(gr, i + 1, j)]
                 + ay[c] * p[IDX(gr, i, j - 1)] + ay[IDX(gr, i, j + 1)] * p[IDX(gr, i, j + 1)];
        }
}

/* ------------------------------------------------------------------ Krylov */
typedef struct {
    const or_grid* gr;
    const double *ax, *ay, *diag;
    int n;
    const int *rowptr, *cols;
    const double* vals;
} or_op;

/* iterate over unknowns in the reference's x-fastest order (mac2d.cpp:1479-1480) */
static double dotv(const or_op* op, const double* a, const double* b) {
    double s = 0.0;
    if (op->gr) {
        const or_grid* gr = op->gr;
        for (int j = gr->g; j < gr->ny + gr->g; j++)
            for (int i = gr->g; i < gr->nx + gr->g; i++) s += a[IDX(gr, i, j)] * b[IDX(gr, i, j)];
    } else {
        for (int k = 0; k < op->n; k++) s += a[k] * b[k];
    }
    return s;
}

static void op_apply(const or_op* op, const double* p, double* q) {
    if (op->gr) { or_stencil_apply(op->gr, op->ax, op->ay, op->diag, p, q); return; }
    for (int r = 0; r < op->n; r++) {
        double s = 0.0;
        for (int k = op->rowptr[r]; k < op->rowptr[r + 1]; k++) s += op->vals[k] * p[op->cols[k]];
        q[r] = s;
    }
}

static int64_t op_len(const or_op* op) { return op->gr ? or_size(op->gr) : op->n; }

/* poisson2d.cpp:285-401 (mode 0) or textbook Jacobi-PCG (mode 1) */
static int pcg_core(const or_op* op, const double* mdiag, const double* b, double* x, double tol,
                    int maxit, int mode, int* iters, double* relres) {
    const int64_t n = op_len(op);
    double* r = (double*)calloc((size_t)n, sizeof(double));
    double* z = (double*)calloc((size_t)n, sizeof(double));
    double* p = (double*)calloc((size_t)n, sizeof(double));
    double* q = (double*)calloc((size_t)n, sizeof(double));
    double* minv = (double*)calloc((size_t)n, sizeof(double));
    for (int64_t k = 0; k < n; k++) { x[k] = 0.0; r[k] = b[k]; minv[k] = (mdiag[k] != 0.0) ? 1.0 / mdiag[k] : 0.0; }
    if (op->gr) { /* only interior entries are unknowns */
        const or_grid* gr = op->gr;
        for (int64_t k = 0; k < n; k++) { r[k] = 0.0; minv[k] = 0.0; }
        for (int i = gr->g; i < gr->nx + gr->g; i++)
            for (int j = gr->g; j < gr->ny + gr->g; j++) {
                const int64_t c = IDX(gr, i, j);
                r[c] = b[c];
                minv[c] = 1.0 / mdiag[c];
            }
    }
    memcpy(p, r, sizeof(double) * (size_t)n);
    double bnorm = sqrt(dotv(op, b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    double rnorm = sqrt(dotv(op, r, r));
    int conv = (rnorm / bnorm <= tol);
    double rho = 1.0, rho1 = 1.0;
    int it = 0;
    while (it < maxit && !conv) {
        for (int64_t k = 0; k < n; k++) z[k] = minv[k] * r[k];
        rho = dotv(op, r, z);
        if (mode == 0) {
            if (it == 1) memcpy(p, z, sizeof(double) * (size_t)n);
            else { const double beta = rho / rho1; for (int64_t k = 0; k < n; k++) p[k] = 1.0 * z[k] + beta * p[k]; }
        } else {
            if (it == 0) memcpy(p, z, sizeof(double) * (size_t)n);
            else { const double beta = rho / rho1; for (int64_t k = 0; k < n; k++) p[k] = z[k] + beta * p[k]; }
        }
        op_apply(op, p, q);
        const double pq = dotv(op, p, q);
        const double alpha = (mode == 0) ? rho / (pq + 1.0e-6 * 1.0e-6) : rho / pq;
        for (int64_t k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] += -alpha * q[k]; }
        rnorm = sqrt(dotv(op, r, r));
        if (rnorm / bnorm <= tol) conv = 1;
        rho1 = rho;
        it++;
    }
    *iters = it;
    *relres = rnorm / bnorm;
    free(r); free(z); free(p); free(q); free(minv);
    return conv ? 0 : 1;
}

int or_pcg_stencil(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                   const double* b, double* x, double tol, int maxit, int mode,
                   int* iters, double* relres) {
    or_op op = {gr, ax, ay, diag, 0, 0, 0, 0};
    return pcg_core(&op, diag, b, x, tol, maxit, mode, iters, relres);
}

int or_pcg_csr(int n, const int* rowptr, const int* cols, const double* vals,
               const double* mdiag, const double* b, double* x, double tol, int maxit,
               int mode, int* iters, double* relres) {
    or_op op = {0, 0, 0, 0, n, rowptr, cols, vals};
    return pcg_core(&op, mdiag, b, x, tol, maxit, mode, iters, relres);
}

/* Jacobi-preconditioned BiCGStab (van der Vorst 1992; IML++ form), the corrected
 * counterpart of poisson2d.cpp:403-575 (rho2 updated every iteration, pin-aware). */
int or_bicgstab_stencil(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                        const double* b, double* x, double tol, int maxit,
                        int* iters, double* relres) {
    or_op op = {gr, ax, ay, diag, 0, 0, 0, 0};
    const int64_t n = or_size(gr);
    double *r = calloc((size_t)n, 8), *rt = calloc((size_t)n, 8), *p = calloc((size_t)n, 8);
    double *v = calloc((size_t)n, 8), *ph = calloc((size_t)n, 8), *s = calloc((size_t)n, 8);
    double *sh = calloc((size_t)n, 8), *t = calloc((size_t)n, 8), *minv = calloc((size_t)n, 8);
    for (int64_t k = 0; k < n; k++) x[k] = 0.0;
    for (int i = gr->g; i < gr->nx + gr->g; i++)
        for (int j = gr->g; j < gr->ny + gr->g; j++) {
            const int64_t c = IDX(gr, i, j);
            r[c] = b[c]; rt[c] = b[c]; minv[c] = 1.0 / diag[c];
        }
    double bnorm = sqrt(dotv(&op, b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    double rnorm = sqrt(dotv(&op, r, r));
    int conv = (rnorm / bnorm <= tol);
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    while (it < maxit && !conv) {
        const double rho = dotv(&op, rt, r);
        if (rho == 0.0) break;
        if (it == 0) memcpy(p, r, sizeof(double) * (size_t)n);
        else {
            const double beta = (rho / rho_old) * (alpha / omega);
            for (int64_t k = 0; k < n; k++) p[k] = r[k] + beta * (p[k] - omega * v[k]);
        }
        for (int64_t k = 0; k < n; k++) ph[k] = minv[k] * p[k];
        op_apply(&op, ph, v);
        alpha = rho / dotv(&op, rt, v);
        for (int64_t k = 0; k < n; k++) s[k] = r[k] - alpha * v[k];
        const double snorm = sqrt(dotv(&op, s, s));
        if (snorm / bnorm <= tol) {
            for (int64_t k = 0; k < n; k++) x[k] += alpha * ph[k];
            rnorm = snorm; conv = 1; it++;
            break;
        }
        for (int64_t k = 0; k < n; k++) sh[k] = minv[k] * s[k];
        op_apply(&op, sh, t);
        omega = dotv(&op, t, s) / dotv(&op, t, t);
        for (int64_t k = 0; k < n; k++) { x[k] += alpha * ph[k] + omega * sh[k]; r[k] = s[k] - omega * t[k]; }
        rnorm = sqrt(dotv(&op, r, r));
        if (rnorm / bnorm <= tol) conv = 1;
        rho_old = rho;
        it++;
    }
    *iters = it;
    *relres = rnorm / bnorm;
    free(r); free(rt); free(p); free(v); free(ph); free(s); free(sh); free(t); free(minv);
    return conv ? 0 : 1;
}

/* Level-set run diagnostics.  Mass count and error norms as the reference's level-set drivers
 * compute them (tests/LargeTests/test_levelset2d_large.cpp:94-102 mass = #(phi > 0) dx dy;
 * :106-127 err_inf = max |phi - ref|, err_l1 = sum |phi - ref|, err_l2 = sqrt(sum (phi - ref)^2)
 * / (nx ny), mass loss = |#phi - #ref| / #ref * 100).  No reference counterpart: centroid of the
 * positive cells' centres, interface length sum delta(phi) |grad phi| dx dy with the cosine delta
 * of half-width 1.5 min(dx, dy) and central differences, circularity 2 sqrt(pi mass) / length. */
void or_ls_diagnostics(const or_grid* gr, const double* ls, const double* ref, double base_x, double base_y,
                       double* out) {
    const int nx = gr->nx, ny = gr->ny, g = gr->g;
    const double dx = gr->dx, dy = gr->dy, eps = 1.5 * smin(dx, dy);
    double cells = 0.0, sx = 0.0, sy = 0.0, len = 0.0, l1 = 0.0, l2 = 0.0, rc = 0.0, linf = 0.0;
    for (int i = g; i < nx + g; i++)
        for (int j = g; j < ny + g; j++) {
            const double p = ls[IDX(gr, i, j)];
            if (p > 0.0) {
                cells += 1.0;
                sx += base_x + ((double)(i - g) + 0.5) * dx;
                sy += base_y + ((double)(j - g) + 0.5) * dy;
            }
            if (fabs(p) < eps) {
                const double gx = (ls[IDX(gr, i + 1, j)] - ls[IDX(gr, i - 1, j)]) / (2.0 * dx);
                const double gy = (ls[IDX(gr, i, j + 1)] - ls[IDX(gr, i, j - 1)]) / (2.0 * dy);
                len += (1.0 + cos(M_PI * p / eps)) / (2.0 * eps) * sqrt(gx * gx + gy * gy);
            }
            if (ref) {
                const double r = ref[IDX(gr, i, j)], e = fabs(p - r);
                l1 += e;
                l2 += e * e;
                if (r > 0.0) rc += 1.0;
                if (e > linf) linf = e;
            }
        }
    const double mass = cells * dx * dy, length = len * dx * dy;
    out[0] = cells;
    out[1] = mass;
    out[2] = cells > 0.0 ? sx / cells : 0.0;
    out[3] = cells > 0.0 ? sy / cells : 0.0;
    out[4] = length;
    out[5] = length > 0.0 ? 2.0 * sqrt(M_PI * mass) / length : 0.0;
    out[6] = ref ? linf : 0.0;
    out[7] = ref ? l1 : 0.0;
    out[8] = ref ? sqrt(l2) / ((double)nx * ny) : 0.0;
    out[9] = ref ? rc : 0.0;
    out[10] = (ref && rc > 0.0) ? fabs(cells - rc) / rc * 100.0 : 0.0;
}

/* PoissonSolver2D::GS_2FUniform_2D (src/poisson2d.cpp:7-88): Gauss-Seidel sweeps in row order
 * (x[c] for c < r already updated in the sweep), then x = x_old + 1.2 (x - x_old) over the whole
 * vector; stop when sqrt(sum (x - x_old)^2 / (sum x_old^2 + 1e-100)) <= 1e-6 or after maxit
 * sweeps.  x holds the start vector.  Unlike the reference, a row without a diagonal entry
 * divides by 0 instead of by the previous row's diagonal. */
int or_gs_csr(int n, const int* rowptr, const int* cols, const double* vals, const double* b,
              double* x, int maxit, int* iters, double* rel_change) {
    double* old = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!old) return -1;
    for (int i = 0; i < n; i++) old[i] = x[i];
    const double eps = 1.0e-6, omega = 1.2;
    double err = 1.0;
    int it = 0;
    while (err > eps && it < maxit) {
        for (int i = 0; i < n; i++) {
            double t = b[i], diag = 0.0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
                const int c = cols[p];
                if (c == i) diag = vals[p];
                else t -= vals[p] * x[c];
            }
            x[i] = t / diag;
        }
        for (int i = 0; i < n; i++) x[i] = old[i] + omega * (x[i] - old[i]);
        double up = 0.0, lo = 0.0;
        for (int i = 0; i < n; i++) {
            up += (x[i] - old[i]) * (x[i] - old[i]);
            lo += old[i] * old[i];
            old[i] = x[i];
        }
        err = sqrt(up / (lo + 1.0e-100));
        it++;
    }
    free(old);
    *iters = it;
    *rel_change = err;
    return 0;
}

/* ------------------------------------------------------------------ axisymmetric (r, z) */
/* MACSolver2DAxisym::SolvePoisson assembly, src/mac2dAxisym.cpp:1564-1870 (x = r, y = z, r of
 * cell i = base_r + (i + 0.5 - g) dr).  Stored like or_gfm_assemble_stencil: ax / ay = CSR value
 * of the W / S neighbour; ae = the row's own E value (the reference's, for checking that it
 * equals the east neighbour's ax up to rounding). */
void or_gfm_assemble_axisym(const or_grid* gr, const or_fluid* fl, double base_r, const double* ls,
                            const double* kappa, const double* H, const double* div, double* ax, double* ay,
                            double* ae, double* diag, double* rhs) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy, dt = fl->dt, sigma = fl->sigma;
    for (int64_t k = 0; k < or_size(gr); k++) { ax[k] = 0.0; ay[k] = 0.0; ae[k] = 0.0; diag[k] = 0.0; rhs[k] = 0.0; }
    int pin = 1;
    for (int s = 0; s < 4; s++) {
        const int k = gr->bc_p[s];
        if (!(k == OR_BC_NEUMANN || k == OR_BC_AXISYM || k == OR_BC_WALL)) pin = 0;
    }
    for (int j = g; j < nz + g; j++)
        for (int i = g; i < nr + g; i++) {
            const int64_t c = IDX(gr, i, j);
            const double lsW = ls[IDX(gr, i - 1, j)], lsE = ls[IDX(gr, i + 1, j)], lsM = ls[c];
            const double lsS = ls[IDX(gr, i, j - 1)], lsN = ls[IDX(gr, i, j + 1)];
            const double rPW = base_r + (i - 0.5 - g) * dr, rPM = base_r + (i + 0.5 - g) * dr;
            const double rPE = base_r + (i + 1.5 - g) * dr;
            const double rPWHalf = base_r + (i - g) * dr, rPEHalf = base_r + (i + 1.0 - g) * dr;
            double thetaH = 0.0, theta = 0.0, rEff = 0.0, rhoEff, kappaEff;
            /* W face */
            if (lsW >= 0.0 && lsM >= 0.0) { thetaH = 1.0; theta = 0.0; rEff = rPWHalf; }
            else if (lsW < 0.0 && lsM < 0.0) { thetaH = 0.0; theta = 0.0; rEff = rPWHalf; }
            else if (lsW >= 0.0 && lsM < 0.0) {
                thetaH = fabs(lsW) / (fabs(lsW) + fabs(lsM));
                theta = fabs(lsW) / (fabs(lsW) + fabs(lsM));
                rEff = smax(rPM * theta + rPW * (1.0 - theta), 0.0);
            } else {
                thetaH = fabs(lsM) / (fabs(lsW) + fabs(lsM));
                theta = fabs(lsW) / (fabs(lsW) + fabs(lsM));
                rEff = smax(rPM * theta + rPW * (1.0 - theta), 0.0);
            }
            rhoEff = fl->rho_h * thetaH + fl->rho_l * (1.0 - thetaH);
            kappaEff = kappa[c] * theta + kappa[IDX(gr, i - 1, j)] * (1.0 - theta);
            double pW = rEff / rhoEff;
            double FW = -dt * (-sigma * kappaEff) * (H[c] - H[IDX(gr, i - 1, j)]);
            FW *= pW / (dr * dr);
            /* E face */
            if (lsM >= 0.0 && lsE >= 0.0) { thetaH = 1.0; theta = 0.0; rEff = rPEHalf; }
            else if (lsM < 0.0 && lsE < 0.0) { thetaH = 0.0; theta = 0.0; rEff = rPEHalf; }
            else if (lsM >= 0.0 && lsE < 0.0) {
                thetaH = fabs(lsM) / (fabs(lsM) + fabs(lsE));
                theta = fabs(lsE) / (fabs(lsM) + fabs(lsE));
                rEff = rPM * theta + rPE * (1.0 - theta);
            } else {
                thetaH = fabs(lsE) / (fabs(lsM) + fabs(lsE));
                theta = fabs(lsE) / (fabs(lsM) + fabs(lsE));
                rEff = rPM * theta + rPE * (1.0 - theta);
            }
            rhoEff = fl->rho_h * thetaH + fl->rho_l * (1.0 - thetaH);
            kappaEff = kappa[c] * theta + kappa[IDX(gr, i + 1, j)] * (1.0 - theta);
            double pE = rEff / rhoEff;
            double FE = dt * (-sigma * kappaEff) * (H[IDX(gr, i + 1, j)] - H[c]);
            FE *= pE / (dr * dr);
            /* S face */
            face_theta(lsS, lsM, 1, &thetaH, &theta);
            rhoEff = fl->rho_h * thetaH + fl->rho_l * (1.0 - thetaH);
            kappaEff = kappa[c] * theta + kappa[IDX(gr, i, j - 1)] * (1.0 - theta);
            double pS = rPM / rhoEff;
            double FS = -dt * (-sigma * kappaEff) * (H[c] - H[IDX(gr, i, j - 1)]);
            FS *= pS / (dz * dz);
            /* N face */
            face_theta(lsM, lsN, 0, &thetaH, &theta);
            rhoEff = fl->rho_h * thetaH + fl->rho_l * (1.0 - thetaH);
            kappaEff = kappa[c] * theta + kappa[IDX(gr, i, j + 1)] * (1.0 - theta);
            double pN = rPM / rhoEff;
            double FN = dt * (-sigma * kappaEff) * (H[IDX(gr, i, j + 1)] - H[c]);
            FN *= pN / (dz * dz);
            rhs[c] -= FW + FE + FS + FN;
            rhs[c] -= div[c];
            /* matrix row, :1756-1870 (ps is zeroed before the assembly) */
            pS *= -1.0; pW *= -1.0; pE *= -1.0; pN *= -1.0;
            double aS = pS / (dz * dz), aW = pW / (dr * dr), aE = pE / (dr * dr);
            double C = -(pW + pE) / (dr * dr) - (pS + pN) / (dz * dz);
            for (int s = 0; s < 4; s++) {
                const int k = gr->bc_p[s];
                const int at = (s == 0 && i == g) || (s == 1 && i == g + nr - 1) || (s == 2 && j == g) ||
                               (s == 3 && j == g + nz - 1);
                if (!at) continue;
                const double p = s == 0 ? pW : s == 1 ? pE : s == 2 ? pS : pN;
                const double h2 = s < 2 ? dr * dr : dz * dz;
                if (k == OR_BC_NEUMANN || k == OR_BC_AXISYM || k == OR_BC_WALL) {
                    if (s == 0) aW = 0.0;
                    if (s == 1) aE = 0.0;
                    if (s == 2) aS = 0.0;
                    C += p / h2;
                } else if (k == OR_BC_DIRICHLET || k == OR_BC_INLET || k == OR_BC_OUTLET || k == OR_BC_PRESSURE) {
                    if (s == 0) aW = 0.0;
                    if (s == 1) aE = 0.0;
                    if (s == 2) aS = 0.0;
                    rhs[c] -= p / h2 * (-0.0 + 2.0 * gr->c_p[s]);
                }
            }
            if (pin) {
                if (i == g + 1 && j == g) aW = 0.0;
                else if (i == g && j == g + 1) aS = 0.0;
                if (i == g && j == g) rhs[c] = 0.0;
            }
            ax[c] = aW;
            ay[c] = aS;
            ae[c] = aE;
            diag[c] = C;
        }
}

/* GetDivergence4Poisson (for_poisson = 1, src/mac2dAxisym.cpp:1974-1989) and GetDivergence
 * (:1991-2008) */
void or_divergence_axisym(const or_grid* gr, double base_r, const double* u, const double* v, double* div,
                          int for_poisson) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) div[k] = 0.0;
    for (int i = g; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const double rPM = base_r + (i + 0.5 - g) * dr;
            const double rPEHalf = base_r + (i + 1.0 - g) * dr, rPWHalf = base_r + (i - g) * dr;
            if (for_poisson)
                div[IDX(gr, i, j)] = (rPEHalf * u[IDX(gr, i + 1, j)] - rPWHalf * u[IDX(gr, i, j)]) / dr
                                   + rPM * (v[IDX(gr, i, j + 1)] - v[IDX(gr, i, j)]) / dz;
            else
                div[IDX(gr, i, j)] = 0.5 * (u[IDX(gr, i, j)] + u[IDX(gr, i + 1, j)]) / rPM
                                   + (u[IDX(gr, i + 1, j)] - u[IDX(gr, i, j)]) / dr
                                   + (v[IDX(gr, i, j + 1)] - v[IDX(gr, i, j)]) / dz;
        }
}

/* MACSolver2DAxisym::UpdateDt, src/mac2dAxisym.cpp:2094-2128 */
double or_update_dt_axisym(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                           const double* kappa) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy;
    double uA = 0.0, vA = 0.0, kA = 0.0;
    for (int i = g; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            uA = smax(uA, fabs((u[IDX(gr, i + 1, j)] * u[IDX(gr, i, j)]) * 0.5));
            vA = smax(vA, fabs((v[IDX(gr, i, j + 1)] * v[IDX(gr, i, j)]) * 0.5));
            kA = smax(kA, fabs(kappa[IDX(gr, i, j)]));
        }
    const double C = uA / dr + vA / dz;
    const double G = smax(sqrt(fabs(fl->g) / dz), sqrt(fabs(fl->g) / dr));
    const double S = sqrt((fl->sigma * kA) / (smin(fl->rho_h, fl->rho_l) * pow(smin(dr, dz), 2.0)));
    double dt = 1.7976931348623157e308;
    dt = smin(dt, fl->cfl / (0.5 * (C + sqrt(pow(C, 2.0) + 4.0 * pow(G, 2.0) + 4.0 * pow(S, 2.0)))));
    return dt;
}

/* MACSolver2DAxisym::SolveTDMA, src/mac2dAxisym.cpp:2255-2304 (n >= 2 unknowns; d <- x) */
void or_tdma(double* a, double* b, double* c, double* d, int n) {
    n--;
    c[0] /= b[0];
    d[0] /= b[0];
    for (int i = 1; i < n; i++) {
        c[i] /= b[i] - a[i] * c[i - 1];
        d[i] = (d[i] - a[i] * d[i - 1]) / (b[i] - a[i] * c[i - 1]);
    }
    d[n] = (d[n] - a[n] * d[n - 1]) / (b[n] - a[n] * c[n - 1]);
    for (int i = n; i-- > 0;) d[i] -= c[i] * d[i + 1];
}

/* MACSolver2DAxisym::UpdateKappa, src/mac2dAxisym.cpp:222-262 (interior cells written) */
void or_update_kappa_axisym(const or_grid* gr, double base_r, const double* ls, double* kappa) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy, eps = 1.0e-100;
    double* dR = zalloc(gr);
    double* dZ = zalloc(gr);
    double* S = zalloc(gr);
    for (int i = g; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const int64_t c = IDX(gr, i, j);
            dR[c] = (ls[IDX(gr, i + 1, j)] - ls[IDX(gr, i - 1, j)]) / (2.0 * dr);
            dZ[c] = (ls[IDX(gr, i, j + 1)] - ls[IDX(gr, i, j - 1)]) / (2.0 * dz);
            if (i == g) dR[c] = (ls[IDX(gr, i + 1, j)] - ls[c]) / dr;
            S[c] = sqrt(dR[c] * dR[c] + dZ[c] * dZ[c] + eps);
        }
    for (int i = g; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const int64_t c = IDX(gr, i, j);
            const double rPM = base_r + (i + 0.5 - g) * dr;
            double k = ((ls[IDX(gr, i - 1, j)] - 2.0 * ls[c] + ls[IDX(gr, i + 1, j)]) / (dr * dr) * dZ[c] * dZ[c]
                        - 2.0 * (dZ[IDX(gr, i + 1, j)] - dZ[IDX(gr, i - 1, j)]) / (2.0 * dr) * dR[c] * dZ[c]
                        + (ls[IDX(gr, i, j - 1)] - 2.0 * ls[c] + ls[IDX(gr, i, j + 1)]) / (dz * dz) * dR[c] * dR[c]
                        + dR[c] * pow(S[c], 2.0) / rPM) / pow(S[c], 3.0);
            kappa[c] = ((k > 0) - (k < 0)) * fabs(smin(fabs(k), 1.0 / smin(dr, dz)));
        }
    free(dR);
    free(dZ);
    free(S);
}

/* ------------------------------------------------------------------ axisymmetric momentum */
/* rho_eff of the axisymmetric viscous terms and line sweeps: rho_H thetaH + rho_L (1 - thetaH),
 * thetaH the heavy share of the segment between the level-set samples a and b (1 or 0 when both
 * lie in one phase); the branch order of src/mac2dAxisym.cpp:687-737, 1024-1045 */
static double rho_eff_axi(const or_fluid* fl, double a, double b) {
    double thH = 0.0;
    if (a >= 0.0 && b >= 0.0) thH = 1.0;
    else if (a <= 0.0 && b <= 0.0) thH = 0.0;
    else if (a >= 0.0 && b < 0.0) thH = fabs(a) / (fabs(a) + fabs(b));
    else if (a < 0.0 && b >= 0.0) thH = fabs(b) / (fabs(a) + fabs(b));
    return fl->rho_h * thH + fl->rho_l * (1.0 - thH);
}

/* mu = mu_L + (mu_H - mu_L) H over the whole padded array (the axisymmetric solver's loops include
 * the ghosts, src/mac2dAxisym.cpp:607-609) */
static double* mu_full_axi(const or_grid* gr, const or_fluid* fl, const double* H) {
    double* mu = zalloc(gr);
    for (int64_t k = 0; k < or_size(gr); k++) mu[k] = fl->mu_l + (fl->mu_h - fl->mu_l) * H[k];
    return mu;
}

/* MACSolver2DAxisym::AddExternalViscosityFU, src/mac2dAxisym.cpp:590-753: the cross term
 * d/dz (mu dv/dr) only, over the arithmetic-mean density of the S-N segment */
void or_viscosity_u_axisym(const or_grid* gr, const or_fluid* fl, const double* v, const double* ls,
                           const double* H, double* dU) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) dU[k] = 0.0;
    if (fl->re <= 0.0) return;
    double* mu = mu_full_axi(gr, fl, H);
#define A_(a, i, j) (a)[IDX(gr, i, j)]
    for (int i = g + 1; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const double vW = A_(v, i - 1, j), vW_N = A_(v, i - 1, j + 1), vM = A_(v, i, j), vN = A_(v, i, j + 1);
            const double lsUNHalf = 0.25 * (A_(ls, i, j) + A_(ls, i - 1, j) + A_(ls, i - 1, j + 1) + A_(ls, i, j + 1));
            const double lsUSHalf = 0.25 * (A_(ls, i, j) + A_(ls, i - 1, j) + A_(ls, i - 1, j - 1) + A_(ls, i, j - 1));
            const double muUNHalf = 0.25 * (A_(mu, i, j) + A_(mu, i - 1, j) + A_(mu, i - 1, j + 1) + A_(mu, i, j + 1));
            const double muUSHalf = 0.25 * (A_(mu, i, j) + A_(mu, i - 1, j) + A_(mu, i - 1, j - 1) + A_(mu, i, j - 1));
            const double rhoSN = rho_eff_axi(fl, lsUSHalf, lsUNHalf);
            const double visZ = (muUNHalf * (vN - vW_N) / dr - muUSHalf * (vM - vW) / dr) / dz;
            A_(dU, i, j) = visZ / rhoSN;
        }
    free(mu);
}

/* MACSolver2DAxisym::AddExternalViscosityFV, src/mac2dAxisym.cpp:755-910: d/dr (mu du/dz) over the
 * W-E segment's density plus mu / r d(u*)/dz from the new U-face velocity (r without the base
 * radius, as the reference computes it) over the S-N density */
void or_viscosity_v_axisym(const or_grid* gr, const or_fluid* fl, const double* u, const double* uhat,
                           const double* ls, const double* H, double* dV) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy;
    for (int64_t k = 0; k < or_size(gr); k++) dV[k] = 0.0;
    if (fl->re <= 0.0) return;
    double* mu = mu_full_axi(gr, fl, H);
    for (int i = g; i < nr + g; i++)
        for (int j = g + 1; j < nz + g; j++) {
            const double rPM = (i + 0.5 - g) * dr;
            const double uM = A_(u, i, j), uE = A_(u, i + 1, j), uS = A_(u, i, j - 1), uS_E = A_(u, i + 1, j - 1);
            const double lsVWHalf = 0.25 * (A_(ls, i, j) + A_(ls, i - 1, j) + A_(ls, i - 1, j - 1) + A_(ls, i, j - 1));
            const double lsVEHalf = 0.25 * (A_(ls, i, j) + A_(ls, i + 1, j) + A_(ls, i + 1, j - 1) + A_(ls, i, j - 1));
            const double lsM = A_(ls, i, j), lsS = A_(ls, i, j - 1);
            const double muVWHalf = 0.25 * (A_(mu, i, j) + A_(mu, i - 1, j) + A_(mu, i - 1, j - 1) + A_(mu, i, j - 1));
            const double muVEHalf = 0.25 * (A_(mu, i, j) + A_(mu, i + 1, j) + A_(mu, i + 1, j - 1) + A_(mu, i, j - 1));
            const double rhoWE = rho_eff_axi(fl, lsVWHalf, lsVEHalf);
            const double rhoSN = rho_eff_axi(fl, lsS, lsM);
            const double visR = (muVEHalf * (uE - uS_E) / dz - muVWHalf * (uM - uS) / dz) / dr;
            const double visZ = 0.5 * (A_(mu, i, j) + A_(mu, i, j - 1)) / rPM
                * (0.5 * (A_(uhat, i, j) + A_(uhat, i + 1, j)) - 0.5 * (A_(uhat, i, j - 1) + A_(uhat, i + 1, j - 1))) / dz;
            A_(dV, i, j) = visR / rhoWE + visZ / rhoSN;
        }
    free(mu);
}

/* MACSolver2DAxisym::GetRHSU (src/mac2dAxisym.cpp:264-292): u + dt (-convection + gravity +
 * viscosity) on the interior U faces, 0 elsewhere, then the U ghost fill.  The convection is the
 * planar AddConvectionFU (:324-405 = src/mac2d.cpp:308-387) with x = r, y = z; the gravity is
 * AddGravityFU (:912-929), -g on the interior U faces when it acts along r. */
void or_rhs_u_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u, const double* v,
                     const double* H, double* rhs) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    double* cU = zalloc(gr);
    double* vU = zalloc(gr);
    or_convection_u(gr, u, v, cU);
    or_viscosity_u_axisym(gr, fl, v, ls, H, vU);
    const double gU = gravity_active(fl, 0) ? -fl->g : 0.0;
    for (int64_t k = 0; k < or_size(gr); k++) rhs[k] = 0.0;
    for (int i = g + 1; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const int64_t c = IDX(gr, i, j);
            rhs[c] = u[c] + fl->dt * (-cU[c] + gU + vU[c]);
        }
    or_bc_u(gr, rhs);
    free(cU);
    free(vU);
}

/* MACSolver2DAxisym::GetRHSV (src/mac2dAxisym.cpp:294-322): v + dt (-convection + gravity +
 * viscosity) on the interior V faces, then the V ghost fill.  The gravity array is AddGravityFU's
 * (the U grid: -g where it acts along r and i >= g + 1), as the reference calls it; gravity_fix
 * selects AddGravityFV (-g where it acts along z) instead. */
void or_rhs_v_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u, const double* uhat,
                     const double* v, const double* H, double* rhs) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    double* cV = zalloc(gr);
    double* vV = zalloc(gr);
    or_convection_v(gr, u, v, cV);
    or_viscosity_v_axisym(gr, fl, u, uhat, ls, H, vV);
    for (int64_t k = 0; k < or_size(gr); k++) rhs[k] = 0.0;
    for (int i = g; i < nr + g; i++)
        for (int j = g + 1; j < nz + g; j++) {
            const int64_t c = IDX(gr, i, j);
            double gv;
            if (fl->gravity_fix) gv = gravity_active(fl, 1) ? -fl->g : 0.0;
            else gv = (gravity_active(fl, 0) && i >= g + 1) ? -fl->g : 0.0;
            rhs[c] = v[c] + fl->dt * (-cV[c] + gv + vV[c]);
        }
    or_bc_v(gr, rhs);
    free(cV);
    free(vV);
}

/* the W / E rows of the U line sweep along r: Neumann, Outlet fold the dropped coupling into the
 * diagonal; Dirichlet, Axisym, Inlet, Wall move coupling x constant to the right-hand side; Periodic
 * moves coupling x the other end's right-hand side value (src/mac2dAxisym.cpp:1052-1078, 1141-1169, 1336-1364,
 * 1423-1449) */
static int sweep_fold(int k, int axis_is_neumann) {
    if (k == OR_BC_NEUMANN || k == OR_BC_OUTLET) return 1;
    if (k == OR_BC_AXISYM) return axis_is_neumann ? 1 : 2;
    if (k == OR_BC_DIRICHLET || k == OR_BC_INLET || k == OR_BC_WALL) return 2;
    if (k == OR_BC_PERIODIC) return 3;
    return 0;
}

/* MACSolver2DAxisym::GetUHat, src/mac2dAxisym.cpp:950-1229: the semi-implicit viscous update of
 * the U faces by a line sweep along r (2 r mu weights over r_{i+1/2} at the face radius), a line
 * sweep along z, and the division by 1 + dt / rho (mu_W + mu_M) / r^2.  Radii without the base
 * radius, as the reference computes them.  Interior U faces written, 0 elsewhere. */
void or_uhat_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* rhs, const double* H,
                    double* uhat) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy, dt = fl->dt;
    const int nu = nr - 1;
    double* mu = mu_full_axi(gr, fl, H);
    double* us3 = (double*)calloc((size_t)nu * nz, sizeof(double));
    double* us2 = (double*)calloc((size_t)nu * nz, sizeof(double));
    double* us1 = (double*)calloc((size_t)nu * nz, sizeof(double));
    const int n = nu > nz ? nu : nz;
    double* a = (double*)calloc((size_t)n, sizeof(double));
    double* b = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)calloc((size_t)n, sizeof(double));
    double* d = (double*)calloc((size_t)n, sizeof(double));
    for (int i = g + 1; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) us3[(j - g) + nz * (i - 1 - g)] = A_(rhs, i, j);
    const int fW = sweep_fold(gr->bc_u[0], 0), fE = sweep_fold(gr->bc_u[1], 0);
    const int fS = sweep_fold(gr->bc_u[2], 1), fN = sweep_fold(gr->bc_u[3], 1);
    for (int j = 0; j < nz; j++) {
        for (int i = 0; i < nu; i++) d[i] = us3[j + nz * i];
        for (int i = 0; i < nu; i++) {
            const double rUM = (i + 1.0) * dr, rPW = (i + 0.5) * dr, rPM = (i + 1.5) * dr;
            const double lsW = A_(ls, i + g, j + g), lsM = A_(ls, i + g + 1, j + g);
            const double muW = A_(mu, i + g, j + g), muM = A_(mu, i + g + 1, j + g);
            const double rho = rho_eff_axi(fl, lsW, lsM);
            a[i] = -dt / (rho * rUM) * (2.0 * rPW * muW) / (dr * dr);
            b[i] = 1.0 + dt / (rho * rUM) * 2.0 * (rPW * muW + rPM * muM) / (dr * dr);
            c[i] = -dt / (rho * rUM) * (2.0 * rPM * muM) / (dr * dr);
            if (i == 0) {
                if (fW == 1) { b[i] += a[i]; a[i] = 0.0; }
                else if (fW == 2) { d[i] -= a[i] * gr->c_u[0]; a[i] = 0.0; }
                else if (fW == 3) { d[i] -= a[i] * us3[j + nz * (nu - 1)]; a[i] = 0.0; }
            }
            if (i == nu - 1) {
                if (fE == 1) { b[i] += c[i]; c[i] = 0.0; }
                else if (fE == 2) { d[i] -= c[i] * gr->c_u[1]; c[i] = 0.0; }
                else if (fE == 3) { d[i] -= c[i] * us3[j + nz * 0]; c[i] = 0.0; }
            }
        }
        or_tdma(a, b, c, d, nu);
        for (int i = 0; i < nu; i++) us2[j + nz * i] = d[i];
    }
    for (int i = 0; i < nu; i++) {
        for (int j = 0; j < nz; j++) d[j] = us2[j + nz * i];
        for (int j = 0; j < nz; j++) {
            const int I = i + g + 1, J = j + g;  /* the U face */
            const double lsUSHalf = 0.25 * (A_(ls, I, J) + A_(ls, I - 1, J) + A_(ls, I - 1, J - 1) + A_(ls, I, J - 1));
            const double lsUNHalf = 0.25 * (A_(ls, I, J) + A_(ls, I - 1, J) + A_(ls, I - 1, J + 1) + A_(ls, I, J + 1));
            const double muUSHalf = 0.25 * (A_(mu, I, J) + A_(mu, I - 1, J) + A_(mu, I - 1, J - 1) + A_(mu, I, J - 1));
            const double muUNHalf = 0.25 * (A_(mu, I, J) + A_(mu, I - 1, J) + A_(mu, I - 1, J + 1) + A_(mu, I, J + 1));
            const double rho = rho_eff_axi(fl, lsUSHalf, lsUNHalf);
            a[j] = -dt / rho * muUSHalf / (dz * dz);
            b[j] = 1.0 + dt / rho * (muUSHalf + muUNHalf) / (dz * dz);
            c[j] = -dt / rho * muUNHalf / (dz * dz);
            if (j == 0) {
                if (fS == 1) { b[j] += a[j]; a[j] = 0.0; }
                else if (fS == 2) { d[j] -= a[j] * 2.0 * gr->c_u[2]; a[j] = 0.0; }
                else if (fS == 3) { d[j] -= a[j] * us2[(nz - 1) + nz * i]; a[j] = 0.0; }
            }
            if (j == nz - 1) {
                if (fN == 1) { b[j] += c[j]; c[j] = 0.0; }
                else if (fN == 2) { d[j] -= c[j] * 2.0 * gr->c_u[3]; c[j] = 0.0; }
                else if (fN == 3) { d[j] -= c[j] * us2[0 + nz * i]; c[j] = 0.0; }
            }
        }
        or_tdma(a, b, c, d, nz);
        for (int j = 0; j < nz; j++) us1[j + nz * i] = d[j];
    }
    for (int64_t k = 0; k < or_size(gr); k++) uhat[k] = 0.0;
    for (int i = g + 1; i < nr + g; i++)
        for (int j = g; j < nz + g; j++) {
            const double rUM = (i - g) * dr;
            const double rho = rho_eff_axi(fl, A_(ls, i - 1, j), A_(ls, i, j));
            A_(uhat, i, j) = us1[(j - g) + nz * (i - g - 1)]
                / (1.0 + dt / rho * (A_(mu, i - 1, j) + A_(mu, i, j)) / (rUM * rUM));
        }
    free(mu); free(us3); free(us2); free(us1); free(a); free(b); free(c); free(d);
}

/* MACSolver2DAxisym::GetVHat, src/mac2dAxisym.cpp:1231-1478: the semi-implicit viscous update of
 * the V faces by a line sweep along r (r mu weights at the half radii over the cell radius) and a
 * line sweep along z (2 mu weights).  Here Axisym folds like Neumann on W / E and acts like
 * Dirichlet on S / N; Dirichlet-family sides add twice the constant along r, once along z.
 * Interior V faces written, 0 elsewhere. */
void or_vhat_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* rhs, const double* H,
                    double* vhat) {
    const int nr = gr->nx, nz = gr->ny, g = gr->g;
    const double dr = gr->dx, dz = gr->dy, dt = fl->dt;
    const int nv = nz - 1;
    double* mu = mu_full_axi(gr, fl, H);
    double* vs2 = (double*)calloc((size_t)nr * nv, sizeof(double));
    double* vs1 = (double*)calloc((size_t)nr * nv, sizeof(double));
    double* vs0 = (double*)calloc((size_t)nr * nv, sizeof(double));
    const int n = nr > nz ? nr : nz;
    double* a = (double*)calloc((size_t)n, sizeof(double));
    double* b = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)calloc((size_t)n, sizeof(double));
    double* d = (double*)calloc((size_t)n, sizeof(double));
    for (int i = g; i < nr + g; i++)
        for (int j = g + 1; j < nz + g; j++) vs2[(j - g - 1) + nv * (i - g)] = A_(rhs, i, j);
    const int fW = sweep_fold(gr->bc_v[0], 1), fE = sweep_fold(gr->bc_v[1], 1);
    const int fS = sweep_fold(gr->bc_v[2], 0), fN = sweep_fold(gr->bc_v[3], 0);
    for (int j = 0; j < nv; j++) {
        for (int i = 0; i < nr; i++) d[i] = vs2[j + nv * i];
        for (int i = 0; i < nr; i++) {
            const double rVM = (i + 0.5) * dr, rVWHalf = i * dr, rVEHalf = (i + 1.0) * dr;
            const int I = i + g, J = j + g + 1;  /* the V face */
            const double lsVWHalf = 0.25 * (A_(ls, I, J) + A_(ls, I - 1, J) + A_(ls, I, J - 1) + A_(ls, I - 1, J - 1));
            const double lsVEHalf = 0.25 * (A_(ls, I, J) + A_(ls, I + 1, J) + A_(ls, I, J - 1) + A_(ls, I + 1, J - 1));
            const double muVWHalf = 0.25 * (A_(mu, I, J) + A_(mu, I - 1, J) + A_(mu, I, J - 1) + A_(mu, I - 1, J - 1));
            const double muVEHalf = 0.25 * (A_(mu, I, J) + A_(mu, I + 1, J) + A_(mu, I, J - 1) + A_(mu, I + 1, J - 1));
            const double rho = rho_eff_axi(fl, lsVWHalf, lsVEHalf);
            a[i] = -dt / (rho * rVM) * rVWHalf * muVWHalf / (dr * dr);
            b[i] = 1.0 + dt / (rho * rVM) * (rVWHalf * muVWHalf + rVEHalf * muVEHalf) / (dr * dr);
            c[i] = -dt / (rho * rVM) * rVEHalf * muVEHalf / (dr * dr);
            if (i == 0) {
                if (fW == 1) { b[i] += a[i]; a[i] = 0.0; }
                else if (fW == 2) { d[i] -= a[i] * 2.0 * gr->c_v[0]; a[i] = 0.0; }
                else if (fW == 3) { d[i] -= a[i] * vs2[j + nv * (nr - 1)]; a[i] = 0.0; }
            }
            if (i == nr - 1) {
                if (fE == 1) { b[i] += c[i]; c[i] = 0.0; }
                else if (fE == 2) { d[i] -= c[i] * 2.0 * gr->c_v[1]; c[i] = 0.0; }
                else if (fE == 3) { d[i] -= c[i] * vs2[j + nv * 0]; c[i] = 0.0; }
            }
        }
        or_tdma(a, b, c, d, nr);
        for (int i = 0; i < nr; i++) vs1[j + nv * i] = d[i];
    }
    for (int i = 0; i < nr; i++) {
        for (int j = 0; j < nv; j++) d[j] = vs1[j + nv * i];
        for (int j = 0; j < nv; j++) {
            const double lsS = A_(ls, i + g, j + g), lsM = A_(ls, i + g, j + g + 1);
            const double muS = A_(mu, i + g, j + g), muM = A_(mu, i + g, j + g + 1);
            const double rho = rho_eff_axi(fl, lsS, lsM);
            a[j] = -dt / rho * (2.0 * muS) / (dz * dz);
            b[j] = 1.0 + dt / rho * 2.0 * (muS + muM) / (dz * dz);
            c[j] = -dt / rho * (2.0 * muM) / (dz * dz);
            if (j == 0) {
                if (fS == 1) { b[j] += a[j]; a[j] = 0.0; }
                else if (fS == 2) { d[j] -= a[j] * gr->c_v[2]; a[j] = 0.0; }
                else if (fS == 3) { d[j] -= a[j] * vs1[(nv - 1) + nv * i]; a[j] = 0.0; }
            }
            if (j == nv - 1) {
                if (fN == 1) { b[j] += c[j]; c[j] = 0.0; }
                else if (fN == 2) { d[j] -= c[j] * gr->c_v[3]; c[j] = 0.0; }
                else if (fN == 3) { d[j] -= c[j] * vs1[0 + nv * i]; c[j] = 0.0; }
            }
        }
        or_tdma(a, b, c, d, nv);
        for (int j = 0; j < nv; j++) vs0[j + nv * i] = d[j];
    }
    for (int64_t k = 0; k < or_size(gr); k++) vhat[k] = 0.0;
    for (int i = g; i < nr + g; i++)
        for (int j = g + 1; j < nz + g; j++) A_(vhat, i, j) = vs0[(j - g - 1) + nv * (i - g)];
#undef A_
    free(mu); free(vs2); free(vs1); free(vs0); free(a); free(b); free(c); free(d);
}
