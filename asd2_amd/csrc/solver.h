// solver.h -- matrix-free pressure system and Krylov state shared by the solver sources.
#pragma once

#include "internal.h"

namespace asd2 {

// Matrix-free 5-point operator on the padded layout (interior cells are unknowns; all
// ghost entries are 0 so ghost neighbours contribute nothing):
//   (A p)_c = diag_c p_c + ax_c p_W + ax_E p_E + ay_c p_S + ay_N p_N
// The FP32 copies apply it in flux form,
//   (A p)_c = ax_c (p_W - p_c) + ax_E (p_E - p_c) + ay_c (p_S - p_c) + ay_N (p_N - p_c) + d_c p_c,
// with diag32 holding the remainder d = diag - sum |couplings| (walls with a Dirichlet value, the
// couplings to a pinned cell; zero in the interior of an assembled system).  For a smooth p the
// differences are exact in FP32, so A p keeps its relative accuracy where the diagonal form
// diag p_c - sum |a| p_nb cancels to rounding noise (the pinned all-Neumann operator has a
// near-null mode at the bottom of its spectrum).
struct System {
    double *ax = nullptr, *ay = nullptr, *diag = nullptr, *rhs = nullptr, *dinv = nullptr;
    float *ax32 = nullptr, *ay32 = nullptr, *diag32 = nullptr, *dinv32 = nullptr;
};

int system_arrays(asd2_ctx* ctx, System* S, bool want32);
int ensure_system32(asd2_ctx* ctx, System* S);
bool all_neumann_family(const asd2_ctx* ctx);

// Device-resident Krylov scalars.  CG kernels advance them themselves (every block
// re-sums the producer's partials; block 0 records the result); BiCGStab and the rare
// bookkeeping steps use a one-block tail kernel.
struct KrylovState {
    double rho, rho_old, alpha, beta, omega;  // recurrence scalars
    double rr;                                // ||r||^2 of the recurrence
    double bnorm;                             // ||b|| (1 if b = 0)
    double tol;
    double rmax;                              // mixed: ||r|| at the last reliable update
    double delta;                             // mixed: replacement factor
    double pq, aux0, aux1, aux2;              // scratch reductions
    double rz[2];                             // CG: r.z of iteration k in rz[k & 1]
    double pqv[8];                            // flexible PCG: p.q of iteration k in pqv[k % slots]
    int iter, maxit, mode, done;              // done: 0 running, 1 converged, 2 maxit, 3 breakdown, 4 paused
    int updates, fallbacks, flags, pad;
};

// flexible PCG with a coarse correction, learned or multilevel (krylov.hip)
int coarse_fcg_solve(asd2_ctx* ctx, const asd2_solve_opts* opts, double* d_ps, asd2_solve_stats* stats, bool learned);

constexpr int DONE_RUNNING = 0, DONE_CONVERGED = 1, DONE_MAXIT = 2, DONE_BREAKDOWN = 3, DONE_PAUSED = 4;

}  // namespace asd2
