/*
 * asd2_oracle.h -- FP64 CPU restatement of the ASD2 2D two-phase hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (asd2_amd/, include/asd2_hip.h) never calls it.
 *
 * Every function restates one reference loop nest, cited as file:line of the
 * appleparan/ASD2 sources.  Arrays use the reference layout: padded with g
 * ghost layers per side, j fastest, idx(i,j) = j + (ny + 2g) * i
 * (src/levelset2d.cpp:1048-1050).  BC kinds follow enum BC2D
 * (include/common.h:5).
 *
 * How each function is pinned (golden vectors, the compiled reference, closed
 * forms) is listed in the header comment of asd2_oracle.c.
 */
#ifndef ASD2_ORACLE_H
#define ASD2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* enum class BC2D order, include/common.h:5 */
enum {
    OR_BC_PERIODIC = 0, OR_BC_NEUMANN, OR_BC_DIRICHLET, OR_BC_AXISYM, OR_BC_WALL,
    OR_BC_INLET, OR_BC_OUTLET, OR_BC_PRESSURE, OR_BC_LSFREE, OR_BC_CUSTOM
};

typedef struct or_grid {
    int nx, ny, g;
    double dx, dy;
    int bc_p[4]; double c_p[4];   /* W, E, S, N */
    int bc_u[4]; double c_u[4];
    int bc_v[4]; double c_v[4];
} or_grid;

typedef struct or_fluid {
    double rho_h, rho_l, mu_h, mu_l;
    double sigma, dt;
    double g;          /* gravity constant kG */
    int g_axis;        /* 0 = X, 1 = Y */
    int gravity_fix;   /* 1: V equation uses the V-grid gravity (deviation from mac2d.cpp:294) */
    double re, we, fr; /* dimensionless groups as built by mac2d.cpp:35 */
    double cfl;
} or_fluid;

int64_t or_size(const or_grid* gr);

/* boundary conditions, src/bc2d.cpp:123-130,239-246,346-353 */
void or_bc_p(const or_grid* gr, double* a);
void or_bc_u(const or_grid* gr, double* a);
void or_bc_v(const or_grid* gr, double* a);

/* level set, src/levelset2d.cpp */
void or_unit_hjweno5(const double* F, double* FP, double* FM, double d, int n, int g);
void or_hjweno5_ls(const or_grid* gr, double* ls, const double* uc, const double* vc, double* L);
void or_solve_levelset(const or_grid* gr, double* ls, const double* u, const double* v, double dt);
void or_eno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out);
void or_subcell_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out);
void or_smoothed_sign(const or_grid* gr, const double* ls, double* out);
void or_heaviside_deriv(const or_grid* gr, const double* ls, double* out);
void or_sussman_constraint(const or_grid* gr, const double* ls, const double* ls0,
                           const double* hd, double* out);
void or_reinit_minrk2(const or_grid* gr, double* ls, int iters);
void or_reinit_original(const or_grid* gr, double* ls, int iters);
void or_reinit_first_time(const or_grid* gr, double* ls, int iters);
void or_weno_absgrad(const or_grid* gr, const double* ls, const double* ls0, double* out);
void or_reinit_sussman(const or_grid* gr, double* ls, int iters);
/* run diagnostics: out = {cells, mass, centroid x, centroid y, length, circularity, err_inf,
 * err_l1, err_l2, ref cells, mass loss %}; ref may be NULL */
void or_ls_diagnostics(const or_grid* gr, const double* ls, const double* ref, double base_x, double base_y,
                       double* out);
int  or_reinit_iterations(double eps, double max_time);

/* MAC helpers, src/mac2d.cpp */
void or_update_kappa(const or_grid* gr, const double* ls, double* kappa);
void or_heaviside(const or_grid* gr, const double* ls, double* H);
void or_smooth_heaviside(const or_grid* gr, const double* ls, double* H);
void or_divergence(const or_grid* gr, const double* u, const double* v, double* div);
void or_update_vel(const or_grid* gr, const or_fluid* fl, double* u, double* v,
                   const double* us, const double* vs, const double* ps,
                   const double* ls, const double* kappa, const double* H);
double or_update_dt(const or_grid* gr, const or_fluid* fl, const double* u, const double* v);

/* momentum predictor, src/mac2d.cpp:254-1239 */
void or_convection_u(const or_grid* gr, const double* u, const double* v, double* cU);
void or_convection_v(const or_grid* gr, const double* u, const double* v, double* cV);
void or_viscosity_u(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                    const double* ls, const double* H, double* dU);
void or_viscosity_v(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                    const double* ls, const double* H, double* dV);
void or_intermediate_vel(const or_grid* gr, const or_fluid* fl, int time_order,
                         const double* ls, const double* u, const double* v,
                         double* uhat, double* vhat, const double* Hs);

/* GFM assembly, src/mac2d.cpp:1249-1656.
 * CSR form: returns the number of rows; vals/cols sized >= 5*nx*ny, rowptr >= nx*ny+1,
 * diag >= nx*ny.  rhs is the padded array.  The all-Neumann pin uses the intended
 * operator (centre column shifted like the others; mac2d.cpp:1601-1604 leaves it
 * unshifted). */
int or_gfm_assemble_csr(const or_grid* gr, const or_fluid* fl, const double* ls,
                        const double* kappa, const double* H, const double* div,
                        double* rhs, double* vals, int* cols, int* rowptr, double* diag,
                        int* nnz_out);
/* Matrix-free form on the padded layout: ax = west-face coefficient of each cell,
 * ay = south-face coefficient (both the CSR off-diagonal values, <= 0), diag = CSR
 * centre value, rhs = right-hand side; the pinned cell keeps its diagonal, rhs = 0. */
void or_gfm_assemble_stencil(const or_grid* gr, const or_fluid* fl, const double* ls,
                             const double* kappa, const double* H, const double* div,
                             double* ax, double* ay, double* diag, double* rhs);
void or_stencil_apply(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                      const double* p, double* q);

/* axisymmetric (r, z): x = r with the cell radius base_r + (i + 0.5 - g) dr, y = z.
 * src/mac2dAxisym.cpp:1564-1870 (assembly; ae = each row's own E value), :1974-2008
 * (divergences), :2094-2128 (time step), :2255-2304 (tridiagonal solve, n >= 2) */
void or_gfm_assemble_axisym(const or_grid* gr, const or_fluid* fl, double base_r, const double* ls,
                            const double* kappa, const double* H, const double* div, double* ax, double* ay,
                            double* ae, double* diag, double* rhs);
void or_divergence_axisym(const or_grid* gr, double base_r, const double* u, const double* v, double* div,
                          int for_poisson);
double or_update_dt_axisym(const or_grid* gr, const or_fluid* fl, const double* u, const double* v,
                           const double* kappa);
void or_tdma(double* a, double* b, double* c, double* d, int n);
void or_update_kappa_axisym(const or_grid* gr, double base_r, const double* ls, double* kappa);
/* axisymmetric momentum (src/mac2dAxisym.cpp:264-322, 590-910, 950-1478); x = r, y = z */
void or_viscosity_u_axisym(const or_grid* gr, const or_fluid* fl, const double* v, const double* ls,
                           const double* H, double* dU);
void or_viscosity_v_axisym(const or_grid* gr, const or_fluid* fl, const double* u, const double* uhat,
                           const double* ls, const double* H, double* dV);
void or_rhs_u_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u, const double* v,
                     const double* H, double* rhs);
void or_rhs_v_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* u, const double* uhat,
                     const double* v, const double* H, double* rhs);
void or_uhat_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* rhs, const double* H,
                    double* uhat);
void or_vhat_axisym(const or_grid* gr, const or_fluid* fl, const double* ls, const double* rhs, const double* H,
                    double* vhat);

/* Krylov solvers.  mode 0 = reference CG iteration (poisson2d.cpp:345-396: iteration-0
 * p = rho*r + z, iteration-1 p = z, alpha denominator + 1e-12), mode 1 = textbook PCG. */
int or_pcg_stencil(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                   const double* b, double* x, double tol, int maxit, int mode,
                   int* iters, double* relres);
int or_pcg_csr(int n, const int* rowptr, const int* cols, const double* vals,
               const double* mdiag, const double* b, double* x, double tol, int maxit,
               int mode, int* iters, double* relres);
/* GS_2FUniform_2D (src/poisson2d.cpp:7-88): Gauss-Seidel sweeps + global relaxation 1.2,
 * relative change 1e-6; x in: start vector, out: result. */
int or_gs_csr(int n, const int* rowptr, const int* cols, const double* vals, const double* b,
              double* x, int maxit, int* iters, double* rel_change);
int or_bicgstab_stencil(const or_grid* gr, const double* ax, const double* ay, const double* diag,
                        const double* b, double* x, double tol, int maxit,
                        int* iters, double* relres);

#ifdef __cplusplus
}
#endif
#endif
