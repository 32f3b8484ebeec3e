/*
 * asd2_hip.h -- C ABI of the MI355X (gfx950) two-phase hot path.
 *
 * The drop-in C++ classes in levelset2d.h / poisson2d.h / mac2d.h / bc2d.h are
 * implemented on top of these entry points; other hosts (ctypes, cgo, JNI) can bind
 * them directly.  Plain pointers and sizes only.  Field pointers named d_* are DEVICE
 * pointers to FP64 arrays in the reference layout: (nx + 2g) x (ny + 2g) doubles, g = 3
 * ghost layers, j fastest, idx(i, j) = j + (ny + 2g) * i (reference
 * src/levelset2d.cpp:1048-1050).  U values sit on west faces, V values on south faces
 * of the cell with the same index (reference src/mac2d.cpp:1729-1735).
 *
 * Every function returns ASD2_OK or an error code; asd2_last_error() gives the message
 * of the last failure on the calling thread.  Work is enqueued on the context's stream;
 * functions that return host scalars synchronise that stream.
 */
#ifndef ASD2_HIP_H
#define ASD2_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASD2_OK 0
#define ASD2_ERR_ARG 1
#define ASD2_ERR_UNSUPPORTED 2
#define ASD2_ERR_HIP 3
#define ASD2_ERR_NOT_CONVERGED 4
#define ASD2_ERR_NONFINITE 5
#define ASD2_ERR_STATE 6

/* boundary-condition kinds, same order as enum class BC2D (reference include/common.h:5) */
#define ASD2_BC_PERIODIC 0
#define ASD2_BC_NEUMANN 1
#define ASD2_BC_DIRICHLET 2
#define ASD2_BC_AXISYM 3
#define ASD2_BC_WALL 4
#define ASD2_BC_INLET 5
#define ASD2_BC_OUTLET 6
#define ASD2_BC_PRESSURE 7
#define ASD2_BC_LSFREE 8
#define ASD2_BC_CUSTOM 9

/* staggered grid a boundary condition applies to */
#define ASD2_GRID_P 0
#define ASD2_GRID_U 1
#define ASD2_GRID_V 2

/* Krylov iteration variants */
#define ASD2_CG_REFERENCE 0   /* reference CG recurrence, src/poisson2d.cpp:345-396 */
#define ASD2_CG_STANDARD 1    /* textbook Jacobi-PCG */

/* arithmetic of the pressure solve */
#define ASD2_PREC_FP64 0      /* FP64 Jacobi-PCG */
#define ASD2_PREC_MIXED 1     /* FP32 PCG with FP64 reliable updates of x and r */

/* preconditioner of the pressure solve */
#define ASD2_PRECOND_JACOBI 0
#define ASD2_PRECOND_LEARNED 1    /* U-Net coarse correction inside flexible PCG, Jacobi fallback */
#define ASD2_PRECOND_MULTILEVEL 2 /* Galerkin multilevel coarse correction inside flexible PCG */

typedef struct asd2_ctx asd2_ctx;

/* grid description, reference LevelSetSolver2D / MACSolver2D constructor arguments */
typedef struct asd2_grid {
    int nx, ny, num_bc_grid;
    double base_x, base_y, dx, dy;
} asd2_grid;

/* two-phase fluid parameters, reference MACSolver2D members (include/mac2d.h:33-51) */
typedef struct asd2_fluid {
    double rho_h, rho_l, mu_h, mu_l;
    double sigma, dt;
    double g;          /* kG */
    int g_axis;        /* 0 = X, 1 = Y (GAXISENUM2D) */
    int gravity_fix;   /* 1: V momentum gets the V-grid gravity (reference mac2d.cpp:294 uses U's) */
    double re, we, fr; /* kRe, kWe, kFr */
    double cfl;
} asd2_fluid;

typedef struct asd2_solve_opts {
    double tol;          /* stop when ||r||_2 / ||b||_2 <= tol */
    int max_iter;        /* iteration cap (inner iterations for MIXED) */
    int mode;            /* ASD2_CG_* */
    int precision;       /* ASD2_PREC_* */
    int precond;         /* ASD2_PRECOND_* */
    double reliable_delta; /* MIXED: FP64 residual replacement when ||r|| drops by this factor
                              (0 -> 1e-4 for the multilevel solver, 0.1 otherwise) */
    int check_every;     /* iterations enqueued between host convergence polls (0 -> 64) */
} asd2_solve_opts;

typedef struct asd2_solve_stats {
    int iterations;          /* Krylov iterations performed */
    int reliable_updates;    /* MIXED: FP64 residual replacements */
    int fallbacks;           /* LEARNED: switches to the Jacobi preconditioner */
    int converged;           /* 1 when the tolerance was met */
    double rel_residual;     /* ||r||/||b|| of the recurrence at exit */
    double true_rel_residual;/* ||b - A x||/||b|| evaluated in FP64 at exit */
} asd2_solve_stats;

/* ---------------------------------------------------------------- runtime plumbing */
const char* asd2_last_error(void);
int asd2_version(void);
int asd2_device_count(int* count);
int asd2_device_name(char* buf, int len);
int asd2_malloc(void** d_ptr, size_t bytes);           /* zero-filled device allocation */
int asd2_free(void* d_ptr);
int asd2_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int asd2_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);
int asd2_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
/* d_dst[0:count] = d_src[0:count] by a plain streaming kernel on the context's stream.  No
 * reference counterpart: it is the known-size copy that HBM byte counters are calibrated on. */
int asd2_stream_copy(asd2_ctx* ctx, double* d_dst, const double* d_src, size_t count);
/* d_dst[0:count] = value on the context's stream (a kernel, no memset node).  Stands for the
 * whole-array host loops of the reference's classes, e.g. the zero initial guess of the pressure
 * CG (src/poisson2d.cpp:299-300), when the drop-in classes keep the arrays on the device. */
int asd2_fill(asd2_ctx* ctx, double* d_dst, size_t count, double value);
int asd2_device_synchronize(void);

/* ---------------------------------------------------------------- context */
/* replaces the solver constructors, reference src/levelset2d.cpp:3-25,
 * src/mac2d.cpp:28-49, src/poisson2d.cpp:3-5; num_bc_grid must be 3 */
int asd2_create(const asd2_grid* grid, asd2_ctx** out);
int asd2_destroy(asd2_ctx* ctx);
int asd2_synchronize(asd2_ctx* ctx);
/* reference BoundaryCondition2D::SetBC_{U,V,P}_2D + SetBCConstant* (src/bc2d.cpp:16-344,630-849);
 * kinds/consts ordered W, E, S, N */
int asd2_set_bc(asd2_ctx* ctx, int grid_kind, const int kinds[4], const double consts[4]);

/* ---------------------------------------------------------------- boundary conditions */
/* BoundaryCondition2D::ApplyBC_{P,U,V}_2D, src/bc2d.cpp:123-130,239-246,346-353 */
int asd2_bc_apply(asd2_ctx* ctx, int grid_kind, double* d_arr);

/* ---------------------------------------------------------------- level set */
/* LevelSetSolver2D::Solve_LevelSet_2D, src/levelset2d.cpp:43-86 (in place) */
int asd2_ls_advect(asd2_ctx* ctx, double* d_ls, const double* d_u, const double* d_v, double dt);
/* No reference counterpart: asd2_ls_advect for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * allocating routine: the two SSP-RK3 stage arrays live in a static device buffer of
 * 2 x ASD2_LS_NOALLOC_MAX_CELLS doubles (68.2 MB) that the library holds on each GPU it runs on
 * until the process ends, instead of the context's workspaces.  Grids of more than
 * ASD2_LS_NOALLOC_MAX_CELLS padded cells, (nx + 6) (ny + 6) (2048 x 2048 fits), return
 * ASD2_ERR_ARG; level-set-free (LSFREE) pressure sides return ASD2_ERR_UNSUPPORTED.  Every call on
 * one GPU shares the buffer, which keeps the last call's values: calls that overlap on one GPU
 * (several streams or host threads at once) race.  Results equal asd2_ls_advect's bit for bit. */
#define ASD2_LS_NOALLOC_MAX_CELLS 4259840
int asd2_ls_advect_noalloc(asd2_ctx* ctx, double* d_ls, const double* d_u, const double* d_v, double dt);
/* LevelSetSolver2D::HJWENO5_LS_2D, src/levelset2d.cpp:88-157 (applies the P BC to d_ls) */
int asd2_ls_weno_rhs(asd2_ctx* ctx, double* d_ls, const double* d_uc, const double* d_vc, double* d_L);
/* LevelSetSolver2D::Reinit_MinRK2_2D, src/levelset2d.cpp:433-486; iterations from
 * asd2_ls_reinit_iterations(min(dx,dy), kMaxATime) */
int asd2_ls_reinit_minrk2(asd2_ctx* ctx, double* d_ls, int iterations);
/* LevelSetSolver2D::Reinit_Sussman_2D, src/levelset2d.cpp:318-375 */
int asd2_ls_reinit_sussman(asd2_ctx* ctx, double* d_ls, int iterations);
/* LevelSetSolver2D::SubcellENO_Min_DerivAbsLS_2D, src/levelset2d.cpp:615-785 */
int asd2_ls_subcell_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out);
/* LevelSetSolver2D::ENO_DerivAbsLS_2D, src/levelset2d.cpp:559-613 */
int asd2_ls_eno_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out);
/* the pseudo-time loop count of src/levelset2d.cpp:447-448 (host only) */
int asd2_ls_reinit_iterations(double eps, double max_time);
/* LevelSetSolver2D::Reinit_Original_2D, src/levelset2d.cpp:264-316 (ENO2, sign(phi0)) */
int asd2_ls_reinit_original(asd2_ctx* ctx, double* d_ls, int iterations);
/* LevelSetSolver2D::FirstTimeOnlyReinit_Sussman_2D, src/levelset2d.cpp:377-431 (ENO2,
 * sign(phi0), second stage restarted from phi; `iterations` from the end time
 * max(dx nx, dy ny), see asd2_ls_reinit_iterations) */
int asd2_ls_reinit_first_time(asd2_ctx* ctx, double* d_ls, int iterations);
/* LevelSetSolver2D::HJWENO5_DerivAbsLS_2D, src/levelset2d.cpp:488-557 (interior written) */
int asd2_ls_weno_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out);
/* LevelSetSolver2D::UpdateHeavisideFuncDeriv, src/levelset2d.cpp:27-41 (padded array) */
int asd2_ls_heaviside_deriv(asd2_ctx* ctx, const double* d_ls, double* d_out);
/* LevelSetSolver2D::GetSmoothedSignFunc, src/levelset2d.cpp:914-923 (padded array) */
int asd2_ls_smoothed_sign(asd2_ctx* ctx, const double* d_ls, double* d_out);
/* LevelSetSolver2D::GetSussmanReinitConstraint, src/levelset2d.cpp:787-893 */
int asd2_ls_sussman_constraint(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, const double* d_hd,
                               double* d_out);
/* LevelSetSolver2D::UnitHJWENO5 on one line of n + 2g values, src/levelset2d.cpp:159-262
 * (FP, FM written on [g, n + g)) */
int asd2_ls_unit_weno5(asd2_ctx* ctx, const double* d_F, double* d_FP, double* d_FM, double d, int n, int g);

/* Level-set run diagnostics.  The mass (interior cells with phi > 0, times dx dy) and, against a
 * reference field, the error norms ||phi - ref||_inf, sum |phi - ref|, sqrt(sum (phi - ref)^2) /
 * (nx ny) and the mass loss |cells - ref_cells| / ref_cells * 100 are the reference drivers'
 * figures (tests/LargeTests/test_levelset2d_large.cpp:94-102, 106-127).  Added, with no
 * reference counterpart: the centroid of the positive cells' centres (base + (i + 0.5) dx), the
 * interface length sum delta(phi) |grad phi| dx dy (delta = (1 + cos(pi phi / e)) / (2 e) for
 * |phi| < e, e = 1.5 min(dx, dy); central differences) and the circularity 2 sqrt(pi mass) /
 * length (1 for a circle).  d_ref may be NULL (then the error fields are 0).  Synchronous. */
typedef struct asd2_ls_diag {
    double cells, mass, centroid_x, centroid_y, length, circularity;
    double err_linf, err_l1, err_l2, ref_cells, mass_loss_pct;
} asd2_ls_diag;
int asd2_ls_diagnostics(asd2_ctx* ctx, const double* d_ls, const double* d_ref, asd2_ls_diag* out);

/* ---------------------------------------------------------------- MAC helpers */
/* MACSolver2D::UpdateKappa, src/mac2d.cpp:221-252 (writes the interior) */
int asd2_mac_kappa(asd2_ctx* ctx, const double* d_ls, double* d_kappa);
/* MACSolver2D::UpdateHeavisideFunc (smooth = 0, src/mac2d.cpp:87-101) and
 * UpdateSmoothHeavisideFunc (smooth = 1, :103-122) */
int asd2_mac_heaviside(asd2_ctx* ctx, const double* d_ls, double* d_H, int smooth);
/* MACSolver2D::GetDivergence, src/mac2d.cpp:1729-1738 */
int asd2_mac_divergence(asd2_ctx* ctx, const double* d_u, const double* d_v, double* d_div);
/* MACSolver2DAxisym::GetDivergence4Poisson (for_poisson = 1, src/mac2dAxisym.cpp:1974-1989):
 * (r_{i+1/2} u_{i+1} - r_{i-1/2} u_i) / dr + r_i (v_{j+1} - v_j) / dz; GetDivergence (for_poisson =
 * 0, :1991-2008): (u_i + u_{i+1}) / (2 r_i) + (u_{i+1} - u_i) / dr + (v_{j+1} - v_j) / dz.  x = r
 * from the grid's base x, y = z; zero outside the interior. */
int asd2_mac_divergence_axisym(asd2_ctx* ctx, const double* d_u, const double* d_v, double* d_div, int for_poisson);
/* MACSolver2DAxisym::UpdateKappa (src/mac2dAxisym.cpp:222-262): the planar curvature plus the hoop
 * term phi_r |grad phi|^2 / r, phi_r one-sided on the axis row, clipped to 1 / min(dr, dz);
 * interior cells written */
int asd2_mac_kappa_axisym(asd2_ctx* ctx, const double* d_ls, double* d_kappa);
/* MACSolver2DAxisym::SolveTDMA (src/mac2dAxisym.cpp:2255-2304), batched: count tridiagonal systems
 * of n >= 2 unknowns (sub-diagonal a, diagonal b, super-diagonal c, right-hand side d), element k of
 * system s at index k * count + s of each device array; the reference's recurrence, solution in d,
 * c overwritten with the eliminated super-diagonal (as the reference leaves them).  The reference's
 * n = 1 reads d[-1]; here n = 1 is refused. */
int asd2_tdma(asd2_ctx* ctx, int n, int count, const double* d_a, const double* d_b, double* d_c, double* d_d);
/* MACSolver2DAxisym::GetRHSU (src/mac2dAxisym.cpp:264-292): u + dt (-convection + gravity +
 * viscosity) on the interior U faces, 0 on the rest, then the U ghost fill.  The convection is the
 * planar AddConvectionFU (x = r, y = z), the viscous term AddExternalViscosityFU (:590-753, zero
 * when re <= 0), the gravity AddGravityFU (:912-929, g_axis 0 = R, 1 = Z).  mu = mu_L + (mu_H -
 * mu_L) H over the whole padded array.  d_H: the smoothed Heaviside the reference driver passes.
 * d_rhs must not alias an input. */
int asd2_mac_rhs_u_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_u,
                          const double* d_v, const double* d_H, double* d_rhs);
/* MACSolver2DAxisym::GetRHSV (src/mac2dAxisym.cpp:294-322): v + dt (-convection + gravity +
 * viscosity) on the interior V faces, then the V ghost fill.  The viscous term
 * AddExternalViscosityFV (:755-910) reads the new U-face velocity d_uhat (GetUHat's result after the
 * U ghost fill); the hoop term's radius omits the base radius, as the reference computes it.  The
 * gravity is AddGravityFU's array as the reference adds it (gravity_fix = 1: AddGravityFV's). */
int asd2_mac_rhs_v_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_u,
                          const double* d_uhat, const double* d_v, const double* d_H, double* d_rhs);
/* one term of the two right-hand sides above (term as for asd2_mac_momentum_term: ASD2_MOM_RHS =
 * the whole right-hand side with its ghost fill, ASD2_MOM_CONVECTION, ASD2_MOM_VISCOSITY = the
 * external viscosity term; interior faces, 0 elsewhere); d_uhat is read for V only */
int asd2_mac_momentum_term_axisym(asd2_ctx* ctx, const asd2_fluid* fl, int grid_kind, int term, const double* d_ls,
                                  const double* d_u, const double* d_uhat, const double* d_v, const double* d_H,
                                  double* d_out);
/* MACSolver2DAxisym::GetUHat (src/mac2dAxisym.cpp:950-1229): the semi-implicit viscous update of
 * the U faces from GetRHSU's d_rhs: a batched tridiagonal line sweep along r (weights 2 r mu at the
 * cell radii over the face radius), one along z, then the division by 1 + dt / rho (mu_W + mu_M) /
 * r^2; rho the arithmetic mean across the interface, radii without the base radius.  W / E:
 * Neumann, Outlet fold; Dirichlet, Axisym, Inlet, Wall add coupling x constant; S / N: Neumann,
 * Axisym, Outlet fold, Dirichlet, Inlet, Wall add twice the constant; Periodic takes the far end's
 * right-hand side.  Interior U faces written, 0 elsewhere (the caller applies the U ghost fill). */
int asd2_mac_uhat_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_rhs,
                         const double* d_H, double* d_uhat);
/* MACSolver2DAxisym::GetVHat (src/mac2dAxisym.cpp:1231-1478): the same for the V faces, weights r mu
 * at the half radii over the cell radius along r and 2 mu along z; W / E: Neumann, Axisym, Outlet
 * fold, Dirichlet, Inlet, Wall add twice the constant; S / N: Neumann, Outlet fold, Dirichlet,
 * Axisym, Inlet, Wall add the constant.  Interior V faces written, 0 elsewhere. */
int asd2_mac_vhat_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_rhs,
                         const double* d_H, double* d_vhat);
/* MACSolver2D::UpdateVel, src/mac2d.cpp:1740-1849 (then U/V BCs) */
int asd2_mac_update_vel(asd2_ctx* ctx, const asd2_fluid* fl, double* d_u, double* d_v,
                        const double* d_us, const double* d_vs, const double* d_ps,
                        const double* d_ls, const double* d_kappa, const double* d_H);
/* the pressure of MACSolver2D::SolvePoisson's last loop, src/mac2d.cpp:1722-1724: p = ps / dt on the
 * interior, the ghosts of p kept */
int asd2_mac_pressure(asd2_ctx* ctx, const double* d_ps, double* d_p, double dt);
/* MACSolver2D::UpdateDt, src/mac2d.cpp:1851-1880 */
int asd2_mac_update_dt(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_u, const double* d_v,
                       double* dt_out);
/* MACSolver2DAxisym::UpdateDt (src/mac2dAxisym.cpp:2094-2128): the planar bound with the capillary
 * term sqrt(sigma max|kappa| / (min rho min(dr, dz)^2)) and no viscous term; x = r, y = z */
int asd2_mac_update_dt_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_u, const double* d_v,
                              const double* d_kappa, double* dt_out);
/* MACSolver2D::GetIntermediateVel, src/mac2d.cpp:1081-1239 (time_order 0/1/2 =
 * EULER/RK2/RK3; d_Hs is the smoothed Heaviside) */
int asd2_mac_intermediate_vel(asd2_ctx* ctx, const asd2_fluid* fl, int time_order,
                              const double* d_ls, const double* d_u, const double* d_v,
                              double* d_uhat, double* d_vhat, const double* d_Hs);
/* one momentum term on the U (grid_kind = ASD2_GRID_U) or V faces, interior faces only:
 * term 0 = UpdateFU/FV (-convection + viscosity + gravity, src/mac2d.cpp:254-306),
 * 1 = AddConvectionFU/FV (:308-463), 2 = AddViscosityFU/FV (:570-1041) */
#define ASD2_MOM_RHS 0
#define ASD2_MOM_CONVECTION 1
#define ASD2_MOM_VISCOSITY 2
int asd2_mac_momentum_term(asd2_ctx* ctx, const asd2_fluid* fl, int grid_kind, int term,
                           const double* d_ls, const double* d_u, const double* d_v,
                           const double* d_Hs, double* d_out);

/* ---------------------------------------------------------------- GFM pressure Poisson */
/* Assembly half of MACSolver2D::SolvePoisson, src/mac2d.cpp:1249-1656: face
 * coefficients, jump right-hand side, BC and pin handling, stored matrix-free in the
 * context.  d_H may be NULL (then H = [ls >= 0]).  The all-Neumann pin uses the
 * intended operator (the reference leaves the centre column unshifted, :1601-1604). */
int asd2_gfm_assemble(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls,
                      const double* d_kappa, const double* d_H, const double* d_div);
/* Assembly half of MACSolver2DAxisym::SolvePoisson, src/mac2dAxisym.cpp:1488-1950, on an (r, z)
 * grid (x = r, base radius = the grid's base x; y = z): the planar assembly with the face
 * coefficients r_eff / rho_eff (r_eff the W / E face radius in one phase, the interface-point
 * radius across the interface, clamped at 0 on W; the cell radius on S / N), WALL and AXISYM
 * sides Neumann-like, Dirichlet-family sides adding twice the wall constant (the reference's
 * -ps + 2c with ps zeroed).  d_div4p: the r-weighted divergence (asd2_mac_divergence_axisym with
 * for_poisson = 1).  The stored operator is the symmetric one (the W value of a cell also serves
 * as the E value of its west neighbour; the reference's own E value differs from it by rounding
 * across the interface).  Solve with asd2_poisson_solve as for the planar system. */
int asd2_gfm_assemble_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_kappa,
                             const double* d_H, const double* d_div4p);
/* copies the assembled system out (any pointer may be NULL): ax/ay = CSR value of the
 * west/south neighbour of each cell, diag = CSR centre value, rhs = right-hand side */
int asd2_gfm_export(asd2_ctx* ctx, double* d_ax, double* d_ay, double* d_diag, double* d_rhs);
/* installs a caller-built system of the same form */
int asd2_gfm_import(asd2_ctx* ctx, const double* d_ax, const double* d_ay, const double* d_diag,
                    const double* d_rhs);
/* y = A x with the assembled operator (interior cells) */
int asd2_gfm_apply(asd2_ctx* ctx, const double* d_x, double* d_y);
/* Solve half of SolvePoisson: PoissonSolver2D::CG_2FUniformP_2D, src/poisson2d.cpp:230-401
 * (x0 = 0; the solution is written to the interior of d_ps, ghosts untouched) */
int asd2_poisson_solve(asd2_ctx* ctx, const asd2_solve_opts* opts, double* d_ps,
                       asd2_solve_stats* stats);
/* PoissonSolver2D::BiCGStab_2FUniform_2D, src/poisson2d.cpp:403-575 (textbook
 * recurrence, pin-aware) */
int asd2_poisson_bicgstab(asd2_ctx* ctx, const asd2_solve_opts* opts, double* d_ps,
                          asd2_solve_stats* stats);
/* CG_2FUniform_2D on a caller-built zero-based CSR matrix (src/poisson2d.cpp:285-401):
 * n rows, Jacobi preconditioner from d_mdiag, x0 = 0 */
int asd2_csr_pcg(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols,
                 const double* d_vals, const double* d_mdiag, const double* d_b, double* d_x,
                 const asd2_solve_opts* opts, asd2_solve_stats* stats);
/* BiCGStab_2FUniform_2D on a caller-built zero-based CSR matrix (src/poisson2d.cpp:403-575,
 * textbook recurrence), Jacobi preconditioner from d_mdiag, x0 = 0 */
int asd2_csr_bicgstab(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols,
                      const double* d_vals, const double* d_mdiag, const double* d_b, double* d_x,
                      const asd2_solve_opts* opts, asd2_solve_stats* stats);

/* PoissonSolver2D::GS_2FUniform_2D on a caller-built zero-based CSR matrix (src/poisson2d.cpp:7-88):
 * row-order Gauss-Seidel sweeps, each followed by x = x_old + 1.2 (x - x_old), until the relative
 * change sqrt(sum (x - x_old)^2 / (sum x_old^2 + 1e-100)) <= 1e-6 or max_iter sweeps.  d_x holds
 * the start vector and receives the result; stats->iterations = sweeps, stats->rel_residual = the
 * last relative change, converged = 1 when it is <= 1e-6 (true_rel_residual = -1: not evaluated).
 * The sweeps equal the sequential ones bit for bit (rows run level by level of their row-order
 * dependencies); a row without a diagonal entry divides by 0. */
int asd2_csr_gs(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols, const double* d_vals,
                const double* d_b, double* d_x, int max_iter, asd2_solve_stats* stats);

/* ---------------------------------------------------------------- learned preconditioner
 * No reference counterpart: the reference preconditions its CG with the Jacobi diagonal only
 * (src/poisson2d.cpp:346-347).  A 4-level U-Net (base 16 channels, 3x3, BF16 weights and
 * activations, FP32 accumulation) runs on the coarse grid of c x c blocks and proposes the
 * direction y of a coarse correction (its linear form, asd2_unet_forward_linear, on the coarse
 * residual); inside flexible PCG the preconditioner is
 *     z = D^-1 r + gamma P y,  gamma = (y.g) / (y.A_c y),  g = P^T (r - A D^-1 r),  A_c = P^T A P
 * (P: piecewise-constant prolongation), i.e. Jacobi plus the energy-optimal step along the
 * network's coarse correction.  Images are coarse grids (nx / c) x (ny / c), channels last. */
typedef struct asd2_unet asd2_unet;
/* network for fine grids nx x ny coarsened by c (a power of two <= 32, nx and ny divisible by
 * 8 c), buffers for up to max_batch images; He-uniform weights from `seed` */
int asd2_unet_create(int nx, int ny, int coarsen, int max_batch, unsigned long long seed, asd2_unet** out);
int asd2_unet_destroy(asd2_unet* net);
int asd2_unet_param_count(const asd2_unet* net, long long* count);
/* FP32 parameter vector (host): per conv layer weights [9][c_out][c_in] (tap = 3 (di + 1) +
 * (dj + 1)), then all biases, then the 16 output weights and the output bias */
int asd2_unet_set_params(asd2_unet* net, const float* host, long long count);
int asd2_unet_get_params(asd2_unet* net, float* host, long long count);
int asd2_unet_get_grads(asd2_unet* net, float* host, long long count);
/* network output y (FP32 [batch][H][W], device) of features (FP32 [batch][H][W][4], device),
 * H x W the coarse grid */
int asd2_unet_forward(asd2_unet* net, int batch, const float* d_feat, float* d_y);
/* parameter gradients (asd2_unet_get_grads) for dL/dy = d_dy after asd2_unet_forward */
int asd2_unet_backward(asd2_unet* net, int batch, const float* d_dy);
/* the preconditioner's form of the network: the ReLU pattern of a pass on channels 1..3 (channel
 * 0 zeroed) frozen, then a pass without biases on channel 0 alone, each conv output multiplied
 * by (pattern > 0); for fixed coefficient channels y is linear in channel 0 */
int asd2_unet_forward_linear(asd2_unet* net, int batch, const float* d_feat, float* d_y);
/* parameter gradients of that pass for dL/dy = d_dy after asd2_unet_forward_linear (the
 * pattern held fixed; bias gradients zero) */
int asd2_unet_backward_linear(asd2_unet* net, int batch, const float* d_dy);
/* one Adam step on the mean energy loss E(z) / r.D^-1 r, E(z) = z.A z - 2 z.r (the A-norm error
 * of z as a solve of A e = r, shifted by a constant of the system), of `batch` fine systems given
 * in the solver's FP32 padded layout ((nx + 6) (ny + 6) values per system, concatenated):
 * residual, Jacobi inverse, W-face and S-face coefficients, diagonal; loss_out = loss before the
 * step */
int asd2_unet_train_step(asd2_unet* net, int batch, const float* d_r, const float* d_dinv, const float* d_ax,
                         const float* d_ay, const float* d_dg, double dx, double dy, float lr, float* loss_out);
/* one Adam step on coarse systems given directly (no reference counterpart): per sample the coarse
 * residual g (FP32 [batch][H][W]) and the coarse operator in the layout of the level arrays (FP32
 * [batch][(H + 1) (W + 1)] each: W-face and S-face coefficients and diagonal, stride W + 1), H x W the
 * network's image size; h2x, h2y = dx^2 / c, dy^2 / c for the coefficient features (c the level's
 * coarsening of the fine grid).  d_energy (FP32 [batch], device, or NULL): g.A_c^-1 g of each sample.
 * Loss with d_energy: the mean of 1 - (y.g)^2 / ((y.A_c y)(g.A_c^-1 g)), the squared A_c-norm error
 * left by the best step along the network's correction y, relative to that of no correction; without
 * it, the mean of -(y.g)^2 / (y.A_c y) / (g.D_c^-1 g).  loss_out = loss before the step */
int asd2_unet_train_step_coarse(asd2_unet* net, int batch, const float* d_g, const float* d_acx, const float* d_acy,
                                const float* d_dc, const float* d_energy, double h2x, double h2y, float lr,
                                float* loss_out);
/* network used by asd2_poisson_solve with precond = ASD2_PRECOND_LEARNED (not owned; NULL detaches).
 * The network solves on the level of the multilevel hierarchy (first coarsening c0 = 8 x 8 blocks,
 * then 4 x 4) whose coarsening equals the network's, with the levels above it smoothed as in
 * ASD2_PRECOND_MULTILEVEL, its output scaled by the energy-optimal step on that level; a network
 * coarsening c0 or one the hierarchy does not reach is the coarse correction of the fine grid
 * itself (z = D^-1 r + gamma P y). */
int asd2_set_preconditioner(asd2_ctx* ctx, asd2_unet* net);
/* Training data from the solver (no reference counterpart): the next asd2_poisson_solve with
 * ASD2_PRECOND_MULTILEVEL or ASD2_PRECOND_LEARNED copies the input g of its coarse correction on the
 * hierarchy level of level_h x level_w cells at iterations iters[0 .. count) into
 * d_out[k * level_h * level_w] (FP64, row-major); slots of iterations the solve does not reach keep
 * their contents.  The request is cleared by that solve; it fails (ASD2_ERR_ARG) when the solve
 * computes no input of that size. */
int asd2_poisson_record_coarse(asd2_ctx* ctx, int level_h, int level_w, int count, const int* iters, double* d_out);
/* Constant-coefficient fast Poisson / Helmholtz solve; replaces the MKL Poisson Library call of
 * PoissonSolver2D::MKL_2FUniform_2D (reference src/poisson2d.cpp:90-228).  Solves
 * -u_xx - u_yy + q u = rhs with the five-point difference on the node mesh of the padded cell
 * centres from the first ghost to the last (node (i, j) = padded cell (i + g - 1, j + g - 1),
 * i = 0..nx+1, spacing (len_x + dx) / (nx + 1); len, d as the reference's lenX, lenY, dx, dy
 * arguments), with the context's pressure sides: Neumann
 * (the side's nodes are unknowns, zero normal derivative by the central difference) or Dirichlet
 * (node values 2c - ps read from d_ps as the reference reads them: ps(g, j+g), ps(nx+g, j+g),
 * ps(i+g, g), ps(i+g, ny+g)).  rhs is read at every node, ghost cells included; the solution is
 * written to d_ps at every node.  q = 0 with four Neumann sides: the compatible projection,
 * solution with zero weighted mean (weights 1/2 on boundary nodes).  Other side kinds:
 * ASD2_ERR_UNSUPPORTED.  Needs g >= 2; q >= 0. */
int asd2_fast_poisson(asd2_ctx* ctx, const double* d_rhs, double* d_ps, double len_x, double len_y, double dx,
                      double dy, double q);

/* ---------------------------------------------------------------- timing */
/* records the start/stop of a region on the context's stream (hipEvents) */
int asd2_timer_start(asd2_ctx* ctx);
int asd2_timer_stop(asd2_ctx* ctx, double* ms);
/* per-launch timing of one kernel family (hipEvents around each of its launches):
 * kernel_id is one of ASD2_KID_*, 0 = off; reading resets the counters */
#define ASD2_KID_WENO_STAGE 1     /* fused HJ-WENO5 + SSP-RK3 stage */
#define ASD2_KID_REINIT 2         /* MinRK2 iterations (one launch each where the P sides allow) */
#define ASD2_KID_PCG64_SPMV 3     /* FP64 stencil SpMV + direction update + dot */
#define ASD2_KID_PCG64_UPDATE 4   /* FP64 x/r/z update + dots */
#define ASD2_KID_PCG32_SPMV 5     /* FP32 counterparts */
#define ASD2_KID_PCG32_UPDATE 6
#define ASD2_KID_GFM 7            /* GFM assembly */
#define ASD2_KID_COARSE 8         /* coarse correction of the flexible PCG (network or V-cycle) */
#define ASD2_KID_FCG_BEGIN 9      /* flexible PCG: fine pass for the coarse residual and dots */
#define ASD2_KID_FCG_SPMV 10      /* flexible PCG: z formed on the fly, direction, SpMV, p.q */
#define ASD2_KID_FCG_UPDATE 11    /* flexible PCG: x/r update + dots */
#define ASD2_KID_FAST_PASS 12     /* fast Poisson: boundary values + gather, scatter */
#define ASD2_KID_FAST_GEMM 13     /* fast Poisson: the two eigenbasis transforms (rocBLAS DGEMM) */
#define ASD2_KID_FAST_TRIDIAG 14  /* fast Poisson: tridiagonal solves along x, one per y mode */
#define ASD2_KID_MOMENTUM 15      /* momentum predictor: one U-face or V-face launch of an RK stage */
int asd2_profile_kernel(asd2_ctx* ctx, int kernel_id);
int asd2_profile_read(asd2_ctx* ctx, int* launches, double* total_ms);
/* the same records one by one: the first min(capacity, *launches) durations in launch order */
int asd2_profile_read_each(asd2_ctx* ctx, float* durations_ms, int capacity, int* launches);

#ifdef __cplusplus
}
#endif
#endif
