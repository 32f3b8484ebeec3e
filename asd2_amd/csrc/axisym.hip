// axisym.hip -- the axisymmetric (r, z) pieces of MACSolver2DAxisym that are not shared with the
// planar solver: the curvature (src/mac2dAxisym.cpp:222-262), the two divergences (:1974-2008)
// and the tridiagonal solve
// of its semi-implicit line sweeps (SolveTDMA, :2255-2304), batched.  The pressure system comes
// from the ghost-fluid assembly kernel in its axisymmetric mode (gfm.hip), the projection from
// the planar UpdateVel kernel (the reference's axisymmetric UpdateVel, :2010-2092, computes the
// same face updates), the level set from the planar level-set kernels.  x = r (base radius =
// the grid's base x), y = z.
#include "internal.h"

namespace asd2 {
namespace {

// FOR_POISSON: (r_{i+1/2} u_{i+1} - r_{i-1/2} u_i) / dr + r_i (v_{j+1} - v_j) / dz
// else:        (u_i + u_{i+1}) / (2 r_i) + (u_{i+1} - u_i) / dr + (v_{j+1} - v_j) / dz
// zero outside the interior
template <bool FOR_POISSON>
__global__ __launch_bounds__(256) void k_div_axisym(Geom G, double base_r, const double* __restrict__ u,
                                                    const double* __restrict__ v, double* __restrict__ div) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + 2 * G.g || j >= G.ny + 2 * G.g) return;
    double d = 0.0;
    if (i >= G.g && i < G.nx + G.g && j >= G.g && j < G.ny + G.g) {
        const double rPM = base_r + (i + 0.5 - G.g) * G.dx;
        const double uW = u[G.idx(i, j)], uE = u[G.idx(i + 1, j)];
        const double vS = v[G.idx(i, j)], vN = v[G.idx(i, j + 1)];
        if (FOR_POISSON) {
            const double rPEHalf = base_r + (i + 1.0 - G.g) * G.dx, rPWHalf = base_r + (i - G.g) * G.dx;
            d = (rPEHalf * uE - rPWHalf * uW) / G.dx + rPM * (vN - vS) / G.dy;
        } else {
            d = 0.5 * (uW + uE) / rPM + (uE - uW) / G.dx + (vN - vS) / G.dy;
        }
    }
    div[G.idx(i, j)] = d;
}

// d phi / dz as the reference stores it: central difference on the interior, 0 elsewhere
__device__ __forceinline__ double dz_at(const Geom& G, const double* __restrict__ ls, int i, int j) {
    if (i < G.g || i >= G.nx + G.g) return 0.0;
    return (ls[G.idx(i, j + 1)] - ls[G.idx(i, j - 1)]) / (2.0 * G.dy);
}

// MACSolver2DAxisym::UpdateKappa (src/mac2dAxisym.cpp:222-262): the planar curvature plus the hoop
// term phi_r |grad phi|^2 / r, the r derivative one-sided on the axis row (i = g), then the clip
// to 1 / min(dr, dz); interior cells only
__global__ __launch_bounds__(256) void k_kappa_axisym(Geom G, double base_r, const double* __restrict__ ls,
                                                      double* __restrict__ kap) {
#pragma clang fp contract(off)
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    const double dr = G.dx, dz = G.dy;
    const double lW = ls[G.idx(i - 1, j)], lE = ls[G.idx(i + 1, j)], lM = ls[c];
    const double dR = (i == G.g) ? (lE - lM) / dr : (lE - lW) / (2.0 * dr);
    const double dZ = dz_at(G, ls, i, j);
    const double S = sqrt(dR * dR + dZ * dZ + 1.0e-100);
    const double rPM = base_r + (i + 0.5 - G.g) * dr;
    const double k = ((lW - 2.0 * lM + lE) / (dr * dr) * dZ * dZ
                      - 2.0 * (dz_at(G, ls, i + 1, j) - dz_at(G, ls, i - 1, j)) / (2.0 * dr) * dR * dZ
                      + (ls[G.idx(i, j - 1)] - 2.0 * lM + ls[G.idx(i, j + 1)]) / (dz * dz) * dR * dR
                      + dR * pow(S, 2.0) / rPM) / pow(S, 3.0);
    kap[c] = isign(k) * fabs(smin(fabs(k), 1.0 / smin(dr, dz)));
}

// one thread per system; element k of system s at k * ks + s * ss (ks = count, ss = 1: coalesced
// across systems; ks = 1, ss = n: each system contiguous).
// The reference's recurrence: c_0 /= b_0, d_0 /= b_0; c_k /= b_k - a_k c_{k-1},
// d_k = (d_k - a_k d_{k-1}) / (b_k - a_k c_{k-1}); back substitution d_k -= c_k d_{k+1}.
__global__ __launch_bounds__(256) void k_tdma(int n, int count, int64_t ks, int64_t ss, const double* __restrict__ a,
                                              const double* __restrict__ b, double* __restrict__ c,
                                              double* __restrict__ d) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= count) return;
    auto at = [&](int k) { return static_cast<int64_t>(k) * ks + static_cast<int64_t>(s) * ss; };
    const int m = n - 1;
    c[at(0)] /= b[at(0)];
    d[at(0)] /= b[at(0)];
    for (int k = 1; k < m; k++) {
        c[at(k)] /= b[at(k)] - a[at(k)] * c[at(k - 1)];
        d[at(k)] = (d[at(k)] - a[at(k)] * d[at(k - 1)]) / (b[at(k)] - a[at(k)] * c[at(k - 1)]);
    }
    d[at(m)] = (d[at(m)] - a[at(m)] * d[at(m - 1)]) / (b[at(m)] - a[at(m)] * c[at(m - 1)]);
    for (int k = m; k-- > 0;) d[at(k)] -= c[at(k)] * d[at(k + 1)];
}

// ------------------------------------------------------------------ semi-implicit viscous sweeps
// GetUHat / GetVHat (src/mac2dAxisym.cpp:950-1478): a line sweep along r and one along z, each a
// batch of tridiagonal systems built by one kernel and solved by k_tdma, then (U only) a diagonal
// factor.  Radii without the base radius, as the reference computes them.

// end-row rule of a sweep side: 1 folds the dropped coupling into the diagonal, 2 moves coupling x
// cst to the right-hand side (cst already doubled where the reference adds twice the constant), 3
// moves coupling x the other end's right-hand side value (periodic), 0 leaves the row
struct SweepArgs {
    double rho_h, rho_l, mu_h, mu_l, dt;
    int fold[4];      // W, E, S, N
    double cst[4];
};

__device__ __forceinline__ double mu_of(const SweepArgs& S, double h) {
#pragma clang fp contract(off)
    return S.mu_l + (S.mu_h - S.mu_l) * h;  // over the whole padded array (src/mac2dAxisym.cpp:996-999)
}

// the first / last row of a system: fold, constant or periodic value
__device__ __forceinline__ void sweep_end(int fold, double cst, double far, double& off, double& diag, double& rhs) {
#pragma clang fp contract(off)
    if (fold == 1) { diag += off; off = 0.0; }
    else if (fold == 2) { rhs -= off * cst; off = 0.0; }
    else if (fold == 3) { rhs -= off * far; off = 0.0; }
}

// U faces along r: system j (0 <= j < nz), unknown k = i - g - 1 (0 <= k < nr - 1) at k * nz + j
__global__ __launch_bounds__(256) void k_uhat_r(Geom G, SweepArgs S, const double* __restrict__ ls,
                                                const double* __restrict__ H, const double* __restrict__ rhs,
                                                double* __restrict__ a, double* __restrict__ b,
                                                double* __restrict__ c, double* __restrict__ d) {
#pragma clang fp contract(off)
    const int nz = G.ny, nu = G.nx - 1, g = G.g;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= static_cast<int64_t>(nu) * nz) return;
    const int k = static_cast<int>(e / nz), j = static_cast<int>(e - static_cast<int64_t>(k) * nz);
    const int J = j + g;
    const double dr = G.dx;
    const double rUM = (k + 1.0) * dr, rPW = (k + 0.5) * dr, rPM = (k + 1.5) * dr;
    const double muW = mu_of(S, H[G.idx(k + g, J)]), muM = mu_of(S, H[G.idx(k + g + 1, J)]);
    const double rho = rho_eff_axi(S.rho_h, S.rho_l, ls[G.idx(k + g, J)], ls[G.idx(k + g + 1, J)]);
    double av = -S.dt / (rho * rUM) * (2.0 * rPW * muW) / (dr * dr);
    double bv = 1.0 + S.dt / (rho * rUM) * 2.0 * (rPW * muW + rPM * muM) / (dr * dr);
    double cv = -S.dt / (rho * rUM) * (2.0 * rPM * muM) / (dr * dr);
    double dv = rhs[G.idx(k + g + 1, J)];
    if (k == 0) sweep_end(S.fold[0], S.cst[0], rhs[G.idx(nu - 1 + g + 1, J)], av, bv, dv);
    if (k == nu - 1) sweep_end(S.fold[1], S.cst[1], rhs[G.idx(g + 1, J)], cv, bv, dv);
    a[e] = av; b[e] = bv; c[e] = cv; d[e] = dv;
}

// U faces along z: system k (0 <= k < nr - 1), unknown j at k * nz + j; us2 is the r sweep's
// solution in the same layout
__global__ __launch_bounds__(256) void k_uhat_z(Geom G, SweepArgs S, const double* __restrict__ ls,
                                                const double* __restrict__ H, const double* __restrict__ us2,
                                                double* __restrict__ a, double* __restrict__ b,
                                                double* __restrict__ c, double* __restrict__ d) {
#pragma clang fp contract(off)
    const int nz = G.ny, nu = G.nx - 1, g = G.g;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= static_cast<int64_t>(nu) * nz) return;
    const int k = static_cast<int>(e / nz), j = static_cast<int>(e - static_cast<int64_t>(k) * nz);
    const int I = k + g + 1, J = j + g;  // the U face
    const double dz = G.dy;
    auto L = [&](int i, int jj) { return ls[G.idx(i, jj)]; };
    auto M = [&](int i, int jj) { return mu_of(S, H[G.idx(i, jj)]); };
    const double lsUSHalf = 0.25 * (L(I, J) + L(I - 1, J) + L(I - 1, J - 1) + L(I, J - 1));
    const double lsUNHalf = 0.25 * (L(I, J) + L(I - 1, J) + L(I - 1, J + 1) + L(I, J + 1));
    const double muUSHalf = 0.25 * (M(I, J) + M(I - 1, J) + M(I - 1, J - 1) + M(I, J - 1));
    const double muUNHalf = 0.25 * (M(I, J) + M(I - 1, J) + M(I - 1, J + 1) + M(I, J + 1));
    const double rho = rho_eff_axi(S.rho_h, S.rho_l, lsUSHalf, lsUNHalf);
    double av = -S.dt / rho * muUSHalf / (dz * dz);
    double bv = 1.0 + S.dt / rho * (muUSHalf + muUNHalf) / (dz * dz);
    double cv = -S.dt / rho * muUNHalf / (dz * dz);
    double dv = us2[e];
    const int64_t row = static_cast<int64_t>(k) * nz;
    if (j == 0) sweep_end(S.fold[2], S.cst[2], us2[row + nz - 1], av, bv, dv);
    if (j == nz - 1) sweep_end(S.fold[3], S.cst[3], us2[row], cv, bv, dv);
    a[e] = av; b[e] = bv; c[e] = cv; d[e] = dv;
}

// u* = us1 / (1 + dt / rho (mu_W + mu_M) / r^2) on the interior U faces, 0 elsewhere
__global__ __launch_bounds__(256) void k_uhat_final(Geom G, SweepArgs S, const double* __restrict__ ls,
                                                    const double* __restrict__ H, const double* __restrict__ us1,
                                                    double* __restrict__ uhat) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int g = G.g;
    if (i >= G.nx + 2 * g || j >= G.ny + 2 * g) return;
    double o = 0.0;
    if (i >= g + 1 && i < G.nx + g && j >= g && j < G.ny + g) {
        const double rUM = (i - g) * G.dx;
        const double rho = rho_eff_axi(S.rho_h, S.rho_l, ls[G.idx(i - 1, j)], ls[G.idx(i, j)]);
        const double muW = mu_of(S, H[G.idx(i - 1, j)]), muM = mu_of(S, H[G.idx(i, j)]);
        o = us1[(j - g) + static_cast<int64_t>(G.ny) * (i - g - 1)] / (1.0 + S.dt / rho * (muW + muM) / (rUM * rUM));
    }
    uhat[G.idx(i, j)] = o;
}

// V faces along r: system j (0 <= j < nz - 1), unknown k = i - g (0 <= k < nr) at k * (nz - 1) + j
__global__ __launch_bounds__(256) void k_vhat_r(Geom G, SweepArgs S, const double* __restrict__ ls,
                                                const double* __restrict__ H, const double* __restrict__ rhs,
                                                double* __restrict__ a, double* __restrict__ b,
                                                double* __restrict__ c, double* __restrict__ d) {
#pragma clang fp contract(off)
    const int nv = G.ny - 1, nr = G.nx, g = G.g;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= static_cast<int64_t>(nr) * nv) return;
    const int k = static_cast<int>(e / nv), j = static_cast<int>(e - static_cast<int64_t>(k) * nv);
    const int I = k + g, J = j + g + 1;  // the V face
    const double dr = G.dx;
    const double rVM = (k + 0.5) * dr, rVWHalf = k * dr, rVEHalf = (k + 1.0) * dr;
    auto L = [&](int i, int jj) { return ls[G.idx(i, jj)]; };
    auto M = [&](int i, int jj) { return mu_of(S, H[G.idx(i, jj)]); };
    const double lsVWHalf = 0.25 * (L(I, J) + L(I - 1, J) + L(I, J - 1) + L(I - 1, J - 1));
    const double lsVEHalf = 0.25 * (L(I, J) + L(I + 1, J) + L(I, J - 1) + L(I + 1, J - 1));
    const double muVWHalf = 0.25 * (M(I, J) + M(I - 1, J) + M(I, J - 1) + M(I - 1, J - 1));
    const double muVEHalf = 0.25 * (M(I, J) + M(I + 1, J) + M(I, J - 1) + M(I + 1, J - 1));
    const double rho = rho_eff_axi(S.rho_h, S.rho_l, lsVWHalf, lsVEHalf);
    double av = -S.dt / (rho * rVM) * rVWHalf * muVWHalf / (dr * dr);
    double bv = 1.0 + S.dt / (rho * rVM) * (rVWHalf * muVWHalf + rVEHalf * muVEHalf) / (dr * dr);
    double cv = -S.dt / (rho * rVM) * rVEHalf * muVEHalf / (dr * dr);
    double dv = rhs[G.idx(I, J)];
    if (k == 0) sweep_end(S.fold[0], S.cst[0], rhs[G.idx(nr - 1 + g, J)], av, bv, dv);
    if (k == nr - 1) sweep_end(S.fold[1], S.cst[1], rhs[G.idx(g, J)], cv, bv, dv);
    a[e] = av; b[e] = bv; c[e] = cv; d[e] = dv;
}

// V faces along z: system k (0 <= k < nr), unknown j (0 <= j < nz - 1) at k * (nz - 1) + j; vs1 is
// the r sweep's solution in the same layout
__global__ __launch_bounds__(256) void k_vhat_z(Geom G, SweepArgs S, const double* __restrict__ ls,
                                                const double* __restrict__ H, const double* __restrict__ vs1,
                                                double* __restrict__ a, double* __restrict__ b,
                                                double* __restrict__ c, double* __restrict__ d) {
#pragma clang fp contract(off)
    const int nv = G.ny - 1, nr = G.nx, g = G.g;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= static_cast<int64_t>(nr) * nv) return;
    const int k = static_cast<int>(e / nv), j = static_cast<int>(e - static_cast<int64_t>(k) * nv);
    const double dz = G.dy;
    const double muS = mu_of(S, H[G.idx(k + g, j + g)]), muM = mu_of(S, H[G.idx(k + g, j + g + 1)]);
    const double rho = rho_eff_axi(S.rho_h, S.rho_l, ls[G.idx(k + g, j + g)], ls[G.idx(k + g, j + g + 1)]);
    double av = -S.dt / rho * (2.0 * muS) / (dz * dz);
    double bv = 1.0 + S.dt / rho * 2.0 * (muS + muM) / (dz * dz);
    double cv = -S.dt / rho * (2.0 * muM) / (dz * dz);
    double dv = vs1[e];
    const int64_t row = static_cast<int64_t>(k) * nv;
    if (j == 0) sweep_end(S.fold[2], S.cst[2], vs1[row + nv - 1], av, bv, dv);
    if (j == nv - 1) sweep_end(S.fold[3], S.cst[3], vs1[row], cv, bv, dv);
    a[e] = av; b[e] = bv; c[e] = cv; d[e] = dv;
}

// v* = vs0 on the interior V faces, 0 elsewhere
__global__ __launch_bounds__(256) void k_vhat_final(Geom G, const double* __restrict__ vs0, double* __restrict__ vhat) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int g = G.g;
    if (i >= G.nx + 2 * g || j >= G.ny + 2 * g) return;
    double o = 0.0;
    if (i >= g && i < G.nx + g && j >= g + 1 && j < G.ny + g) o = vs0[(j - g - 1) + static_cast<int64_t>(G.ny - 1) * (i - g)];
    vhat[G.idx(i, j)] = o;
}

}  // namespace
}  // namespace asd2

using namespace asd2;

namespace {

// sweep end rules of the U (grid 1) or V (grid 2) faces.  U: W / E take Axisym as a wall (u = 0 on
// the axis) and add the constant once, S / N fold Axisym and add twice the constant; V: W / E fold
// Axisym and add twice the constant, S / N take Axisym as a wall and add the constant once
// (src/mac2dAxisym.cpp:1052-1078, 1141-1169, 1336-1364, 1423-1449)
SweepArgs sweep_args(const asd2_ctx* ctx, const asd2_fluid* fl, int grid) {
    SweepArgs S{};
    S.rho_h = fl->rho_h; S.rho_l = fl->rho_l; S.mu_h = fl->mu_h; S.mu_l = fl->mu_l; S.dt = fl->dt;
    for (int s = 0; s < 4; s++) {
        const bool radial = s < 2;
        const bool axis_folds = (grid == 1) ? !radial : radial;
        const bool twice = (grid == 1) ? !radial : radial;
        const int k = ctx->bc[grid][s];
        int f = 0;
        if (k == ASD2_BC_NEUMANN || k == ASD2_BC_OUTLET) f = 1;
        else if (k == ASD2_BC_AXISYM) f = axis_folds ? 1 : 2;
        else if (k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET || k == ASD2_BC_WALL) f = 2;
        else if (k == ASD2_BC_PERIODIC) f = 3;
        S.fold[s] = f;
        S.cst[s] = twice ? 2.0 * ctx->bcc[grid][s] : ctx->bcc[grid][s];
    }
    return S;
}

int sweep_inputs_ok(asd2_ctx* ctx, const asd2_fluid* fl, const void* ls, const void* rhs, const void* H,
                    const void* out) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !ls || !rhs || !H || !out) { set_error("null argument"); return ASD2_ERR_ARG; }
    if (out == rhs) { set_error("the output must not alias the right-hand side"); return ASD2_ERR_ARG; }
    if (ctx->geo.nx < 3 || ctx->geo.ny < 3) {
        set_error("axisymmetric sweeps: needs nr >= 3 and nz >= 3");
        return ASD2_ERR_ARG;
    }
    return ASD2_OK;
}

}  // namespace

int asd2_mac_divergence_axisym(asd2_ctx* ctx, const double* d_u, const double* d_v, double* d_div, int for_poisson) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_u || !d_v || !d_div) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    const dim3 grid(cdiv(G.ny + 2 * G.g, 64), cdiv(G.nx + 2 * G.g, 4));
    if (for_poisson)
        hipLaunchKernelGGL(k_div_axisym<true>, grid, dim3(256), 0, ctx->stream, G, ctx->base_x, d_u, d_v, d_div);
    else
        hipLaunchKernelGGL(k_div_axisym<false>, grid, dim3(256), 0, ctx->stream, G, ctx->base_x, d_u, d_v, d_div);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_kappa_axisym(asd2_ctx* ctx, const double* d_ls, double* d_kappa) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_kappa) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_kappa_axisym, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G,
                       ctx->base_x, d_ls, d_kappa);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_tdma(asd2_ctx* ctx, int n, int count, const double* d_a, const double* d_b, double* d_c, double* d_d) {
    ASD2_TRY(check_ctx(ctx));
    if (n < 2 || count < 1 || !d_a || !d_b || !d_c || !d_d) {
        set_error("asd2_tdma: needs n >= 2 unknowns, count >= 1 systems and four arrays");
        return ASD2_ERR_ARG;
    }
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(count, 256)), dim3(256), 0, ctx->stream, n, count,
                       static_cast<int64_t>(count), int64_t{1}, d_a, d_b, d_c, d_d);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_uhat_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_rhs,
                         const double* d_H, double* d_uhat) {
    ASD2_TRY(sweep_inputs_ok(ctx, fl, d_ls, d_rhs, d_H, d_uhat));
    const Geom& G = ctx->geo;
    const int nz = G.ny, nu = G.nx - 1;
    const size_t m = static_cast<size_t>(nu) * nz;
    const size_t cap = static_cast<size_t>(G.nx) * G.ny;  // shared with the V sweeps, which need nr (nz - 1)
    double *a, *b, *c, *d, *d2;
    ASD2_TRY(ws(ctx, "axi_a", cap, &a));
    ASD2_TRY(ws(ctx, "axi_b", cap, &b));
    ASD2_TRY(ws(ctx, "axi_c", cap, &c));
    ASD2_TRY(ws(ctx, "axi_d", cap, &d));
    ASD2_TRY(ws(ctx, "axi_d2", cap, &d2));
    const SweepArgs S = sweep_args(ctx, fl, 1);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_uhat_r, dim3(cdiv(m, 256)), dim3(256), 0, st, G, S, d_ls, d_H, d_rhs, a, b, c, d);
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(nz, 256)), dim3(256), 0, st, nu, nz, static_cast<int64_t>(nz), int64_t{1},
                       a, b, c, d);
    hipLaunchKernelGGL(k_uhat_z, dim3(cdiv(m, 256)), dim3(256), 0, st, G, S, d_ls, d_H, d, a, b, c, d2);
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(nu, 256)), dim3(256), 0, st, nz, nu, int64_t{1}, static_cast<int64_t>(nz),
                       a, b, c, d2);
    hipLaunchKernelGGL(k_uhat_final, dim3(cdiv(G.ny + 2 * G.g, 64), cdiv(G.nx + 2 * G.g, 4)), dim3(256), 0, st, G, S,
                       d_ls, d_H, d2, d_uhat);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_vhat_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_rhs,
                         const double* d_H, double* d_vhat) {
    ASD2_TRY(sweep_inputs_ok(ctx, fl, d_ls, d_rhs, d_H, d_vhat));
    const Geom& G = ctx->geo;
    const int nv = G.ny - 1, nr = G.nx;
    const size_t m = static_cast<size_t>(nr) * nv;
    const size_t cap = static_cast<size_t>(G.nx) * G.ny;  // shared with the U sweeps, which need (nr - 1) nz
    double *a, *b, *c, *d, *d2;
    ASD2_TRY(ws(ctx, "axi_a", cap, &a));
    ASD2_TRY(ws(ctx, "axi_b", cap, &b));
    ASD2_TRY(ws(ctx, "axi_c", cap, &c));
    ASD2_TRY(ws(ctx, "axi_d", cap, &d));
    ASD2_TRY(ws(ctx, "axi_d2", cap, &d2));
    const SweepArgs S = sweep_args(ctx, fl, 2);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_vhat_r, dim3(cdiv(m, 256)), dim3(256), 0, st, G, S, d_ls, d_H, d_rhs, a, b, c, d);
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(nv, 256)), dim3(256), 0, st, nr, nv, static_cast<int64_t>(nv), int64_t{1},
                       a, b, c, d);
    hipLaunchKernelGGL(k_vhat_z, dim3(cdiv(m, 256)), dim3(256), 0, st, G, S, d_ls, d_H, d, a, b, c, d2);
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(nr, 256)), dim3(256), 0, st, nv, nr, int64_t{1}, static_cast<int64_t>(nv),
                       a, b, c, d2);
    hipLaunchKernelGGL(k_vhat_final, dim3(cdiv(G.ny + 2 * G.g, 64), cdiv(G.nx + 2 * G.g, 4)), dim3(256), 0, st, G,
                       d2, d_vhat);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}
